// sht_plan.hip -- plan construction: Gauss-Legendre points (host, FP64,
// reference Newton loop), index stacks, and the FP64 Schmidt P_l^m /
// dP_l^m/dtheta tables generated on the GPU.
//
// Reference: initialize_legendre_trans (SS/init_sph_trans.f90:79-168),
// const_gauss_points_coefs (MHD/Fortran_src/CORE_src/gauss_integration/
// gauss_integration.f90:60-133), schmidt_legendres_m
// (MHD/Fortran_src/CORE_src/spherical_harmonics/schmidt_fix_m.f90:62-175).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "sht_internal.h"

namespace ksht {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char *last_error() { return g_err; }

// Gauss-Legendre nodes x (ascending) and weights with the reference's
// Newton iteration: only the final pass n = n_point of the reference's
// n = 2..n_point loop determines the result; the stopping tolerance is the
// REAL(4) literal 3.0e-12 and the weight uses P' at the pre-update iterate.
// Operation order is kept (no contraction) so that the weights, whose
// error depends on the Newton step count, match the reference bit for bit.
#pragma clang fp contract(off)
static void gauss_points_reference(int n, double *x, double *w) {
  const double one = 1.0, two = 2.0, half = 0.5, quad = 0.25;
  const double pi = 4.0 * atan(one);
  const double tol = (double)3.0e-12f;
  const double xm = half * (one + 0.0), xl = half * (one - 0.0);
  if (n == 1) {
    x[0] = half;
    w[0] = one;
  } else {
    const int m = (n + 1 - (n + 1) % 2) / 2;
    for (int i = 1; i <= m; i++) {
      double z = cos(pi * ((double)i - quad) / ((double)n + half));
      double z1, pp, p1, p2, p3;
      for (;;) {
        p1 = one;
        p2 = 0.0;
        for (int j = 1; j <= n; j++) {
          p3 = p2;
          p2 = p1;
          p1 = ((two * (double)j - one) * z * p2 - ((double)j - one) * p3) /
               (double)j;
        }
        pp = (double)n * (z * p1 - p2) / (z * z - one);
        z1 = z;
        z = z1 - p1 / pp;
        if (fabs(z - z1) < tol) break;
      }
      x[i - 1] = xm - xl * z;
      x[n - i] = xm + xl * z;
      w[i - 1] = two * xl / ((one - z * z) * pp * pp);
      w[n - i] = w[i - 1];
    }
  }
  for (int i = 0; i < n; i++) {
    x[i] = x[i] * two - one;
    w[i] = w[i] * two;
  }
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Schmidt recurrence, streamed over l for one order m (no contraction, same
// operation order as schmidt_legendre_m).  For m > 0 the value at l is
// df(l) * s^m with s^m formed by m successive multiplications.
struct SchmidtStream {
  int m, l;
  double c, sm, cur, nxt;
  __device__ void init(int m_, double c_, double s_) {
#pragma clang fp contract(off)
    m = m_;
    l = m_;
    c = c_;
    if (m == 0) {
      cur = 1.0;
      nxt = c;
      sm = 1.0;
    } else {
      double d = 1.0;
      for (int k = 1; k <= m; k++) d = d * (double)(2 * k - 1) / (double)(2 * k);
      nxt = sqrt(2.0 * d * (double)(2 * m + 1)) * c;
      cur = sqrt(2.0 * d);
      sm = 1.0;
      for (int k = 1; k <= m; k++) sm = sm * s_;
    }
  }
  __device__ double value() const { return m == 0 ? cur : cur * sm; }
  __device__ void advance() {
#pragma clang fp contract(off)
    const int ln = l + 2;  // degree of the new value
    double nw;
    if (m == 0) {
      nw = nxt * (double)(2 * ln - 1) / (double)ln * c -
           cur * (double)(ln - 1) / (double)ln;
    } else {
      nw = (c * (double)(2 * ln - 1) * nxt -
            sqrt((double)((ln - 1) * (ln - 1) - m * m)) * cur) /
           sqrt((double)(ln * ln - m * m));
    }
    cur = nxt;
    nxt = nw;
    l++;
  }
};

// One thread per (hemisphere row t, order M).  Writes the even/odd (l-M)
// blocks of both table layouts; padded entries stay zero.
__global__ void __launch_bounds__(256) k_gen_p(
    int L, int nth_h, int nth_hp, const double *__restrict__ cos_t,
    const double *__restrict__ sin_t, const long long *__restrict__ offb,
    const long long *__restrict__ offf, const int *__restrict__ ipad,
    double *__restrict__ Pb, double *__restrict__ dPb, double *__restrict__ Pf,
    double *__restrict__ dPf) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int M = blockIdx.y;
  if (t >= nth_h) return;
  const double c = cos_t[t], s = sin_t[t];
  SchmidtStream rm, rn, rp;  // orders M, M-1, M+1
  rm.init(M, c, s);
  if (M >= 1) {
    rn.init(M - 1, c, s);
    rn.advance();  // now at l = M
  }
  const bool have_p = (M + 1 <= L);
  if (have_p) rp.init(M + 1, c, s);
  for (int l = M; l <= L; l++) {
    const double p = rm.value();
    double dp;
    if (M == 0) {
      dp = (l == 0) ? 0.0 : -sqrt((double)(l * (l + 1) / 2)) * rp.value();
    } else if (M == 1) {
      dp = (l == 1) ? rn.value()
                    : 0.5 * (sqrt((double)(2 * l * (l + 1))) * rn.value() -
                             sqrt((double)((l - 1) * (l + 2))) * rp.value());
    } else {
      dp = (l == M) ? 0.5 * sqrt((double)(2 * M)) * rn.value()
                    : 0.5 * (sqrt((double)((l + M) * (l - M + 1))) * rn.value() -
                             sqrt((double)((l - M) * (l + M + 1))) * rp.value());
    }
    const int par = (l - M) & 1, i = (l - M) >> 1;
    const long long ob = offb[2 * M + par] + (long long)i * nth_hp + t;
    const long long of = offf[2 * M + par] + (long long)t * ipad[2 * M + par] + i;
    Pb[ob] = p;
    dPb[ob] = dp;
    Pf[of] = p;
    dPf[of] = dp;
    rm.advance();
    if (M >= 1) rn.advance();
    if (have_p && l >= M + 1) rp.advance();
  }
}

// One thread per (hemisphere row t < nth_hp, order M0 + blockIdx.y): one
// tile of one layout (fwd: theta-major, i fastest; else i-major, theta
// fastest), every entry of the tile's blocks written, padding included, so
// the tile buffer needs no clearing between tiles (and no memset enters a
// captured graph).  Same recurrence and operation order as k_gen_p.
__global__ void __launch_bounds__(256) k_gen_p_tile(
    int L, int M0, int nth_h, int nth_hp, int fwd, const double *__restrict__ cos_t,
    const double *__restrict__ sin_t, const long long *__restrict__ off,
    const int *__restrict__ ipad, double *__restrict__ P, double *__restrict__ dP) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int M = M0 + blockIdx.y;
  if (t >= nth_hp) return;
  const int ipp[2] = {ipad[2 * M], ipad[2 * M + 1]};
  auto at = [&](int par, int i) -> long long {
    return fwd ? off[2 * M + par] + (long long)t * ipp[par] + i
               : off[2 * M + par] + (long long)i * nth_hp + t;
  };
  const int n = L - M + 1;
  const int np[2] = {(n + 1) / 2, n / 2};
  if (t >= nth_h) {  // padded rows
    for (int par = 0; par < 2; par++)
      for (int i = 0; i < ipp[par]; i++) {
        P[at(par, i)] = 0.0;
        dP[at(par, i)] = 0.0;
      }
    return;
  }
  const double c = cos_t[t], s = sin_t[t];
  SchmidtStream rm, rn, rp;
  rm.init(M, c, s);
  if (M >= 1) {
    rn.init(M - 1, c, s);
    rn.advance();
  }
  const bool have_p = (M + 1 <= L);
  if (have_p) rp.init(M + 1, c, s);
  for (int l = M; l <= L; l++) {
    const double pv = rm.value();
    double dp;
    if (M == 0) {
      dp = (l == 0) ? 0.0 : -sqrt((double)(l * (l + 1) / 2)) * rp.value();
    } else if (M == 1) {
      dp = (l == 1) ? rn.value()
                    : 0.5 * (sqrt((double)(2 * l * (l + 1))) * rn.value() -
                             sqrt((double)((l - 1) * (l + 2))) * rp.value());
    } else {
      dp = (l == M) ? 0.5 * sqrt((double)(2 * M)) * rn.value()
                    : 0.5 * (sqrt((double)((l + M) * (l - M + 1))) * rn.value() -
                             sqrt((double)((l - M) * (l + M + 1))) * rp.value());
    }
    const long long o = at((l - M) & 1, (l - M) >> 1);
    P[o] = pv;
    dP[o] = dp;
    rm.advance();
    if (M >= 1) rn.advance();
    if (have_p && l >= M + 1) rp.advance();
  }
  for (int par = 0; par < 2; par++)
    for (int i = np[par]; i < ipp[par]; i++) {
      P[at(par, i)] = 0.0;
      dP[at(par, i)] = 0.0;
    }
}

// Orders [M0, M1) per tile: about an eighth of the table elements each.
static void partition_mtiles(sht_plan *p) {
  const int L = p->L;
  long long total = 0;
  for (int M = 0; M <= L; M++)
    total += (long long)(p->h_ipad[2 * M] + p->h_ipad[2 * M + 1]) * p->nth_hp;
  const long long target = (total + 7) / 8;
  p->mtiles.clear();
  int M0 = 0;
  long long acc = 0;
  for (int M = 0; M <= L; M++) {
    acc += (long long)(p->h_ipad[2 * M] + p->h_ipad[2 * M + 1]) * p->nth_hp;
    if (acc >= target || M == L) {
      p->mtiles.push_back({M0, M + 1});
      M0 = M + 1;
      acc = 0;
    }
  }
}

int ptile_count(const sht_plan *p) { return p->p_otf ? (int)p->mtiles.size() : 1; }

int ptile_prepare(sht_plan *p, int i, bool fwd, PTile *t) {
  if (!p->p_otf) {
    *t = PTile{0, p->L + 1, p->t.offb, p->t.offf, p->t.Pb, p->t.dPb, p->t.Pf, p->t.dPf};
    return SHT_OK;
  }
  const int M0 = p->mtiles[i].first, M1 = p->mtiles[i].second;
  const long long *off = fwd ? p->offf_r : p->offb_r;
  hipLaunchKernelGGL(k_gen_p_tile, dim3((p->nth_hp + 255) / 256, M1 - M0), dim3(256), 0,
                     p->stream, p->L, M0, p->nth_h, p->nth_hp, fwd ? 1 : 0, p->t.cos_t,
                     p->t.sin_t, off, p->t.ipad, p->Pt, p->dPt);
  KSHT_HIP(hipGetLastError());
  *t = PTile{M0, M1, p->offb_r, p->offf_r, p->Pt, p->dPt, p->Pt, p->dPt};
  return SHT_OK;
}

int generate_p_tables(sht_plan *p) {
  const int L = p->L;
  p->h_offb.assign(2 * (L + 1), 0);
  p->h_offf.assign(2 * (L + 1), 0);
  p->h_ipad.assign(2 * (L + 1), 0);
  long long ob = 0, of = 0;
  for (int M = 0; M <= L; M++) {
    const int n = L - M + 1;
    const int np[2] = {(n + 1) / 2, n / 2};
    for (int par = 0; par < 2; par++) {
      const int ip = (np[par] + kTile - 1) / kTile * kTile;
      p->h_ipad[2 * M + par] = ip;
      p->h_offb[2 * M + par] = ob;
      p->h_offf[2 * M + par] = of;
      ob += (long long)ip * p->nth_hp;
      of += (long long)ip * p->nth_hp;
    }
  }
  p->p_table_elems_b = ob;
  p->p_table_elems_f = of;
  const size_t bb = sizeof(double) * (size_t)ob, bf = sizeof(double) * (size_t)of;
  if (p->p_otf) {
    // offsets rebased to each tile's first block, tile buffers for the largest tile
    partition_mtiles(p);
    std::vector<long long> rb(2 * (L + 1)), rf(2 * (L + 1));
    long long most = 0;
    for (const auto &tl : p->mtiles) {
      const long long b0 = p->h_offb[2 * tl.first], f0 = p->h_offf[2 * tl.first];
      for (int q = 2 * tl.first; q < 2 * tl.second; q++) {
        rb[q] = p->h_offb[q] - b0;
        rf[q] = p->h_offf[q] - f0;
      }
      const long long end = tl.second <= L ? p->h_offb[2 * tl.second] : ob;
      most = std::max(most, end - b0);
    }
    p->tile_elems = most;
    KSHT_HIP(hipMalloc(&p->offb_r, sizeof(long long) * 2 * (L + 1)));
    KSHT_HIP(hipMalloc(&p->offf_r, sizeof(long long) * 2 * (L + 1)));
    KSHT_HIP(hipMemcpy(p->offb_r, rb.data(), sizeof(long long) * 2 * (L + 1), hipMemcpyHostToDevice));
    KSHT_HIP(hipMemcpy(p->offf_r, rf.data(), sizeof(long long) * 2 * (L + 1), hipMemcpyHostToDevice));
  }
  KSHT_HIP(hipMalloc(&p->t.offb, sizeof(long long) * 2 * (L + 1)));
  KSHT_HIP(hipMalloc(&p->t.offf, sizeof(long long) * 2 * (L + 1)));
  KSHT_HIP(hipMalloc(&p->t.ipad, sizeof(int) * 2 * (L + 1)));
  KSHT_HIP(hipMemcpy(p->t.offb, p->h_offb.data(), sizeof(long long) * 2 * (L + 1), hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(p->t.offf, p->h_offf.data(), sizeof(long long) * 2 * (L + 1), hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(p->t.ipad, p->h_ipad.data(), sizeof(int) * 2 * (L + 1), hipMemcpyHostToDevice));
  if (p->p_otf) {
    const size_t tb = sizeof(double) * (size_t)p->tile_elems;
    if (hipMalloc(&p->Pt, tb) != hipSuccess || hipMalloc(&p->dPt, tb) != hipSuccess) {
      set_error("P tiles: device allocation of %zu bytes failed", 2 * tb);
      return SHT_ERR_OOM;
    }
    p->device_bytes += 2 * (long long)tb;
    return SHT_OK;
  }
  if (hipMalloc(&p->t.Pb, bb) != hipSuccess || hipMalloc(&p->t.dPb, bb) != hipSuccess ||
      hipMalloc(&p->t.Pf, bf) != hipSuccess || hipMalloc(&p->t.dPf, bf) != hipSuccess) {
    set_error("P tables: device allocation of %zu bytes failed", 2 * (bb + bf));
    return SHT_ERR_OOM;
  }
  p->device_bytes += 2 * (long long)(bb + bf);
  KSHT_HIP(hipMemsetAsync(p->t.Pb, 0, bb, p->stream));
  KSHT_HIP(hipMemsetAsync(p->t.dPb, 0, bb, p->stream));
  KSHT_HIP(hipMemsetAsync(p->t.Pf, 0, bf, p->stream));
  KSHT_HIP(hipMemsetAsync(p->t.dPf, 0, bf, p->stream));
  dim3 grid((p->nth_h + 255) / 256, L + 1);
  hipLaunchKernelGGL(k_gen_p, grid, dim3(256), 0, p->stream, L, p->nth_h, p->nth_hp,
                     p->t.cos_t, p->t.sin_t, p->t.offb, p->t.offf, p->t.ipad, p->t.Pb,
                     p->t.dPb, p->t.Pf, p->t.dPf);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

template <class T>
static int upload(T **dst, const T *src, size_t n, long long *bytes) {
  KSHT_HIP(hipMalloc(dst, sizeof(T) * (n ? n : 1)));
  if (n) KSHT_HIP(hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
  *bytes += (long long)(sizeof(T) * n);
  return SHT_OK;
}

}  // namespace ksht

using namespace ksht;

static void free_plan(sht_plan *p) {
  if (!p) return;
  void *ptrs[] = {p->t.offb, p->t.offf, p->t.ipad, p->t.jst, p->t.Pb, p->t.dPb,
                  p->t.Pf, p->t.dPf, p->t.radius, p->t.a_r, p->t.weight, p->t.asin_t,
                  p->t.cos_t, p->t.sin_t, p->t.tw, p->t.twn, p->t.ztab, p->t.slot_off_X, p->t.slot_off_rtm,
                  p->t.pole, p->X, p->offb_r, p->offf_r, p->Pt, p->dPt, p->nonfinite};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  split_free(p);
  dft_free(p);
  for (auto &v : p->tev)
    for (auto &e : v) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  if (p->own_stream) (void)hipStreamDestroy(p->own_stream);
  if (p->side) (void)hipStreamDestroy(p->side);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  delete p;
}

extern "C" int sht_plan_create(const sht_desc *d, sht_plan **out) {
  if (!d || !out) {
    set_error("sht_plan_create: null argument");
    return SHT_ERR_ARG;
  }
  *out = nullptr;
  const int L = d->l_truncation, nr = d->nidx_r, nth = d->nidx_theta, nphi = d->nidx_phi;
  // one message per field, naming it and the value received
#define BAD(...)                                  \
  do {                                            \
    set_error("sht_plan_create: " __VA_ARGS__);   \
    return SHT_ERR_ARG;                           \
  } while (0)
  if (L < 1) BAD("l_truncation = %d: need >= 1", L);
  if (nr < 1) BAD("nidx_r = %d: need >= 1", nr);
  if (nth < L + 1 || nth < 2) BAD("nidx_theta = %d: need >= max(2, L+1) = %d", nth, L + 1 > 2 ? L + 1 : 2);
  if (nphi < 2 * L + 2 || (nphi & 1)) BAD("nidx_phi = %d: need even and >= 2L+2 = %d", nphi, 2 * L + 2);
  if (!d->radius) BAD("radius: null pointer");
  if (d->max_ncomp < 1 || d->max_ncomp > 64) BAD("max_ncomp = %d: need 1..64", d->max_ncomp);
  if (d->precision != SHT_PREC_FP64 && d->precision != SHT_PREC_SPLIT_INT8)
    BAD("precision = %d: need SHT_PREC_FP64 (0) or SHT_PREC_SPLIT_INT8 (1)", d->precision);
  if (d->precision == SHT_PREC_SPLIT_INT8 && d->n_slices != 0 &&
      (d->n_slices < 4 || d->n_slices > 7))
    BAD("n_slices = %d: need 0 (default) or 4..7 in split mode", d->n_slices);
  if (d->flags & ~(SHT_FLAG_P_ON_THE_FLY | SHT_FLAG_CHECK_FINITE))
    BAD("flags = %d: unknown bits (SHT_FLAG_P_ON_THE_FLY = %d, SHT_FLAG_CHECK_FINITE = %d)",
        d->flags, SHT_FLAG_P_ON_THE_FLY, SHT_FLAG_CHECK_FINITE);
#undef BAD
  for (int k = 0; k < nr; k++)
    if (!(d->radius[k] > 0.0)) {
      set_error("sht_plan_create: radius[%d] must be positive", k);
      return SHT_ERR_ARG;
    }
  sht_plan *p = new sht_plan();
  p->L = L;
  p->nr = nr;
  p->nth = nth;
  p->nphi = nphi;
  p->nslot = 2 * L + 1;
  p->jmax = (L + 1) * (L + 1);
  p->nth_h = (nth + 1) / 2;
  p->nth_hp = (p->nth_h + 63) / 64 * 64;
  p->nr_p = (nr + kTile - 1) / kTile * kTile;
  p->max_ncomp = d->max_ncomp;
  p->precision = d->precision;
  p->n_slices = d->n_slices;
  p->p_otf = (d->flags & SHT_FLAG_P_ON_THE_FLY) ? 1 : 0;
  p->own_stream = nullptr;
  int rc = SHT_OK;
#define CHECK(x)          \
  do {                    \
    rc = (x);             \
    if (rc != SHT_OK) {   \
      free_plan(p);       \
      return rc;          \
    }                     \
  } while (0)
  // the plan's own stream is a blocking stream: it is ordered with the legacy
  // default stream, so callers that work on the default stream (torch's
  // default stream included) need no explicit synchronisation around calls
  if (hipGetDevice(&p->device) != hipSuccess ||
      hipStreamCreateWithFlags(&p->own_stream, hipStreamDefault) != hipSuccess) {
    set_error("sht_plan_create: no HIP device available");
    p->own_stream = nullptr;
    free_plan(p);
    return SHT_ERR_HIP;
  }
  p->stream = p->own_stream;
  // side stream for overlapping launches; without it calls run in order
  if (hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) != hipSuccess) {
    if (p->side) (void)hipStreamDestroy(p->side);
    p->side = nullptr;
    (void)hipGetLastError();
  }
  // the fixed workspace of sht_backward_noalloc on this device; a failure
  // here is reported by that entry point, which tries again
  (void)scratch_attach(p->device);
  // host tables
  p->radius.assign(d->radius, d->radius + nr);
  p->gpoint.resize(nth);
  p->weight.resize(nth);
  p->colat.resize(nth);
  p->asin_t.resize(nth);
  gauss_points_reference(nth, p->gpoint.data(), p->weight.data());
  std::vector<double> cos_t(nth), sin_t(nth), a_r(nr);
  for (int t = 0; t < nth; t++) {
    p->colat[t] = acos(p->gpoint[t]);
    p->asin_t[t] = 1.0 / sin(p->colat[t]);
    cos_t[t] = cos(p->colat[t]);
    sin_t[t] = sin(p->colat[t]);
  }
  for (int k = 0; k < nr; k++) a_r[k] = 1.0 / p->radius[k];
  p->lstack.assign(p->nslot + 1, 0);
  for (int s = 0; s < p->nslot; s++) p->lstack[s + 1] = p->lstack[s] + (L - abs(s - L) + 1);
  // slot offset tables: slot s (m = s - L) sits at Fourier row q of the
  // internal layout (q = 0 for m = 0, 2m - 1 for cos m, 2|m| for sin m)
  const long long nline = (long long)nr * nth;
  std::vector<long long> so_X(p->nslot), so_rtm(p->nslot);
  for (int s = 0; s < p->nslot; s++) {
    const int m = s - L;
    const int q = m == 0 ? 0 : (m > 0 ? 2 * m - 1 : 2 * (-m));
    so_X[s] = (long long)q * nline;
  }
  for (int s = 0; s < p->nslot; s++) so_rtm[s] = (long long)s * nth * nr;  // x ncomp
  CHECK(upload(&p->t.jst, p->lstack.data(), p->nslot, &p->device_bytes));
  if (d->flags & SHT_FLAG_CHECK_FINITE) {
    if (hipMalloc(&p->nonfinite, sizeof(unsigned long long)) != hipSuccess) {
      p->nonfinite = nullptr;
      set_error("sht_plan_create: device allocation failed (finite-check counter)");
      CHECK(SHT_ERR_OOM);
    }
    p->device_bytes += (long long)sizeof(unsigned long long);
  }
  CHECK(upload(&p->t.radius, p->radius.data(), nr, &p->device_bytes));
  CHECK(upload(&p->t.a_r, a_r.data(), nr, &p->device_bytes));
  CHECK(upload(&p->t.weight, p->weight.data(), nth, &p->device_bytes));
  CHECK(upload(&p->t.asin_t, p->asin_t.data(), nth, &p->device_bytes));
  CHECK(upload(&p->t.cos_t, cos_t.data(), nth, &p->device_bytes));
  CHECK(upload(&p->t.sin_t, sin_t.data(), nth, &p->device_bytes));
  CHECK(upload(&p->t.slot_off_X, so_X.data(), p->nslot, &p->device_bytes));
  CHECK(upload(&p->t.slot_off_rtm, so_rtm.data(), p->nslot, &p->device_bytes));
  CHECK(generate_p_tables(p));
  CHECK(pole_tables_init(p));
  CHECK(fft_plan_init(p));
  p->x_elems = (long long)p->max_ncomp * p->nslot * nline;
  if (hipMalloc(&p->X, sizeof(double) * (size_t)p->x_elems) != hipSuccess) {
    set_error("sht_plan_create: workspace of %lld bytes: out of device memory",
              (long long)(sizeof(double) * p->x_elems));
    free_plan(p);
    return SHT_ERR_OOM;
  }
  p->device_bytes += (long long)sizeof(double) * p->x_elems;
  if (p->precision == SHT_PREC_SPLIT_INT8) CHECK(split_init(p));
  if (hipStreamSynchronize(api_stream(p->stream)) != hipSuccess) {
    set_error("sht_plan_create: table generation failed: %s",
              hipGetErrorString(hipGetLastError()));
    free_plan(p);
    return SHT_ERR_HIP;
  }
#undef CHECK
  *out = p;
  return SHT_OK;
}

extern "C" int sht_plan_destroy(sht_plan *p) {
  if (!p) return SHT_OK;
  (void)hipStreamSynchronize(api_stream(p->stream));
  free_plan(p);
  return SHT_OK;
}

extern "C" int sht_plan_set_stream(sht_plan *p, void *s) {
  if (!p) {
    set_error("sht_plan_set_stream: null plan");
    return SHT_ERR_ARG;
  }
  p->stream = s ? (hipStream_t)s : p->own_stream;
  return SHT_OK;
}

extern "C" long long sht_plan_device_bytes(const sht_plan *p) { return p ? p->device_bytes : 0; }

// Kernel variants: name -> the plan field and its allowed values.
static int *variant_slot(sht_plan *p, const char *name, const int **allowed, int *nallowed) {
  static const int mfma[] = {16, 32}, form[] = {0, 1, 2, 3}, onoff[] = {0, 1};
  if (!strcmp(name, "split_fwd_mfma")) {
    *allowed = mfma;
    *nallowed = 2;
    return &p->var_fwd_mfma;
  }
  if (!strcmp(name, "split_bwd_fused")) {
    *allowed = onoff;
    *nallowed = 2;
    return &p->var_bwd_fused;
  }
  if (!strcmp(name, "fft_form")) {
    *allowed = form;
    *nallowed = 4;
    return &p->var_fft_form;
  }
  return nullptr;
}

extern "C" int sht_plan_set_variant(sht_plan *p, const char *name, int value) {
  if (!p || !name) {
    set_error("sht_plan_set_variant: null argument");
    return SHT_ERR_ARG;
  }
  const int *allowed = nullptr;
  int n = 0;
  int *slot = variant_slot(p, name, &allowed, &n);
  if (!slot) {
    set_error("sht_plan_set_variant: unknown variant '%s'", name);
    return SHT_ERR_ARG;
  }
  for (int i = 0; i < n; i++)
    if (allowed[i] == value) {
      // the DFT and chirp-z forms set up their tables on first selection
      if (slot == &p->var_fft_form && (value == 2 || value == 3)) {
        const int rc = value == 2 ? dft_init(p) : czt_init(p);
        if (rc) return rc;
      }
      *slot = value;
      return SHT_OK;
    }
  set_error("sht_plan_set_variant: %s = %d is not an allowed value", name, value);
  return SHT_ERR_ARG;
}

extern "C" int sht_plan_get_variant(const sht_plan *p, const char *name, int *value) {
  if (!p || !name || !value) {
    set_error("sht_plan_get_variant: null argument");
    return SHT_ERR_ARG;
  }
  const int *allowed = nullptr;
  int n = 0;
  int *slot = variant_slot(const_cast<sht_plan *>(p), name, &allowed, &n);
  if (!slot) {
    set_error("sht_plan_get_variant: unknown variant '%s'", name);
    return SHT_ERR_ARG;
  }
  *value = *slot;
  return SHT_OK;
}

extern "C" int sht_plan_get_table(const sht_plan *p, const char *name, void *out, long long n) {
  if (!p || !name || !out) {
    set_error("sht_plan_get_table: null argument");
    return SHT_ERR_ARG;
  }
  const std::vector<double> *v = nullptr;
  if (!strcmp(name, "gpoint")) v = &p->gpoint;
  else if (!strcmp(name, "weight")) v = &p->weight;
  else if (!strcmp(name, "colat")) v = &p->colat;
  else if (!strcmp(name, "asin_t")) v = &p->asin_t;
  else if (!strcmp(name, "radius")) v = &p->radius;
  if (v) {
    if (n < (long long)v->size()) {
      set_error("sht_plan_get_table: capacity %lld < %zu", n, v->size());
      return SHT_ERR_ARG;
    }
    memcpy(out, v->data(), sizeof(double) * v->size());
    return SHT_OK;
  }
  if (!strcmp(name, "lstack")) {
    if (n < (long long)p->lstack.size()) {
      set_error("sht_plan_get_table: capacity %lld < %zu", n, p->lstack.size());
      return SHT_ERR_ARG;
    }
    memcpy(out, p->lstack.data(), sizeof(int) * p->lstack.size());
    return SHT_OK;
  }
  set_error("sht_plan_get_table: unknown table '%s'", name);
  return SHT_ERR_ARG;
}

extern "C" int sht_get_last_error(char *buf, int n) {
  if (!buf || n <= 0) return SHT_ERR_ARG;
  snprintf(buf, (size_t)n, "%s", ksht::last_error());
  return SHT_OK;
}

extern "C" const char *sht_version(void) { return "kemorin_sht 0.1 (gfx950)"; }
