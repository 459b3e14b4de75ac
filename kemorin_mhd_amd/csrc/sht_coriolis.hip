// sht_coriolis.hip -- Coriolis term of a rotating run in rlm space, between
// the stages of the MHD step (MHD_src/sph_MHD/sph_trans_w_coriols.f90:146-149
// after the rj -> rlm exchange, :233-235 after the forward Legendre
// transform).
//
// init_sum_coriolis_rlm (MHD_src/sph_MHD/t_coriolis_terms_rlm.f90:133-183)
// tabulates, per rlm mode j = (l, m), the Gaunt integrals coupling it to
// (l-1, m) and (l+1, m) and the Elsasser integral coupling it to (l, -m)
// (interact_coriolis_rlm.f90:42-118), and from them the coefficients sw(2,3),
// tw(2,4) of the curl of the Coriolis force (:122-205).  The reference
// evaluates the integrals from prime factorisations; for these index sets
// they are, with Schmidt-normalised P_l^m,
//   K(1,0; l-1,m; l,m) = 4 sqrt(l^2 - m^2) / ((2l-1)(2l+1)),
//   K(1,0; l+1,m; l,m) = 4 sqrt((l+1)^2 - m^2) / ((2l+1)(2l+3)),
//   L(1,0; l,-m; l,m)  = -4 m / (2l+1),
// which the tables here take in closed form on the host.
//
// sum_rot_coriolis_rlm_10 (sum_coriolis_terms_rlm.f90:96-186) is then a
// three-point stencil in j per radius: a thread takes one j and a run of
// radii, so that its 14 coefficients and 3 partner addresses are read once
// per run; contraction off and the reference's operation order, so that the
// sums round as the reference's do.  HBM-bound: per node it reads the
// velocity and vorticity records of three partner modes (up to 11 doubles)
// and writes 2.
#include "sht_internal.h"

#include <cmath>
#include <cstdlib>

struct sht_coriolis {
  sht_plan *plan;
  int kr_icb;
  int *jp = nullptr;        // device: partner rows j(l-1), j(l+1), j(l,-m), 0-based, 3*jmax
  double *coef = nullptr;   // device: sw(2,3,j), tw(2,4,j) as 14 per j, then g3 (jmax)
  double *omega = nullptr;  // device: r^2/2, r, 1 per radius (3*nr)
};

namespace ksht {

constexpr int kCorThreads = 256;
constexpr int kCorRadii = 8;  // radii per thread

__global__ void __launch_bounds__(kCorThreads) k_sum_coriolis(
    int jmax, int nr, int ncomp, int iv, int iw, double coef_cor, const int *__restrict__ jp,
    const double *__restrict__ cf, const double *__restrict__ omega,
    const double *__restrict__ a_r, const double *__restrict__ WR, long long nnod_recv,
    const int *__restrict__ irev, double *__restrict__ out) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * kCorThreads + threadIdx.x;
  if (j >= jmax) return;
  const int k0 = blockIdx.y * kCorRadii;
  const int k1 = min(nr, k0 + kCorRadii);
  const long long nnod = (long long)jmax * nr;
  const int j11 = jp[j], j21 = jp[jmax + j], j12 = jp[2 * jmax + j];
  double s[6], t[8];
#pragma unroll
  for (int a = 0; a < 6; a++) s[a] = cf[14 * (long long)j + a];
#pragma unroll
  for (int a = 0; a < 8; a++) t[a] = cf[14 * (long long)j + 6 + a];
  const double g3 = cf[14 * (long long)jmax + j];
  const double two = 2.0;
  for (int k = k0; k < k1; k++) {
    const long long ko = (long long)jmax * k;
    long long r11 = ko + j11, r21 = ko + j21, r12 = ko + j12;
    if (irev) {
      r11 = (long long)irev[r11] - 1;
      r21 = (long long)irev[r21] - 1;
      r12 = (long long)irev[r12] - 1;
    }
    // records outside the buffer (the dummy record) read zero
    auto rec = [&](int c, long long r) -> double {
      return (r >= 0 && r < nnod_recv) ? WR[c + (long long)ncomp * r] : 0.0;
    };
    const double ar = a_r[k];
    const double om0 = omega[3 * k], om1 = omega[3 * k + 1], om2 = omega[3 * k + 2];
    const double dvp_k1 = rec(iv + 1, r11);
    const double dvp_k2 = rec(iv + 1, r21);
    const double wp_l1 = rec(iw, r12);
    const double vp_k1 = rec(iv, r11);
    const double vp_k2 = rec(iv, r21);
    const double vp_l1 = rec(iv, r12);
    const double d2vp_l1 = ar * ar * g3 * vp_l1 - rec(iw + 2, r12);
    const double wp_k1 = rec(iw, r11);
    const double wp_k2 = rec(iw, r21);
    const double dwp_k1 = rec(iw + 1, r11) - two * ar * wp_k1;
    const double dwp_k2 = rec(iw + 1, r21) - two * ar * wp_k2;
    const double rp = s[0] * om0 * dvp_k1 + s[1] * om0 * dvp_k2 + s[4] * om0 * wp_l1 +
                      s[2] * om1 * vp_k1 + s[3] * om1 * vp_k2;
    const double rt = t[0] * om0 * d2vp_l1 - t[2] * om2 * vp_l1 + t[4] * om1 * wp_k1 +
                      t[5] * om1 * wp_k2 + t[6] * om0 * dwp_k1 + t[7] * om0 * dwp_k2;
    out[ko + j] = -coef_cor * rp * ar * ar;
    out[nnod + ko + j] = -coef_cor * rt * ar * ar;
  }
  (void)nnod;
}

// inner_core_rot_z_coriolis_rlm: the degree-one poloidal terms at the inner
// core boundary radius k (0-based), written after the sum (one thread).
__global__ void k_icb_coriolis(int jmax, int ncomp, int iw, double coef_cor, int k,
                               long long i11s, long long i10c, long long i11c, double rc,
                               const double *__restrict__ omega, const double *__restrict__ WR,
                               long long nnod_recv, const int *__restrict__ irev,
                               double *__restrict__ out) {
#pragma clang fp contract(off)
  (void)jmax;
  (void)k;
  const long long r11s = irev ? (long long)irev[i11s] - 1 : i11s;
  const long long r11c = irev ? (long long)irev[i11c] - 1 : i11c;
  const double two = 2.0;
  const double om0 = omega[3 * k];
  const double wp11s = (r11s >= 0 && r11s < nnod_recv) ? WR[iw + (long long)ncomp * r11s] : 0.0;
  const double wp11c = (r11c >= 0 && r11c < nnod_recv) ? WR[iw + (long long)ncomp * r11c] : 0.0;
  out[i10c] = 0.0;
  out[i11s] = -two * coef_cor * rc * om0 * wp11c;
  out[i11c] = two * coef_cor * rc * om0 * wp11s;
}

// Host tables in the reference's operation order (contraction off).
#pragma clang fp contract(off)
static void coriolis_tables(const sht_plan *p, std::vector<int> &jp, std::vector<double> &cf) {
  const int L = p->L, jmax = p->jmax;
  jp.assign(3 * (size_t)jmax, 0);
  cf.assign(15 * (size_t)jmax, 0.0);
  auto jaddr = [&](int l, int m) { return p->lstack[m + L] + (l - std::abs(m)); };
  const double two = 2.0, half = 0.5;
  for (int s = 0; s < p->nslot; s++) {
    const int m = s - L;
    for (int l = std::abs(m); l <= L; l++) {
      const int j = jaddr(l, m);
      double gi1 = 0.0, gi2 = 0.0, e1 = 0.0;
      int a1 = j, a2 = j, ae = j;
      if (l - 1 >= 1 && std::abs(m) <= l - 1) {
        gi1 = 4.0 * std::sqrt((double)(l * l - m * m)) /
              ((double)(2 * l - 1) * (double)(2 * l + 1));
        a1 = jaddr(l - 1, m);
      }
      if (l + 1 <= L) {
        gi2 = 4.0 * std::sqrt((double)((l + 1) * (l + 1) - m * m)) /
              ((double)(2 * l + 1) * (double)(2 * l + 3));
        a2 = jaddr(l + 1, m);
      }
      if (l >= 1) {
        if (m != 0) e1 = -4.0 * (double)m / (double)(2 * l + 1);
        ae = jaddr(l, -m);
      }
      if (l == 0) {  // j3 = 0: every partner is (0, 0)
        a1 = a2 = ae = jaddr(0, 0);
        gi2 = 0.0;
        e1 = 0.0;
      }
      jp[j] = a1;
      jp[jmax + j] = a2;
      jp[2 * jmax + j] = ae;
      // g3 = l(l+1) (1/2 at l = m = 0), g4 = (l-1) l, g5 = (l+1)(l+2),
      // g17 = (2l+1) / (4 l (l+1)) (0 at l = 0): spherical_harmonics.f90:106-138
      const double g3 = l == 0 ? half : (double)(l * (l + 1));
      const double g4 = (double)((l - 1) * l), g5 = (double)((l + 1) * (l + 2));
      const double g17 = l == 0 ? 0.0 : (double)(2 * l + 1) / (double)(4 * l * (l + 1));
      double *sw = &cf[14 * (size_t)j], *tw = sw + 6;
      sw[0] = (two - g4 - g3) * gi1 * g17;
      sw[1] = (two - g5 - g3) * gi2 * g17;
      sw[2] = g4 * (two - g4 + g3) * gi1 * g17 * half;
      sw[3] = g5 * (two - g5 + g3) * gi2 * g17 * half;
      sw[4] = two * e1 * g17;
      sw[5] = 0.0;
      tw[4] = -(g3 * (two + g4 - g3) + two * (-two + g4 + g3)) * gi1 * g17 * half;
      tw[5] = -(g3 * (two + g5 - g3) + two * (-two + g5 + g3)) * gi2 * g17 * half;
      tw[6] = -(-two + g4 + g3) * gi1 * g17;
      tw[7] = -(-two + g5 + g3) * gi2 * g17;
      tw[0] = -two * e1 * g17;
      tw[1] = 0.0;
      tw[2] = -g3 * e1 * g17;
      tw[3] = 0.0;
      cf[14 * (size_t)jmax + j] = g3;
    }
  }
}
#pragma clang fp contract(on)

// scalar copies of the two parts of d_cor to the send buffer (sht_copy.hip)
int launch_to_send_pairs(sht_plan *p, int n, const int *fld, const int *trn, long long nnod_fld,
                         const double *d_fld, const int *d_inod, long long nnod_send, int nb,
                         double *d_WS);

}  // namespace ksht

using namespace ksht;

extern "C" int sht_coriolis_create(sht_plan *p, int kr_icb, sht_coriolis **out) {
  if (!p || !out) {
    set_error("sht_coriolis_create: null argument");
    return SHT_ERR_ARG;
  }
  *out = nullptr;
  if (kr_icb < 0 || kr_icb > p->nr) {
    set_error("sht_coriolis_create: kr_icb = %d outside 0..%d", kr_icb, p->nr);
    return SHT_ERR_ARG;
  }
  std::vector<int> jp;
  std::vector<double> cf;
  coriolis_tables(p, jp, cf);
  std::vector<double> om(3 * (size_t)p->nr);
  {
#pragma clang fp contract(off)
    for (int k = 0; k < p->nr; k++) {
      const double r = p->radius[k];
      om[3 * k] = 0.5 * 1.0 * (r * r);
      om[3 * k + 1] = 1.0 * r;
      om[3 * k + 2] = 1.0;
    }
  }
  sht_coriolis *c = new sht_coriolis();
  c->plan = p;
  c->kr_icb = kr_icb;
  auto fail = [&](hipError_t e, const char *what) {
    set_error("sht_coriolis_create: %s: %s", what, hipGetErrorString(e));
    sht_coriolis_destroy(c);
    return e == hipErrorOutOfMemory ? SHT_ERR_OOM : SHT_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipMalloc(&c->jp, sizeof(int) * jp.size())) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&c->coef, sizeof(double) * cf.size())) != hipSuccess)
    return fail(e, "hipMalloc");
  if ((e = hipMalloc(&c->omega, sizeof(double) * om.size())) != hipSuccess)
    return fail(e, "hipMalloc");
  if ((e = hipMemcpy(c->jp, jp.data(), sizeof(int) * jp.size(), hipMemcpyHostToDevice)) !=
          hipSuccess ||
      (e = hipMemcpy(c->coef, cf.data(), sizeof(double) * cf.size(), hipMemcpyHostToDevice)) !=
          hipSuccess ||
      (e = hipMemcpy(c->omega, om.data(), sizeof(double) * om.size(), hipMemcpyHostToDevice)) !=
          hipSuccess)
    return fail(e, "hipMemcpy");
  *out = c;
  return SHT_OK;
}

extern "C" int sht_coriolis_destroy(sht_coriolis *c) {
  if (!c) return SHT_OK;
  if (c->jp) (void)hipFree(c->jp);
  if (c->coef) (void)hipFree(c->coef);
  if (c->omega) (void)hipFree(c->omega);
  delete c;
  return SHT_OK;
}

extern "C" int sht_sum_coriolis_rlm(sht_coriolis *c, int ncomp, int i_velo, int i_vort,
                                    double coef_cor, const double *d_WR, long long nnod_recv,
                                    const int *d_irev, double *d_cor) {
  static const char *who = "sht_sum_coriolis_rlm";
  if (!c || !c->plan) {
    set_error("%s: null handle", who);
    return SHT_ERR_ARG;
  }
  sht_plan *p = c->plan;
  const long long nnod = (long long)p->jmax * p->nr;
  if (ncomp < 1 || i_velo < 0 || i_velo + 3 > ncomp || i_vort < 0 || i_vort + 3 > ncomp) {
    set_error("%s: velocity at %d and vorticity at %d need 3 of %d components", who, i_velo,
              i_vort, ncomp);
    return SHT_ERR_ARG;
  }
  if (nnod_recv < 0 || (!d_irev && nnod_recv != nnod)) {
    set_error("%s: nnod_recv = %lld (identity table needs %lld)", who, nnod_recv, nnod);
    return SHT_ERR_ARG;
  }
  if (!d_WR || !d_cor) {
    set_error("%s: null buffer", who);
    return SHT_ERR_ARG;
  }
  StageTimer tm(p, 4);
  dim3 grid((unsigned)((p->jmax + kCorThreads - 1) / kCorThreads),
            (unsigned)((p->nr + kCorRadii - 1) / kCorRadii));
  hipLaunchKernelGGL(k_sum_coriolis, grid, dim3(kCorThreads), 0, p->stream, p->jmax, p->nr,
                     ncomp, i_velo, i_vort, coef_cor, c->jp, c->coef, c->omega, p->t.a_r, d_WR,
                     nnod_recv, d_irev, d_cor);
  KSHT_HIP(hipGetLastError());
  if (c->kr_icb > 0 && p->L >= 1) {
    const int L = p->L, k = c->kr_icb - 1;
    const long long ko = (long long)p->jmax * k;
    const long long i11s = ko + p->lstack[L - 1];      // (1, -1)
    const long long i10c = ko + p->lstack[L] + 1;      // (1, 0)
    const long long i11c = ko + p->lstack[L + 1];      // (1, 1)
    hipLaunchKernelGGL(k_icb_coriolis, dim3(1), dim3(1), 0, p->stream, p->jmax, ncomp, i_vort,
                       coef_cor, k, i11s, i10c, i11c, p->radius[k], c->omega, d_WR, nnod_recv,
                       d_irev, d_cor);
    KSHT_HIP(hipGetLastError());
  }
  return SHT_OK;
}

extern "C" int sht_copy_coriolis_terms_rlm(sht_coriolis *c, int ncomp, int i_rot_cor,
                                           const double *d_cor, const int *d_item,
                                           long long nnod_send, double *d_WS) {
  static const char *who = "sht_copy_coriolis_terms_rlm";
  if (!c || !c->plan) {
    set_error("%s: null handle", who);
    return SHT_ERR_ARG;
  }
  sht_plan *p = c->plan;
  const long long nnod = (long long)p->jmax * p->nr;
  if (ncomp < 1 || i_rot_cor < 0 || i_rot_cor + 3 > ncomp) {
    set_error("%s: rot Coriolis at %d needs 3 of %d components", who, i_rot_cor, ncomp);
    return SHT_ERR_ARG;
  }
  if (nnod_send < 0 || (!d_item && nnod_send != nnod)) {
    set_error("%s: nnod_send = %lld (identity table needs %lld)", who, nnod_send, nnod);
    return SHT_ERR_ARG;
  }
  if (!d_cor || !d_WS) {
    set_error("%s: null buffer", who);
    return SHT_ERR_ARG;
  }
  const int fld[2] = {0, 1}, trn[2] = {i_rot_cor, i_rot_cor + 2};
  return launch_to_send_pairs(p, 2, fld, trn, nnod, d_cor, d_item, nnod_send, ncomp, d_WS);
}
