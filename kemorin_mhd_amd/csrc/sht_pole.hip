// sht_pole.hip -- backward transform at the poles and the centre.
//
// The grid of the transform has no point at theta = 0 or pi; the solver's
// FEM mesh does, and pole_b_transform / sph_b_trans_w_poles
// (SS/t_sph_transforms.f90:127-245) fill those points from the spectra
// directly: only orders 0 and +-1 survive at a pole, so the value there is
// a sum over l of one spectral column (SS/pole_sph_transform.f90:107-162,
// SS/schmidt_b_trans_at_pole.f90:88-289, SS/schmidt_b_trans_at_center.f90:
// 45-115).  Output rows: 0..Nr-1 north pole, Nr..2Nr-1 south pole, 2Nr the
// centre; v_pole[i + (2 Nr + 1) nd].  Vector components come out Cartesian,
// (x, y, z) = (v_theta, v_phi, v_r) at phi = 0 in the north and
// (-v_theta, v_phi, -v_r) in the south.
//
// The work is O(Nr L ncomp) and latency-bound: one wave per output, whose
// lanes form the per-degree products and whose lane 0 adds them in the
// reference's order with contraction off, so the result is the reference's
// operation sequence on the same tables.
#include <cmath>
#include <vector>

#include "sht_internal.h"

namespace ksht {

// P_l^0 and dP_l^1/dtheta at one colatitude, l = 0..L, by the Schmidt
// recurrence of SH/schmidt_fix_m.f90:89-175 (order 0 by three terms; the
// derivative of order 1 from orders 0 and 2, order 2 as df(l) sin^2), the
// tables set_lagender_pole_rlm fills (SP/schmidt_poly_on_rtm_grid.f90:
// 201-266) and the pole transform reads.
static void pole_schmidt(int L, double theta, double *P0, double *dP1) {
#pragma clang fp contract(off)
  const double c = cos(theta), s = sin(theta);
  std::vector<double> df(L + 3, 0.0), p2(L + 3, 0.0);
  P0[0] = 1.0;
  if (L >= 1) P0[1] = c;
  for (int l = 2; l <= L; l++)
    P0[l] = P0[l - 1] * (double)(2 * l - 1) / (double)l * c - P0[l - 2] * (double)(l - 1) / (double)l;
  double d = 1.0;
  for (int k = 1; k <= 2; k++) d = d * (double)(2 * k - 1) / (double)(2 * k);
  df[3] = sqrt(2.0 * d * 5.0) * c;
  df[2] = sqrt(2.0 * d);
  for (int l = 4; l <= L; l++)
    df[l] = (c * (double)(2 * l - 1) * df[l - 1] - sqrt((double)((l - 1) * (l - 1) - 4)) * df[l - 2]) /
            sqrt((double)(l * l - 4));
  double sm = 1.0;
  for (int k = 1; k <= 2; k++) sm = sm * s;
  for (int l = 2; l <= L; l++) p2[l] = df[l] * sm;
  dP1[0] = 0.0;
  if (L >= 1) dP1[1] = P0[1];
  for (int l = 2; l <= L; l++)
    dP1[l] = 0.5 * (sqrt((double)(2 * l * (l + 1))) * P0[l] - sqrt((double)((l - 1) * (l + 2))) * p2[l]);
}

// tab: [P0 north | P0 south | dP1 north | dP1 south], L + 1 each
int pole_tables_init(sht_plan *p) {
  const int L = p->L, n = L + 1;
  std::vector<double> h(4 * n);
  const double pi = 4.0 * atan(1.0);
  pole_schmidt(L, 0.0, h.data(), h.data() + 2 * n);
  pole_schmidt(L, pi, h.data() + n, h.data() + 3 * n);
  KSHT_HIP(hipMalloc(&p->t.pole, sizeof(double) * 4 * n));
  KSHT_HIP(hipMemcpy(p->t.pole, h.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice));
  p->device_bytes += (long long)sizeof(double) * 4 * n;
  return SHT_OK;
}

// Sum over degrees l = lo..hi of term(l), in ascending l: the lanes of the
// (single-wave) work-group form the terms of a 64-degree chunk together, and
// lane 0 adds them in order with add(acc, t), so the loads run in parallel
// while the rounding sequence stays the sequential one.
template <class Term, class Add>
__device__ __forceinline__ double ordered_sum(double acc, int lo, int hi, double *term, Term tf,
                                              Add add) {
  const int lane = threadIdx.x;
  for (int l0 = lo; l0 <= hi; l0 += 64) {
    const int l = l0 + lane;
    if (l <= hi) term[lane] = tf(l);
    __syncthreads();
    if (lane == 0) {
      const int n = min(64, hi - l0 + 1);
      for (int q = 0; q < n; q++) acc = add(acc, term[q]);
    }
    __syncthreads();
  }
  return acc;
}

// One wave per output row i (blockIdx.x) of component nd (blockIdx.y).
// j0, jm, jp: spectral index of (l, m) = (0, 0), (1, -1), (1, 1) (degrees
// ascending within an order).
__global__ void __launch_bounds__(64) k_pole_bwd(int L, int nr, int nv, int ns, int center,
                                                 int j0, int jm, int jp, long long jmax,
                                                 const double *__restrict__ sp,
                                                 const double *__restrict__ tab,
                                                 const double *__restrict__ a_r,
                                                 double *__restrict__ vp) {
#pragma clang fp contract(off)
  __shared__ double term[64];
  const int ncomp = 3 * nv + ns, npole = 2 * nr + 1;
  const int i = blockIdx.x, nd = blockIdx.y, lane = threadIdx.x;
  double *o = vp + i + (long long)npole * nd;
  const long long sj = ncomp;  // stride of one degree/order step
  if (i == 2 * nr) {
    // centre (radius index 0): the degree-1 poloidal terms; scalars and
    // every centre entry without a centre are left as they are
    if (!center || nd >= 3 * nv || lane != 0) return;
    const int f = nd / 3, c = nd % 3;
    const double a = a_r[0];
    const double *pol = sp + 3 * f;
    if (c == 2)
      *o = 2.0 * pol[sj * (j0 + 1)] * a * a;
    else
      *o = -(2.0 * pol[sj * (c == 0 ? jm : jp)] * a * a);
    return;
  }
  const bool north = i < nr;
  const int k = north ? i : i - nr;
  const double *P = tab + (north ? 0 : L + 1);
  const double *dP = tab + 2 * (L + 1) + (north ? 0 : L + 1);
  const double *s = sp + (long long)ncomp * jmax * k;
  auto plus = [](double acc, double t) { return acc + t; };
  auto minus = [](double acc, double t) { return acc - t; };
  auto sgn = [north](double acc, double t) { return north ? acc + t : acc - t; };
  double acc = 0.0;
  if (nd >= 3 * nv) {
    acc = ordered_sum(acc, 0, L, term, [&](int l) { return s[nd + sj * (j0 + l)] * P[l]; }, plus);
    if (lane == 0) *o = acc;
    return;
  }
  const int f = nd / 3, c = nd % 3;
  const double a = a_r[k];
  const double *pol = s + 3 * f, *dpol = pol + 1, *tor = pol + 2;
  if (c == 2) {
    // z: order 0 of the poloidal term times l(l+1) (1/2 at l = 0, the
    // normalisation table's value, SP/schmidt_poly_on_rtm_grid.f90:109)
    acc = ordered_sum(acc, 0, L, term, [&](int l) {
      const double g3 = l == 0 ? 0.5 : (double)(l * (l + 1));
      return pol[sj * (j0 + l)] * P[l] * g3;
    }, sgn);
    if (lane == 0) *o = acc * a * a;
  } else if (c == 0) {
    acc = ordered_sum(acc, 1, L, term, [&](int l) { return tor[sj * (jm + l - 1)] * dP[l]; }, plus);
    acc = ordered_sum(acc, 1, L, term, [&](int l) { return dpol[sj * (jp + l - 1)] * dP[l]; }, sgn);
    if (lane == 0) *o = acc * a;
  } else {
    acc = ordered_sum(acc, 1, L, term, [&](int l) { return dpol[sj * (jm + l - 1)] * dP[l]; }, sgn);
    acc = ordered_sum(acc, 1, L, term, [&](int l) { return tor[sj * (jp + l - 1)] * dP[l]; }, minus);
    if (lane == 0) *o = acc * a;
  }
}

static int launch_pole_bwd(sht_plan *p, int nv, int ns, const double *sp, double *vp,
                           int center) {
  const int ncomp = 3 * nv + ns, npole = 2 * p->nr + 1, L = p->L;
  dim3 grid(npole, ncomp);
  hipLaunchKernelGGL(k_pole_bwd, grid, dim3(64), 0, p->stream, L, p->nr, nv, ns, center,
                     p->lstack[L], p->lstack[L - 1], p->lstack[L + 1], (long long)p->jmax, sp,
                     p->t.pole, p->t.a_r, vp);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

}  // namespace ksht

using namespace ksht;

static int check_pole(const sht_plan *p, int nv, int ns, const void *a, const void *b,
                      const char *who) {
  if (!p) {
    set_error("%s: null plan", who);
    return SHT_ERR_ARG;
  }
  if (nv < 0 || ns < 0 || 3 * nv + ns < 1 || 3 * nv + ns > p->max_ncomp) {
    set_error("%s: ncomp = 3*%d + %d outside 1..%d (max_ncomp of the plan)", who, nv, ns,
              p->max_ncomp);
    return SHT_ERR_ARG;
  }
  if (!a || !b) {
    set_error("%s: null buffer", who);
    return SHT_ERR_ARG;
  }
  return SHT_OK;
}

extern "C" int sht_pole_backward(sht_plan *p, int nv, int ns, const double *WS_rj,
                                 double *v_pole, int with_center) {
  const int rc = check_pole(p, nv, ns, WS_rj, v_pole, "sht_pole_backward");
  if (rc) return rc;
  return launch_pole_bwd(p, nv, ns, WS_rj, v_pole, with_center ? 1 : 0);
}

extern "C" int sht_backward_with_poles(sht_plan *p, int nv, int ns, const double *WS_rj,
                                       double *v_rtp, double *v_pole, int with_center) {
  int rc = check_pole(p, nv, ns, WS_rj, v_pole, "sht_backward_with_poles");
  if (rc) return rc;
  if (!v_rtp) {
    set_error("sht_backward_with_poles: null buffer");
    return SHT_ERR_ARG;
  }
  // the reference's order: poles from the rlm buffer, then the grid
  rc = launch_pole_bwd(p, nv, ns, WS_rj, v_pole, with_center ? 1 : 0);
  if (rc) return rc;
  return sht_backward(p, nv, ns, WS_rj, v_rtp);
}
