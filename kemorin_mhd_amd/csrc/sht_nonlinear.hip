// sht_nonlinear.hip -- grid-space products of one MHD step.
//
// Between the backward and the forward transform the solver forms its
// nonlinear terms point by point on the rtp grid
// (MS/cal_nonlinear_sph_MHD.f90:64-128, nonlinear_terms_in_rtp): the
// advection w x u, the Lorentz force J x B, the induction u x B, the heat and
// composition fluxes u T and u C (each times its coefficient,
// CORE_src/BASE/cal_products_smp.f90:153-170, 218-233) and the Coriolis term
// -c (omega x u) with omega = (cos theta, -sin theta, 0)
// (MHD_src/sph_MHD/const_wz_coriolis_rtp.f90:117-156).  The reference runs
// one loop per term, reading u up to five times; here one pass reads each
// input component once and writes every enabled output: the work is pure
// HBM streaming (no reuse), 16-byte accesses per lane where the buffers
// allow, contraction off so the products round as the reference's do.
// Inputs come straight from sht_backward's grid layout and the outputs feed
// sht_forward, so the step stays on the device.
#include "sht_internal.h"

namespace ksht {

struct NlArgs {
  int idx[12];     // velo vort magne current temp light | m_advect lorentz vp_induct h_flux c_flux coriolis
  double coef[6];  // per output
};

template <int W>
__device__ __forceinline__ void ldw(const double *p, double (&v)[W]) {
  if constexpr (W == 2) {
    const double2 t = *reinterpret_cast<const double2 *>(p);
    v[0] = t.x;
    v[1] = t.y;
  } else {
    v[0] = *p;
  }
}
template <int W>
__device__ __forceinline__ void stw(double *p, const double (&v)[W]) {
  if constexpr (W == 2)
    *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
  else
    *p = v[0];
}

// o = (a x b) c, in the reference's operation order
template <int W>
__device__ __forceinline__ void cross_c(const double (&a)[3][W], const double (&b)[3][W], double c,
                                        double (&o)[3][W]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int n = 0; n < W; n++) {
    o[0][n] = (a[1][n] * b[2][n] - a[2][n] * b[1][n]) * c;
    o[1][n] = (a[2][n] * b[0][n] - a[0][n] * b[2][n]) * c;
    o[2][n] = (a[0][n] * b[1][n] - a[1][n] * b[0][n]) * c;
  }
}

// W consecutive nodes per thread (W = 2: 16-byte loads and stores)
template <int W>
__global__ void __launch_bounds__(256) k_nonlinear(unsigned nnod, unsigned nr, unsigned nth, NlArgs a,
                                                   const double *__restrict__ fld,
                                                   double *__restrict__ frc,
                                                   const double *__restrict__ cos_t,
                                                   const double *__restrict__ sin_t) {
#pragma clang fp contract(off)
  const unsigned i = (blockIdx.x * blockDim.x + threadIdx.x) * W;
  if (i >= nnod) return;
  const long long n = nnod;
  auto in = [&](int c) { return fld + (long long)c * n + i; };
  auto out = [&](int c) { return frc + (long long)c * n + i; };
  double u[3][W], w[3][W], b[3][W], j[3][W], s[W], o[3][W];
  const bool adv = a.idx[6] >= 0, lor = a.idx[7] >= 0, ind = a.idx[8] >= 0;
  const bool hf = a.idx[9] >= 0, cf = a.idx[10] >= 0, cor = a.idx[11] >= 0;
  if (adv || ind || hf || cf || cor)
#pragma unroll
    for (int c = 0; c < 3; c++) ldw<W>(in(a.idx[0] + c), u[c]);
  if (adv)
#pragma unroll
    for (int c = 0; c < 3; c++) ldw<W>(in(a.idx[1] + c), w[c]);
  if (lor || ind)
#pragma unroll
    for (int c = 0; c < 3; c++) ldw<W>(in(a.idx[2] + c), b[c]);
  if (lor)
#pragma unroll
    for (int c = 0; c < 3; c++) ldw<W>(in(a.idx[3] + c), j[c]);
  if (adv) {
    cross_c<W>(w, u, a.coef[0], o);
#pragma unroll
    for (int c = 0; c < 3; c++) stw<W>(out(a.idx[6] + c), o[c]);
  }
  if (lor) {
    cross_c<W>(j, b, a.coef[1], o);
#pragma unroll
    for (int c = 0; c < 3; c++) stw<W>(out(a.idx[7] + c), o[c]);
  }
  if (ind) {
    cross_c<W>(u, b, a.coef[2], o);
#pragma unroll
    for (int c = 0; c < 3; c++) stw<W>(out(a.idx[8] + c), o[c]);
  }
#pragma unroll
  for (int f = 0; f < 2; f++) {
    if (a.idx[9 + f] < 0) continue;
    ldw<W>(in(a.idx[4 + f]), s);
    const double cc = a.coef[3 + f];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int q = 0; q < W; q++) o[c][q] = u[c][q] * s[q] * cc;
#pragma unroll
    for (int c = 0; c < 3; c++) stw<W>(out(a.idx[9 + f] + c), o[c]);
  }
  if (cor) {
    const double cc = a.coef[5];
#pragma unroll
    for (int q = 0; q < W; q++) {
      const unsigned t = ((i + q) / nr) % nth;
      const double o1 = cos_t[t], o2 = -sin_t[t], o3 = 0.0;
      o[0][q] = -(cc * (o2 * u[2][q] - o3 * u[1][q]));
      o[1][q] = -(cc * (o3 * u[0][q] - o1 * u[2][q]));
      o[2][q] = -(cc * (o1 * u[1][q] - o2 * u[0][q]));
    }
#pragma unroll
    for (int c = 0; c < 3; c++) stw<W>(out(a.idx[11] + c), o[c]);
  }
}

}  // namespace ksht

using namespace ksht;

extern "C" int sht_nonlinear_rtp(sht_plan *p, const sht_nonlinear_desc *d, int ncomp_fld,
                                 const double *fld, int ncomp_frc, double *frc) {
  if (!p || !d || !fld || !frc) {
    set_error("sht_nonlinear_rtp: null argument");
    return SHT_ERR_ARG;
  }
  if (ncomp_fld < 1 || ncomp_frc < 1) {
    set_error("sht_nonlinear_rtp: component counts %d, %d must be positive", ncomp_fld, ncomp_frc);
    return SHT_ERR_ARG;
  }
  const NlArgs a = {{d->i_velo, d->i_vort, d->i_magne, d->i_current, d->i_temp, d->i_light,
                     d->i_m_advect, d->i_lorentz, d->i_vp_induct, d->i_h_flux, d->i_c_flux,
                     d->i_coriolis},
                    {d->coef_velo, d->coef_lor, d->coef_induct, d->coef_h_advect,
                     d->coef_c_advect, d->coef_cor}};
  static const char *names[12] = {"velo",     "vort",    "magne",     "current",
                                  "temp",     "light",   "m_advect",  "lorentz",
                                  "vp_induct", "h_flux", "c_flux",    "coriolis"};
  // inputs each output needs
  static const int needs[6][2] = {{0, 1}, {3, 2}, {0, 2}, {0, 4}, {0, 5}, {0, -1}};
  bool used[6] = {false, false, false, false, false, false};
  for (int o = 0; o < 6; o++) {
    const int io = a.idx[6 + o];
    if (io < -1 || (io >= 0 && io + 3 > ncomp_frc)) {
      set_error("sht_nonlinear_rtp: output %s at component %d outside -1 or 0..%d", names[6 + o],
                io, ncomp_frc - 3);
      return SHT_ERR_ARG;
    }
    if (io < 0) continue;
    for (int q = 0; q < 2; q++)
      if (needs[o][q] >= 0) used[needs[o][q]] = true;
  }
  for (int k = 0; k < 6; k++) {
    const int w = k < 4 ? 3 : 1;
    if (!used[k]) continue;
    if (a.idx[k] < 0 || a.idx[k] + w > ncomp_fld) {
      set_error("sht_nonlinear_rtp: input %s at component %d (needed) outside 0..%d", names[k],
                a.idx[k], ncomp_fld - w);
      return SHT_ERR_ARG;
    }
  }
  const long long nnod = (long long)p->nr * p->nth * p->nphi;
  if (nnod >= (1LL << 31) - 256) {
    set_error("sht_nonlinear_rtp: %lld grid points per component exceed 2^31", nnod);
    return SHT_ERR_UNSUPPORTED;
  }
  const bool vec = (((unsigned long long)fld | (unsigned long long)frc) & 15) == 0 && (nnod & 1) == 0;
  const int W = vec ? 2 : 1;
  const long long nthreads = (nnod + W - 1) / W;
  dim3 grid((unsigned)((nthreads + 255) / 256));
  if (vec)
    hipLaunchKernelGGL(k_nonlinear<2>, grid, dim3(256), 0, p->stream, (unsigned)nnod,
                       (unsigned)p->nr, (unsigned)p->nth, a, fld, frc, p->t.cos_t, p->t.sin_t);
  else
    hipLaunchKernelGGL(k_nonlinear<1>, grid, dim3(256), 0, p->stream, (unsigned)nnod,
                       (unsigned)p->nr, (unsigned)p->nth, a, fld, frc, p->t.cos_t, p->t.sin_t);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}
