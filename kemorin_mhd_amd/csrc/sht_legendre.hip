// sht_legendre.hip -- FP64-MFMA Legendre transforms (gfx950).
//
// Per zonal order M = |m| the reference's symmetric_BLAS_big mode
// (SS/leg_b_trans_sym_matmul_big.f90:183-293, SS/leg_f_trans_sym_matmul_big.f90:
// 187-305) does four DGEMMs against hemisphere tables of P and dP/dtheta,
// split by the parity of l-m.  Here both slots +M and -M share the tables,
// so one work-group handles both; the pack (SS/set_sp_rlm_for_leg_matmul.f90:
// 281-398, SS/set_vr_rtm_for_leg_matmul.f90:369-565) is fused into the LDS
// staging of the operands and the unpack (SS/cal_vr_rtm_by_matmul.f90:357-540,
// SS/cal_sp_rlm_by_matmul.f90:365-545) into the accumulator epilogue.  The
// 1/r, 1/r^2, m and 1/sin(theta) factors are applied in the epilogue rather
// than to the operands (same products, different rounding order).
//
// MFMA: v_mfma_f64_16x16x4f64.  A(16x4): lane holds A[lane%16][lane/16];
// B(4x16): lane holds B[lane/16][lane%16]; D(16x16): lane holds
// D[4*r + lane/16][lane%16] in element r (measured on gfx950, checked by
// sht_selftest).
#include "sht_internal.h"

namespace ksht {

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// g3 as packed by the reference: l(l+1), with 1/2 for l = m = 0
// (SP/schmidt_poly_on_rtm_grid.f90:109).
__device__ __forceinline__ double g3_of(int l) { return l == 0 ? 0.5 : (double)(l * (l + 1)); }
// g7, g6 of SH/spherical_harmonics.f90:118-137.
__device__ __forceinline__ double g7_of(int l, int M) {
  if (l == 0) return 1.0;
  return M == 0 ? (double)(2 * l + 1) / (double)(2 * l * (l + 1))
                : (double)(2 * l + 1) / (double)(4 * l * (l + 1));
}
__device__ __forceinline__ double g6_of(int l, int M) {
  if (l == 0 && M == 0) return 0.5;
  return M == 0 ? (double)(2 * l + 1) / 2.0 : (double)(2 * l + 1) / 4.0;
}

struct LegDims {
  int L, nr, nth, nth_h, nth_hp, nv, ns, n_rt, ncg, nblk;
};

// ---------------------------------------------------------------------------
// Backward (spectra -> Legendre-side grid).  Work-group = (M, radius tile,
// component group, theta block); 4 waves, wave w owns hemisphere rows
// [16*(4*tb+w), +16).  Group cg < nv: vector cg; otherwise scalars
// 4*(cg-nv) .. +3.
constexpr int kVecQ = 5;  // P*pol*g3, P*tor, P*dpol, dP*dpol, dP*tor
constexpr int kSclPer = 4;

__global__ void __launch_bounds__(256) k_leg_bwd(
    LegDims d, const double *__restrict__ sp, SpecAddr sa, double *__restrict__ out,
    GridAddr ga, const int *__restrict__ jst, const long long *__restrict__ offb,
    const int *__restrict__ ipad, const double *__restrict__ Pb,
    const double *__restrict__ dPb, const double *__restrict__ a_r,
    const double *__restrict__ asin_t) {
  __shared__ double sB[2 * 4 * 2 * 16 * 16];  // [sig][c][par][ii][kk]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  long long b = blockIdx.x;
  const int tb = (int)(b % d.nblk);
  b /= d.nblk;
  const int cg = (int)(b % d.ncg);
  b /= d.ncg;
  const int rt = (int)(b % d.n_rt);
  const int M = (int)(b / d.n_rt);
  const int L = d.L;
  const int n = L - M + 1, ne = (n + 1) / 2;
  const int nsig = (M == 0) ? 1 : 2;
  const int k0 = rt * 16;
  const int t0 = (tb * 4 + wave) * 16;
  const bool active = t0 < d.nth_h;
  const bool is_vec = cg < d.nv;
  const int nc = is_vec ? 3 : min(kSclPer, d.ns - kSclPer * (cg - d.nv));
  const int nd0 = is_vec ? 3 * cg : 3 * d.nv + kSclPer * (cg - d.nv);
  const int ip[2] = {ipad[2 * M], ipad[2 * M + 1]};
  const long long ob[2] = {offb[2 * M], offb[2 * M + 1]};
  const int slot[2] = {L + M, L - M};
  const int jb[2] = {jst[L + M], jst[L - M]};

  d4 acc[2][2][kVecQ];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int q = 0; q < kVecQ; q++) acc[s][p][q] = d4{0.0, 0.0, 0.0, 0.0};

  const int nchunk = (ne + 15) / 16;
  for (int ic = 0; ic < nchunk; ic++) {
    __syncthreads();
    for (int idx = tid; idx < 2 * 16 * 32; idx += 256) {
      const int lo = idx & 31, kk = (idx >> 5) & 15, sg = idx >> 9;
      const int par = lo & 1, ii = lo >> 1;
      const int l = M + 2 * (ic * 16 + ii) + par;
      const int k = k0 + kk;
      const bool valid = sg < nsig && l <= L && k < d.nr;
      const int jbs = sg == 0 ? jb[0] : jb[1];
      const double *src = sp + (long long)(jbs + l - M) * sa.s_j + (long long)k * sa.s_k +
                          (long long)nd0 * sa.s_nd;
      double *dst = sB + ((sg * 4 * 2 + par) * 16 + ii) * 16 + kk;  // + c*512
      if (is_vec) {
        const double pol = valid ? src[0] * g3_of(l) : 0.0;
        const double dpol = valid ? src[sa.s_nd] : 0.0;
        const double tor = valid ? src[2 * sa.s_nd] : 0.0;
        dst[0 * 512] = pol;
        dst[1 * 512] = dpol;
        dst[2 * 512] = tor;
      } else {
        for (int c = 0; c < kSclPer; c++)
          dst[c * 512] = (valid && c < nc) ? src[(long long)c * sa.s_nd] : 0.0;
      }
    }
    __syncthreads();
    if (!active) continue;
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
      const int ii = ks * 4 + (lane >> 4);
      const int ia = ic * 16 + ii;
#pragma unroll
      for (int par = 0; par < 2; par++) {
        if (ic * 16 >= ip[par]) continue;
        const long long ao = ob[par] + (long long)ia * d.nth_hp + t0 + (lane & 15);
        const double aP = Pb[ao];
        const double adP = is_vec ? dPb[ao] : 0.0;
#pragma unroll
        for (int sg = 0; sg < 2; sg++) {
          if (sg >= nsig) continue;
          const double *bb = sB + ((sg * 4 * 2 + par) * 16 + ii) * 16 + (lane & 15);
          if (is_vec) {
            const double bpol = bb[0], bdpol = bb[512], btor = bb[1024];
            acc[sg][par][0] = mfma64(aP, bpol, acc[sg][par][0]);
            acc[sg][par][1] = mfma64(aP, btor, acc[sg][par][1]);
            acc[sg][par][2] = mfma64(aP, bdpol, acc[sg][par][2]);
            acc[sg][par][3] = mfma64(adP, bdpol, acc[sg][par][3]);
            acc[sg][par][4] = mfma64(adP, btor, acc[sg][par][4]);
          } else {
#pragma unroll
            for (int c = 0; c < kSclPer; c++) acc[sg][par][c] = mfma64(aP, bb[c * 512], acc[sg][par][c]);
          }
        }
      }
    }
  }
  if (!active) return;
  // epilogue: N/S recombination, conjugate-slot terms, radial factors
  const int k = k0 + (lane & 15);
  if (k >= d.nr) return;
  const double a1 = a_r[k], a2 = a1 * a1;
  const bool odd_nth = (d.nth & 1) != 0;
#pragma unroll
  for (int sg = 0; sg < 2; sg++) {
    if (sg >= nsig) continue;
    const bool self = (nsig == 1);
    const double mc = (sg == 0) ? -(double)M : (double)M;  // m of the conjugate slot
    const long long base = ga.slot_off[slot[sg]] * ga.slot_mul + (long long)k * ga.s_k;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int t = t0 + 4 * r + (lane >> 4);
      if (t >= d.nth_h) continue;
      const bool eq = odd_nth && t == d.nth / 2;
      const long long on = base + (long long)t * ga.s_t;
      const long long os = base + (long long)(d.nth - 1 - t) * ga.s_t;
      if (is_vec) {
        const double c1 = -asin_t[t] * mc * a1;
        const double E0 = acc[sg][0][0][r], O0 = acc[sg][1][0][r];
        const double E3 = acc[sg][0][3][r], O3 = acc[sg][1][3][r];
        const double E4 = acc[sg][0][4][r], O4 = acc[sg][1][4][r];
        // conjugate slot: the other sign of m, or the slot itself for M = 0
        const double Ec1 = self ? acc[sg][0][1][r] : acc[1 - sg][0][1][r];
        const double Oc1 = self ? acc[sg][1][1][r] : acc[1 - sg][1][1][r];
        const double Ec2 = self ? acc[sg][0][2][r] : acc[1 - sg][0][2][r];
        const double Oc2 = self ? acc[sg][1][2][r] : acc[1 - sg][1][2][r];
        double *o = out + (long long)nd0 * ga.s_nd;
        if (eq) {
          o[on] = a2 * E0;
          o[on + ga.s_nd] = a1 * O3 + c1 * Ec1;
          o[on + 2 * ga.s_nd] = -a1 * O4 + c1 * Ec2;
        } else {
          o[on] = a2 * (E0 + O0);
          o[on + ga.s_nd] = a1 * (E3 + O3) + c1 * (Ec1 + Oc1);
          o[on + 2 * ga.s_nd] = -a1 * (E4 + O4) + c1 * (Ec2 + Oc2);
          o[os] = a2 * (E0 - O0);
          o[os + ga.s_nd] = a1 * (O3 - E3) + c1 * (Ec1 - Oc1);
          o[os + 2 * ga.s_nd] = -a1 * (O4 - E4) + c1 * (Ec2 - Oc2);
        }
      } else {
#pragma unroll
        for (int c = 0; c < kSclPer; c++) {
          if (c >= nc) continue;
          const double E = acc[sg][0][c][r], O = acc[sg][1][c][r];
          double *o = out + (long long)(nd0 + c) * ga.s_nd;
          if (eq) {
            o[on] = E;
          } else {
            o[on] = E + O;
            o[os] = E - O;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Forward (Legendre-side grid -> spectra).  Work-group = (M, radius tile,
// component group, degree block); 4 waves, wave w owns pair indices
// i in [64*ib + 16*w, +16) of both parities (l = M + 2i + parity).  The
// reduction runs over hemisphere rows in chunks of 16 staged in LDS as
// weighted symmetric / antisymmetric combinations of the mirrored rows.
__global__ void __launch_bounds__(256) k_leg_fwd(
    LegDims d, const double *__restrict__ in, GridAddr ga, double *__restrict__ sp,
    SpecAddr sa, const int *__restrict__ jst, const long long *__restrict__ offf,
    const int *__restrict__ ipad, const double *__restrict__ Pf,
    const double *__restrict__ dPf, const double *__restrict__ radius,
    const double *__restrict__ weight, const double *__restrict__ asin_t) {
  __shared__ double sA[2 * 4 * 2 * 16 * 16];  // [sig][f][sym/asym][tt][kk]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  long long b = blockIdx.x;
  const int ib = (int)(b % d.nblk);
  b /= d.nblk;
  const int cg = (int)(b % d.ncg);
  b /= d.ncg;
  const int rt = (int)(b % d.n_rt);
  const int M = (int)(b / d.n_rt);
  const int L = d.L;
  const int ip[2] = {ipad[2 * M], ipad[2 * M + 1]};
  if (ib * 64 >= ip[0]) return;  // whole work-group beyond the degrees of M
  const int nsig = (M == 0) ? 1 : 2;
  const int k0 = rt * 16;
  const int i0 = ib * 64 + wave * 16;
  const bool active = i0 < ip[0];
  const bool is_vec = cg < d.nv;
  const int nc = is_vec ? 3 : min(kSclPer, d.ns - kSclPer * (cg - d.nv));
  const int nd0 = is_vec ? 3 * cg : 3 * d.nv + kSclPer * (cg - d.nv);
  const long long of[2] = {offf[2 * M], offf[2 * M + 1]};
  const int slot[2] = {L + M, L - M};
  const bool odd_nth = (d.nth & 1) != 0;

  d4 acc[2][2][kVecQ];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int q = 0; q < kVecQ; q++) acc[s][p][q] = d4{0.0, 0.0, 0.0, 0.0};

  const int nchunk = (d.nth_h + 15) / 16;
  for (int tc = 0; tc < nchunk; tc++) {
    __syncthreads();
    for (int idx = tid; idx < 2 * 4 * 16 * 16; idx += 256) {
      const int kk = idx & 15, tt = (idx >> 4) & 15, f = (idx >> 8) & 3, sg = idx >> 10;
      const int t = tc * 16 + tt, k = k0 + kk;
      const bool valid = sg < nsig && f < nc && t < d.nth_h && k < d.nr;
      double sym = 0.0, asym = 0.0;
      if (valid) {
        const int sl = sg == 0 ? slot[0] : slot[1];
        const long long base = ga.slot_off[sl] * ga.slot_mul +
                               (long long)(nd0 + f) * ga.s_nd + (long long)k * ga.s_k;
        const double w = weight[t];
        const double vn = in[base + (long long)t * ga.s_t];
        if (odd_nth && t == d.nth / 2) {
          sym = vn * w;
        } else {
          const double vs = in[base + (long long)(d.nth - 1 - t) * ga.s_t];
          sym = (vn + vs) * w;
          asym = (vn - vs) * w;
        }
      }
      double *dst = sA + (((sg * 4 + f) * 2) * 16 + tt) * 16 + kk;
      dst[0] = sym;
      dst[256] = asym;
    }
    __syncthreads();
    if (!active) continue;
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
      const int tt = ks * 4 + (lane >> 4);
      const int t = tc * 16 + tt;
      const double asn = (t < d.nth_h) ? asin_t[t] : 0.0;
#pragma unroll
      for (int par = 0; par < 2; par++) {
        if (i0 >= ip[par]) continue;
        const long long bo = of[par] + (long long)t * ip[par] + i0 + (lane & 15);
        const double bP = Pf[bo];
        const double bdP = is_vec ? dPf[bo] : 0.0;
        const int sP = par, sdP = 1 - par;  // P: sym for even, dP: asym for even
#pragma unroll
        for (int sg = 0; sg < 2; sg++) {
          if (sg >= nsig) continue;
          const int sc = (nsig == 1) ? 0 : 1 - sg;
          const double *aa = sA + ((sg * 4 * 2) * 16 + tt) * 16 + (lane & 15);
          const double *ac = sA + ((sc * 4 * 2) * 16 + tt) * 16 + (lane & 15);
          if (is_vec) {
            const double avr = aa[(0 * 2 + sP) * 256];
            const double avt = aa[(1 * 2 + sdP) * 256];
            const double avp = aa[(2 * 2 + sdP) * 256];
            const double avpc = ac[(2 * 2 + sP) * 256] * asn;
            const double avtc = ac[(1 * 2 + sP) * 256] * asn;
            acc[sg][par][0] = mfma64(avr, bP, acc[sg][par][0]);
            acc[sg][par][1] = mfma64(avt, bdP, acc[sg][par][1]);
            acc[sg][par][2] = mfma64(avpc, bP, acc[sg][par][2]);
            acc[sg][par][3] = mfma64(avp, bdP, acc[sg][par][3]);
            acc[sg][par][4] = mfma64(avtc, bP, acc[sg][par][4]);
          } else {
#pragma unroll
            for (int c = 0; c < kSclPer; c++)
              acc[sg][par][c] = mfma64(aa[(c * 2 + sP) * 256], bP, acc[sg][par][c]);
          }
        }
      }
    }
  }
  if (!active) return;
  const int ii = i0 + (lane & 15);
#pragma unroll
  for (int par = 0; par < 2; par++) {
    const int l = M + 2 * ii + par;
    if (l > L) continue;
    const double g7 = g7_of(l, M), g6 = g6_of(l, M);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int k = k0 + 4 * r + (lane >> 4);
      if (k >= d.nr) continue;
      const double r1 = radius[k], r2 = r1 * r1;
#pragma unroll
      for (int sg = 0; sg < 2; sg++) {
        if (sg >= nsig) continue;
        const double ms = (sg == 0) ? (double)M : -(double)M;
        const long long o = (long long)(jst[slot[sg]] + l - M) * sa.s_j + (long long)k * sa.s_k +
                            (long long)nd0 * sa.s_nd;
        if (is_vec) {
          sp[o] = r2 * g7 * acc[sg][par][0][r];
          sp[o + sa.s_nd] = r1 * g7 * (acc[sg][par][1][r] - ms * acc[sg][par][2][r]);
          sp[o + 2 * sa.s_nd] = r1 * g7 * (-acc[sg][par][3][r] - ms * acc[sg][par][4][r]);
        } else {
#pragma unroll
          for (int c = 0; c < kSclPer; c++)
            if (c < nc) sp[o + (long long)c * sa.s_nd] = g6 * acc[sg][par][c][r];
        }
      }
    }
  }
}

static LegDims make_dims(const sht_plan *p, int nv, int ns, int nblk) {
  LegDims d;
  d.L = p->L;
  d.nr = p->nr;
  d.nth = p->nth;
  d.nth_h = p->nth_h;
  d.nth_hp = p->nth_hp;
  d.nv = nv;
  d.ns = ns;
  d.n_rt = p->nr_p / 16;
  d.ncg = nv + (ns + kSclPer - 1) / kSclPer;
  d.nblk = nblk;
  return d;
}

int launch_leg_bwd(sht_plan *p, int nv, int ns, const double *sp, SpecAddr sa, double *out,
                   GridAddr ga) {
  const LegDims d = make_dims(p, nv, ns, p->nth_hp / 64);
  const long long nblocks = (long long)(p->L + 1) * d.n_rt * d.ncg * d.nblk;
  if (nblocks == 0) return SHT_OK;
  hipLaunchKernelGGL(k_leg_bwd, dim3((unsigned)nblocks), dim3(256), 0, p->stream, d, sp, sa,
                     out, ga, p->t.jst, p->t.offb, p->t.ipad, p->t.Pb, p->t.dPb, p->t.a_r,
                     p->t.asin_t);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

int launch_leg_fwd(sht_plan *p, int nv, int ns, const double *in, GridAddr ga, double *sp,
                   SpecAddr sa) {
  const int ne_max = (p->L + 2) / 2;
  const LegDims d = make_dims(p, nv, ns, (ne_max + 63) / 64);
  const long long nblocks = (long long)(p->L + 1) * d.n_rt * d.ncg * d.nblk;
  if (nblocks == 0) return SHT_OK;
  hipLaunchKernelGGL(k_leg_fwd, dim3((unsigned)nblocks), dim3(256), 0, p->stream, d, in, ga,
                     sp, sa, p->t.jst, p->t.offf, p->t.ipad, p->t.Pf, p->t.dPf, p->t.radius,
                     p->t.weight, p->t.asin_t);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

}  // namespace ksht
