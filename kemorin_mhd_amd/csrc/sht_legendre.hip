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
  int M0, nM;  // orders M0 .. M0+nM-1 of this launch (one tile of the P tables)
};

// ---------------------------------------------------------------------------
// Backward (spectra -> Legendre-side grid).  Work-group = (M, radius tile,
// component group, theta block); 4 waves, wave w owns hemisphere rows
// [16*(4*tb+w), +16).  VEC: one vector (pol, dpol, tor); otherwise up to 4
// scalars.  Software pipeline: the spectra of degree chunk ic+1 (32 degrees,
// 16 per parity) and the P/dP fragments of chunk ic+1 are loaded into
// registers while the MFMAs of chunk ic run on the other LDS buffer.
constexpr int kSclPer = 4;

// Column of (radius kk) within the 16-double row (ii, par) of the backward's
// staging: XOR-swizzled by 3 bits of ii and the parity, so that the stores
// (lanes over 16 ii x 2 parities at one radius) reach 32 different banks;
// the MFMA reads (16 radii of one row per 16 lanes) still cover whole rows.
__device__ __forceinline__ int bwd_col(int ii, int par, int kk) {
  return kk ^ (((ii >> 1) & 7) | (par << 3));
}

template <bool VEC>
__global__ void __launch_bounds__(256) k_leg_bwd(
    LegDims d, int cg0, const double *__restrict__ sp, SpecAddr sa, double *__restrict__ out,
    GridAddr ga, const int *__restrict__ jst, const long long *__restrict__ offb,
    const int *__restrict__ ipad, const double *__restrict__ Pb,
    const double *__restrict__ dPb, const double *__restrict__ a_r,
    const double *__restrict__ asin_t) {
  constexpr int NC = VEC ? 3 : kSclPer;  // components per group
  constexpr int NQ = VEC ? 5 : kSclPer;  // P*pol*g3, P*tor, P*dpol, dP*dpol, dP*tor
  constexpr int SBUF = 2 * NC * 2 * 256;
  __shared__ double sB[2][SBUF];  // [buf][sig][c][par][ii][kk]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // logical id = ((M n_rt + rt) ncg + cgl) nblk + tb in runs of a whole
  // order per XCD (xcd_group_order): the theta blocks that read the same
  // spectra are resident together on one L2
  long long b = xcd_group_order(blockIdx.x, gridDim.x, (long long)d.n_rt * d.ncg * d.nblk);
  const int tb = (int)(b % d.nblk);
  b /= d.nblk;
  const int cgl = (int)(b % d.ncg);
  b /= d.ncg;
  const int rt = (int)(b % d.n_rt);
  const int M = d.M0 + (int)(b / d.n_rt);
  const int L = d.L;
  const int n = L - M + 1, ne = (n + 1) / 2;
  const int nsig = (M == 0) ? 1 : 2;
  const int k0 = rt * 16;
  const int t0 = (tb * 4 + wave) * 16;
  const bool active = t0 < d.nth_h;
  const int cg = cg0 + cgl;
  const int nc = VEC ? 3 : min(kSclPer, d.ns - kSclPer * (cg - d.nv));
  const int nd0 = VEC ? 3 * cg : 3 * d.nv + kSclPer * (cg - d.nv);
  const int ip0 = ipad[2 * M], ip1 = ipad[2 * M + 1];
  const long long ob0 = offb[2 * M], ob1 = offb[2 * M + 1];
  const int jb0 = jst[L + M], jb1 = jst[L - M];

  d4 acc[2][2][NQ];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int q = 0; q < NQ; q++) acc[s][p][q] = d4{0.0, 0.0, 0.0, 0.0};

  // staging: 1024 items (sig, degree offset lo, radius kk), 4 per thread
  double sreg[4][NC];
  auto stage_load = [&](int ic) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int idx = tid + 256 * u;
      const int lo = idx & 31, kk = (idx >> 5) & 15, sg = idx >> 9;
      const int l = M + 2 * (ic * 16 + (lo >> 1)) + (lo & 1);
      const int k = k0 + kk;
      const bool valid = sg < nsig && l <= L && k < d.nr;
      const int jbs = sg == 0 ? jb0 : jb1;
      const double *src = sp + (long long)(jbs + l - M) * sa.s_j + (long long)k * sa.s_k +
                          (long long)nd0 * sa.s_nd;
#pragma unroll
      for (int c = 0; c < NC; c++)
        sreg[u][c] = (valid && c < nc) ? src[(long long)c * sa.s_nd] : 0.0;
      if (VEC) sreg[u][0] *= g3_of(l);
    }
  };
  auto stage_store = [&](double *buf) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int idx = tid + 256 * u;
      const int lo = idx & 31, kk = (idx >> 5) & 15, sg = idx >> 9;
      const int par = lo & 1, ii = lo >> 1;
#pragma unroll
      for (int c = 0; c < NC; c++)
        buf[(((sg * NC + c) * 2 + par) * 16 + ii) * 16 + bwd_col(ii, par, kk)] = sreg[u][c];
    }
  };
  double pc[4][2], dpc[4][2], pn[4][2], dpn[4][2];
  auto pload = [&](int ic, double (&a)[4][2], double (&da)[4][2]) {
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
      const long long row = (long long)(ic * 16 + ks * 4 + (lane >> 4)) * d.nth_hp + t0 + (lane & 15);
      const bool v0 = ic * 16 < ip0, v1 = ic * 16 < ip1;
      a[ks][0] = v0 ? Pb[ob0 + row] : 0.0;
      a[ks][1] = v1 ? Pb[ob1 + row] : 0.0;
      da[ks][0] = (VEC && v0) ? dPb[ob0 + row] : 0.0;
      da[ks][1] = (VEC && v1) ? dPb[ob1 + row] : 0.0;
    }
  };

  const int nchunk = (ne + 15) / 16;
  if (nchunk > 0) {
    stage_load(0);
    if (active) pload(0, pc, dpc);
  }
  for (int ic = 0; ic < nchunk; ic++) {
    double *buf = sB[ic & 1];
    stage_store(buf);
    __syncthreads();
    const bool more = ic + 1 < nchunk;
    if (more) {
      stage_load(ic + 1);
      if (active) pload(ic + 1, pn, dpn);
    }
    if (active) {
#pragma unroll
      for (int ks = 0; ks < 4; ks++) {
        const int ii = ks * 4 + (lane >> 4);
#pragma unroll
        for (int par = 0; par < 2; par++) {
          if (ic * 16 >= (par ? ip1 : ip0)) continue;
          const double aP = pc[ks][par], adP = dpc[ks][par];
#pragma unroll
          for (int sg = 0; sg < 2; sg++) {
            if (sg >= nsig) continue;
            const double *bb =
                buf + ((sg * NC * 2 + par) * 16 + ii) * 16 + bwd_col(ii, par, lane & 15);
            if (VEC) {
              const double bpol = bb[0], bdpol = bb[512], btor = bb[1024];
              acc[sg][par][0] = mfma64(aP, bpol, acc[sg][par][0]);
              acc[sg][par][1] = mfma64(aP, btor, acc[sg][par][1]);
              acc[sg][par][2] = mfma64(aP, bdpol, acc[sg][par][2]);
              acc[sg][par][3] = mfma64(adP, bdpol, acc[sg][par][3]);
              acc[sg][par][4] = mfma64(adP, btor, acc[sg][par][4]);
            } else {
              // nc is uniform over the work-group: a group of fewer than
              // kSclPer scalars skips the products of its empty components
#pragma unroll
              for (int c = 0; c < NQ; c++)
                if (c < nc) acc[sg][par][c] = mfma64(aP, bb[c * 512], acc[sg][par][c]);
            }
          }
        }
      }
    }
    if (more && active) {
#pragma unroll
      for (int ks = 0; ks < 4; ks++)
#pragma unroll
        for (int par = 0; par < 2; par++) {
          pc[ks][par] = pn[ks][par];
          dpc[ks][par] = dpn[ks][par];
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
    const long long base = ga.slot_off[sg == 0 ? L + M : L - M] * ga.slot_mul + (long long)k * ga.s_k;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int t = t0 + 4 * r + (lane >> 4);
      if (t >= d.nth_h) continue;
      const bool eq = odd_nth && t == d.nth / 2;
      const long long on = base + (long long)t * ga.s_t;
      const long long os = base + (long long)(d.nth - 1 - t) * ga.s_t;
      if (VEC) {
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
        for (int c = 0; c < NQ; c++) {
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
// weighted symmetric / antisymmetric combinations of the mirrored rows; the
// rows of chunk tc+1 and the P/dP fragments are prefetched into registers
// while chunk tc is reduced.
template <bool VEC>
__global__ void __launch_bounds__(256) k_leg_fwd(
    LegDims d, int cg0, const double *__restrict__ in, GridAddr ga, double *__restrict__ sp,
    SpecAddr sa, const int *__restrict__ jst, const long long *__restrict__ offf,
    const int *__restrict__ ipad, const double *__restrict__ Pf,
    const double *__restrict__ dPf, const double *__restrict__ radius,
    const double *__restrict__ weight, const double *__restrict__ asin_t) {
  constexpr int NC = VEC ? 3 : kSclPer;
  constexpr int NQ = VEC ? 5 : kSclPer;
  constexpr int NU = NC * 2;  // staging items per thread: 2 sig x NC x 256 / 256
  constexpr int SBUF = 2 * NC * 2 * 256;
  __shared__ double sA[2][SBUF];  // [buf][sig][f][sym/asym][tt][kk]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // logical id = ((ib (L+1) + M) n_rt + rt) ncg + cgl in runs of one (M, row
  // block) per XCD (xcd_group_order), sharing the P rows on one L2; the run
  // index ib (L+1) + M puts consecutive orders on consecutive XCDs (row
  // blocks past the degrees of M are empty, and the work of row block ib
  // would otherwise all land on two XCDs)
  long long b = xcd_group_order(blockIdx.x, gridDim.x, (long long)d.n_rt * d.ncg);
  const int cgl = (int)(b % d.ncg);
  b /= d.ncg;
  const int rt = (int)(b % d.n_rt);
  b /= d.n_rt;
  const int M = d.M0 + (int)(b % d.nM);
  const int ib = (int)(b / d.nM);
  const int L = d.L;
  const int ip0 = ipad[2 * M], ip1 = ipad[2 * M + 1];
  if (ib * 64 >= ip0) return;  // whole work-group beyond the degrees of M
  const int nsig = (M == 0) ? 1 : 2;
  const int k0 = rt * 16;
  const int i0 = ib * 64 + wave * 16;
  const bool active = i0 < ip0;
  const int cg = cg0 + cgl;
  const int nc = VEC ? 3 : min(kSclPer, d.ns - kSclPer * (cg - d.nv));
  const int nd0 = VEC ? 3 * cg : 3 * d.nv + kSclPer * (cg - d.nv);
  const long long of0 = offf[2 * M], of1 = offf[2 * M + 1];
  const int slot0 = L + M, slot1 = L - M;
  const bool odd_nth = (d.nth & 1) != 0;
  const int t_eq = odd_nth ? d.nth / 2 : -1;

  d4 acc[2][2][NQ];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int q = 0; q < NQ; q++) acc[s][p][q] = d4{0.0, 0.0, 0.0, 0.0};

  double vn[NU], vs[NU];
  auto stage_load = [&](int tc) {
#pragma unroll
    for (int u = 0; u < NU; u++) {
      const int idx = tid + 256 * u;
      const int kk = idx & 15, tt = (idx >> 4) & 15, f = (idx >> 8) % NC, sg = idx / (256 * NC);
      const int t = tc * 16 + tt, k = k0 + kk;
      const bool valid = sg < nsig && f < nc && t < d.nth_h && k < d.nr;
      const long long base = ga.slot_off[sg == 0 ? slot0 : slot1] * ga.slot_mul +
                             (long long)(nd0 + f) * ga.s_nd + (long long)k * ga.s_k;
      vn[u] = valid ? in[base + (long long)t * ga.s_t] : 0.0;
      vs[u] = (valid && t != t_eq) ? in[base + (long long)(d.nth - 1 - t) * ga.s_t] : 0.0;
    }
  };
  auto stage_store = [&](int tc, double *buf) {
#pragma unroll
    for (int u = 0; u < NU; u++) {
      const int idx = tid + 256 * u;
      const int kk = idx & 15, tt = (idx >> 4) & 15, f = (idx >> 8) % NC, sg = idx / (256 * NC);
      const int t = tc * 16 + tt;
      const double w = t < d.nth_h ? weight[t] : 0.0;
      double *dst = buf + (((sg * NC + f) * 2) * 16 + tt) * 16 + kk;
      dst[0] = (vn[u] + vs[u]) * w;
      dst[256] = (t == t_eq) ? 0.0 : (vn[u] - vs[u]) * w;
    }
  };
  double pc[4][2], dpc[4][2], pn[4][2], dpn[4][2], asc[4], asn_[4];
  auto pload = [&](int tc, double (&a)[4][2], double (&da)[4][2], double (&as)[4]) {
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
      const int t = tc * 16 + ks * 4 + (lane >> 4);
      const long long c = i0 + (lane & 15);
      const bool v1 = i0 < ip1;
      a[ks][0] = Pf[of0 + (long long)t * ip0 + c];
      a[ks][1] = v1 ? Pf[of1 + (long long)t * ip1 + c] : 0.0;
      da[ks][0] = VEC ? dPf[of0 + (long long)t * ip0 + c] : 0.0;
      da[ks][1] = (VEC && v1) ? dPf[of1 + (long long)t * ip1 + c] : 0.0;
      as[ks] = (VEC && t < d.nth_h) ? asin_t[t] : 0.0;
    }
  };

  const int nchunk = (d.nth_h + 15) / 16;
  stage_load(0);
  if (active) pload(0, pc, dpc, asc);
  for (int tc = 0; tc < nchunk; tc++) {
    double *buf = sA[tc & 1];
    stage_store(tc, buf);
    __syncthreads();
    const bool more = tc + 1 < nchunk;
    if (more) {
      stage_load(tc + 1);
      if (active) pload(tc + 1, pn, dpn, asn_);
    }
    if (active) {
#pragma unroll
      for (int ks = 0; ks < 4; ks++) {
        const int tt = ks * 4 + (lane >> 4);
        const double asn = asc[ks];
#pragma unroll
        for (int par = 0; par < 2; par++) {
          if (i0 >= (par ? ip1 : ip0)) continue;
          const double bP = pc[ks][par], bdP = dpc[ks][par];
          const int sP = par, sdP = 1 - par;  // P: sym for even, dP: asym for even
#pragma unroll
          for (int sg = 0; sg < 2; sg++) {
            if (sg >= nsig) continue;
            const int sc = (nsig == 1) ? 0 : 1 - sg;
            const double *aa = buf + ((sg * NC * 2) * 16 + tt) * 16 + (lane & 15);
            const double *ac = buf + ((sc * NC * 2) * 16 + tt) * 16 + (lane & 15);
            if (VEC) {
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
              for (int c = 0; c < NQ; c++)
                if (c < nc) acc[sg][par][c] = mfma64(aa[(c * 2 + sP) * 256], bP, acc[sg][par][c]);
            }
          }
        }
      }
    }
    if (more && active) {
#pragma unroll
      for (int ks = 0; ks < 4; ks++) {
        asc[ks] = asn_[ks];
#pragma unroll
        for (int par = 0; par < 2; par++) {
          pc[ks][par] = pn[ks][par];
          dpc[ks][par] = dpn[ks][par];
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
        const long long o = (long long)(jst[sg == 0 ? slot0 : slot1] + l - M) * sa.s_j +
                            (long long)k * sa.s_k + (long long)nd0 * sa.s_nd;
        if (VEC) {
          sp[o] = r2 * g7 * acc[sg][par][0][r];
          sp[o + sa.s_nd] = r1 * g7 * (acc[sg][par][1][r] - ms * acc[sg][par][2][r]);
          sp[o + 2 * sa.s_nd] = r1 * g7 * (-acc[sg][par][3][r] - ms * acc[sg][par][4][r]);
        } else {
#pragma unroll
          for (int c = 0; c < NQ; c++)
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
  d.M0 = 0;
  d.nM = p->L + 1;
  return d;
}

int launch_leg_bwd(sht_plan *p, int nv, int ns, const double *sp, SpecAddr sa, double *out,
                   GridAddr ga) {
  const int nsg = (ns + kSclPer - 1) / kSclPer;
  for (int ti = 0; ti < ptile_count(p); ti++) {
    PTile pt;
    int rc = ptile_prepare(p, ti, false, &pt);
    if (rc) return rc;
    // the scalar groups run beside the vector groups (their work-groups
    // fill the tail of the vector launch)
    ForkJoin fj(p, nv > 0 && nsg > 0);
    for (int vec = 1; vec >= 0; vec--) {
      const int ncg = vec ? nv : nsg;
      if (ncg == 0) continue;
      const hipStream_t st = fj.stream(!vec);
      LegDims d = make_dims(p, nv, ns, p->nth_hp / 64);
      d.ncg = ncg;
      d.M0 = pt.M0;
      d.nM = pt.M1 - pt.M0;
      const long long nblocks = (long long)d.nM * d.n_rt * d.ncg * d.nblk;
      const int cg0 = vec ? 0 : nv;
      if (vec) {
        StageTimer tk(p, 5);
        hipLaunchKernelGGL(k_leg_bwd<true>, dim3((unsigned)nblocks), dim3(256), 0, st, d,
                           cg0, sp, sa, out, ga, p->t.jst, pt.offb, p->t.ipad, pt.Pb, pt.dPb,
                           p->t.a_r, p->t.asin_t);
      } else
        hipLaunchKernelGGL(k_leg_bwd<false>, dim3((unsigned)nblocks), dim3(256), 0, st, d,
                           cg0, sp, sa, out, ga, p->t.jst, pt.offb, p->t.ipad, pt.Pb, pt.dPb,
                           p->t.a_r, p->t.asin_t);
      KSHT_HIP(hipGetLastError());
    }
  }
  return SHT_OK;
}

int launch_leg_fwd(sht_plan *p, int nv, int ns, const double *in, GridAddr ga, double *sp,
                   SpecAddr sa) {
  const int ne_max = (p->L + 2) / 2;
  const int nsg = (ns + kSclPer - 1) / kSclPer;
  for (int ti = 0; ti < ptile_count(p); ti++) {
    PTile pt;
    int rc = ptile_prepare(p, ti, true, &pt);
    if (rc) return rc;
    ForkJoin fj(p, nv > 0 && nsg > 0);  // scalar groups beside the vector groups
    for (int vec = 1; vec >= 0; vec--) {
      const int ncg = vec ? nv : nsg;
      if (ncg == 0) continue;
      const hipStream_t st = fj.stream(!vec);
      LegDims d = make_dims(p, nv, ns, (ne_max + 63) / 64);
      d.ncg = ncg;
      d.M0 = pt.M0;
      d.nM = pt.M1 - pt.M0;
      const long long nblocks = (long long)d.nM * d.n_rt * d.ncg * d.nblk;
      const int cg0 = vec ? 0 : nv;
      if (vec) {
        StageTimer tk(p, 6);
        hipLaunchKernelGGL(k_leg_fwd<true>, dim3((unsigned)nblocks), dim3(256), 0, st, d,
                           cg0, in, ga, sp, sa, p->t.jst, pt.offf, p->t.ipad, pt.Pf, pt.dPf,
                           p->t.radius, p->t.weight, p->t.asin_t);
      } else
        hipLaunchKernelGGL(k_leg_fwd<false>, dim3((unsigned)nblocks), dim3(256), 0, st, d,
                           cg0, in, ga, sp, sa, p->t.jst, pt.offf, p->t.ipad, pt.Pf, pt.dPf,
                           p->t.radius, p->t.weight, p->t.asin_t);
      KSHT_HIP(hipGetLastError());
    }
  }
  return SHT_OK;
}

}  // namespace ksht
