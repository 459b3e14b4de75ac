// sht_split.hip -- split-precision Legendre transforms on INT8 MFMA.
//
// Same contractions as sht_legendre.hip (per order M, both slots +-M, even /
// odd l-M, fused pack and unpack), but every 64-wide chunk of the reduction
// index runs as an Ozaki-style sliced product:
//   A = 2^(eA(row)+2) sum_s a_s 2^-8(s+1),  B = 2^(eB(col)+2) sum_t b_t 2^-8(t+1),
// a_s, b_t balanced base-256 digits in [-128, 127] (round to nearest), with
// per-row exponents of the P tables per chunk (precomputed in FP64 at plan
// time) and per-column exponents of the data operand per chunk (computed on
// the fly).  Products with s + t = g are summed exactly in INT32 by
// v_mfma_i32_16x16x64_i8 (|sum| <= S * 64 * 128^2 < 2^31) for the S groups
// g < S (triangular truncation), and each group is folded into the FP64
// accumulator as ldexp((double)acc_g, eA + eB + 4 - 8(g+2)) at the end of
// the chunk.  FP64 remains on the GPU for P generation, Gauss weights, the
// operand exponents and this recombination.
//
// INT8 MFMA layout (sht_selftest): A(16x64) lane holds row lane%16, bytes
// K = 16*(lane/16) .. +15; B(64x16) lane holds column lane%16, the same K
// bytes; D lane holds D[4*(lane/16)+r][lane%16] in element r.
#include <cmath>

#include "sht_internal.h"

namespace ksht {

constexpr int kEmpty = -4000;  // exponent of an all-zero row / column

struct SplitState {
  int S;
  // backward tables: rows theta (nth_hp), K = degree pair index i, chunk 64
  std::vector<long long> hb_off, hb_eoff;  // per (M,par): byte / exponent offsets
  std::vector<long long> hf_off, hf_eoff;  // forward: rows i (ipad), K = theta
  long long *b_off = nullptr, *b_eoff = nullptr, *f_off = nullptr, *f_eoff = nullptr;
  signed char *bP = nullptr, *bdP = nullptr;            // backward slices
  int *beP = nullptr, *bedP = nullptr;                  // backward row exponents
  signed char *fP = nullptr, *fdP = nullptr, *fPa = nullptr;  // forward slices
  int *feP = nullptr, *fedP = nullptr, *fePa = nullptr;
};

__device__ __forceinline__ i4 mfma8(i4 a, i4 b, i4 c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// frexp-style exponent: |v| < 2^e; kEmpty for 0, subnormals treated as 0.
__device__ __forceinline__ int exp_of(double v) {
  const long long bits = __double_as_longlong(v);
  const int ex = (int)((bits >> 52) & 0x7ff);
  return ex == 0 ? kEmpty : ex - 1022;
}

// Signed fixed point F = round(v 2^(8S-2-e)) for |v| < 2^e, so |F| <= 2^(8S-2),
// from the IEEE fields with integer operations only; the bias turns it into
// balanced base-256 digits: F + bias = sum_k u_k 256^(S-1-k) with digits
// d_k = u_k - 128 in [-128, 127] for k >= 1 and d_0 = (F + bias) >> 8(S-1) in
// [-64, 64].  Then v ~= 2^(e+2) sum_k d_k 2^-8(k+1) with error <= 2^(e+1-8S),
// unbiased (round to nearest), and every digit fits in an int8.
template <int S>
__device__ __forceinline__ long long fix_of(double v, int e) {
  constexpr long long kBias = 128LL * (((1LL << (8 * (S - 1))) - 1) / 255);
  const long long bits = __double_as_longlong(v);
  const int ex = (int)((bits >> 52) & 0x7ff);
  unsigned long long mag = 0;
  if (ex != 0) {
    const unsigned long long mant = (unsigned long long)(bits & 0xFFFFFFFFFFFFFll) | (1ull << 52);
    const int sh = ex - 1077 - e + 8 * S;  // mag = mant 2^sh; sh <= 8S - 55 when |v| < 2^e
    if (sh >= 0)
      mag = sh < 10 ? (mant << sh) : 0;
    else if (sh > -64)
      mag = (mant + (1ull << (-sh - 1))) >> (-sh);
  }
  const long long F = bits < 0 ? -(long long)mag : (long long)mag;
  return F + kBias;
}

// digit k (0 = most significant) of a biased fixed-point value, as a byte
template <int S>
__device__ __forceinline__ unsigned digit_byte(long long Fb, int k) {
  if (k == 0) return (unsigned)(int)(Fb >> (8 * (S - 1))) & 0xffu;
  return (unsigned)(((int)((Fb >> (8 * (S - 1 - k))) & 255)) - 128) & 0xffu;
}

template <int S>
__device__ __forceinline__ unsigned pack4(const long long *Fb, int k) {
  return digit_byte<S>(Fb[0], k) | (digit_byte<S>(Fb[1], k) << 8) |
         (digit_byte<S>(Fb[2], k) << 16) | (digit_byte<S>(Fb[3], k) << 24);
}

// ---------------------------------------------------------------------------
// Plan-time slicing of the P tables.  One thread per (row, chunk): reads the
// 64 values of the chunk, takes their exponent maximum, writes S rows of 64
// bytes.  Backward: value (row t, K i) = Pb[offb + i*nth_hp + t].  Forward:
// value (row i, K t) = Pf[offf + t*ipad + i] (* 1/sin(theta_t) for Pa).
template <int S>
__global__ void __launch_bounds__(256) k_slice_b(int L, int nth_hp, const long long *offb,
                                                 const int *ipad, const double *P,
                                                 const long long *soff, const long long *eoff,
                                                 signed char *out, int *eout) {
  const int M = blockIdx.z >> 1, par = blockIdx.z & 1;
  const int n = L - M + 1, np = par ? n / 2 : (n + 1) / 2;
  const int nch = (np + 63) / 64;
  const int c = blockIdx.y;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (c >= nch || t >= nth_hp) return;
  const int ip = ipad[2 * M + par];
  const double *src = P + offb[2 * M + par] + t;
  double v[64];
  int e = kEmpty;
#pragma unroll
  for (int q = 0; q < 64; q++) {
    const int i = c * 64 + q;
    v[q] = i < ip ? src[(long long)i * nth_hp] : 0.0;
    e = max(e, exp_of(v[q]));
  }
  eout[eoff[2 * M + par] + (long long)c * nth_hp + t] = e;
  signed char *o = out + soff[2 * M + par] + ((long long)c * S * nth_hp + t) * 64;
  long long Fb[64];
#pragma unroll
  for (int q = 0; q < 64; q++) Fb[q] = fix_of<S>(v[q], e);
#pragma unroll
  for (int s = 0; s < S; s++) {
    int4 pk[4];
    int *w = reinterpret_cast<int *>(pk);
#pragma unroll
    for (int q4 = 0; q4 < 16; q4++) w[q4] = (int)pack4<S>(Fb + 4 * q4, s);
    int4 *dst = reinterpret_cast<int4 *>(o + (long long)s * nth_hp * 64);
#pragma unroll
    for (int q = 0; q < 4; q++) dst[q] = pk[q];
  }
}

template <int S>
__global__ void __launch_bounds__(256) k_slice_f(int L, int nth_h, int nth_hp, const long long *offf,
                                                 const int *ipad, const double *P,
                                                 const double *asin_t, int use_asin,
                                                 const long long *soff, const long long *eoff,
                                                 signed char *out, int *eout) {
  const int M = blockIdx.z >> 1, par = blockIdx.z & 1;
  const int ip = ipad[2 * M + par];
  const int c = blockIdx.y;  // theta chunk
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ip) return;
  const double *src = P + offf[2 * M + par] + i;
  double v[64];
  int e = kEmpty;
#pragma unroll
  for (int q = 0; q < 64; q++) {
    const int t = c * 64 + q;
    double x = src[(long long)t * ip];
    if (use_asin) x = t < nth_h ? x * asin_t[t] : 0.0;
    v[q] = x;
    e = max(e, exp_of(x));
  }
  eout[eoff[2 * M + par] + (long long)c * ip + i] = e;
  signed char *o = out + soff[2 * M + par] + ((long long)c * S * ip + i) * 64;
  long long Fb[64];
#pragma unroll
  for (int q = 0; q < 64; q++) Fb[q] = fix_of<S>(v[q], e);
#pragma unroll
  for (int s = 0; s < S; s++) {
    int4 pk[4];
    int *w = reinterpret_cast<int *>(pk);
#pragma unroll
    for (int q4 = 0; q4 < 16; q4++) w[q4] = (int)pack4<S>(Fb + 4 * q4, s);
    int4 *dst = reinterpret_cast<int4 *>(o + (long long)s * ip * 64);
#pragma unroll
    for (int q = 0; q < 4; q++) dst[q] = pk[q];
  }
}

struct SplitTabs {
  const long long *off, *eoff;
  const signed char *P, *dP, *Pa;
  const int *eP, *edP, *ePa;
};

__device__ __forceinline__ double g3s_of(int l) { return l == 0 ? 0.5 : (double)(l * (l + 1)); }
__device__ __forceinline__ double g7s_of(int l, int M) {
  if (l == 0) return 1.0;
  return M == 0 ? (double)(2 * l + 1) / (double)(2 * l * (l + 1))
                : (double)(2 * l + 1) / (double)(4 * l * (l + 1));
}
__device__ __forceinline__ double g6s_of(int l, int M) {
  if (l == 0 && M == 0) return 0.5;
  return M == 0 ? (double)(2 * l + 1) / 2.0 : (double)(2 * l + 1) / 4.0;
}

struct SplitDims {
  int L, nr, nth, nth_h, nth_hp, nv, ns, n_rt, ncg, nblk;
};
constexpr int kSplitScl = 4;

// ---------------------------------------------------------------------------
// Backward.  Work-group = (M, radius tile, component group, theta block of
// 32 rows), 4 waves: wave w owns rows [16*(2 tb + w%2), +16) and parity
// w/2 of l-M.  Per chunk of 64 pair indices i: a first pass over the
// spectra takes per-column exponents (LDS atomicMax), a second pass slices
// them into LDS, then S groups of INT8 MFMAs.  The odd-parity waves hand
// their FP64 sums to the even-parity waves through LDS for the N/S
// recombination.
constexpr int kSplitThreads = 256;  // 4 waves: 2 tiles x 2 parities

template <int S, bool VEC>
__global__ void __launch_bounds__(kSplitThreads) k_leg_bwd_split(
    SplitDims d, int cg0, const double *__restrict__ sp, SpecAddr sa, double *__restrict__ out,
    GridAddr ga, const int *__restrict__ jst, SplitTabs tb, const int *__restrict__ ipad,
    const double *__restrict__ a_r, const double *__restrict__ asin_t) {
  constexpr int NC = VEC ? 3 : kSplitScl;
  constexpr int NQ = VEC ? 5 : kSplitScl;
  constexpr int SB_I4 = 2 * NC * 2 * S * 16 * 4;  // [sig][c][par][s][k][64 bytes]
  constexpr int EX_I4 = (2 * 2 * NQ * 4 * 64 * 8 + 15) / 16;  // epilogue exchange (doubles)
  __shared__ i4 sB[SB_I4 > EX_I4 ? SB_I4 : EX_I4];
  __shared__ int sE[2 * NC * 2 * 16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  long long bidx = blockIdx.x;
  const int tbk = (int)(bidx % d.nblk);
  bidx /= d.nblk;
  const int cgl = (int)(bidx % d.ncg);
  bidx /= d.ncg;
  const int rt = (int)(bidx % d.n_rt);
  const int M = (int)(bidx / d.n_rt);
  const int L = d.L;
  const int n = L - M + 1;
  const int np0 = (n + 1) / 2, np1 = n / 2;
  const int nsig = (M == 0) ? 1 : 2;
  const int k0 = rt * 16;
  const int wt = wave & 1, par = wave >> 1;
  const int t0 = (tbk * 2 + wt) * 16;
  const bool active = t0 < d.nth_h;
  const int npar = par ? np1 : np0;
  const int cg = cg0 + cgl;
  const int nc = VEC ? 3 : min(kSplitScl, d.ns - kSplitScl * (cg - d.nv));
  const int nd0 = VEC ? 3 * cg : 3 * d.nv + kSplitScl * (cg - d.nv);
  const int jb0 = jst[L + M], jb1 = jst[L - M];
  // staging role: (sg, par, kk, iq): 16 pair indices per thread
  const int iq = tid & 3, kk_s = (tid >> 2) & 15, par_s = (tid >> 6) & 1, sg_s = tid >> 7;
  const int npar_s = par_s ? np1 : np0;
  const int k_s = k0 + kk_s;
  const double *src_s = sp + (long long)((sg_s == 0 ? jb0 : jb1) + par_s) * sa.s_j +
                        (long long)k_s * sa.s_k + (long long)nd0 * sa.s_nd;
  const bool sv = sg_s < nsig && k_s < d.nr;

  double C[2][NQ][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
      for (int r = 0; r < 4; r++) C[a][q][r] = 0.0;

  const int nchunk = (np0 + 63) / 64;
  for (int c = 0; c < nchunk; c++) {
    __syncthreads();
    for (int idx = tid; idx < 2 * NC * 2 * 16; idx += kSplitThreads) sE[idx] = kEmpty;
    __syncthreads();
    {
      int ecol[NC];
#pragma unroll
      for (int cc = 0; cc < NC; cc++) ecol[cc] = kEmpty;
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int i = c * 64 + iq * 16 + u;
        const bool valid = sv && i < npar_s;
        const double *src = src_s + (long long)(2 * i) * sa.s_j;
#pragma unroll
        for (int cc = 0; cc < NC; cc++) {
          double x = (valid && cc < nc) ? src[(long long)cc * sa.s_nd] : 0.0;
          if (VEC && cc == 0) x *= g3s_of(M + 2 * i + par_s);
          ecol[cc] = max(ecol[cc], exp_of(x));
        }
      }
#pragma unroll
      for (int cc = 0; cc < NC; cc++)
        atomicMax(&sE[((sg_s * NC + cc) * 2 + par_s) * 16 + kk_s], ecol[cc]);
    }
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < NC; cc++) {
      const int e = sE[((sg_s * NC + cc) * 2 + par_s) * 16 + kk_s];
      long long F[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int i = c * 64 + iq * 16 + u;
        const bool valid = sv && i < npar_s && cc < nc;
        double x = valid ? src_s[(long long)(2 * i) * sa.s_j + (long long)cc * sa.s_nd] : 0.0;
        if (VEC && cc == 0) x *= g3s_of(M + 2 * i + par_s);
        F[u] = fix_of<S>(x, e);
      }
#pragma unroll
      for (int s = 0; s < S; s++)
        sB[((((sg_s * NC + cc) * 2 + par_s) * S + s) * 16 + kk_s) * 4 + iq] =
            i4{(int)pack4<S>(F, s), (int)pack4<S>(F + 4, s), (int)pack4<S>(F + 8, s),
               (int)pack4<S>(F + 12, s)};
    }
    __syncthreads();
    if (!active || c * 64 >= npar) continue;
    const int row0 = t0 + 4 * (lane >> 4);
    const long long so = tb.off[2 * M + par];
    const long long eo = tb.eoff[2 * M + par] + (long long)c * d.nth_hp;
    i4 aP[S], adP[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
      const long long ab = so + (((long long)c * S + s) * d.nth_hp + t0 + (lane & 15)) * 64 +
                           16 * (lane >> 4);
      aP[s] = *reinterpret_cast<const i4 *>(tb.P + ab);
      if (VEC) adP[s] = *reinterpret_cast<const i4 *>(tb.dP + ab);
    }
    int eAP[4], eAD[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      eAP[r] = tb.eP[eo + row0 + r];
      eAD[r] = VEC ? tb.edP[eo + row0 + r] : kEmpty;
    }
    int eBc[2][NC];
#pragma unroll
    for (int sg = 0; sg < 2; sg++)
#pragma unroll
      for (int cc = 0; cc < NC; cc++) eBc[sg][cc] = sE[((sg * NC + cc) * 2 + par) * 16 + (lane & 15)];
#pragma unroll
    for (int g = 0; g < S; g++) {
      i4 acc[2][NQ];
#pragma unroll
      for (int sg = 0; sg < 2; sg++)
#pragma unroll
        for (int q = 0; q < NQ; q++) acc[sg][q] = i4{0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s <= g; s++) {
        const int t = g - s;
#pragma unroll
        for (int sg = 0; sg < 2; sg++) {
          if (sg >= nsig) continue;
          const i4 *bb = sB + ((((sg * NC) * 2 + par) * S + t) * 16 + (lane & 15)) * 4 + (lane >> 4);
          if (VEC) {
            const i4 bpol = bb[0], bdpol = bb[2 * S * 64], btor = bb[2 * 2 * S * 64];
            acc[sg][0] = mfma8(aP[s], bpol, acc[sg][0]);
            acc[sg][1] = mfma8(aP[s], btor, acc[sg][1]);
            acc[sg][2] = mfma8(aP[s], bdpol, acc[sg][2]);
            acc[sg][3] = mfma8(adP[s], bdpol, acc[sg][3]);
            acc[sg][4] = mfma8(adP[s], btor, acc[sg][4]);
          } else {
#pragma unroll
            for (int q = 0; q < NQ; q++) acc[sg][q] = mfma8(aP[s], bb[q * 2 * S * 64], acc[sg][q]);
          }
        }
      }
      const int sh = 4 - 8 * (g + 2);  // 2^(eA+2) 2^(eB+2) 2^-8(g+2)
#pragma unroll
      for (int sg = 0; sg < 2; sg++) {
        if (sg >= nsig) continue;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          // component of quantity q: pol*g3, tor, dpol, dpol, tor
          const int cq = VEC ? (q == 0 ? 0 : (q == 1 || q == 4) ? 2 : 1) : q;
          const int eb = eBc[sg][cq];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int ea = (VEC && q >= 3) ? eAD[r] : eAP[r];
            C[sg][q][r] += ldexp((double)acc[sg][q][r], ea + eb + sh);
          }
        }
      }
    }
  }
  // odd-parity waves hand their sums to the even-parity waves
  __syncthreads();
  double *X = reinterpret_cast<double *>(sB);
  if (par == 1) {
#pragma unroll
    for (int sg = 0; sg < 2; sg++)
#pragma unroll
      for (int q = 0; q < NQ; q++)
#pragma unroll
        for (int r = 0; r < 4; r++) X[(((wt * 2 + sg) * NQ + q) * 4 + r) * 64 + lane] = C[sg][q][r];
  }
  __syncthreads();
  if (par == 1 || !active) return;
  const int k = k0 + (lane & 15);
  if (k >= d.nr) return;
  const double a1 = a_r[k], a2 = a1 * a1;
  const bool odd_nth = (d.nth & 1) != 0;
#define ODD(sg, q, r) X[(((wt * 2 + (sg)) * NQ + (q)) * 4 + (r)) * 64 + lane]
#pragma unroll
  for (int sg = 0; sg < 2; sg++) {
    if (sg >= nsig) continue;
    const bool self = (nsig == 1);
    const int so_ = self ? sg : 1 - sg;
    const double mc = (sg == 0) ? -(double)M : (double)M;
    const long long base = ga.slot_off[sg == 0 ? L + M : L - M] * ga.slot_mul + (long long)k * ga.s_k;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int t = t0 + 4 * (lane >> 4) + r;
      if (t >= d.nth_h) continue;
      const bool eq = odd_nth && t == d.nth / 2;
      const long long on = base + (long long)t * ga.s_t;
      const long long os = base + (long long)(d.nth - 1 - t) * ga.s_t;
      if (VEC) {
        const double c1 = -asin_t[t] * mc * a1;
        const double E0 = C[sg][0][r], O0 = ODD(sg, 0, r);
        const double E3 = C[sg][3][r], O3 = ODD(sg, 3, r);
        const double E4 = C[sg][4][r], O4 = ODD(sg, 4, r);
        const double Ec1 = self ? C[sg][1][r] : C[1 - sg][1][r], Oc1 = ODD(so_, 1, r);
        const double Ec2 = self ? C[sg][2][r] : C[1 - sg][2][r], Oc2 = ODD(so_, 2, r);
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
        for (int q = 0; q < NQ; q++) {
          if (q >= nc) continue;
          const double E = C[sg][q][r], O = ODD(sg, q, r);
          double *o = out + (long long)(nd0 + q) * ga.s_nd;
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
#undef ODD
}

// ---------------------------------------------------------------------------
// Forward.  Work-group = (M, radius tile, component group, degree block of
// 32 pair indices), 4 waves: wave w owns rows i in [32 ib + 16 (w%2), +16)
// of parity w/2.  Per chunk of 64 hemisphere rows: a first pass over the
// grid coefficients takes the per-column exponents of the weighted
// symmetric / antisymmetric combinations, a second pass slices them into
// LDS, then S groups of INT8 MFMAs against the P, dP and P/sin(theta)
// slices.
template <int S, bool VEC>
__global__ void __launch_bounds__(kSplitThreads) k_leg_fwd_split(
    SplitDims d, int cg0, const double *__restrict__ in, GridAddr ga, double *__restrict__ sp,
    SpecAddr sa, const int *__restrict__ jst, SplitTabs tb, const int *__restrict__ ipad,
    const double *__restrict__ radius, const double *__restrict__ weight) {
  constexpr int NC = VEC ? 3 : kSplitScl;
  constexpr int NQ = VEC ? 5 : kSplitScl;
  __shared__ i4 sB[2 * NC * 2 * S * 16 * 4];  // [sig][f][sym/asym][s][k][64 bytes]
  __shared__ int sE[2 * NC * 2 * 16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  long long bidx = blockIdx.x;
  const int ib = (int)(bidx % d.nblk);
  bidx /= d.nblk;
  const int cgl = (int)(bidx % d.ncg);
  bidx /= d.ncg;
  const int rt = (int)(bidx % d.n_rt);
  const int M = (int)(bidx / d.n_rt);
  const int L = d.L;
  const int ip0 = ipad[2 * M], ip1 = ipad[2 * M + 1];
  if (ib * 32 >= ip0) return;
  const int nsig = (M == 0) ? 1 : 2;
  const int k0 = rt * 16;
  const int par = wave >> 1;
  const int i0 = ib * 32 + (wave & 1) * 16;
  const int ip = par ? ip1 : ip0;
  const bool active = i0 < ip;
  const int cg = cg0 + cgl;
  const int nc = VEC ? 3 : min(kSplitScl, d.ns - kSplitScl * (cg - d.nv));
  const int nd0 = VEC ? 3 * cg : 3 * d.nv + kSplitScl * (cg - d.nv);
  const int slot0 = L + M, slot1 = L - M;
  const bool odd_nth = (d.nth & 1) != 0;
  const int t_eq = odd_nth ? d.nth / 2 : -1;
  // staging role: (sg, kk, tq): 8 theta rows per thread
  const int tq = tid & 7, kk_s = (tid >> 3) & 15, sg_s = tid >> 7;
  const int k_s = k0 + kk_s;
  const bool sv = sg_s < nsig && k_s < d.nr;
  const long long base_s = ga.slot_off[sg_s == 0 ? slot0 : slot1] * ga.slot_mul +
                           (long long)nd0 * ga.s_nd + (long long)k_s * ga.s_k;

  double C[2][NQ][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
      for (int r = 0; r < 4; r++) C[a][q][r] = 0.0;

  const int nchunk = d.nth_hp / 64;
  for (int c = 0; c < nchunk; c++) {
    if (c * 64 >= d.nth_h) break;
    __syncthreads();
    for (int idx = tid; idx < 2 * NC * 2 * 16; idx += kSplitThreads) sE[idx] = kEmpty;
    __syncthreads();
    {
      int es[NC], ea[NC];
#pragma unroll
      for (int f = 0; f < NC; f++) {
        es[f] = kEmpty;
        ea[f] = kEmpty;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int t = c * 64 + tq * 8 + u;
        const bool valid = sv && t < d.nth_h;
        const double w = valid ? weight[t] : 0.0;
#pragma unroll
        for (int f = 0; f < NC; f++) {
          const bool vf = valid && f < nc;
          const long long b = base_s + (long long)f * ga.s_nd;
          const double xn = vf ? in[b + (long long)t * ga.s_t] : 0.0;
          const double xs = (vf && t != t_eq) ? in[b + (long long)(d.nth - 1 - t) * ga.s_t] : 0.0;
          es[f] = max(es[f], exp_of((xn + xs) * w));
          ea[f] = max(ea[f], (t == t_eq) ? kEmpty : exp_of((xn - xs) * w));
        }
      }
#pragma unroll
      for (int f = 0; f < NC; f++) {
        atomicMax(&sE[((sg_s * NC + f) * 2 + 0) * 16 + kk_s], es[f]);
        atomicMax(&sE[((sg_s * NC + f) * 2 + 1) * 16 + kk_s], ea[f]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NC; f++) {
      const int e0 = sE[((sg_s * NC + f) * 2 + 0) * 16 + kk_s];
      const int e1 = sE[((sg_s * NC + f) * 2 + 1) * 16 + kk_s];
      long long Fs[8], Fa[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int t = c * 64 + tq * 8 + u;
        const bool vf = sv && t < d.nth_h && f < nc;
        const double w = vf ? weight[t] : 0.0;
        const long long b = base_s + (long long)f * ga.s_nd;
        const double xn = vf ? in[b + (long long)t * ga.s_t] : 0.0;
        const double xs = (vf && t != t_eq) ? in[b + (long long)(d.nth - 1 - t) * ga.s_t] : 0.0;
        Fs[u] = fix_of<S>((xn + xs) * w, e0);
        Fa[u] = fix_of<S>((t == t_eq) ? 0.0 : (xn - xs) * w, e1);
      }
#pragma unroll
      for (int s = 0; s < S; s++) {
        const unsigned ws0 = pack4<S>(Fs, s), ws1 = pack4<S>(Fs + 4, s);
        const unsigned wa0 = pack4<S>(Fa, s), wa1 = pack4<S>(Fa + 4, s);
        int *ds = reinterpret_cast<int *>(sB + ((((sg_s * NC + f) * 2 + 0) * S + s) * 16 + kk_s) * 4);
        int *da = reinterpret_cast<int *>(sB + ((((sg_s * NC + f) * 2 + 1) * S + s) * 16 + kk_s) * 4);
        ds[2 * tq] = (int)ws0;
        ds[2 * tq + 1] = (int)ws1;
        da[2 * tq] = (int)wa0;
        da[2 * tq + 1] = (int)wa1;
      }
    }
    __syncthreads();
    if (!active) continue;
    const int row0 = i0 + 4 * (lane >> 4);
    const long long so = tb.off[2 * M + par];
    const long long eo = tb.eoff[2 * M + par] + (long long)c * ip;
    i4 aP[S], adP[S], aPa[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
      const long long ab = so + (((long long)c * S + s) * ip + i0 + (lane & 15)) * 64 + 16 * (lane >> 4);
      aP[s] = *reinterpret_cast<const i4 *>(tb.P + ab);
      if (VEC) {
        adP[s] = *reinterpret_cast<const i4 *>(tb.dP + ab);
        aPa[s] = *reinterpret_cast<const i4 *>(tb.Pa + ab);
      }
    }
    int eAP[4], eAD[4], eAA[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      eAP[r] = tb.eP[eo + row0 + r];
      eAD[r] = VEC ? tb.edP[eo + row0 + r] : kEmpty;
      eAA[r] = VEC ? tb.ePa[eo + row0 + r] : kEmpty;
    }
    const int sP = par, sdP = 1 - par;  // P uses sym for even l-M, dP asym
    constexpr int FS = 2 * S * 64;      // stride of f (in i4)
    constexpr int HS = S * 64;          // stride of sym/asym
#pragma unroll
    for (int g = 0; g < S; g++) {
      i4 acc[2][NQ];
#pragma unroll
      for (int sg = 0; sg < 2; sg++)
#pragma unroll
        for (int q = 0; q < NQ; q++) acc[sg][q] = i4{0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s <= g; s++) {
        const int t = g - s;
#pragma unroll
        for (int sg = 0; sg < 2; sg++) {
          if (sg >= nsig) continue;
          const int sc = (nsig == 1) ? 0 : 1 - sg;
          const i4 *bb = sB + (((sg * NC * 2) * S + t) * 16 + (lane & 15)) * 4 + (lane >> 4);
          const i4 *bc = sB + (((sc * NC * 2) * S + t) * 16 + (lane & 15)) * 4 + (lane >> 4);
          if (VEC) {
            const i4 bvr = bb[0 * FS + sP * HS], bvt = bb[1 * FS + sdP * HS], bvp = bb[2 * FS + sdP * HS];
            const i4 bvpc = bc[2 * FS + sP * HS], bvtc = bc[1 * FS + sP * HS];
            acc[sg][0] = mfma8(aP[s], bvr, acc[sg][0]);
            acc[sg][1] = mfma8(adP[s], bvt, acc[sg][1]);
            acc[sg][2] = mfma8(aPa[s], bvpc, acc[sg][2]);
            acc[sg][3] = mfma8(adP[s], bvp, acc[sg][3]);
            acc[sg][4] = mfma8(aPa[s], bvtc, acc[sg][4]);
          } else {
#pragma unroll
            for (int q = 0; q < NQ; q++) acc[sg][q] = mfma8(aP[s], bb[q * FS + sP * HS], acc[sg][q]);
          }
        }
      }
      const int sh = 4 - 8 * (g + 2);  // 2^(eA+2) 2^(eB+2) 2^-8(g+2)
#pragma unroll
      for (int sg = 0; sg < 2; sg++) {
        if (sg >= nsig) continue;
        const int sc = (nsig == 1) ? 0 : 1 - sg;
        const int lk = lane & 15;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          int eb;
          if (VEC) {
            eb = q == 0 ? sE[((sg * NC + 0) * 2 + sP) * 16 + lk]
               : q == 1 ? sE[((sg * NC + 1) * 2 + sdP) * 16 + lk]
               : q == 2 ? sE[((sc * NC + 2) * 2 + sP) * 16 + lk]
               : q == 3 ? sE[((sg * NC + 2) * 2 + sdP) * 16 + lk]
                        : sE[((sc * NC + 1) * 2 + sP) * 16 + lk];
          } else {
            eb = sE[((sg * NC + q) * 2 + sP) * 16 + lk];
          }
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int ea2 = VEC ? (q == 0 ? eAP[r] : (q == 1 || q == 3) ? eAD[r] : eAA[r]) : eAP[r];
            C[sg][q][r] += ldexp((double)acc[sg][q][r], ea2 + eb + sh);
          }
        }
      }
    }
  }
  if (!active) return;
  const int k = k0 + (lane & 15);
  if (k >= d.nr) return;
  const double r1 = radius[k], r2 = r1 * r1;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int ii = i0 + 4 * (lane >> 4) + r;
    const int l = M + 2 * ii + par;
    if (l > L) continue;
    const double g7 = g7s_of(l, M), g6 = g6s_of(l, M);
#pragma unroll
    for (int sg = 0; sg < 2; sg++) {
      if (sg >= nsig) continue;
      const double ms = (sg == 0) ? (double)M : -(double)M;
      const long long o = (long long)(jst[sg == 0 ? slot0 : slot1] + l - M) * sa.s_j +
                          (long long)k * sa.s_k + (long long)nd0 * sa.s_nd;
      if (VEC) {
        sp[o] = r2 * g7 * C[sg][0][r];
        sp[o + sa.s_nd] = r1 * g7 * (C[sg][1][r] - ms * C[sg][2][r]);
        sp[o + 2 * sa.s_nd] = r1 * g7 * (-C[sg][3][r] - ms * C[sg][4][r]);
      } else {
#pragma unroll
        for (int q = 0; q < NQ; q++)
          if (q < nc) sp[o + (long long)q * sa.s_nd] = g6 * C[sg][q][r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
static int table_layouts(sht_plan *p, SplitState *st) {
  const int L = p->L, S = st->S;
  st->hb_off.assign(2 * (L + 1), 0);
  st->hb_eoff.assign(2 * (L + 1), 0);
  st->hf_off.assign(2 * (L + 1), 0);
  st->hf_eoff.assign(2 * (L + 1), 0);
  long long bo = 0, be = 0, fo = 0, fe = 0;
  const int nchf = p->nth_hp / 64;
  for (int M = 0; M <= L; M++) {
    const int n = L - M + 1;
    for (int par = 0; par < 2; par++) {
      const int np = par ? n / 2 : (n + 1) / 2;
      const int nchb = (np + 63) / 64;
      const int ip = p->h_ipad[2 * M + par];
      st->hb_off[2 * M + par] = bo;
      st->hb_eoff[2 * M + par] = be;
      bo += (long long)nchb * S * p->nth_hp * 64;
      be += (long long)nchb * p->nth_hp;
      st->hf_off[2 * M + par] = fo;
      st->hf_eoff[2 * M + par] = fe;
      fo += (long long)nchf * S * ip * 64;
      fe += (long long)nchf * ip;
    }
  }
  auto alloc = [&](void **ptr, size_t bytes) -> int {
    if (hipMalloc(ptr, bytes ? bytes : 64) != hipSuccess) {
      set_error("split tables: allocation of %zu bytes failed", bytes);
      return SHT_ERR_OOM;
    }
    p->device_bytes += (long long)bytes;
    return SHT_OK;
  };
  int rc = SHT_OK;
  const size_t nb = 2 * (L + 1) * sizeof(long long);
  if ((rc = alloc((void **)&st->b_off, nb)) || (rc = alloc((void **)&st->b_eoff, nb)) ||
      (rc = alloc((void **)&st->f_off, nb)) || (rc = alloc((void **)&st->f_eoff, nb)) ||
      (rc = alloc((void **)&st->bP, bo)) || (rc = alloc((void **)&st->bdP, bo)) ||
      (rc = alloc((void **)&st->beP, be * 4)) || (rc = alloc((void **)&st->bedP, be * 4)) ||
      (rc = alloc((void **)&st->fP, fo)) || (rc = alloc((void **)&st->fdP, fo)) ||
      (rc = alloc((void **)&st->fPa, fo)) || (rc = alloc((void **)&st->feP, fe * 4)) ||
      (rc = alloc((void **)&st->fedP, fe * 4)) || (rc = alloc((void **)&st->fePa, fe * 4)))
    return rc;
  KSHT_HIP(hipMemcpy(st->b_off, st->hb_off.data(), nb, hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(st->b_eoff, st->hb_eoff.data(), nb, hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(st->f_off, st->hf_off.data(), nb, hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(st->f_eoff, st->hf_eoff.data(), nb, hipMemcpyHostToDevice));
  return SHT_OK;
}

template <int S>
static int slice_tables(sht_plan *p, SplitState *st) {
  const int L = p->L;
  const int nchb_max = ((L + 2) / 2 + 63) / 64;
  const dim3 gb((p->nth_hp + 255) / 256, nchb_max, 2 * (L + 1));
  hipLaunchKernelGGL(k_slice_b<S>, gb, dim3(256), 0, p->stream, L, p->nth_hp, p->t.offb,
                     p->t.ipad, p->t.Pb, st->b_off, st->b_eoff, st->bP, st->beP);
  hipLaunchKernelGGL(k_slice_b<S>, gb, dim3(256), 0, p->stream, L, p->nth_hp, p->t.offb,
                     p->t.ipad, p->t.dPb, st->b_off, st->b_eoff, st->bdP, st->bedP);
  int ipmax = 0;
  for (int v : p->h_ipad) ipmax = v > ipmax ? v : ipmax;
  const dim3 gf((ipmax + 255) / 256, p->nth_hp / 64, 2 * (L + 1));
  hipLaunchKernelGGL(k_slice_f<S>, gf, dim3(256), 0, p->stream, L, p->nth_h, p->nth_hp, p->t.offf,
                     p->t.ipad, p->t.Pf, p->t.asin_t, 0, st->f_off, st->f_eoff, st->fP, st->feP);
  hipLaunchKernelGGL(k_slice_f<S>, gf, dim3(256), 0, p->stream, L, p->nth_h, p->nth_hp, p->t.offf,
                     p->t.ipad, p->t.dPf, p->t.asin_t, 0, st->f_off, st->f_eoff, st->fdP, st->fedP);
  hipLaunchKernelGGL(k_slice_f<S>, gf, dim3(256), 0, p->stream, L, p->nth_h, p->nth_hp, p->t.offf,
                     p->t.ipad, p->t.Pf, p->t.asin_t, 1, st->f_off, st->f_eoff, st->fPa, st->fePa);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

int split_init(sht_plan *p) {
  int S = p->n_slices == 0 ? 6 : p->n_slices;
  if (S < 4 || S > 7) {
    set_error("split precision: n_slices must be 4..7 (got %d)", p->n_slices);
    return SHT_ERR_ARG;
  }
  SplitState *st = new SplitState();
  st->S = S;
  p->split = st;
  p->n_slices = S;
  int rc = table_layouts(p, st);
  if (rc) return rc;
  // the FP64 tables must be complete before slicing (same stream)
  switch (S) {
    case 4: return slice_tables<4>(p, st);
    case 5: return slice_tables<5>(p, st);
    case 6: return slice_tables<6>(p, st);
    default: return slice_tables<7>(p, st);
  }
}

void split_free(sht_plan *p) {
  SplitState *st = (SplitState *)p->split;
  if (!st) return;
  void *ptrs[] = {st->b_off, st->b_eoff, st->f_off, st->f_eoff, st->bP, st->bdP, st->beP, st->bedP,
                  st->fP, st->fdP, st->fPa, st->feP, st->fedP, st->fePa};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete st;
  p->split = nullptr;
}

static SplitDims split_dims(const sht_plan *p, int nv, int ns, int nblk) {
  SplitDims d;
  d.L = p->L;
  d.nr = p->nr;
  d.nth = p->nth;
  d.nth_h = p->nth_h;
  d.nth_hp = p->nth_hp;
  d.nv = nv;
  d.ns = ns;
  d.n_rt = p->nr_p / 16;
  d.ncg = 0;
  d.nblk = nblk;
  return d;
}

template <int S>
static int bwd_split_S(sht_plan *p, int nv, int ns, const double *sp, SpecAddr sa, double *out,
                       GridAddr ga) {
  SplitState *st = (SplitState *)p->split;
  SplitTabs tb{st->b_off, st->b_eoff, st->bP, st->bdP, nullptr, st->beP, st->bedP, nullptr};
  const int nsg = (ns + kSplitScl - 1) / kSplitScl;
  for (int vec = 1; vec >= 0; vec--) {
    SplitDims d = split_dims(p, nv, ns, p->nth_hp / 32);
    d.ncg = vec ? nv : nsg;
    if (d.ncg == 0) continue;
    const long long nb = (long long)(p->L + 1) * d.n_rt * d.ncg * d.nblk;
    if (vec)
      hipLaunchKernelGGL((k_leg_bwd_split<S, true>), dim3((unsigned)nb), dim3(kSplitThreads), 0, p->stream, d,
                         0, sp, sa, out, ga, p->t.jst, tb, p->t.ipad, p->t.a_r, p->t.asin_t);
    else
      hipLaunchKernelGGL((k_leg_bwd_split<S, false>), dim3((unsigned)nb), dim3(kSplitThreads), 0, p->stream,
                         d, nv, sp, sa, out, ga, p->t.jst, tb, p->t.ipad, p->t.a_r, p->t.asin_t);
    KSHT_HIP(hipGetLastError());
  }
  return SHT_OK;
}

template <int S>
static int fwd_split_S(sht_plan *p, int nv, int ns, const double *in, GridAddr ga, double *sp,
                       SpecAddr sa) {
  SplitState *st = (SplitState *)p->split;
  SplitTabs tb{st->f_off, st->f_eoff, st->fP, st->fdP, st->fPa, st->feP, st->fedP, st->fePa};
  const int nsg = (ns + kSplitScl - 1) / kSplitScl;
  const int ne_max = (p->L + 2) / 2;
  for (int vec = 1; vec >= 0; vec--) {
    SplitDims d = split_dims(p, nv, ns, (ne_max + 31) / 32);
    d.ncg = vec ? nv : nsg;
    if (d.ncg == 0) continue;
    const long long nb = (long long)(p->L + 1) * d.n_rt * d.ncg * d.nblk;
    if (vec)
      hipLaunchKernelGGL((k_leg_fwd_split<S, true>), dim3((unsigned)nb), dim3(kSplitThreads), 0, p->stream, d,
                         0, in, ga, sp, sa, p->t.jst, tb, p->t.ipad, p->t.radius, p->t.weight);
    else
      hipLaunchKernelGGL((k_leg_fwd_split<S, false>), dim3((unsigned)nb), dim3(kSplitThreads), 0, p->stream,
                         d, nv, in, ga, sp, sa, p->t.jst, tb, p->t.ipad, p->t.radius, p->t.weight);
    KSHT_HIP(hipGetLastError());
  }
  return SHT_OK;
}

int launch_leg_bwd_split(sht_plan *p, int nv, int ns, const double *sp, SpecAddr sa, double *out,
                         GridAddr ga) {
  if (!p->split) {
    set_error("split precision: plan was not created in split mode");
    return SHT_ERR_STATE;
  }
  switch (((SplitState *)p->split)->S) {
    case 4: return bwd_split_S<4>(p, nv, ns, sp, sa, out, ga);
    case 5: return bwd_split_S<5>(p, nv, ns, sp, sa, out, ga);
    case 6: return bwd_split_S<6>(p, nv, ns, sp, sa, out, ga);
    default: return bwd_split_S<7>(p, nv, ns, sp, sa, out, ga);
  }
}

int launch_leg_fwd_split(sht_plan *p, int nv, int ns, const double *in, GridAddr ga, double *sp,
                         SpecAddr sa) {
  if (!p->split) {
    set_error("split precision: plan was not created in split mode");
    return SHT_ERR_STATE;
  }
  switch (((SplitState *)p->split)->S) {
    case 4: return fwd_split_S<4>(p, nv, ns, in, ga, sp, sa);
    case 5: return fwd_split_S<5>(p, nv, ns, in, ga, sp, sa);
    case 6: return fwd_split_S<6>(p, nv, ns, in, ga, sp, sa);
    default: return fwd_split_S<7>(p, nv, ns, in, ga, sp, sa);
  }
}

}  // namespace ksht
