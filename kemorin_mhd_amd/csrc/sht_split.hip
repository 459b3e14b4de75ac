// sht_split.hip -- split-precision Legendre transforms on INT8 MFMA.
//
// Same contractions as sht_legendre.hip (per order M, both slots +-M, even /
// odd l-M, fused pack and unpack), with every 64-wide chunk of the reduction
// index run as an Ozaki-style sliced product:
//   A = 2^(eA(row)+2) sum_s a_s 2^-8(s+1),  B = 2^(eB(col)+2) sum_t b_t 2^-8(t+1),
// a_s, b_t balanced base-256 digits (round to nearest; a_0 in [-64, 64], the
// others in [-128, 127]), with per-row exponents of the P tables per chunk
// (sliced from the FP64 tables at plan time) and per-column exponents of the
// data operand per chunk (sliced once per call by k_bslice / k_fslice).
// Products with s + t = g are summed exactly in INT32 by
// v_mfma_i32_16x16x64_i8 for the groups g < S (triangular truncation).  With
// |a_0|, |b_0| <= 64 and the other digits <= 128, |acc_g| <= 2^20 g for g >= 1
// (2^18 for g = 0), so two groups combine exactly in INT32 as
// 256 acc_g + acc_(g+1) (< 2^28 * 6 + 2^23 < 2^31 for g <= 6), and each such
// pair is folded into the FP64 accumulators as
// ldexp((double)pair, eA + eB + 4 - 8(g+3)).  FP64 remains on the GPU for P
// generation, Gauss weights, the operand exponents and this recombination.
//
// Slices are stored as 1 KiB blocks of 16 rows (P tables) or 16 columns
// (data operand) x 64 K bytes, one 64-byte row each; a wave moves a block
// straight from global memory into LDS with one 16-byte load per lane, lane L
// taking row L%16, K bytes 16 (L/16) .. +15, which lands the block in MFMA
// fragment order (lane L's bytes at 16 L).
//
// INT8 MFMA layout (sht_selftest): A(16x64) lane holds row lane%16, bytes
// K = 16*(lane/16) .. +15; B(64x16) lane holds column lane%16, the same K
// bytes; D lane holds D[4*(lane/16)+r][lane%16] in element r.
#include <cmath>

#include "sht_internal.h"

namespace ksht {

constexpr int kEmpty = -4000;  // exponent of an all-zero row / column

// Backward reduction chunks over the pair index i (l = M + 2i + parity): 64
// wide, except that orders M < kNarrowM start with [0, 16) and [16, 64)
// (zero-padded to 64 K lanes).  At low orders the spectra fall and P/sin, dP
// grow across l = M .. M+127 by orders of magnitude; per-chunk exponents over
// narrower degree ranges keep the sliced products accurate there.
constexpr int kNarrowM = 64;
__host__ __device__ __forceinline__ int bchunk_count(int M, int np) {
  if (np <= 0) return 0;
  if (M >= kNarrowM) return (np + 63) / 64;
  return np <= 16 ? 1 : np <= 64 ? 2 : 1 + (np + 63) / 64;
}
__host__ __device__ __forceinline__ int bchunk_start(int M, int c) {
  if (M >= kNarrowM) return 64 * c;
  return c == 0 ? 0 : c == 1 ? 16 : 64 * (c - 1);
}
__host__ __device__ __forceinline__ int bchunk_end(int M, int c) {
  if (M >= kNarrowM) return 64 * c + 64;
  return c == 0 ? 16 : c == 1 ? 64 : 64 * c;
}

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
  // data-operand slices (one transform at a time per plan, so the backward
  // and forward share them)
  std::vector<long long> hcb;  // backward: first chunk of (slot, parity)
  long long *cb = nullptr;
  long long nchb = 0;          // backward chunks over all (slot, parity)
  int nchf = 0;                // forward chunks of 64 hemisphere rows
  signed char *dig = nullptr;  // digits, fragment order
  int *dexp = nullptr;         // per (chunk, component, radius) exponents
  // P computed per call (plan flag): slices of one m-tile at a time, at
  // offsets rebased to the tile's first block; the tile buffers above (bP ..
  // fePa) then hold one tile, the backward and forward sharing them
  long long *b_off_r = nullptr, *b_eoff_r = nullptr, *f_off_r = nullptr, *f_eoff_r = nullptr;
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
// For S <= 6 the shift is a right shift for every non-zero v (sh <= 8S - 55
// < 0), and a zero mantissa rounds to zero for any right shift: no branches.
template <int S>
__device__ __forceinline__ long long fix_of(double v, int e) {
  constexpr long long kBias = 128LL * (((1LL << (8 * (S - 1))) - 1) / 255);
  const long long bits = __double_as_longlong(v);
  const int ex = (int)((bits >> 52) & 0x7ff);
  const int sh = ex - 1077 - e + 8 * S;  // mag = mant 2^sh; sh <= 8S - 55 when |v| < 2^e
  unsigned long long mag;
  if constexpr (8 * S - 55 < 0) {
    const unsigned long long mant =
        ex != 0 ? ((unsigned long long)(bits & 0xFFFFFFFFFFFFFll) | (1ull << 52)) : 0ull;
    const int rs = min(max(-sh, 1), 63);
    mag = (mant + (1ull << (rs - 1))) >> rs;
  } else {
    mag = 0;
    if (ex != 0) {
      const unsigned long long mant =
          (unsigned long long)(bits & 0xFFFFFFFFFFFFFll) | (1ull << 52);
      if (sh >= 0)
        mag = sh < 10 ? (mant << sh) : 0;
      else if (sh > -64)
        mag = (mant + (1ull << (-sh - 1))) >> (-sh);
    }
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

// digit k of four values, one byte each: the byte of the 32-bit half that
// holds it (u_k - 128 = u_k ^ 0x80 as a byte for k >= 1)
template <int S>
__device__ __forceinline__ unsigned pack4(const long long *Fb, int k) {
  const int bit = 8 * (S - 1 - k);
  unsigned r = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const unsigned w = (unsigned)((unsigned long long)Fb[i] >> (bit & ~31));
    r |= ((w >> (bit & 31)) & 0xffu) << (8 * i);
  }
  return k == 0 ? r : r ^ 0x80808080u;
}

// ---------------------------------------------------------------------------
// Plan-time slicing of the P tables.  One thread per (row, chunk): reads the
// 64 values of the chunk, takes their exponent maximum, writes S rows of 64
// bytes.  Backward: value (row t, K i) = Pb[offb + i*nth_hp + t].  Forward:
// value (row i, K t) = Pf[offf + t*ipad + i] (* 1/sin(theta_t) for Pa).
template <int S>
__global__ void __launch_bounds__(256) k_slice_b(int L, int M0, int nth_hp, const long long *offb,
                                                 const int *ipad, const double *P,
                                                 const long long *soff, const long long *eoff,
                                                 signed char *out, int *eout) {
  const int M = M0 + (blockIdx.z >> 1), par = blockIdx.z & 1;
  const int n = L - M + 1, np = par ? n / 2 : (n + 1) / 2;
  const int nch = bchunk_count(M, np);
  const int c = blockIdx.y;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (c >= nch || t >= nth_hp) return;
  const int ip = ipad[2 * M + par];
  const int i0 = bchunk_start(M, c), i1 = min(bchunk_end(M, c), ip);
  const double *src = P + offb[2 * M + par] + t;
  double v[64];
  int e = kEmpty;
#pragma unroll
  for (int q = 0; q < 64; q++) {
    const int i = i0 + q;
    v[q] = i < i1 ? src[(long long)i * nth_hp] : 0.0;
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
__global__ void __launch_bounds__(256) k_slice_f(int L, int M0, int nth_h, int nth_hp, const long long *offf,
                                                 const int *ipad, const double *P,
                                                 const double *asin_t, int use_asin,
                                                 const long long *soff, const long long *eoff,
                                                 signed char *out, int *eout) {
  const int M = M0 + (blockIdx.z >> 1), par = blockIdx.z & 1;
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

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// byte offset of (column c16, K index q) in a 1 KiB slice block: row-major,
// one 64-byte row per column, so a slicing wave (lane = K index) writes whole
// rows; the GEMM copies a block into LDS in MFMA fragment order by giving
// lane L the source bytes (row L%16, K 16 (L/16) .. +15)
__device__ __forceinline__ int slice_off(int c16, int q) { return c16 * 64 + q; }
__device__ __forceinline__ int frag_src(int lane) { return (lane & 15) * 64 + 16 * (lane >> 4); }

template <int S>
__device__ __forceinline__ void put_digits(signed char *blk, long long sstride, int off,
                                           long long Fb) {
#pragma unroll
  for (int s = 0; s < S; s++) blk[s * sstride + off] = (signed char)digit_byte<S>(Fb, s);
}

constexpr int kSliceThreads = 256;

// ---------------------------------------------------------------------------
// Backward data operand.  Work-group = (nrad <= 4 consecutive radii, slot
// m + L; nrad = 4 unless the staging would pass 64 KiB of LDS); per
// chunk of pair indices [i0, i1) the slot's spectra at those radii for
// degrees l - M in [2 i0, 2 i1) (one contiguous run of at most 128 ncomp
// doubles per radius) are staged in LDS with coalesced loads.  Wave task =
// (parity, component): lane (radius r4 = lane / 16, pair indices
// 4 (lane % 16) .. +3 of l = M + 2i + parity) takes the chunk exponent over
// the 16 lanes of its radius and writes each digit's 4 bytes with one store
// (4 rows of a block per instruction) plus the exponent.  The first component
// of each vector (pol) is scaled by g3 = l(l+1) (1/2 at l = 0) as in the FP64
// kernels; radii nr <= k < nr_p get zero digits.
constexpr int kBsRadii = 4;  // lanes per wave: 4 radii x 16
// Staging offset of element e of radius rr's run: one double of padding per
// 32, so that the slicing lanes (8 ncomp doubles apart, 512 bytes for 8
// components) fall into different LDS banks instead of one.
__device__ __forceinline__ int bs_at(int rr, int e, int ncomp) {
  return rr * 132 * ncomp + e + (e >> 5);
}
template <int S>
__global__ void __launch_bounds__(kSliceThreads) k_bslice(
    int L, int nr, int nr_p, int nrad, int nv, int ncomp, const double *__restrict__ sp, SpecAddr sa,
    const int *__restrict__ jst, const long long *__restrict__ cb, signed char *__restrict__ dig,
    int *__restrict__ dexp) {
  extern __shared__ double run[];  // [nrad][128 degrees][ncomp], padded (bs_at)
  const int kq = blockIdx.x, sl = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int M = sl >= L ? sl - L : L - sl;
  const int n = L - M + 1, np0 = (n + 1) / 2, np1 = n / 2;
  const int nch0 = bchunk_count(M, np0), nch1 = bchunk_count(M, np1);
  const int ntile = nr_p >> 4;
  const long long sstride = (long long)ntile * 1024;
  const int k0 = kq * nrad;
  const double *src = sp + (long long)jst[sl] * sa.s_j;
  const int r4 = lane >> 4, q16 = lane & 15;
  const int k = k0 + r4;  // this lane's radius in the slicing phase
  for (int c = 0; c < nch0; c++) {
    const int i0 = bchunk_start(M, c), ni = bchunk_end(M, c) - i0;
    const int d0 = 2 * i0, nd_run = min(2 * ni, n - d0) * ncomp;
    const int nall = nrad * nd_run;
    __syncthreads();
    if (sa.s_nd == 1 && sa.s_j == ncomp && nd_run <= 4 * kSliceThreads && nrad <= kBsRadii) {
      // components contiguous within a degree's record and a run of at
      // most 4 doubles per thread (ncomp <= 8): the loads of every radius
      // are issued before any LDS store, so up to 16 per thread are in
      // flight together
      double val[kBsRadii][4];
#pragma unroll
      for (int rr = 0; rr < kBsRadii; rr++) {
        const int kr = min(k0 + (rr < nrad ? rr : 0), nr - 1);  // radii >= nr read nr - 1
        const double *rs = src + (long long)kr * sa.s_k + (long long)d0 * ncomp;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int e = tid + u * kSliceThreads;
          val[rr][u] = rs[e < nd_run ? e : 0];
        }
      }
#pragma unroll
      for (int rr = 0; rr < kBsRadii; rr++) {
        const bool live = k0 + rr < nr;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int e = tid + u * kSliceThreads;
          if (rr < nrad && e < nd_run) run[bs_at(rr, e, ncomp)] = live ? val[rr][u] : 0.0;
        }
      }
    } else if (sa.s_nd == 1 && sa.s_j == ncomp) {
      // components contiguous within a degree's record: a radius's run is
      // nd_run consecutive doubles, staged radius by radius without any
      // index division
      for (int rr = 0; rr < nrad; rr++) {
        const int kr = min(k0 + rr, nr - 1);  // radii >= nr read radius nr - 1
        const double *rs = src + (long long)kr * sa.s_k + (long long)d0 * ncomp;
        const bool live = k0 + rr < nr;
        for (int base = 0; base < nd_run; base += 8 * kSliceThreads) {
          double val[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int e = base + tid + u * kSliceThreads;
            val[u] = rs[e < nd_run ? e : 0];
          }
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int e = base + tid + u * kSliceThreads;
            if (e < nd_run) run[bs_at(rr, e, ncomp)] = live ? val[u] : 0.0;
          }
        }
      }
    } else
    for (int base = 0; base < nall; base += 8 * kSliceThreads) {
      // independent loads first, then the LDS stores; the loads are
      // unconditional (clamped to a valid element, masked at the store) so
      // the eight stay in flight together.  Radii >= nr read radius nr - 1.
      double val[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int idx = base + tid + u * kSliceThreads;
        const int ii = idx < nall ? idx : 0;
        const int rr = ii / nd_run, e = ii - rr * nd_run;
        const int dd = e / ncomp, cq = e - dd * ncomp;
        const int kr = min(k0 + rr, nr - 1);
        val[u] = src[(long long)kr * sa.s_k + (long long)(d0 + dd) * sa.s_j + (long long)cq * sa.s_nd];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int idx = base + tid + u * kSliceThreads;
        if (idx < nall) {
          const int rr = idx / nd_run, e = idx - rr * nd_run;
          run[bs_at(rr, e, ncomp)] = k0 + rr < nr ? val[u] : 0.0;
        }
      }
    }
    __syncthreads();
    const int ntask = (c < nch1 ? 2 : 1) * ncomp;
    for (int task = wave; task < ntask; task += kSliceThreads / 64) {
      const int par = task >= ncomp ? 1 : 0, cq = task - par * ncomp;
      const int npar = par ? np1 : np0;
      double x[4];
      int e = kEmpty;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int il = 4 * q16 + j, i = i0 + il;
        x[j] = 0.0;
        if (r4 < nrad && il < ni && i < npar) {
          x[j] = run[bs_at(r4, (2 * il + par) * ncomp + cq, ncomp)];
          if (cq < 3 * nv && cq % 3 == 0) x[j] *= g3s_of(M + 2 * i + par);
        }
        e = max(e, exp_of(x[j]));
      }
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) e = max(e, __shfl_xor(e, o));
      long long F[4];
#pragma unroll
      for (int j = 0; j < 4; j++) F[j] = fix_of<S>(x[j], e);
      const long long gb = (cb[2 * sl + par] + c) * ncomp + cq;
      signed char *blk = dig + (gb * S * ntile + (k >> 4)) * 1024 + slice_off(k & 15, 4 * q16);
      if (r4 < nrad) {
#pragma unroll
        for (int sd = 0; sd < S; sd++)
          *reinterpret_cast<unsigned *>(blk + sd * sstride) = pack4<S>(F, sd);
        if (q16 == 0) dexp[gb * nr_p + k] = e;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Forward data operand.  Work-group = (radius tile of 16, slot m + L).  Per
// chunk of 64 hemisphere rows and group of up to kFsGroup components the
// tile's north and mirrored south rows are staged in LDS (loads ordered along
// the smaller of the radius and colatitude strides) as the weighted symmetric
// and antisymmetric combinations (v(t) +- v(nth-1-t)) w(t), as in the FP64
// kernel (an odd grid's equator row is its own mirror: (v w, 0)); each wave
// then slices one (component, sym/asym) block of 64 rows x 16 radii: lane
// (radius c16 = lane % 16, rows 16 (lane / 16) .. +15) takes the column
// exponent over its 16 rows and the 4 lanes of its radius, and stores each
// digit's 16 bytes with one 16-byte store (one whole block per instruction).
constexpr int kFsGroup = 2;  // 2 components x sym/asym = one block per wave
template <int S>
__global__ void __launch_bounds__(kSliceThreads) k_fslice(
    int nr, int nr_p, int nth, int nth_h, int nchf, int ncomp, const double *__restrict__ in,
    GridAddr ga, const double *__restrict__ weight, signed char *__restrict__ dig,
    int *__restrict__ dexp) {
  __shared__ double sC[kFsGroup][2][64][17];  // [component][sym/asym][row][radius]
  // 1-D grid over (slot, radius tile), radius tile fastest.  Work-groups are
  // dealt round-robin to the 8 XCDs; neighbouring radius tiles share the
  // 128-byte lines their 16-radius rows straddle (Nr need not be a multiple
  // of 16), so consecutive logical tiles are kept on one XCD and its L2.
  const int ntile_g = nr_p >> 4;
  const unsigned nwg = gridDim.x, per = nwg / 8;
  const unsigned lb = blockIdx.x < per * 8 ? (blockIdx.x % 8) * per + blockIdx.x / 8 : blockIdx.x;
  const int kt = (int)(lb % ntile_g), sl = (int)(lb / ntile_g);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ntile = nr_p >> 4;
  const int t_eq = (nth & 1) ? nth / 2 : -1;
  const bool k_fast = ga.s_k <= ga.s_t;
  const long long base = ga.slot_off[sl] * ga.slot_mul;
  const long long sstride = (long long)ntile * 1024;
  // (chunk, component group) steps.  Loading the next step's rows while a
  // step is sliced was measured slower (5.67 against 5.39 ms at L=511, with
  // more HBM bytes fetched): the loads stay in the step
  const int ncg = (ncomp + kFsGroup - 1) / kFsGroup, nit = nchf * ncg;
  double vn[kFsGroup * 4], vs[kFsGroup * 4];
  auto load_rows = [&](int it) {
    const int c = it / ncg, cg = (it % ncg) * kFsGroup;
    const int ng = min(kFsGroup, ncomp - cg);
#pragma unroll
    for (int u = 0; u < kFsGroup * 4; u++) {
      const int f = u / 4, idx = tid + kSliceThreads * (u % 4);
      const int kk = k_fast ? (idx & 15) : (idx >> 6);
      const int tt = k_fast ? (idx >> 4) : (idx & 63);
      const int t = c * 64 + tt, k = kt * 16 + kk;
      // unconditional loads (clamped into range, zeroed at the LDS store) so
      // that all of them stay in flight together (radii past the last read
      // the last one: the line the tile reads anyway)
      const int fc = f < ng ? f : 0, kc = k < nr ? k : nr - 1, tc = t < nth_h ? t : 0;
      const long long b = base + (long long)(cg + fc) * ga.s_nd + (long long)kc * ga.s_k;
      vn[u] = in[b + (long long)tc * ga.s_t];
      vs[u] = in[b + (long long)(nth - 1 - tc) * ga.s_t];
    }
  };
  for (int it = 0; it < nit; it++) {
    {
      const int c = it / ncg, cg = (it % ncg) * kFsGroup;
      const int ng = min(kFsGroup, ncomp - cg);
      __syncthreads();  // the previous step's slicing done with sC
      load_rows(it);
#pragma unroll
      for (int u = 0; u < kFsGroup * 4; u++) {
        const int f = u / 4, idx = tid + kSliceThreads * (u % 4);
        const int kk = k_fast ? (idx & 15) : (idx >> 6);
        const int tt = k_fast ? (idx >> 4) : (idx & 63);
        const int t = c * 64 + tt, k = kt * 16 + kk;
        const bool ok = f < ng && k < nr && t < nth_h;
        const double w = weight[t < nth_h ? t : 0];
        const double xn = ok ? vn[u] : 0.0, xs = (ok && t != t_eq) ? vs[u] : 0.0;
        sC[f][0][tt][kk] = (xn + xs) * w;
        sC[f][1][tt][kk] = (xn - xs) * w;
      }
      __syncthreads();
      // wave task = (component, sym/asym) block
      for (int task = wave; task < ng * 2; task += kSliceThreads / 64) {
        const int h = task & 1, f = task >> 1;
        const int c16 = lane & 15, r0 = 16 * (lane >> 4);
        double x[16];
        int e = kEmpty;
#pragma unroll
        for (int i = 0; i < 16; i++) {
          x[i] = sC[f][h][r0 + i][c16];
          e = max(e, exp_of(x[i]));
        }
        e = max(e, __shfl_xor(e, 16));
        e = max(e, __shfl_xor(e, 32));
        long long F[16];
#pragma unroll
        for (int i = 0; i < 16; i++) F[i] = fix_of<S>(x[i], e);
        const long long g = (((long long)sl * nchf + c) * 2 + h) * ncomp + cg + f;
        signed char *blk = dig + (g * S * ntile + kt) * 1024 + slice_off(c16, r0);
#pragma unroll
        for (int sd = 0; sd < S; sd++) {
          uint4 w4;
          w4.x = pack4<S>(F + 0, sd);
          w4.y = pack4<S>(F + 4, sd);
          w4.z = pack4<S>(F + 8, sd);
          w4.w = pack4<S>(F + 12, sd);
          *reinterpret_cast<uint4 *>(blk + sd * sstride) = w4;
        }
        if (lane < 16) dexp[g * nr_p + kt * 16 + c16] = e;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// GEMM side.  Work-groups of 8 waves.  Per step the shared operands (data
// slices, the backward's P slices, and all exponents) are copied global ->
// LDS without passing through registers (global_load_lds: lane L's 16 or 4
// bytes land at block base + 16 L or 4 L) into one of two buffers while the
// other is reduced.  These copies and the A-slice register loads are issued
// as inline assembly, outside the compiler's s_waitcnt bookkeeping: the
// compiler cannot tell that a copy writes the other buffer and would wait for
// it before every LDS read.  The kernels count their vector-memory operations
// themselves (issue order = completion order for vmcnt): one wait for
// everything at the top of a step, and partial waits that leave only the
// next step's copies in flight before the A slices of this step are used.
constexpr int kGemmThreads = 512;

__device__ __forceinline__ unsigned lds_addr(const void *p) {
  return (unsigned)(size_t)(const __attribute__((address_space(3))) void *)p;
}
// wave-uniform base (SGPRs) + 32-bit lane offset, as the hardware addresses
// them: keeps the address arithmetic scalar and one VGPR per lane offset
__device__ __forceinline__ const void *uniform_ptr(const void *p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (const void *)(((unsigned long long)hi << 32) | lo);
}
// dst: LDS byte address of the destination (lane i lands at dst + 16 i or
// dst + 4 i), computed from the work-group's LDS base as an integer.
// The compiler's hazard recognizer does not look inside inline assembly, so
// the statements provide the wait states themselves: an LDS DMA needs one
// after an SALU write of M0 (which the compiler places right before the
// statement); a vector-memory instruction needs five after a VALU write
// (v_readfirstlane) of an SGPR it reads, which the exponent copies' runtime-
// selected bases get.  The build checks the generated code for both.
__device__ __forceinline__ void copy_block(const void *base, unsigned voff, unsigned dst) {
  asm volatile("s_nop 0\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(voff), "s"(uniform_ptr(base)),
               "{m0}"(__builtin_amdgcn_readfirstlane(dst))
               : "memory");
}
__device__ __forceinline__ void copy_dword(const void *base, unsigned voff, unsigned dst) {
  asm volatile("s_nop 4\n\tglobal_load_lds_dword %0, %1" ::"v"(voff), "s"(uniform_ptr(base)),
               "{m0}"(__builtin_amdgcn_readfirstlane(dst))
               : "memory");
}
__device__ __forceinline__ i4 load_slice(const void *base, unsigned voff) {
  i4 v;
  asm volatile("global_load_dwordx4 %0, %1, %2" : "=v"(v) : "v"(voff), "s"(uniform_ptr(base)) : "memory");
  return v;
}
// load_slice with the five wait states of a base SGPR just written by a VALU
// instruction (the 32x32 kernel's bases can come back from SGPR spill lanes
// through v_readlane right before the load)
__device__ __forceinline__ i4 load_slice_ws(const void *base, unsigned voff) {
  i4 v;
  asm volatile("s_nop 4\n\tglobal_load_dwordx4 %0, %1, %2"
               : "=v"(v)
               : "v"(voff), "s"(uniform_ptr(base))
               : "memory");
  return v;
}
template <int N>
__device__ __forceinline__ void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}
__device__ __forceinline__ void wait_copies() { wait_vm<0>(); }
// registers loaded by load_slice are used only after a wait_vm that covers
// them; this ties each use to the wait
template <int S>
__device__ __forceinline__ void landed(i4 (&a)[S]) {
#pragma unroll
  for (int s = 0; s < S; s++) asm volatile("" : "+v"(a[s]));
}

struct SplitDims {
  int L, nr, nr_p, nth, nth_h, nth_hp, ncomp, n_rt, ncg, nblk, nchf;
  int M0, nM;  // orders M0 .. M0+nM-1 of this launch (one tile of the P slices)
};

// Relative scale of group pair (g0, g0+1) (or of the single last group g0)
// against the first pair: the pair value 256 acc_g0 + acc_(g0+1) carries
// 2^(eA + eB + 4 - 8(g0+3)), the first pair 2^(eA + eB - 20).
__device__ __forceinline__ double pair_scale(int g0, bool two) {
  const int e = two ? -8 * g0 : 8 - 8 * g0;
  return __longlong_as_double((long long)(1023 + e) << 52);
}

// Triangular sliced products of up to two quantities QA, QB (NQB of them)
// with A slices AOF(q, s) and data slices BOF(q, t); T[j][r] collects
// sum over pairs of pair_value * pair_scale, exactly as INT32 -> FP64.
#define KSHT_SLICED_PRODUCTS(QA, QB, NQB, T)                                             \
  {                                                                                    \
    _Pragma("unroll") for (int j_ = 0; j_ < 2; j_++)                                   \
      _Pragma("unroll") for (int r_ = 0; r_ < 4; r_++) T[j_][r_] = 0.0;                \
    /* all data slices of the quantities first (one LDS latency, not one */            \
    /* per product group; occupancy is bounded by LDS, not registers), */              \
    /* unless the kernel has no registers to spare (BATCH_B) */                        \
    i4 bv_[2][S];                                                                      \
    if constexpr (BATCH_B) {                                                           \
      _Pragma("unroll") for (int j_ = 0; j_ < NQB; j_++)                               \
        _Pragma("unroll") for (int t = 0; t < S; t++) bv_[j_][t] = BOF((j_ ? (QB) : (QA)), t); \
      __builtin_amdgcn_sched_group_barrier(0x100, NQB * S, 0);                         \
    }                                                                                  \
    /* groups from the smallest scale up (the conversions of one group */             \
    /* overlap the products of the next) */                                            \
    _Pragma("unroll") for (int gg = 0; gg < (S + 1) / 2; gg++) {                       \
      const int g0 = 2 * ((S + 1) / 2 - 1 - gg);                                       \
      const bool two = g0 + 1 < S;                                                     \
      i4 a0[2], a1[2];                                                                 \
      _Pragma("unroll") for (int j_ = 0; j_ < 2; j_++) {                               \
        a0[j_] = i4{0, 0, 0, 0};                                                       \
        a1[j_] = i4{0, 0, 0, 0};                                                       \
      }                                                                                \
      _Pragma("unroll") for (int t = 0; t < S; t++) {                                  \
        if (t > g0 + 1) continue;                                                      \
        const int s0 = g0 - t, s1 = g0 + 1 - t;                                        \
        _Pragma("unroll") for (int j_ = 0; j_ < NQB; j_++) {                           \
          const int q = j_ ? (QB) : (QA);                                              \
          const i4 b = BATCH_B ? bv_[j_][t] : BOF(q, t);                               \
          if (s0 >= 0) a0[j_] = mfma8(AOF(q, s0), b, a0[j_]);                          \
          if (two && s1 >= 0) a1[j_] = mfma8(AOF(q, s1), b, a1[j_]);                   \
        }                                                                              \
      }                                                                                \
      const double sc = pair_scale(g0, two);                                           \
      _Pragma("unroll") for (int j_ = 0; j_ < NQB; j_++)                               \
        _Pragma("unroll") for (int r_ = 0; r_ < 4; r_++) {                             \
          const int v = two ? a0[j_][r_] * 256 + a1[j_][r_] : a0[j_][r_];              \
          T[j_][r_] = fma((double)v, sc, T[j_][r_]);                                   \
        }                                                                              \
    }                                                                                  \
    /* issue pattern: each product followed by two conversion steps */                 \
    _Pragma("unroll") for (int i_ = 0; i_ < NQB * S * (S + 1) / 2; i_++) {             \
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                               \
      __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);                               \
    }                                                                                  \
  }

// ---------------------------------------------------------------------------
// Backward.  Work-group = (M, radius tile, component group, theta block of
// 64 rows); wave w owns theta tile w%4 for output slot w/4 (slot 0: m = +M,
// slot 1: m = -M), over the steps (chunk c, parity) of both parities.  LDS
// per step: the data slices of both slots, the P slices of the four theta
// tiles (fragment order) and the exponents; dP slices go straight to
// registers, loaded at the start of the step and first used after the P
// quantities.  The FP64 accumulators hold the north / south rows directly:
//   even l-M: P terms add to both, dP terms (antisymmetric) to N and -S;
//   odd l-M:  P terms N and -S, dP terms both,
// and per output slot: vr = a^2 [P pol g3], vtheta = a ([dP dpol] - asin m' [P tor']),
// vphi = a (-[dP tor] - asin m' [P dpol']), with primes on the conjugate slot.
// VEC with NB = 4: the vector and one scalar (component nd_scl + group)
// in one work-group: the scalar's product [P s] shares the P slices already
// in LDS, so a separate scalar launch, which would load them again for a
// fifth of the products, is not needed.
template <int S, int NB>
struct BwdLds {
  static constexpr int NBB = 2 * NB * S;  // data blocks [slot][component][digit]
  static constexpr int NPB = 4 * S;       // P blocks [theta tile][digit]
  // every wave issues the same number of copies (spare ones into dummy
  // space), so the compiler can count them in s_waitcnt vmcnt
  static constexpr int KB = (NBB + NPB + 7) / 8;
  static constexpr int EX = 8 * KB * 64;  // i4 offset of the exponents
  static constexpr int I4 = EX + 64 + 64;  // 256 exponents + 256 dummy
  static constexpr int bytes() { return 2 * I4 * 16; }
};

template <int S, int NB, bool VEC>
__global__ void __launch_bounds__(kGemmThreads) k_leg_bwd_split(
    SplitDims d, int nd_first, SplitTabs tb, const long long *__restrict__ cb,
    const signed char *__restrict__ dig, const int *__restrict__ dexp, double *__restrict__ out,
    GridAddr ga, const double *__restrict__ a_r, const double *__restrict__ asin_t, int nd_scl) {
  extern __shared__ i4 smem[];
  using LY = BwdLds<S, NB>;
  constexpr bool BATCH_B = true;  // data slices of a quantity read together
  constexpr bool SCL = VEC && NB == 4;  // vector plus one scalar
  constexpr int NQ = VEC ? (SCL ? 6 : 5) : NB;
  constexpr int NO = NB;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform per wave
  // logical id = ((M n_rt + rt) ncg + cgl) nblk + tbk: a run is a whole
  // order, so the theta blocks that read the same data slices (and the
  // component groups that write the same output lines) are resident together
  long long bidx = xcd_group_order(blockIdx.x, gridDim.x, (long long)d.n_rt * d.ncg * d.nblk);
  const int tbk = (int)(bidx % d.nblk);
  bidx /= d.nblk;
  const int cgl = (int)(bidx % d.ncg);
  bidx /= d.ncg;
  const int rt = (int)(bidx % d.n_rt);
  const int M = d.M0 + (int)(bidx / d.n_rt);
  const int L = d.L;
  const int n = L - M + 1, np0 = (n + 1) / 2, np1 = n / 2;
  const int nch0 = bchunk_count(M, np0), nch1 = bchunk_count(M, np1);
  const int nsig = M == 0 ? 1 : 2;
  const int tt = wave & 3, sg = wave >> 2;
  const int t0 = (tbk * 4 + tt) * 16;
  const bool active = t0 < d.nth_h && sg < nsig;
  const int nd0 = nd_first + (VEC ? 3 : NB) * cgl;
  // component of block f: the vector's three, then the group's scalar
  auto comp = [&](int f) { return (SCL && f == 3) ? nd_scl + cgl : nd0 + f; };
  const int ntile = d.nr_p >> 4;
  const int sgc = nsig == 1 ? 0 : 1 - sg;  // conjugate slot
  const int sl_own = sg == 0 ? L + M : L - M;
  const int kcol = rt * 16 + (lane & 15);

  // chunk bases of the data slices of both slots and parities (the
  // block and row indices below fit in 32 bits)
  const int cb00 = (int)cb[2 * (L + M)], cb01 = (int)cb[2 * (L + M) + 1];
  const int cb10 = (int)cb[2 * (L - M)], cb11 = (int)cb[2 * (L - M) + 1];
  // chunk base of slot sgb (0/1) and parity par, as arithmetic (nested
  // selects of four values compile to a table in scratch)
  auto cbase = [&](int sgb, int par) {
    const int c0 = par ? cb01 : cb00, c1 = par ? cb11 : cb10;
    return c0 + sgb * (c1 - c0);
  };
  // P-slice 64-byte row and exponent index of (M, parity), chunk 0 (scalars:
  // a runtime-indexed array would live in scratch)
  const int prow0 = (int)(tb.off[2 * M] >> 6), prow1 = (int)(tb.off[2 * M + 1] >> 6);
  const int erow0 = (int)tb.eoff[2 * M], erow1 = (int)tb.eoff[2 * M + 1];
  constexpr int KCOPY = LY::KB + 1;  // copies per wave per step
  const unsigned fsrc = (unsigned)frag_src(lane);
  const unsigned lbase = lds_addr(smem);
  auto issue = [&](int c, int par, int b) {
    const unsigned bufa = lbase + 16u * (unsigned)(b * LY::I4);
    const int pr0 = (par ? prow1 : prow0) + c * S * d.nth_hp + tbk * 64;
#pragma unroll
    for (int j = 0; j < LY::KB; j++) {
      const int blk = wave + 8 * j;
      const signed char *base;
      if (blk < LY::NBB) {
        const int sgb = blk / (NB * S), f = (blk / S) % NB, s = blk % S;
        const int gb = (cbase(sgb, par) + c) * d.ncomp + comp(f);
        base = dig + (size_t)((gb * S + s) * ntile + rt) * 1024;
      } else if (blk < LY::NBB + LY::NPB) {
        const int t2 = (blk - LY::NBB) / S, s = (blk - LY::NBB) % S;
        base = tb.P + (size_t)(pr0 + s * d.nth_hp + t2 * 16) * 64;
      } else {
        base = tb.P + (size_t)pr0 * 64;  // spare copy
      }
      copy_block(base, fsrc, bufa + 1024u * blk);
    }
    const unsigned exa = bufa + 16u * LY::EX;  // exponents, 4 bytes each
    const int eo = (par ? erow1 : erow0) + c * d.nth_hp + tbk * 64;
    const int *ebase = tb.eP + eo;
    unsigned eoff = 4u * lane;
    unsigned edst = exa + 4u * (256 + (wave & 3) * 64);  // spare copies
    if (wave == 0) {
      edst = exa;
    } else if (wave == 1) {
      if (VEC) ebase = tb.edP + eo;
      edst = exa + 4u * 64;
    } else if (wave < 4) {
      // four (slot, component) groups of 16 radii per instruction
      const int sgf = (wave - 2) * 4 + (lane >> 4);
      ebase = dexp;
      eoff = 0;
      if (sgf < 2 * NB) {
        const int sgb = sgf / NB, f = sgf % NB;
        eoff = 4u * (unsigned)(((cbase(sgb, par) + c) * d.ncomp + comp(f)) * d.nr_p + rt * 16 +
                               (lane & 15));
      }
      edst = exa + 4u * (128 + (wave - 2) * 64);
    }
    copy_dword(ebase, eoff, edst);
  };

  double C[NO][2][4];
#pragma unroll
  for (int o = 0; o < NO; o++)
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int r = 0; r < 4; r++) C[o][h][r] = 0.0;
  double fr[4];  // -asin(theta) m of the conjugate slot
  {
    const double mc = sg == 0 ? -(double)M : (double)M;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int t = t0 + 4 * (lane >> 4) + r;
      fr[r] = (VEC && t < d.nth_h) ? -asin_t[t] * mc : 0.0;
    }
  }

  issue(0, 0, 0);
  int st = 0;
  for (int c = 0; c < nch0; c++) {
#pragma unroll
    for (int par = 0; par < 2; par++) {
      if (par == 1 && c >= nch1) continue;
      wait_copies();
      __syncthreads();
      const i4 *buf = smem + (st & 1) * LY::I4;
      const int *ex = reinterpret_cast<const int *>(buf + LY::EX);
      // dP slices of this step first (register loads), then the next step's
      // copies, so waiting for dP does not wait for the copies
      // Every wave issues the same vector-memory operations on every path
      // (rows t0 < nth_hp exist for idle waves too; the last step copies
      // itself again into the idle buffer), so the counted waits below hold
      // on all paths.
      i4 aP[S], adP[S];
      if (VEC) {
        const int r0 = (par ? prow1 : prow0) + c * S * d.nth_hp + t0;
#pragma unroll
        for (int s = 0; s < S; s++) adP[s] = load_slice(tb.dP + (size_t)(r0 + s * d.nth_hp) * 64, fsrc);
      }
      // next step: (c, 1), else (c + 1, 0), else this one again
      const bool nx1 = par == 0 && c < nch1, nx0 = !nx1 && c + 1 < nch0;
      issue(nx0 ? c + 1 : c, nx1 ? 1 : nx0 ? 0 : par, (st + 1) & 1);
      {
        if (active) {
#pragma unroll
          for (int s = 0; s < S; s++) aP[s] = buf[(LY::NBB + tt * S + s) * 64 + lane];
        }
        const int rb = tt * 16 + 4 * (lane >> 4);
        const int lk = lane & 15;
        auto ecol = [&](int sgb, int f) { return ex[128 + (sgb * NB + f) * 16 + lk]; };
        const i4 *bo = buf + sg * NB * S * 64 + lane;
        const i4 *bc = buf + sgc * NB * S * 64 + lane;
        // quantity q: A table, data block
#define AOF(q, s) ((VEC && ((q) == 3 || (q) == 4)) ? adP[s] : aP[s])
#define BOF(q, t)                                                            \
  (VEC ? ((q) == 0   ? bo[(0 * S + (t)) * 64]                                \
          : (q) == 1 ? bc[(2 * S + (t)) * 64]                                \
          : (q) == 2 ? bc[(1 * S + (t)) * 64]                                \
          : (q) == 3 ? bo[(1 * S + (t)) * 64]                                \
          : (q) == 4 ? bo[(2 * S + (t)) * 64]                                \
                     : bo[(3 * S + (t)) * 64])                               \
       : bo[((q) * S + (t)) * 64])
        // fold of quantity q with the collected T: x = ldexp(T, eA + eB - 20)
        auto fold = [&](int q, const double (&T)[4]) {
          const int eb = VEC ? (q == 0 ? ecol(sg, 0) : q == 1 ? ecol(sgc, 2) : q == 2 ? ecol(sgc, 1)
                                : q == 3 ? ecol(sg, 1) : q == 4 ? ecol(sg, 2) : ecol(sg, 3))
                             : ecol(sg, q);
          const int o = VEC ? (q == 0 ? 0 : (q == 1 || q == 3) ? 1 : q == 5 ? 3 : 2) : q;
          const bool dq = VEC && (q == 3 || q == 4);  // a dP quantity
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int ea = ex[(dq ? 64 : 0) + rb + r];
            const double x = ldexp(T[r], ea + eb - 20);
            // P terms even in theta about the equator for even l-M, dP terms odd
            const bool sym = dq ? (par == 1) : (par == 0);
            if (!VEC || q == 0 || q == 3 || q == 5) {
              C[o][0][r] += x;
              C[o][1][r] += sym ? x : -x;
            } else if (q == 4) {
              C[o][0][r] -= x;
              C[o][1][r] -= sym ? x : -x;
            } else {
              C[o][0][r] = fma(x, fr[r], C[o][0][r]);
              C[o][1][r] = fma(x, sym ? fr[r] : -fr[r], C[o][1][r]);
            }
          }
        };
        double T[2][4];
        // P quantities first (dP slices arrive meanwhile), then dP quantities
        if (active) {
          __builtin_amdgcn_sched_barrier(0);
          KSHT_SLICED_PRODUCTS(0, 1 % NQ, (NQ > 1 ? 2 : 1), T)
          fold(0, T[0]);
          if (NQ > 1) fold(1 % NQ, T[1]);
          if constexpr (SCL) {
            // the third P quantity together with the scalar's
            __builtin_amdgcn_sched_barrier(0);
            KSHT_SLICED_PRODUCTS(2, 5 % NQ, 2, T)
            fold(2, T[0]);
            fold(5 % NQ, T[1]);
          } else if (NQ > 2) {
            __builtin_amdgcn_sched_barrier(0);
            KSHT_SLICED_PRODUCTS(2 % NQ, 2 % NQ, 1, T)
            fold(2 % NQ, T[0]);
          }
        }
        if (VEC) {
          // on every path (idle waves too): the registers stay allocated
          // to the loads until they have landed
          wait_vm<KCOPY>();  // dP slices landed; the copies stay in flight
          landed(adP);
          if (active) {
            __builtin_amdgcn_sched_barrier(0);
            KSHT_SLICED_PRODUCTS(3 % NQ, 4 % NQ, 2, T)
            fold(3 % NQ, T[0]);
            fold(4 % NQ, T[1]);
          }
        }
#undef AOF
#undef BOF
      }
      st++;
    }
  }
  wait_vm<0>();
  if (!active) return;
  const int k = kcol;
  if (k >= d.nr) return;
  const double a1 = a_r[k], a2 = a1 * a1;
  const bool odd_nth = (d.nth & 1) != 0;
  const long long base = ga.slot_off[sl_own] * ga.slot_mul + (long long)k * ga.s_k;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int t = t0 + 4 * (lane >> 4) + r;
    if (t >= d.nth_h) continue;
    const bool eq = odd_nth && t == d.nth / 2;
    const long long on = base + (long long)t * ga.s_t;
    const long long os = base + (long long)(d.nth - 1 - t) * ga.s_t;
#pragma unroll
    for (int o = 0; o < NO; o++) {
      const double scl = VEC ? (o == 0 ? a2 : o < 3 ? a1 : 1.0) : 1.0;
      double *dst = out + (long long)comp(o) * ga.s_nd;
      if (eq) {
        dst[on] = scl * (0.5 * (C[o][0][r] + C[o][1][r]));
      } else {
        dst[on] = scl * C[o][0][r];
        dst[os] = scl * C[o][1][r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Forward.  Work-group = (M, radius tile, component group, block of 64 pair
// indices); wave w owns rows i in [64 ib + 16 (w/2), +16) of parity w%2 for
// both slots, over the chunks of 64 hemisphere rows.  Rows of even l-M take
// the symmetric combinations against P and P/sin, the antisymmetric ones
// against dP; odd l-M the other way round.  Per slot (m = +-M, primes on the
// conjugate slot):
//   pol = r^2 g7 [P vr], dpol = r g7 ([dP vt] - m [P/sin vp']),
//   tor = r g7 (-[dP vp] - m [P/sin vt']); scalars g6 [P s].
// LDS per step: the data slices of both slots and the exponents; the wave's
// own P, dP, P/sin slices go to registers, P and dP prefetched one step
// ahead and P/sin loaded at the top of its step.
template <int S, int NB>
struct FwdLds {
  static constexpr int NBB = 4 * NB * S;  // [slot][sym/asym][component][digit]
  // column exponents (NB tasks of 4 groups) and row exponents (2 or 6
  // tasks); every wave issues KB block and KE exponent copies (spares into
  // dummy space) so that s_waitcnt vmcnt can count them
  static constexpr int NET = NB + 6;
  static constexpr int KB = (NBB + 7) / 8, KE = (NET + 7) / 8;
  static constexpr int EX = 8 * KB * 64;       // i4 offset of the exponents
  static constexpr int I4 = EX + 8 * KE * 16;  // 64 ints per exponent task
  static constexpr int NBUF = (2 * (I4 * 16) <= 160 * 1024) ? 2 : 1;
  static constexpr int bytes() { return NBUF * I4 * 16; }
};

template <int S, int NB, bool VEC>
__global__ void __launch_bounds__(kGemmThreads) k_leg_fwd_split(
    SplitDims d, int nd_first, SplitTabs tb, const signed char *__restrict__ dig,
    const int *__restrict__ dexp, double *__restrict__ sp, SpecAddr sa,
    const int *__restrict__ jst, const int *__restrict__ ipad, const double *__restrict__ radius) {
  extern __shared__ i4 smem[];
  using LY = FwdLds<S, NB>;
  constexpr int NBUF = LY::NBUF;
  // data slices of a quantity read together (not for 7-slice vectors: no
  // registers to spare)
  constexpr bool BATCH_B = !(VEC && S >= 7);
  constexpr int NQ = VEC ? 5 : NB;
  constexpr int NO = VEC ? 3 : NB;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform per wave
  // logical id = ((ib (L+1) + M) n_rt + rt) ncg + cgl: a run is one (M, row
  // block); the run index ib (L+1) + M puts consecutive orders, not row
  // blocks, on consecutive XCDs (row blocks past the degrees of M are empty,
  // and the work of row block ib would otherwise all land on two XCDs)
  long long bidx = xcd_group_order(blockIdx.x, gridDim.x, (long long)d.n_rt * d.ncg);
  const int cgl = (int)(bidx % d.ncg);
  bidx /= d.ncg;
  const int rt = (int)(bidx % d.n_rt);
  bidx /= d.n_rt;
  const int M = d.M0 + (int)(bidx % d.nM);
  const int ib = (int)(bidx / d.nM);
  const int L = d.L;
  const int ip0 = ipad[2 * M], ip1 = ipad[2 * M + 1];
  if (ib * 64 >= ip0) return;  // whole work-group beyond the degrees of M
  const int nsig = M == 0 ? 1 : 2;
  // wave w: parity w & 1, row tile w >> 1.  Waves go to the four SIMDs
  // round-robin, so the idle tiles of a last, partly filled row block fall
  // on different SIMDs instead of leaving two SIMDs without work
  const int par = wave & 1, tile = wave >> 1;
  const int i0 = ib * 64 + tile * 16;
  const int ip = par ? ip1 : ip0;
  const bool active = i0 < ip;
  const int nd0 = nd_first + NB * cgl;
  const int ntile = d.nr_p >> 4;
  const int kcol = rt * 16 + (lane & 15);
  const int nchf = d.nchf;

  // 32-bit block / row indices (they fit): slice group of (slot, chunk 0, sym/asym)
  // is ((sl * nchf + c) * 2 + h) * ncomp + component
  const int gsl[2] = {(L + M) * nchf * 2 * d.ncomp, (L - M) * nchf * 2 * d.ncomp};
  const int erow0 = (int)tb.eoff[2 * M], erow1 = (int)tb.eoff[2 * M + 1];
  const unsigned fsrc = (unsigned)frag_src(lane);
  const unsigned lbase = lds_addr(smem);
  auto issue = [&](int c, int b) {
    const unsigned bufa = lbase + 16u * (unsigned)(b * LY::I4);
#pragma unroll
    for (int j = 0; j < LY::KB; j++) {
      const int b = wave + 8 * j;
      const int bb = b < LY::NBB ? b : 0;  // spare copies repeat block 0
      const int sgb = bb / (2 * NB * S), h = (bb / (NB * S)) & 1, f = bb / S % NB, s = bb % S;
      const int gb = gsl[sgb] + (c * 2 + h) * d.ncomp + nd0 + f;
      copy_block(dig + (size_t)((gb * S + s) * ntile + rt) * 1024, fsrc, bufa + 1024u * b);
    }
    const unsigned exa = bufa + 16u * LY::EX;  // exponents, 4 bytes each
    // column exponents [slot][sym/asym][component][16 radii] (four groups per
    // task) and row exponents [P, dP, P/sin][parity][64 rows]
#pragma unroll
    for (int j = 0; j < LY::KE; j++) {
      const int task = wave + 8 * j;
      const int *base = tb.eP;  // spare copies
      unsigned off = 0;
      if (task < NB) {
        const int g = task * 4 + (lane >> 4);
        const int sgb = g / (2 * NB), h = (g / NB) & 1, f = g % NB;
        const int gb = gsl[sgb] + (c * 2 + h) * d.ncomp + nd0 + f;
        base = dexp;
        off = 4u * (unsigned)(gb * d.nr_p + rt * 16 + (lane & 15));
      } else if (task < NB + (VEC ? 6 : 2)) {
        const int tw = task - NB, tab = tw >> 1, pr = tw & 1;
        const int ipp = pr ? ip1 : ip0;
        const int row = ib * 64 + lane;
        base = tab == 0 ? tb.eP : tab == 1 ? tb.edP : tb.ePa;
        off = row < ipp ? 4u * (unsigned)((pr ? erow1 : erow0) + c * ipp + row) : 0u;
      }
      copy_dword(base, off, exa + 256u * task);
    }
  };

  double C[2][NO][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int o = 0; o < NO; o++)
#pragma unroll
      for (int r = 0; r < 4; r++) C[a][o][r] = 0.0;

  const int hx = par, hy = 1 - par;  // P: symmetric for even l-M; dP the other one
  // 64-byte rows; idle waves load row 0 of the table, so that every wave
  // issues the same vector-memory operations on every path and the counted
  // waits hold on all paths
  const int arow = active ? (int)(tb.off[2 * M + par] >> 6) + i0 : 0;
  const int ipa = active ? ip : 0;
  auto load_a = [&](const signed char *tab, int c, i4 (&a)[S]) {
#pragma unroll
    for (int s = 0; s < S; s++) a[s] = load_slice(tab + (size_t)(arow + (c * S + s) * ipa) * 64, fsrc);
  };
  constexpr int KCOPY = LY::KB + LY::KE;  // copies per wave per step
  i4 aP[S], adP[S], aPa[S];
  // P and dP slices are prefetched one step ahead (issued after their last
  // use), P/sin at the end of the step before (first used after three
  // quantities, so it stays in flight across the top-of-step wait)
  load_a(tb.P, 0, aP);
  if (VEC) load_a(tb.dP, 0, adP);
  issue(0, 0);
  if (VEC) load_a(tb.Pa, 0, aPa);
  for (int c = 0; c < nchf; c++) {
    // everything but the P/sin slices of this step (issued last)
    wait_vm<VEC ? S : 0>();
    __syncthreads();
    const i4 *buf = smem + (c % NBUF) * LY::I4;
    const int *ex = reinterpret_cast<const int *>(buf + LY::EX);
    // The last step loads and copies itself again (into registers and the
    // idle buffer).
    const int cn = c + 1 < nchf ? c + 1 : c;
    if (NBUF == 2) issue(cn, (c + 1) % NBUF);
    {
      const int sgc0 = nsig == 1 ? 0 : 1;  // conjugate of slot 0 (slot 1's is 0)
      const int rb = 64 * NB + par * 64 + tile * 16 + 4 * (lane >> 4);
      const int lk = lane & 15;
      // quantity q of slot sg: data block (slot, sym/asym, component)
#define QSLOT(q, sg) ((VEC && ((q) == 2 || (q) == 4)) ? ((sg) == 0 ? sgc0 : 0) : (sg))
#define QH(q) ((!VEC || (q) == 0 || (q) == 2 || (q) == 4) ? hx : hy)
#define QF(q) (!VEC ? (q) : (q) == 0 ? 0 : ((q) == 1 || (q) == 4) ? 1 : 2)
#define BOFS(q, sg, t) buf[(((QSLOT(q, sg) * 2 + QH(q)) * NB + QF(q)) * S + (t)) * 64 + lane]
      // one quantity for both slots: the two slots play the role of the two
      // quantities of KSHT_SLICED_PRODUCTS; M = 0 has one slot (slot 1 then
      // repeats slot 0 and is not stored)
#define KSHT_FWD_QUANTITY(Q, TAB)                                                           \
  {                                                                                         \
    constexpr int qq = (Q);                                                                 \
    __builtin_amdgcn_sched_barrier(0);                                                      \
    double T[2][4];                                                                         \
    KSHT_SLICED_PRODUCTS(0, 1, 2, T)                                                        \
    constexpr int o = VEC ? (qq == 0 ? 0 : (qq <= 2) ? 1 : 2) : qq;                         \
    _Pragma("unroll") for (int sg = 0; sg < 2; sg++) {                                      \
      const double ms = sg == 0 ? (double)M : -(double)M;                                   \
      const double cf = (!VEC || qq <= 1) ? 1.0 : (qq == 3 ? -1.0 : -ms);                   \
      const int eb = ex[((QSLOT(qq, sg) * 2 + QH(qq)) * NB + QF(qq)) * 16 + lk];            \
      _Pragma("unroll") for (int r = 0; r < 4; r++) {                                       \
        const int ea = ex[rb + (TAB) * 128 + r];                                            \
        C[sg][o][r] = fma(ldexp(T[sg][r], ea + eb - 20), cf, C[sg][o][r]);                  \
      }                                                                                     \
    }                                                                                       \
  }
#define AOF(q, s) AA_[s]
#define BOF(q, t) BOFS(qq, (q), t)
      landed(aP);  // loaded during the previous step, covered by the wait above
#define AA_ aP
      if (active) {
        KSHT_FWD_QUANTITY(0, 0)
        if (!VEC) {
          if (NQ > 1) KSHT_FWD_QUANTITY(1 % NQ, 0)
          if (NQ > 2) KSHT_FWD_QUANTITY(2 % NQ, 0)
        }
      }
#undef AA_
      load_a(tb.P, cn, aP);  // prefetch for the next step
      if (VEC) {
        landed(adP);  // loaded during the previous step
#define AA_ adP
        if (active) {
          KSHT_FWD_QUANTITY(1 % NQ, 1)
          KSHT_FWD_QUANTITY(3 % NQ, 1)
        }
#undef AA_
        load_a(tb.dP, cn, adP);  // prefetch for the next step
        // P/sin landed; the next step's copies, P and dP stay in flight
        wait_vm<(NBUF == 2 ? KCOPY : 0) + 2 * S>();
        landed(aPa);
#define AA_ aPa
        if (active) {
          KSHT_FWD_QUANTITY(2 % NQ, 2)
          KSHT_FWD_QUANTITY(4 % NQ, 2)
        }
#undef AA_
      }
#undef KSHT_FWD_QUANTITY
#undef AOF
#undef BOF
#undef BOFS
#undef QSLOT
#undef QH
#undef QF
    }
    if (NBUF == 1 && c + 1 < nchf) {
      __syncthreads();
      issue(c + 1, 0);
    }
    if (VEC) load_a(tb.Pa, cn, aPa);  // the next step's, in flight across its barrier
  }
  // the last step's repeated loads: keep their registers until they land
  wait_vm<0>();
  landed(aP);
  if (VEC) {
    landed(adP);
    landed(aPa);
  }
  if (!active) return;
  const int k = kcol;
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
      const long long o = (long long)(jst[sg == 0 ? L + M : L - M] + l - M) * sa.s_j +
                          (long long)k * sa.s_k + (long long)nd0 * sa.s_nd;
      if (VEC) {
        sp[o] = r2 * g7 * C[sg][0][r];
        sp[o + sa.s_nd] = r1 * g7 * C[sg][1 % NO][r];
        sp[o + 2 * sa.s_nd] = r1 * g7 * C[sg][2 % NO][r];
      } else {
#pragma unroll
        for (int q = 0; q < NO; q++) sp[o + (long long)q * sa.s_nd] = g6 * C[sg][q][r];
      }
    }
  }
}
#undef KSHT_SLICED_PRODUCTS

// ---------------------------------------------------------------------------
// Forward on v_mfma_i32_32x32x32_i8.  On MI355X this form sustains about 1.7x
// the INT8 rate of the 16x16x64 form (measured with independent accumulator
// chains: 4.0 against 2.3 POP/s at the occupancies these kernels run at), so
// the same products take a 32 x 32 output tile per instruction: 32 degree
// rows of one parity by the 16 radii of BOTH slots +-M (the two slots share
// the P slices, so they sit side by side in N).  A work-group of 4 waves
// (one per SIMD; the tile's accumulators need the whole register file)
// covers the same 64-row block of both parities as the 16x16 kernel and
// keeps its LDS layout, copies and exponents: only the lanes' fragment reads
// change.
//   A (32 rows x 32 K): lane l holds row l%32, K 16(l/32) .. +15 of a K half;
//   B (32 K x 32 cols): lane l holds column l%32 (slot (l/16)%2, radius l%16),
//     the same K bytes; read from the slot's 1 KiB LDS block, whose 16 bytes
//     of (radius r, K 16 g .. +15) sit at 16 (r + 16 g);
//   D: lane l, element e: column l%32, row 8 (e/4) + 4 (l/32) + e%4.
// A 64-wide exponent chunk is two K halves into the same INT32 accumulators
// (the bounds of the 16x16 form hold: the sum still runs over 64 products).
template <int S, int NB>
struct FwdLds32 {
  static constexpr int W = 4;             // waves
  static constexpr int NBB = 4 * NB * S;  // [slot][sym/asym][component][digit]
  static constexpr int NET = NB + 6;
  static constexpr int KB = (NBB + W - 1) / W, KE = (NET + W - 1) / W;
  static constexpr int EX = W * KB * 64;       // i4 offset of the exponents
  static constexpr int I4 = EX + W * KE * 16;  // 64 ints per exponent task
  static constexpr int NBUF = (2 * (I4 * 16) <= 160 * 1024) ? 2 : 1;
  static constexpr int bytes() { return NBUF * I4 * 16; }
};
constexpr int kGemm32Threads = 256;

typedef int i16 __attribute__((ext_vector_type(16)));
__device__ __forceinline__ i16 mfma32(i4 a, i4 b, i16 c) {
  return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// Triangular sliced products of quantity Q: A slices AOF(s, kh), data slices
// BOF(t, kh).  Every digit group g = s + t < S has its own INT32 accumulator,
// so each data slice feeds up to S independent MFMA chains back to back; the
// groups are then folded pairwise into FP64, T[e] = sum of pair value times
// pair_scale, from the smallest scale up (as in the 16x16 form, so the sums
// round identically).
#define KSHT_SLICED32(Q, T)                                                              \
  {                                                                                      \
    i16 acc_[S];                                                                         \
    _Pragma("unroll") for (int g_ = 0; g_ < S; g_++)                                     \
      _Pragma("unroll") for (int e_ = 0; e_ < 16; e_++) acc_[g_][e_] = 0;                \
    _Pragma("unroll") for (int t = 0; t < S; t++) {                                      \
      _Pragma("unroll") for (int kh = 0; kh < 2; kh++) {                                 \
        const i4 b = BOF(Q, t, kh);                                                      \
        _Pragma("unroll") for (int s_ = 0; s_ + t < S; s_++)                             \
          acc_[s_ + t] = mfma32(AOF(s_, kh), b, acc_[s_ + t]);                           \
      }                                                                                  \
    }                                                                                    \
    _Pragma("unroll") for (int e_ = 0; e_ < 16; e_++) T[e_] = 0.0;                       \
    _Pragma("unroll") for (int gg = 0; gg < (S + 1) / 2; gg++) {                         \
      const int g0 = 2 * ((S + 1) / 2 - 1 - gg);                                         \
      const bool two = g0 + 1 < S;                                                       \
      const double sc = pair_scale(g0, two);                                             \
      _Pragma("unroll") for (int e_ = 0; e_ < 16; e_++) {                                \
        const int v = two ? acc_[g0][e_] * 256 + acc_[two ? g0 + 1 : g0][e_] : acc_[g0][e_]; \
        T[e_] = fma((double)v, sc, T[e_]);                                               \
      }                                                                                  \
    }                                                                                    \
  }

template <int S, int NB, bool VEC>
__global__ void __launch_bounds__(kGemm32Threads) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_leg_fwd_split32(
    SplitDims d, int nd_first, SplitTabs tb, const signed char *__restrict__ dig,
    const int *__restrict__ dexp, double *__restrict__ sp, SpecAddr sa,
    const int *__restrict__ jst, const int *__restrict__ ipad, const double *__restrict__ radius) {
  extern __shared__ i4 smem[];
  using LY = FwdLds32<S, NB>;
  constexpr int NBUF = LY::NBUF;
  constexpr int W = LY::W;
  constexpr int NO = VEC ? 3 : NB;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform per wave
  // same logical order as the 16x16 kernel: runs of (M, row block)
  long long bidx = xcd_group_order(blockIdx.x, gridDim.x, (long long)d.n_rt * d.ncg);
  const int cgl = (int)(bidx % d.ncg);
  bidx /= d.ncg;
  const int rt = (int)(bidx % d.n_rt);
  bidx /= d.n_rt;
  const int M = d.M0 + (int)(bidx % d.nM);
  const int ib = (int)(bidx / d.nM);
  const int L = d.L;
  const int ip0 = ipad[2 * M], ip1 = ipad[2 * M + 1];
  if (ib * 64 >= ip0) return;  // whole work-group beyond the degrees of M
  const int nsig = M == 0 ? 1 : 2;
  // wave w: parity w & 1, 32-row tile w >> 1
  const int par = wave & 1, tile = wave >> 1;
  const int i0 = ib * 64 + tile * 32;
  const int ip = par ? ip1 : ip0;
  const bool active = i0 < ip;
  const int nd0 = nd_first + NB * cgl;
  const int ntile = d.nr_p >> 4;
  const int nchf = d.nchf;
  const int sgl = (lane >> 4) & 1;  // this lane's slot (column l%32 / 16)
  const int lk = lane & 15;         // and radius within the tile

  const int gsl[2] = {(L + M) * nchf * 2 * d.ncomp, (L - M) * nchf * 2 * d.ncomp};
  const int erow0 = (int)tb.eoff[2 * M], erow1 = (int)tb.eoff[2 * M + 1];
  const unsigned fsrc = (unsigned)frag_src(lane);
  const unsigned lbase = lds_addr(smem);
  auto issue = [&](int c, int b) {
    const unsigned bufa = lbase + 16u * (unsigned)(b * LY::I4);
#pragma unroll
    for (int j = 0; j < LY::KB; j++) {
      const int b = wave + W * j;
      const int bb = b < LY::NBB ? b : 0;  // spare copies repeat block 0
      const int sgb = bb / (2 * NB * S), h = (bb / (NB * S)) & 1, f = bb / S % NB, s = bb % S;
      const int gb = gsl[sgb] + (c * 2 + h) * d.ncomp + nd0 + f;
      copy_block(dig + (size_t)((gb * S + s) * ntile + rt) * 1024, fsrc, bufa + 1024u * b);
    }
    const unsigned exa = bufa + 16u * LY::EX;
#pragma unroll
    for (int j = 0; j < LY::KE; j++) {
      const int task = wave + W * j;
      const int *base = tb.eP;  // spare copies
      unsigned off = 0;
      if (task < NB) {
        const int g = task * 4 + (lane >> 4);
        const int sgb = g / (2 * NB), h = (g / NB) & 1, f = g % NB;
        const int gb = gsl[sgb] + (c * 2 + h) * d.ncomp + nd0 + f;
        base = dexp;
        off = 4u * (unsigned)(gb * d.nr_p + rt * 16 + (lane & 15));
      } else if (task < NB + (VEC ? 6 : 2)) {
        const int tw = task - NB, tab = tw >> 1, pr = tw & 1;
        const int ipp = pr ? ip1 : ip0;
        const int row = ib * 64 + lane;
        base = tab == 0 ? tb.eP : tab == 1 ? tb.edP : tb.ePa;
        off = row < ipp ? 4u * (unsigned)((pr ? erow1 : erow0) + c * ipp + row) : 0u;
      }
      copy_dword(base, off, exa + 256u * task);
    }
  };

  double C[NO][16];
#pragma unroll
  for (int o = 0; o < NO; o++)
#pragma unroll
    for (int e = 0; e < 16; e++) C[o][e] = 0.0;

  const int hx = par, hy = 1 - par;
  // A rows: the wave's 32 rows, clamped into the (M, parity) table (its
  // rows are padded to 16 only); idle waves load from row 0 of the table so
  // that every wave issues the same vector-memory operations
  const int arow = active ? (int)(tb.off[2 * M + par] >> 6) + i0 : 0;
  const int ipa = active ? ip : 0;
  const int rmax = active ? ip - 1 - i0 : 31;
  const int arl = (lane & 31) < rmax ? (lane & 31) : rmax;
  const unsigned asrc0 = (unsigned)(arl * 64 + 16 * (lane >> 5));
  auto load_a = [&](const signed char *tab, int c, i4 (&a)[S][2]) {
#pragma unroll
    for (int s = 0; s < S; s++) {
      const signed char *rowp = tab + (size_t)(arow + (c * S + s) * ipa) * 64;
      a[s][0] = load_slice_ws(rowp, asrc0);
      a[s][1] = load_slice_ws(rowp, asrc0 + 32u);
    }
  };
  auto landed2 = [&](i4 (&a)[S][2]) {
#pragma unroll
    for (int s = 0; s < S; s++) {
      asm volatile("" : "+v"(a[s][0]));
      asm volatile("" : "+v"(a[s][1]));
    }
  };
  constexpr int KCOPY = LY::KB + LY::KE;
  i4 aP[S][2], adP[S][2], aPa[S][2];
  load_a(tb.P, 0, aP);
  if (VEC) load_a(tb.dP, 0, adP);
  issue(0, 0);
  if (VEC) load_a(tb.Pa, 0, aPa);
  // this lane's 16-byte offset inside a 1 KiB data block for K half kh
  const int boff0 = lk + 16 * (lane >> 5);
  const int sgc0 = nsig == 1 ? 0 : 1;  // conjugate of slot 0 (slot 1's is 0)
  const double msl = sgl == 0 ? (double)M : -(double)M;
  for (int c = 0; c < nchf; c++) {
    wait_vm<VEC ? 2 * S : 0>();
    __syncthreads();
    const i4 *buf = smem + (c % NBUF) * LY::I4;
    const int *ex = reinterpret_cast<const int *>(buf + LY::EX);
    const int cn = c + 1 < nchf ? c + 1 : c;
    if (NBUF == 2) issue(cn, (c + 1) % NBUF);
    {
#define QSLOT(q) ((VEC && ((q) == 2 || (q) == 4)) ? (sgl == 0 ? sgc0 : 0) : sgl)
#define QH(q) ((!VEC || (q) == 0 || (q) == 2 || (q) == 4) ? hx : hy)
#define QF(q) (!VEC ? (q) : (q) == 0 ? 0 : ((q) == 1 || (q) == 4) ? 1 : 2)
#define BOF(q, t, kh) buf[(((QSLOT(q) * 2 + QH(q)) * NB + QF(q)) * S + (t)) * 64 + boff0 + 32 * (kh)]
#define KSHT_FWD32_QUANTITY(Q, TAB)                                                        \
  {                                                                                        \
    constexpr int qq = (Q);                                                                \
    double T[16];                                                                          \
    KSHT_SLICED32(qq, T)                                                                   \
    constexpr int o = VEC ? (qq == 0 ? 0 : (qq <= 2) ? 1 : 2) : qq;                        \
    const double cf = (!VEC || qq <= 1) ? 1.0 : (qq == 3 ? -1.0 : -msl);                   \
    const int eb = ex[((QSLOT(qq) * 2 + QH(qq)) * NB + QF(qq)) * 16 + lk];                  \
    const int *er = ex + 64 * NB + (TAB) * 128 + par * 64 + tile * 32 + 4 * (lane >> 5);    \
    _Pragma("unroll") for (int e4 = 0; e4 < 4; e4++) {                                     \
      const i4 ea = *reinterpret_cast<const i4 *>(er + 8 * e4);                            \
      _Pragma("unroll") for (int r = 0; r < 4; r++) {                                      \
        const int e = 4 * e4 + r;                                                          \
        C[o][e] = fma(ldexp(T[e], ea[r] + eb - 20), cf, C[o][e]);                          \
      }                                                                                    \
    }                                                                                      \
  }
#define AOF(s, kh) AA_[s][kh]
      landed2(aP);
#define AA_ aP
      if (active) {
        KSHT_FWD32_QUANTITY(0, 0)
        if (!VEC) {
          if (NB > 1) KSHT_FWD32_QUANTITY(1 % NB, 0)
          if (NB > 2) KSHT_FWD32_QUANTITY(2 % NB, 0)
        }
      }
#undef AA_
      load_a(tb.P, cn, aP);
      if (VEC) {
        landed2(adP);
#define AA_ adP
        if (active) {
          KSHT_FWD32_QUANTITY(1, 1)
          KSHT_FWD32_QUANTITY(3, 1)
        }
#undef AA_
        load_a(tb.dP, cn, adP);
        wait_vm<(NBUF == 2 ? KCOPY : 0) + 4 * S>();
        landed2(aPa);
#define AA_ aPa
        if (active) {
          KSHT_FWD32_QUANTITY(2, 2)
          KSHT_FWD32_QUANTITY(4, 2)
        }
#undef AA_
      }
#undef KSHT_FWD32_QUANTITY
#undef AOF
#undef BOF
#undef QSLOT
#undef QH
#undef QF
    }
    if (NBUF == 1 && c + 1 < nchf) {
      __syncthreads();
      issue(c + 1, 0);
    }
    if (VEC) load_a(tb.Pa, cn, aPa);
  }
  wait_vm<0>();
  landed2(aP);
  if (VEC) {
    landed2(adP);
    landed2(aPa);
  }
  if (!active) return;
  const int k = rt * 16 + lk;
  if (k >= d.nr || sgl >= nsig) return;
  const double r1 = radius[k], r2 = r1 * r1;
  const long long obase = (long long)jst[sgl == 0 ? L + M : L - M] * sa.s_j +
                          (long long)k * sa.s_k + (long long)nd0 * sa.s_nd;
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const int ii = i0 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
    const int l = M + 2 * ii + par;
    if (l > L) continue;
    const double g7 = g7s_of(l, M), g6 = g6s_of(l, M);
    const long long o = obase + (long long)(l - M) * sa.s_j;
    if (VEC) {
      sp[o] = r2 * g7 * C[0][e];
      sp[o + sa.s_nd] = r1 * g7 * C[1 % NO][e];
      sp[o + 2 * sa.s_nd] = r1 * g7 * C[2 % NO][e];
    } else {
#pragma unroll
      for (int q = 0; q < NO; q++) sp[o + (long long)q * sa.s_nd] = g6 * C[q][e];
    }
  }
}
#undef KSHT_SLICED32

// ---------------------------------------------------------------------------
// Host side.
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
      const int nchb = bchunk_count(M, np);
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
  // data-operand slices: backward chunks per (slot, parity), forward
  // (slot, chunk, sym/asym); sized for max_ncomp components
  st->hcb.assign(2 * p->nslot, 0);
  long long ch = 0;
  for (int sl = 0; sl < p->nslot; sl++) {
    const int M = sl >= L ? sl - L : L - sl, n = L - M + 1;
    for (int par = 0; par < 2; par++) {
      st->hcb[2 * sl + par] = ch;
      ch += bchunk_count(M, par ? n / 2 : (n + 1) / 2);
    }
  }
  st->nchb = ch;
  st->nchf = nchf;
  const long long ntile = p->nr_p / 16;
  const long long gb = ch * p->max_ncomp, gf = (long long)p->nslot * nchf * 2 * p->max_ncomp;
  const long long groups = gb > gf ? gb : gf;
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
  if (p->p_otf) {
    // one tile's slices, both directions in the same buffers; offsets rebased
    std::vector<long long> rbo(2 * (L + 1)), rbe(2 * (L + 1)), rfo(2 * (L + 1)), rfe(2 * (L + 1));
    long long mb = 0, mbe = 0, mf = 0, mfe = 0;
    for (const auto &tl : p->mtiles) {
      const int q0 = 2 * tl.first, q1 = 2 * tl.second;
      for (int q = q0; q < q1; q++) {
        rbo[q] = st->hb_off[q] - st->hb_off[q0];
        rbe[q] = st->hb_eoff[q] - st->hb_eoff[q0];
        rfo[q] = st->hf_off[q] - st->hf_off[q0];
        rfe[q] = st->hf_eoff[q] - st->hf_eoff[q0];
      }
      const bool last = q1 >= 2 * (L + 1);
      mb = std::max(mb, (last ? bo : st->hb_off[q1]) - st->hb_off[q0]);
      mbe = std::max(mbe, (last ? be : st->hb_eoff[q1]) - st->hb_eoff[q0]);
      mf = std::max(mf, (last ? fo : st->hf_off[q1]) - st->hf_off[q0]);
      mfe = std::max(mfe, (last ? fe : st->hf_eoff[q1]) - st->hf_eoff[q0]);
    }
    if ((rc = alloc((void **)&st->b_off_r, nb)) || (rc = alloc((void **)&st->b_eoff_r, nb)) ||
        (rc = alloc((void **)&st->f_off_r, nb)) || (rc = alloc((void **)&st->f_eoff_r, nb)))
      return rc;
    KSHT_HIP(hipMemcpy(st->b_off_r, rbo.data(), nb, hipMemcpyHostToDevice));
    KSHT_HIP(hipMemcpy(st->b_eoff_r, rbe.data(), nb, hipMemcpyHostToDevice));
    KSHT_HIP(hipMemcpy(st->f_off_r, rfo.data(), nb, hipMemcpyHostToDevice));
    KSHT_HIP(hipMemcpy(st->f_eoff_r, rfe.data(), nb, hipMemcpyHostToDevice));
    const long long so = std::max(mb, mf), se = std::max(mbe, mfe);
    // the backward uses bP, bdP (beP, bedP), the forward fP, fdP, fPa: three
    // buffers serve both
    if ((rc = alloc((void **)&st->fP, so)) || (rc = alloc((void **)&st->fdP, so)) ||
        (rc = alloc((void **)&st->fPa, so)) || (rc = alloc((void **)&st->feP, se * 4)) ||
        (rc = alloc((void **)&st->fedP, se * 4)) || (rc = alloc((void **)&st->fePa, se * 4)))
      return rc;
    st->bP = st->fP;
    st->bdP = st->fdP;
    st->beP = st->feP;
    st->bedP = st->fedP;
  }
  if ((rc = alloc((void **)&st->b_off, nb)) || (rc = alloc((void **)&st->b_eoff, nb)) ||
      (rc = alloc((void **)&st->f_off, nb)) || (rc = alloc((void **)&st->f_eoff, nb)) ||
      (!p->p_otf && ((rc = alloc((void **)&st->bP, bo)) || (rc = alloc((void **)&st->bdP, bo)) ||
                     (rc = alloc((void **)&st->beP, be * 4)) ||
                     (rc = alloc((void **)&st->bedP, be * 4)) ||
                     (rc = alloc((void **)&st->fP, fo)) || (rc = alloc((void **)&st->fdP, fo)) ||
                     (rc = alloc((void **)&st->fPa, fo)) ||
                     (rc = alloc((void **)&st->feP, fe * 4)) ||
                     (rc = alloc((void **)&st->fedP, fe * 4)) ||
                     (rc = alloc((void **)&st->fePa, fe * 4)))) ||
      (rc = alloc((void **)&st->cb, st->hcb.size() * sizeof(long long))) ||
      (rc = alloc((void **)&st->dig, (size_t)(groups * S * ntile * 1024))) ||
      (rc = alloc((void **)&st->dexp, (size_t)(groups * p->nr_p * 4))))
    return rc;
  KSHT_HIP(hipMemcpy(st->b_off, st->hb_off.data(), nb, hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(st->b_eoff, st->hb_eoff.data(), nb, hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(st->f_off, st->hf_off.data(), nb, hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(st->f_eoff, st->hf_eoff.data(), nb, hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(st->cb, st->hcb.data(), st->hcb.size() * sizeof(long long),
                     hipMemcpyHostToDevice));
  return SHT_OK;
}

template <int S>
static constexpr int bwd_lds(int nb) {
  return nb == 4 ? BwdLds<S, 4>::bytes() : nb == 3 ? BwdLds<S, 3>::bytes()
         : nb == 2 ? BwdLds<S, 2>::bytes() : BwdLds<S, 1>::bytes();
}
template <int S>
static constexpr int fwd_lds(int nb) {
  return nb == 3 ? FwdLds<S, 3>::bytes() : nb == 2 ? FwdLds<S, 2>::bytes() : FwdLds<S, 1>::bytes();
}

template <int S>
static constexpr int fwd32_lds(int nb) {
  return nb == 3 ? FwdLds32<S, 3>::bytes() : nb == 2 ? FwdLds32<S, 2>::bytes()
                                                    : FwdLds32<S, 1>::bytes();
}

// Slices of the P tables of orders [M0, M1) for one direction, from the FP64
// tables of tile pt into the slice tables at offsets (off, eoff).
template <int S>
static int slice_range(sht_plan *p, SplitState *st, const PTile &pt, bool fwd,
                       const long long *off, const long long *eoff) {
  const int L = p->L, M0 = pt.M0, nM = pt.M1 - pt.M0;
  if (!fwd) {
    const int nchb_max = bchunk_count(M0, (L - M0 + 2) / 2);
    const dim3 gb((p->nth_hp + 255) / 256, nchb_max, 2 * nM);
    hipLaunchKernelGGL(k_slice_b<S>, gb, dim3(256), 0, p->stream, L, M0, p->nth_hp, pt.offb,
                       p->t.ipad, pt.Pb, off, eoff, st->bP, st->beP);
    hipLaunchKernelGGL(k_slice_b<S>, gb, dim3(256), 0, p->stream, L, M0, p->nth_hp, pt.offb,
                       p->t.ipad, pt.dPb, off, eoff, st->bdP, st->bedP);
  } else {
    int ipmax = 0;
    for (int q = 2 * pt.M0; q < 2 * pt.M1; q++) ipmax = std::max(ipmax, p->h_ipad[q]);
    const dim3 gf((ipmax + 255) / 256, p->nth_hp / 64, 2 * nM);
    hipLaunchKernelGGL(k_slice_f<S>, gf, dim3(256), 0, p->stream, L, M0, p->nth_h, p->nth_hp,
                       pt.offf, p->t.ipad, pt.Pf, p->t.asin_t, 0, off, eoff, st->fP, st->feP);
    hipLaunchKernelGGL(k_slice_f<S>, gf, dim3(256), 0, p->stream, L, M0, p->nth_h, p->nth_hp,
                       pt.offf, p->t.ipad, pt.dPf, p->t.asin_t, 0, off, eoff, st->fdP, st->fedP);
    hipLaunchKernelGGL(k_slice_f<S>, gf, dim3(256), 0, p->stream, L, M0, p->nth_h, p->nth_hp,
                       pt.offf, p->t.ipad, pt.Pf, p->t.asin_t, 1, off, eoff, st->fPa, st->fePa);
  }
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

// Tile ti of the P slices for one direction: with stored tables the whole
// sliced table; with P per call, the FP64 tile generated and sliced.
template <int S>
static int split_tile(sht_plan *p, SplitState *st, int ti, bool fwd, SplitTabs *tb, int *M0,
                      int *nM) {
  PTile pt;
  int rc = ptile_prepare(p, ti, fwd, &pt);
  if (rc) return rc;
  *M0 = pt.M0;
  *nM = pt.M1 - pt.M0;
  if (p->p_otf) {
    rc = fwd ? slice_range<S>(p, st, pt, true, st->f_off_r, st->f_eoff_r)
             : slice_range<S>(p, st, pt, false, st->b_off_r, st->b_eoff_r);
    if (rc) return rc;
  }
  if (fwd)
    *tb = SplitTabs{p->p_otf ? st->f_off_r : st->f_off, p->p_otf ? st->f_eoff_r : st->f_eoff,
                    st->fP, st->fdP, st->fPa, st->feP, st->fedP, st->fePa};
  else
    *tb = SplitTabs{p->p_otf ? st->b_off_r : st->b_off, p->p_otf ? st->b_eoff_r : st->b_eoff,
                    st->bP, st->bdP, nullptr, st->beP, st->bedP, nullptr};
  return SHT_OK;
}

template <int S>
static int slice_tables(sht_plan *p, SplitState *st) {
  if (!p->p_otf) {
    PTile pt;
    int rc = ptile_prepare(p, 0, false, &pt);
    if (rc) return rc;
    if ((rc = slice_range<S>(p, st, pt, false, st->b_off, st->b_eoff)) ||
        (rc = slice_range<S>(p, st, pt, true, st->f_off, st->f_eoff)))
      return rc;
  }
  // LDS limits of the GEMM kernels (dynamic shared memory above 64 KiB)
#define KSHT_LDS(K, BYTES) \
  KSHT_HIP(hipFuncSetAttribute((const void *)(K), hipFuncAttributeMaxDynamicSharedMemorySize, (BYTES)))
  KSHT_LDS((k_leg_bwd_split<S, 3, true>), bwd_lds<S>(3));
  if constexpr (bwd_lds<S>(4) <= 160 * 1024) KSHT_LDS((k_leg_bwd_split<S, 4, true>), bwd_lds<S>(4));
  KSHT_LDS((k_leg_bwd_split<S, 3, false>), bwd_lds<S>(3));
  KSHT_LDS((k_leg_bwd_split<S, 2, false>), bwd_lds<S>(2));
  KSHT_LDS((k_leg_bwd_split<S, 1, false>), bwd_lds<S>(1));
  KSHT_LDS((k_leg_fwd_split<S, 3, true>), fwd_lds<S>(3));
  KSHT_LDS((k_leg_fwd_split<S, 3, false>), fwd_lds<S>(3));
  KSHT_LDS((k_leg_fwd_split<S, 2, false>), fwd_lds<S>(2));
  KSHT_LDS((k_leg_fwd_split<S, 1, false>), fwd_lds<S>(1));
  KSHT_LDS((k_leg_fwd_split32<S, 3, true>), fwd32_lds<S>(3));
  KSHT_LDS((k_leg_fwd_split32<S, 3, false>), fwd32_lds<S>(3));
  KSHT_LDS((k_leg_fwd_split32<S, 2, false>), fwd32_lds<S>(2));
  KSHT_LDS((k_leg_fwd_split32<S, 1, false>), fwd32_lds<S>(1));
  // one radius of 64 components: 132 * 64 doubles of padded staging
  KSHT_LDS(k_bslice<S>, (int)sizeof(double) * 132 * 64);
#undef KSHT_LDS
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
  if (p->p_otf) {  // the backward's tile buffers are the forward's
    st->bP = st->bdP = nullptr;
    st->beP = st->bedP = nullptr;
  }
  void *ptrs[] = {st->b_off, st->b_eoff, st->f_off, st->f_eoff, st->bP,  st->bdP,
                  st->beP,   st->bedP,   st->fP,    st->fdP,    st->fPa, st->feP,
                  st->fedP,  st->fePa,   st->cb,    st->dig,    st->dexp, st->b_off_r,
                  st->b_eoff_r, st->f_off_r, st->f_eoff_r};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete st;
  p->split = nullptr;
}

void split_ws_bytes(const sht_plan *p, int ncomp, bool backward, long long *dig_bytes,
                    long long *dexp_bytes) {
  const SplitState *st = (const SplitState *)p->split;
  const long long groups =
      backward ? st->nchb * ncomp : (long long)p->nslot * st->nchf * 2 * ncomp;
  *dig_bytes = groups * st->S * (p->nr_p / 16) * 1024;
  *dexp_bytes = groups * p->nr_p * 4;
}

Workspace plan_workspace(const sht_plan *p) {
  const SplitState *st = (const SplitState *)p->split;
  return Workspace{p->X, st ? st->dig : nullptr, st ? st->dexp : nullptr};
}

static SplitDims split_dims(const sht_plan *p, int ncomp, int nblk) {
  SplitDims d;
  d.L = p->L;
  d.nr = p->nr;
  d.nr_p = p->nr_p;
  d.nth = p->nth;
  d.nth_h = p->nth_h;
  d.nth_hp = p->nth_hp;
  d.ncomp = ncomp;
  d.n_rt = p->nr_p / 16;
  d.ncg = 0;
  d.nblk = nblk;
  d.nchf = p->nth_hp / 64;
  d.M0 = 0;
  d.nM = p->L + 1;
  return d;
}

// component groups of a call: vectors as (pol, dpol, tor) triples, scalars in
// threes plus a remainder group of one or two
struct GroupLaunch {
  int nb, vec, ncg, nd_first, nd_scl;  // nd_scl: first fused scalar (nb = 4)
};
static int component_groups(int nv, int ns, GroupLaunch g[3]) {
  int n = 0;
  if (nv) g[n++] = GroupLaunch{3, 1, nv, 0, 0};
  if (ns / 3) g[n++] = GroupLaunch{3, 0, ns / 3, 3 * nv, 0};
  if (ns % 3) g[n++] = GroupLaunch{ns % 3, 0, 1, 3 * nv + 3 * (ns / 3), 0};
  return n;
}
// backward groups with the first nf scalars fused into the first nf vector
// groups (one each), then the remaining vectors and scalars
static int component_groups_fused(int nv, int ns, int nf, GroupLaunch g[4]) {
  int n = 0;
  if (nf) g[n++] = GroupLaunch{4, 1, nf, 0, 3 * nv};
  if (nv > nf) g[n++] = GroupLaunch{3, 1, nv - nf, 3 * nf, 0};
  const int r = ns - nf, s0 = 3 * nv + nf;
  if (r / 3) g[n++] = GroupLaunch{3, 0, r / 3, s0, 0};
  if (r % 3) g[n++] = GroupLaunch{r % 3, 0, 1, s0 + 3 * (r / 3), 0};
  return n;
}

template <int S>
static int bwd_split_S(sht_plan *p, const Workspace &ws, int nv, int ns, const double *spec,
                       SpecAddr sa, double *out, GridAddr ga) {
  SplitState *st = (SplitState *)p->split;
  const int ncomp = 3 * nv + ns;
  // radii per work-group: 4, unless the padded staging (132 ncomp doubles
  // per radius) would pass 64 KiB of LDS (nr_p is a multiple of 16)
  const int nrad = ncomp <= 15 ? 4 : ncomp <= 31 ? 2 : 1;
  hipLaunchKernelGGL(k_bslice<S>, dim3(p->nr_p / nrad, p->nslot), dim3(kSliceThreads),
                     sizeof(double) * nrad * 132 * ncomp, p->stream,
                     p->L, p->nr, p->nr_p, nrad, nv, ncomp, spec, sa, p->t.jst, st->cb, ws.dig,
                     ws.dexp);
  KSHT_HIP(hipGetLastError());
  GroupLaunch g[4];
  constexpr bool can_fuse = bwd_lds<S>(4) <= 160 * 1024;
  const int nf = can_fuse && p->var_bwd_fused ? (nv < ns ? nv : ns) : 0;
  const int ng = component_groups_fused(nv, ns, nf, g);
  for (int ti = 0; ti < ptile_count(p); ti++) {
    SplitTabs tb;
    int M0 = 0, nM = 0;
    int rc = split_tile<S>(p, st, ti, false, &tb, &M0, &nM);
    if (rc) return rc;
    // the groups run in order: beside the vector group (ForkJoin, as the
    // FP64 kernels do) the scalar group only slowed it by as much as it
    // saved (15.5 against 15.3 ms at L=511), the split GEMMs keeping the
    // chip busy to their last work-groups
    const hipStream_t sm = p->stream;
    for (int i = 0; i < ng; i++) {
      SplitDims d = split_dims(p, ncomp, p->nth_hp / 64);
      d.ncg = g[i].ncg;
      d.M0 = M0;
      d.nM = nM;
      const long long nblocks = (long long)nM * d.n_rt * d.ncg * d.nblk;
      const dim3 grid((unsigned)nblocks), blk(kGemmThreads);
#define KSHT_BWD(NB, VEC)                                                                      \
  hipLaunchKernelGGL((k_leg_bwd_split<S, NB, VEC>), grid, blk, bwd_lds<S>(NB), sm, d,        \
                     g[i].nd_first, tb, st->cb, ws.dig, ws.dexp, out, ga, p->t.a_r,          \
                     p->t.asin_t, g[i].nd_scl)
      if (g[i].vec && g[i].nb == 4) {
        if constexpr (can_fuse) {
          StageTimer tk(p, 5);
          KSHT_BWD(4, true);
        }
      } else if (g[i].vec) {
        StageTimer tk(p, 5);
        KSHT_BWD(3, true);
      } else if (g[i].nb == 3)
        KSHT_BWD(3, false);
      else if (g[i].nb == 2)
        KSHT_BWD(2, false);
      else
        KSHT_BWD(1, false);
#undef KSHT_BWD
      KSHT_HIP(hipGetLastError());
    }
  }
  return SHT_OK;
}

template <int S>
static int fwd_split_S(sht_plan *p, const Workspace &ws, int nv, int ns, const double *in,
                       GridAddr ga, double *sp, SpecAddr sa) {
  SplitState *st = (SplitState *)p->split;
  const int ncomp = 3 * nv + ns;
  hipLaunchKernelGGL(k_fslice<S>, dim3((p->nr_p / 16) * p->nslot), dim3(kSliceThreads), 0,
                     p->stream, p->nr, p->nr_p, p->nth, p->nth_h, st->nchf, ncomp, in, ga,
                     p->t.weight, ws.dig, ws.dexp);
  KSHT_HIP(hipGetLastError());
  const int ne_max = (p->L + 2) / 2;
  GroupLaunch g[3];
  const int ng = component_groups(nv, ns, g);
  for (int ti = 0; ti < ptile_count(p); ti++) {
    SplitTabs tb;
    int M0 = 0, nM = 0;
    int rc = split_tile<S>(p, st, ti, true, &tb, &M0, &nM);
    if (rc) return rc;
    const hipStream_t sm = p->stream;  // groups in order, as in the backward
    for (int i = 0; i < ng; i++) {
      SplitDims d = split_dims(p, ncomp, (ne_max + 63) / 64);
      d.ncg = g[i].ncg;
      d.M0 = M0;
      d.nM = nM;
      const long long nblocks = (long long)nM * d.n_rt * d.ncg * d.nblk;
      const dim3 grid((unsigned)nblocks);
      // INT8 MFMA form of the products: 32x32x32 (4-wave work-groups) or
      // 16x16x64 (8 waves), the plan's "split_fwd_mfma" variant
#define KSHT_FWD(NB, VEC)                                                                      \
  if (p->var_fwd_mfma == 32)                                                                   \
    hipLaunchKernelGGL((k_leg_fwd_split32<S, NB, VEC>), grid, dim3(kGemm32Threads),            \
                       fwd32_lds<S>(NB), sm, d, g[i].nd_first, tb, ws.dig, ws.dexp, sp,        \
                       sa, p->t.jst, p->t.ipad, p->t.radius);                                  \
  else                                                                                         \
    hipLaunchKernelGGL((k_leg_fwd_split<S, NB, VEC>), grid, dim3(kGemmThreads),                \
                       fwd_lds<S>(NB), sm, d, g[i].nd_first, tb, ws.dig, ws.dexp, sp,          \
                       sa, p->t.jst, p->t.ipad, p->t.radius)
      if (g[i].vec) {
        StageTimer tk(p, 6);
        KSHT_FWD(3, true);
      } else if (g[i].nb == 3)
        KSHT_FWD(3, false);
      else if (g[i].nb == 2)
        KSHT_FWD(2, false);
      else
        KSHT_FWD(1, false);
#undef KSHT_FWD
      KSHT_HIP(hipGetLastError());
    }
  }
  return SHT_OK;
}

int launch_leg_bwd_split(sht_plan *p, const Workspace &ws, int nv, int ns, const double *sp,
                         SpecAddr sa, double *out, GridAddr ga) {
  if (!p->split) {
    set_error("split precision: plan was not created in split mode");
    return SHT_ERR_STATE;
  }
  if (3 * nv + ns > p->max_ncomp) {
    set_error("split precision: %d components exceed the plan's max_ncomp %d", 3 * nv + ns,
              p->max_ncomp);
    return SHT_ERR_ARG;
  }
  switch (((SplitState *)p->split)->S) {
    case 4: return bwd_split_S<4>(p, ws, nv, ns, sp, sa, out, ga);
    case 5: return bwd_split_S<5>(p, ws, nv, ns, sp, sa, out, ga);
    case 6: return bwd_split_S<6>(p, ws, nv, ns, sp, sa, out, ga);
    default: return bwd_split_S<7>(p, ws, nv, ns, sp, sa, out, ga);
  }
}

int launch_leg_fwd_split(sht_plan *p, const Workspace &ws, int nv, int ns, const double *in,
                         GridAddr ga, double *sp, SpecAddr sa) {
  if (!p->split) {
    set_error("split precision: plan was not created in split mode");
    return SHT_ERR_STATE;
  }
  if (3 * nv + ns > p->max_ncomp) {
    set_error("split precision: %d components exceed the plan's max_ncomp %d", 3 * nv + ns,
              p->max_ncomp);
    return SHT_ERR_ARG;
  }
  switch (((SplitState *)p->split)->S) {
    case 4: return fwd_split_S<4>(p, ws, nv, ns, in, ga, sp, sa);
    case 5: return fwd_split_S<5>(p, ws, nv, ns, in, ga, sp, sa);
    case 6: return fwd_split_S<6>(p, ws, nv, ns, in, ga, sp, sa);
    default: return fwd_split_S<7>(p, ws, nv, ns, in, ga, sp, sa);
  }
}

}  // namespace ksht
