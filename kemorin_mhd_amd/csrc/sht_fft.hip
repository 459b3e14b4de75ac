// sht_fft.hip -- zonal real FFT along phi in the FFTPACK5.1 convention
// (MHD/External_libs/FFTPACK5.1D/fftpack5.1d.f90:9288-9335; slot use of
// SS/t_sph_FFTPACK5.f90:183-331):
//   forward  a0 = (1/N) sum x_j,  a_m = (2/N) sum x_j cos(2 pi j m/N),
//            b_m = (2/N) sum x_j sin(2 pi j m/N)
//   backward x_j = a0 + sum_{m>=1} (a_m cos + b_m sin)   (m <= L only).
// Two real lines are packed as the real and imaginary parts of one complex
// line; a work-group runs several complex Stockham FFTs (mixed radix 8/4/2/3/
// 5, small DFTs for 7, 11, 13 and direct sums for larger primes) in LDS.  The Fourier side
// holds orders m <= L in the compact order q = 0 (a0), 2m-1 (a_m), 2m (b_m);
// the grid side is the reference's field-major v_rtp with phi slowest.
#include <algorithm>
#include <cmath>

#include "sht_internal.h"

namespace ksht {

constexpr int kFftThreads = 256;
// Complex elements per work-group held in LDS: 3072 (48 KiB, three
// work-groups per CU) for Nphi <= 3072, 6144 (96 KiB) for longer lines.
constexpr int kFftElemsSmall = 3072, kFftElemsLarge = 6144;
// primes from this size up take a Bluestein pass when its work arrays fit
constexpr int kBluesteinMin = 48;

// Work-groups are dispatched round-robin over the 8 XCDs; neighbouring line
// blocks share 128-byte lines of the phi-major grid, so consecutive logical
// blocks are kept on one XCD (and its L2).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned per = nb / 8;
  if (b >= per * 8) return b;  // tail blocks keep their index
  return (b % 8) * per + b / 8;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by s*i (s = +1 or -1)
__device__ __forceinline__ double2 cmuli(double2 a, double s) { return make_double2(-s * a.y, s * a.x); }

template <int R>
__device__ __forceinline__ void dft_small(double2 *v, double s) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const double2 t2 = cadd(v[1], v[3]), t3 = cmuli(csub(v[1], v[3]), s);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else if constexpr (R == 8) {
    double2 a[4] = {v[0], v[2], v[4], v[6]}, b[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4>(a, s);
    dft_small<4>(b, s);
    const double h = 0.70710678118654752440;
    // w^k b_k with w = exp(s*2*pi*i/8)
    const double2 b1 = make_double2(h * (b[1].x - s * b[1].y), h * (b[1].y + s * b[1].x));
    const double2 b2 = cmuli(b[2], s);
    const double2 b3 = make_double2(h * (-b[3].x - s * b[3].y), h * (-b[3].y + s * b[3].x));
    v[0] = cadd(a[0], b[0]);
    v[4] = csub(a[0], b[0]);
    v[1] = cadd(a[1], b1);
    v[5] = csub(a[1], b1);
    v[2] = cadd(a[2], b2);
    v[6] = csub(a[2], b2);
    v[3] = cadd(a[3], b3);
    v[7] = csub(a[3], b3);
  } else if constexpr (R == 3) {
    const double c = -0.5, sn = 0.86602540378443864676 * s;
    const double2 t = cadd(v[1], v[2]), u = csub(v[1], v[2]);
    const double2 m = make_double2(v[0].x + c * t.x, v[0].y + c * t.y);
    const double2 iu = make_double2(-sn * u.y, sn * u.x);
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, iu);
    v[2] = csub(m, iu);
  } else {
    double2 o[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
      double2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; r++) {
        const int e = (r * q) % R;
        double sv, cv;
        sincospi(2.0 * e / R, &sv, &cv);
        acc = cadd(acc, cmul(v[r], make_double2(cv, s * sv)));
      }
      o[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; q++) v[q] = o[q];
  }
}

// One Stockham pass over F complex FFTs of length n held in LDS.
template <int E, int R>
__device__ void fft_pass(double2 *buf, int n, int F, int Ns, const double2 *__restrict__ tw,
                         double s) {
  constexpr int NBT = (E / kFftThreads + R - 1) / R;
  const int nR = n / R, nb = F * nR;
  double2 v[NBT][R];
#pragma unroll
  for (int it = 0; it < NBT; it++) {
    const int bi = threadIdx.x + it * kFftThreads;
    if (bi < nb) {
      const int f = bi / nR, j = bi - f * nR;
      const int k = j % Ns;
      const double2 *src = buf + (long)f * n + j;
#pragma unroll
      for (int r = 0; r < R; r++) v[it][r] = src[r * nR];
#pragma unroll
      for (int r = 1; r < R; r++) {
        double2 w = tw[k * R + r];
        w.y *= -s;  // table holds exp(-2 pi i k r / (Ns R))
        v[it][r] = cmul(v[it][r], w);
      }
      dft_small<R>(v[it], s);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NBT; it++) {
    const int bi = threadIdx.x + it * kFftThreads;
    if (bi < nb) {
      const int f = bi / nR, j = bi - f * nR;
      const int k = j % Ns;
      double2 *dst = buf + (long)f * n + (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) dst[r * Ns] = v[it][r];
    }
  }
  __syncthreads();
}

// One Stockham pass of a prime radix R above 13 (FFTPACK accepts any length;
// Nphi = 3 (L+1) has factors up to L+1).  Too many points per butterfly for
// registers: every output is a direct sum over its R inputs, written out of
// place to scr (F n elements past buf) and copied back.  Output q of the
// butterfly j = (j / Ns) Ns + k is sum_r x_r W^{r (k + Ns q)}, W = exp(s 2 pi
// i / (Ns R)), the twiddle of the pass and the R-point DFT in one root; tb
// holds exp(-2 pi i e / (Ns R)), e < Ns R.
__device__ void fft_pass_big(double2 *buf, double2 *scr, int n, int F, int Ns, int R,
                             const double2 *__restrict__ tb, double s) {
  const int nR = n / R, NR = Ns * R, tot = F * n;
  if (Ns == 1) {
    // first pass (the plan orders these primes first): no twiddle, output q
    // of butterfly j is sum_r x_{j + r nR} w^{r q}, w = exp(s 2 pi i / R).
    // Consecutive lanes take consecutive q of one butterfly, so each input
    // is one broadcast LDS read; the powers w^{r q} come from a recurrence,
    // re-anchored on the exact table every 16 terms.
    for (int o = threadIdx.x; o < tot; o += kFftThreads) {
      const int f = o / n, pos = o - f * n;
      const int j = pos / R, q = pos - j * R;
      const double2 *src = buf + (long)f * n + j;
      const double2 wq = make_double2(tb[q].x, -s * tb[q].y);
      double2 acc = make_double2(0.0, 0.0), t = make_double2(1.0, 0.0);
      int e = 0;  // r q mod R
      for (int r = 0; r < R; r++) {
        if ((r & 15) == 0) t = make_double2(tb[e].x, -s * tb[e].y);
        const double2 x = src[r * nR];
        acc.x += x.x * t.x - x.y * t.y;
        acc.y += x.x * t.y + x.y * t.x;
        t = cmul(t, wq);
        e += q;
        if (e >= R) e -= R;
      }
      scr[o] = acc;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < tot; o += kFftThreads) buf[o] = scr[o];
    __syncthreads();
    return;
  }
  for (int o = threadIdx.x; o < tot; o += kFftThreads) {
    const int f = o / n, pos = o - f * n;
    const int jb = pos / NR, w = pos - jb * NR, q = w / Ns, k = w - q * Ns;
    const double2 *src = buf + (long)f * n + jb * Ns + k;
    const int e0 = k + Ns * q;
    double2 acc = make_double2(0.0, 0.0);
    int e = 0;
    for (int r = 0; r < R; r++) {
      const double2 t = tb[e];
      const double2 x = src[r * nR];
      acc.x += x.x * t.x + s * x.y * t.y;  // x * (t.x, -s t.y)
      acc.y += x.y * t.x - s * x.x * t.y;
      e += e0;
      if (e >= NR) e -= NR;
    }
    scr[o] = acc;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < tot; o += kFftThreads) buf[o] = scr[o];
  __syncthreads();
}

// M-point FFTs (M a power of two) over nseq sequences held in LDS.
template <int E>
__device__ void fft_pow2(double2 *w, const FftPlan &fp, const FftPass &ps, int nseq,
                         const double2 *__restrict__ tw, double s) {
  for (int ip = 0; ip < ps.nm; ip++) {
    const FftPass q = fp.pass[ps.m0 + ip];
    switch (q.R) {
      case 8: fft_pass<E, 8>(w, ps.M, nseq, q.Ns, tw + q.tw_off, s); break;
      case 4: fft_pass<E, 4>(w, ps.M, nseq, q.Ns, tw + q.tw_off, s); break;
      default: fft_pass<E, 2>(w, ps.M, nseq, q.Ns, tw + q.tw_off, s); break;
    }
  }
}

// First pass of a large prime radix R as Bluestein's chirp convolution:
// with c_k = exp(s i pi k^2 / R) (2 r q = r^2 + q^2 - (q-r)^2),
//   y_q = sum_r x_r w^{r q} = c_q sum_r (x_r c_r) conj(c_{q-r}),
// a cyclic convolution of length M >= 2R - 1 done by M-point FFTs; the
// filter's transform (1/M included) is precomputed.  The F (n/R) sequences
// of the pass are transformed together in wrk (F (n/R) M elements past buf).
template <int E>
__device__ void fft_pass_bluestein(double2 *buf, double2 *wrk, const FftPlan &fp,
                                   const FftPass &ps, int F, const double2 *__restrict__ tw,
                                   double s) {
  const int n = fp.n, R = ps.R, M = ps.M, nR = n / R, nseq = F * nR;
  const int si = s > 0 ? 0 : 1;
  const double2 *c = tw + ps.chirp_off + si * R, *hat = tw + ps.hat_off + si * M;
  for (int o = threadIdx.x; o < nseq * M; o += kFftThreads) {
    const int sq = o / M, m = o - sq * M;
    const int f = sq / nR, j = sq - f * nR;
    wrk[o] = m < R ? cmul(buf[(long)f * n + j + m * nR], c[m]) : make_double2(0.0, 0.0);
  }
  __syncthreads();
  fft_pow2<E>(wrk, fp, ps, nseq, tw, -1.0);
  for (int o = threadIdx.x; o < nseq * M; o += kFftThreads) wrk[o] = cmul(wrk[o], hat[o % M]);
  __syncthreads();
  fft_pow2<E>(wrk, fp, ps, nseq, tw, 1.0);
  // output q of butterfly j goes to Stockham position j R + q (Ns = 1)
  for (int o = threadIdx.x; o < nseq * R; o += kFftThreads) {
    const int sq = o / R, q = o - sq * R;
    const int f = sq / nR, j = sq - f * nR;
    buf[(long)f * n + j * R + q] = cmul(wrk[(long)sq * M + q], c[q]);
  }
  __syncthreads();
}

template <int E, bool GEN>
__device__ void fft_run(double2 *buf, const FftPlan &fp, int F, const double2 *__restrict__ tw,
                        double s) {
  for (int ip = 0; ip < fp.npass; ip++) {
    const FftPass ps = fp.pass[ip];
    const double2 *t = tw + ps.tw_off;
    if constexpr (GEN) {
      if (ps.M) {
        fft_pass_bluestein<E>(buf, buf + (long)F * fp.n, fp, ps, F, tw, s);
        continue;
      }
    }
    switch (ps.R) {
      case 8: fft_pass<E, 8>(buf, fp.n, F, ps.Ns, t, s); break;
      case 4: fft_pass<E, 4>(buf, fp.n, F, ps.Ns, t, s); break;
      case 2: fft_pass<E, 2>(buf, fp.n, F, ps.Ns, t, s); break;
      case 3: fft_pass<E, 3>(buf, fp.n, F, ps.Ns, t, s); break;
      default:
        // radices 5, 7, 11, 13 and the larger primes only in the GEN
        // instantiation (register and LDS cost)
        if constexpr (GEN) {
          switch (ps.R) {
            case 5: fft_pass<E, 5>(buf, fp.n, F, ps.Ns, t, s); break;
            case 7: fft_pass<E, 7>(buf, fp.n, F, ps.Ns, t, s); break;
            case 11: fft_pass<E, 11>(buf, fp.n, F, ps.Ns, t, s); break;
            case 13: fft_pass<E, 13>(buf, fp.n, F, ps.Ns, t, s); break;
            default: fft_pass_big(buf, buf + (long)F * fp.n, fp.n, F, ps.Ns, ps.R, t, s); break;
          }
        }
        break;
    }
  }
}

// Fourier coefficients -> grid.  blockIdx.x: block of B lines, blockIdx.y: nd.
template <int E, bool GEN>
__global__ void __launch_bounds__(kFftThreads) k_fft_bwd(
    FftPlan fp, int L, long long nline, const double *__restrict__ X, FourierAddr xa,
    double *__restrict__ v, const double2 *__restrict__ tw) {
  extern __shared__ double2 buf[];
  const int n = fp.n, B = fp.lines_per_wg, F = B / 2;
  const int nd = blockIdx.y;
  const long long line0 = (long long)xcd_block(blockIdx.x, gridDim.x) * B;
  const long long nnod = nline * n;
  // zero the orders above L
  for (int idx = threadIdx.x; idx < F * (n - 2 * L - 1); idx += kFftThreads) {
    const int f = idx / (n - 2 * L - 1), k = L + 1 + idx % (n - 2 * L - 1);
    buf[f * n + k] = make_double2(0.0, 0.0);
  }
  for (int idx = threadIdx.x; idx < F * (L + 1); idx += kFftThreads) {
    const int m = idx / F, f = idx - m * F;
    const long long la = line0 + 2 * f, lb = la + 1;
    const long long base = (long long)nd * xa.s_nd;
    double aa = 0, ba = 0, ab = 0, bb = 0;
    if (m == 0) {
      const long long o = q_offset(xa, 0) + base;
      if (la < nline) aa = X[o + la * xa.s_line];
      if (lb < nline) ab = X[o + lb * xa.s_line];
      buf[f * n] = make_double2(aa, ab);
    } else {
      const long long oc = q_offset(xa, 2 * m - 1) + base;
      const long long os = q_offset(xa, 2 * m) + base;
      if (la < nline) {
        aa = X[oc + la * xa.s_line];
        ba = X[os + la * xa.s_line];
      }
      if (lb < nline) {
        ab = X[oc + lb * xa.s_line];
        bb = X[os + lb * xa.s_line];
      }
      // C = (a - i b)/2 per line; Z_m = C_a + i C_b, Z_{n-m} = conj(C_a) + i conj(C_b)
      buf[f * n + m] = make_double2(0.5 * (aa + bb), 0.5 * (ab - ba));
      buf[f * n + n - m] = make_double2(0.5 * (aa - bb), 0.5 * (ab + ba));
    }
  }
  __syncthreads();
  fft_run<E, GEN>(buf, fp, F, tw, +1.0);
  for (int idx = threadIdx.x; idx < F * n; idx += kFftThreads) {
    const int j = idx / F, f = idx - j * F;
    const long long la = line0 + 2 * f;
    const double2 z = buf[f * n + j];
    double *o = v + (long long)nd * nnod + (long long)j * nline + la;
    if (la + 1 < nline) {
      if ((nline & 1) == 0)
        *reinterpret_cast<double2 *>(o) = z;  // 16-byte aligned: la and nline even
      else {
        o[0] = z.x;
        o[1] = z.y;
      }
    } else if (la < nline) {
      o[0] = z.x;
    }
  }
}

// grid -> Fourier coefficients (orders m <= L).
template <int E, bool GEN>
__global__ void __launch_bounds__(kFftThreads) k_fft_fwd(
    FftPlan fp, int L, long long nline, const double *__restrict__ v, double *__restrict__ X,
    FourierAddr xa, const double2 *__restrict__ tw) {
  extern __shared__ double2 buf[];
  const int n = fp.n, B = fp.lines_per_wg, F = B / 2;
  const int nd = blockIdx.y;
  const long long line0 = (long long)xcd_block(blockIdx.x, gridDim.x) * B;
  const long long nnod = nline * n;
  for (int idx = threadIdx.x; idx < F * n; idx += kFftThreads) {
    const int j = idx / F, f = idx - j * F;
    const long long la = line0 + 2 * f;
    const double *src = v + (long long)nd * nnod + (long long)j * nline + la;
    double2 z = make_double2(0.0, 0.0);
    if (la + 1 < nline) {
      if ((nline & 1) == 0)
        z = *reinterpret_cast<const double2 *>(src);
      else
        z = make_double2(src[0], src[1]);
    } else if (la < nline) {
      z.x = src[0];
    }
    buf[f * n + j] = z;
  }
  __syncthreads();
  fft_run<E, GEN>(buf, fp, F, tw, -1.0);
  const double sn = 1.0 / (double)n, tsn = 2.0 / (double)n;
  for (int idx = threadIdx.x; idx < F * (L + 1); idx += kFftThreads) {
    const int m = idx / F, f = idx - m * F;
    const long long la = line0 + 2 * f, lb = la + 1;
    const double2 zm = buf[f * n + m];
    const double2 zn = buf[f * n + (m == 0 ? 0 : n - m)];
    // X_a = (Z_m + conj Z_{n-m})/2, X_b = -i (Z_m - conj Z_{n-m})/2
    const double xar = 0.5 * (zm.x + zn.x), xai = 0.5 * (zm.y - zn.y);
    const double xbr = 0.5 * (zm.y + zn.y), xbi = -0.5 * (zm.x - zn.x);
    const long long base = (long long)nd * xa.s_nd;
    if (m == 0) {
      const long long o = q_offset(xa, 0) + base;
      if (la < nline) X[o + la * xa.s_line] = sn * xar;
      if (lb < nline) X[o + lb * xa.s_line] = sn * xbr;
    } else {
      const long long oc = q_offset(xa, 2 * m - 1) + base;
      const long long os = q_offset(xa, 2 * m) + base;
      if (la < nline) {
        X[oc + la * xa.s_line] = tsn * xar;
        X[os + la * xa.s_line] = -tsn * xai;
      }
      if (lb < nline) {
        X[oc + lb * xa.s_line] = tsn * xbr;
        X[os + lb * xa.s_line] = -tsn * xbi;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Split form for Nphi = R0 * NP with NP a power of two, L < NP <= 1024 (the
// usual Nphi = 3 * 2^k grids).  With j = R0 j2 + j1:
//   backward  x_j = Re/Im of sum_m Z_m e^{+2 pi i m j/N}
//                 = sum_{m'<NP} D^{(j1)}_m' e^{+2 pi i m' j2/NP},
//             D^{(j1)}_m' = e^{2 pi i m' j1/N} Z_m' + e^{2 pi i (m'+N-NP) j1/N} Z_{m'+N-NP}
//   (Z_m is non-zero only for m <= L and m >= N-L, so only these two of the
//   R0 aliases of m' survive), i.e. R0 NP-point FFTs per line; the forward
//   runs the same steps in reverse and accumulates Z_m (m <= L) and Z_{N-m}.
// The NP-point FFTs are Stockham radix-8 passes (radix 4 / 2 to finish) with
// eight points per thread in registers and padded LDS between passes.
//
// A work-group owns NC complex lines (2 NC consecutive real lines packed in
// pairs), T = NP/8 threads per complex line, so a row of the line-contiguous
// layouts is 16 NC contiguous bytes.  NC = 4: 64-byte half rows and about
// 72 KiB of LDS, so two work-groups share a CU and one's copies and stores
// overlap the other's passes.  The two halves of a 128-byte row belong to
// neighbouring logical blocks, which xcd_block keeps on one XCD (one L2)
// at about the same time.
// The backward runs NC = 4 (launch bounds: at least two waves per SIMD, i.e.
// at most 256 VGPRs), NC = 2 at NP = 1024 (fft_ncb).  The forward keeps NC = 8, one work-group per CU, and
// overlaps instead the copy of the next phi decimation with the passes of
// the current one inside the work-group (NC = 4 with two work-groups per CU
// measured the same).  Both take their twiddles from the NP roots and the
// R0 roots.  NP = 1024 (L up to 1023): a complex line spans two waves
// (T = 128), the passes synchronise the work-group, and the forward runs
// NC = 4 with the same in-kernel copy (about 156 KB of LDS).
// backward complex lines per work-group: 4, and 2 at NP = 1024, where the
// whole-block staging of 2 NP rows then leaves room for two work-groups per
// CU once the twiddles are two-level factors (NP/64 + 64 entries)
template <int NP>
constexpr int fft_ncb() { return NP >= 1024 ? 2 : 4; }
template <int NP>
constexpr bool fft_bwd_2l() { return NP >= 1024; }
template <int NP>
constexpr int fft_bwd_tw_entries() { return fft_bwd_2l<NP>() ? NP / 64 + 64 : NP; }
template <int NP>
constexpr int fft_ncf() { return NP <= 512 ? 8 : 4; }
template <int NP>
constexpr bool fft_fwd_dma() { return NP <= 1024; }

__device__ __forceinline__ int fpad(int i) { return i + (i >> 3); }

// LDS layout, in double2 units.  The twiddles (NP roots of order NP, then the
// R0 roots exp(-2 pi i b / N), b < R0; copied from global memory at the
// start) sit past every other area, so that the passes and the alias sums
// read no global memory: a global load there would be counted with the
// copies or stores in flight (vmcnt completes in issue order) and wait for
// them.
//   backward: coefficient block (2 NP rows of NC slots) or staging of
//             256-order chunks (rows of 2 NC + 1 doubles), then the pass
//             areas and the output staging (NC + 1 slots per row), reused;
//   forward:  pass areas NC (NP + NP/8), then a stage of NP rows x NC slots;
//             the output staging (rows of 2 NC + 1 doubles) reuses both.
template <int NP, int NC>
constexpr int fft3_tw_off_bwd() {
  constexpr int a = 2 * NP * NC, b = (2 * NC + 1) * (NP < 256 ? NP : 256);
  constexpr int c = NC * (NP + NP / 8), d = (NC + 1) * NP;
  constexpr int ab = a > b ? a : b, cd = c > d ? c : d;
  return ab > cd ? ab : cd;
}
template <int NP, int NC>
constexpr int fft3_tw_off_fwd() {
  // with the copy stage past the pass areas, or (no in-kernel copy) the
  // global-load stage of NP rows x (NC + 1) slots over them
  constexpr int c = NC * (NP + NP / 8), d = (NC + 1) * NP;
  constexpr int a = fft_fwd_dma<NP>() ? c + NC * NP : (c > d ? c : d);
  constexpr int b = (2 * NC + 1) * (NP < 256 ? NP : 256);
  return a > b ? a : b;
}

// For NP <= 512 the pass area of a complex line belongs to the wave that
// transforms it (T = NP/8 <= 64 threads per line, lines never straddle a
// wave), and one wave's LDS operations complete in issue order: between the
// passes only the compiler has to keep the stores before the loads and the
// loads before the next stores.  Callers put a work-group barrier between
// the passes and any use of the area by other waves.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
// ordering between passes: wave-local when a complex line fits one wave,
// else a work-group barrier (every thread runs the passes)
template <int NP>
__device__ __forceinline__ void pass_lds_order() {
  if constexpr (NP / 8 <= 64)
    wave_lds_order();
  else
    __syncthreads();
}

// exp(-2 pi i k / MP) for k < MP as hi[k / 64] lo[k % 64] (MP / 64 + 64
// entries instead of MP): the chirp-z kernels keep their LDS for the lines
struct TwoLevelRoots {
  const double2 *hi, *lo;
  __device__ __forceinline__ double2 operator[](int k) const { return cmul(hi[k >> 6], lo[k & 63]); }
};

// tw: roots of unity in LDS, of order NP at stride ts (the NP-entry table
// with ts = 1).  A butterfly's twiddles
// w^r, r < R, are powers of its one root w: one table read and R - 2 complex
// products instead of R - 1 reads, whose strides (k r entries apart across
// the lanes) made the LDS serve them in two to four passes.
template <int NP, int Ns>
__device__ __forceinline__ void f3_passes(double2 (&v)[8], double2 *ln, int j,
                                          const double2 *__restrict__ tw, int ts, double s) {
  constexpr int T = NP / 8;
  constexpr int R = (NP / Ns >= 8) ? 8 : NP / Ns;
#pragma unroll
  for (int i = 0; i < 8 / R; i++) {
    const int b = j + T * i;
    if constexpr (Ns > 1) {
      const int k = b % Ns;
      double2 w1 = tw[k * (NP / (Ns * R)) * ts];  // exp(-2 pi i k / (Ns R))
      w1.y *= -s;
      double2 w = w1;
#pragma unroll
      for (int r = 1; r < R; r++) {
        v[i * R + r] = cmul(v[i * R + r], w);
        if (r + 1 < R) w = cmul(w, w1);
      }
    }
    dft_small<R>(v + i * R, s);
  }
  if constexpr (Ns * R < NP) {
#pragma unroll
    for (int i = 0; i < 8 / R; i++) {
      const int b = j + T * i;
#pragma unroll
      for (int r = 0; r < R; r++) ln[fpad((b / Ns) * Ns * R + b % Ns + Ns * r)] = v[i * R + r];
    }
    pass_lds_order<NP>();
    constexpr int R2 = (NP / (Ns * R) >= 8) ? 8 : NP / (Ns * R);
#pragma unroll
    for (int i = 0; i < 8 / R2; i++)
#pragma unroll
      for (int r = 0; r < R2; r++) v[i * R2 + r] = ln[fpad(j + T * i + (NP / R2) * r)];
    pass_lds_order<NP>();
    f3_passes<NP, Ns * R>(v, ln, j, tw, ts, s);
  }
}

// The same passes with the roots composed from two-level factors (order NP,
// stride 1): for the chirp-z kernels, whose LDS holds the lines instead of
// an NP-entry table.  A butterfly's twiddles w^r, r < R, are powers of its
// one root w (one table lookup, R - 2 complex products, error growing as r).
template <int NP, int Ns>
__device__ __forceinline__ void f3_passes_2l(double2 (&v)[8], double2 *ln, int j,
                                             const TwoLevelRoots &tw, double s) {
  constexpr int T = NP / 8;
  constexpr int R = (NP / Ns >= 8) ? 8 : NP / Ns;
#pragma unroll
  for (int i = 0; i < 8 / R; i++) {
    const int b = j + T * i;
    if constexpr (Ns > 1) {
      const int k = b % Ns;
      double2 w1 = tw[k * (NP / (Ns * R))];  // exp(-2 pi i k / (Ns R))
      w1.y *= -s;
      double2 w = w1;
#pragma unroll
      for (int r = 1; r < R; r++) {
        v[i * R + r] = cmul(v[i * R + r], w);
        if (r + 1 < R) w = cmul(w, w1);
      }
    }
    dft_small<R>(v + i * R, s);
  }
  if constexpr (Ns * R < NP) {
#pragma unroll
    for (int i = 0; i < 8 / R; i++) {
      const int b = j + T * i;
#pragma unroll
      for (int r = 0; r < R; r++) ln[fpad((b / Ns) * Ns * R + b % Ns + Ns * r)] = v[i * R + r];
    }
    pass_lds_order<NP>();
    constexpr int R2 = (NP / (Ns * R) >= 8) ? 8 : NP / (Ns * R);
#pragma unroll
    for (int i = 0; i < 8 / R2; i++)
#pragma unroll
      for (int r = 0; r < R2; r++) v[i * R2 + r] = ln[fpad(j + T * i + (NP / R2) * r)];
    pass_lds_order<NP>();
    f3_passes_2l<NP, Ns * R>(v, ln, j, tw, s);
  }
}

// radix of the last pass, and the register holding output slot q (index j + T q)
template <int NP, int Ns = 1>
constexpr int f3_last_radix() {
  constexpr int R = (NP / Ns >= 8) ? 8 : NP / Ns;
  if constexpr (Ns * R < NP)
    return f3_last_radix<NP, Ns * R>();
  else
    return R;
}
template <int NP>
__device__ __forceinline__ constexpr int f3_pos(int q) {
  constexpr int R = f3_last_radix<NP>();
  return (q % (8 / R)) * R + q / (8 / R);
}

// exp(-2 pi i k / N), k < N = R0 NP, as exp(-2 pi i a / NP) exp(-2 pi i b / N)
// with k = R0 a + b; a = k / R0 by a multiply-high with r0m = ceil(2^32 / R0)
// (exact for k < 2^32 / R0; r0m is unused for R0 = 1)
// (twp: the NP roots as a table or as two-level factors)
template <class TP>
__device__ __forceinline__ double2 tw_n(const TP &twp, const double2 *twr, int k, int R0,
                                        unsigned r0m) {
  const int a = R0 == 1 ? k : (int)__umulhi((unsigned)k, r0m);
  return cmul(twp[a], twr[k - a * R0]);
}

// LDS DMA of 16 bytes per lane (global_load_lds_dwordx4) from a wave-uniform
// base plus a 32-bit lane offset, as inline assembly: the compiler would wait
// for a builtin DMA before every later LDS access, which would serialise the
// copy with the FFT passes it is meant to overlap.  The statement carries its
// own wait states (the compiler's hazard recognizer does not look inside):
// s_nop 4 covers both an SALU write of M0 and a VALU write of the base SGPRs.
__device__ __forceinline__ void fft_dma16(const void *base, unsigned voff, unsigned lds_byte) {
  const unsigned long long b = (unsigned long long)base;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  const void *ub = (const void *)(((unsigned long long)hi << 32) | lo);
  asm volatile("s_nop 4\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(voff), "s"(ub),
               "{m0}"(__builtin_amdgcn_readfirstlane(lds_byte))
               : "memory");
}

__device__ __forceinline__ double2 cfma(double2 a, double2 w, double2 acc) {
  return make_double2(acc.x + a.x * w.x - a.y * w.y, acc.y + a.x * w.y + a.y * w.x);
}

template <int NP, int NC>
__global__ void __launch_bounds__(NC * NP / 8, 2) k_fft3_bwd(int N, int R0, unsigned r0m, int L,
                                                          long long nline,
                                                          const double *__restrict__ X,
                                                          FourierAddr xa, double *__restrict__ v,
                                                          const double2 *__restrict__ twn) {
  extern __shared__ double2 lds[];
  constexpr int T = NP / 8, NT = NC * T, W = 2 * NC, SS = W + 1;
  const int tid = threadIdx.x, cl = tid / T, j = tid % T;
  const int nd = blockIdx.y;
  const long long line0 = (long long)xcd_block(blockIdx.x, gridDim.x) * W;
  double *stage = reinterpret_cast<double *>(lds);
  constexpr bool TL = fft_bwd_2l<NP>();
  constexpr int NTW = fft_bwd_tw_entries<NP>();
  double2 *twp = lds + fft3_tw_off_bwd<NP, NC>(), *twr = twp + NTW;
  // the NP roots exp(-2 pi i a / NP) (as hi[a / 64] lo[a % 64] when TL),
  // then the R0 roots; read after the staging barrier below
  for (int i = tid; i < NTW + R0; i += NT) {
    const int src = i >= NTW ? i - NTW
                    : !TL    ? R0 * i
                    : i < NP / 64 ? R0 * 64 * i : R0 * (i - NP / 64);
    twp[i] = twn[src];
  }
  const TwoLevelRoots tl{twp, twp + NP / 64};
  double2 zlo[8], zhi[8];
#pragma unroll
  for (int q = 0; q < 8; q++) zlo[q] = zhi[q] = make_double2(0.0, 0.0);
  const long long xbase = (long long)nd * xa.s_nd;
  if (xa.s_line == 1 && (nline & 1) == 0 && (xa.s_nd & 1) == 0 && (xa.q_stride & 1) == 0 &&
      ((unsigned long long)X & 15) == 0) {
    // Lines contiguous, rows 16-byte aligned (even strides, aligned base):
    // the whole coefficient block (rows q = 0 .. 2L of the W lines) goes
    // straight to LDS with global_load_lds_dwordx4, one wait for all of it.
    // Row r's 16-byte slot u holds lines 2 (u ^ sw(r)) and +1, sw(r) =
    // (r >> 1) & (NC - 1), so the reads below (rows 2f - 1, 2f of
    // consecutive orders f) spread over the slots.  Rows are padded to 2 NP;
    // lines past the end read a valid pair (their results are never stored).
    double2 *st2 = lds;
    const int nrow = 2 * L + 1;
    for (int i0 = tid & ~63; i0 < 2 * NP * NC; i0 += NT) {
      const int i = i0 + (tid & 63);
      const int r = i / NC, u = i % NC;
      const int c = u ^ ((r >> 1) & (NC - 1));
      long long line = line0 + 2 * c;
      if (line + 1 >= nline) line = line0;
      const double *src = X + q_offset(xa, r < nrow ? r : 0) + xbase + line;
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                       (__attribute__((address_space(3))) void *)(st2 + i0), 16,
                                       0, 0);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    auto row2 = [&](int r) { return st2[r * NC + (cl ^ ((r >> 1) & (NC - 1)))]; };
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int f = j + T * q;  // order of Z_f; Z_{f+N-NP} comes from order NP - f
      if (f <= L) {
        if (f == 0) {
          zlo[q] = row2(0);
        } else {
          const double2 a = row2(2 * f - 1), b = row2(2 * f);
          zlo[q] = make_double2(0.5 * (a.x + b.y), 0.5 * (a.y - b.x));
        }
      }
      const int g = NP - f;
      if (g <= L && g >= 1) {
        const double2 a = row2(2 * g - 1), b = row2(2 * g);
        zhi[q] = make_double2(0.5 * (a.x - b.y), 0.5 * (a.y + b.x));
      }
    }
    __syncthreads();
  } else
  // coefficients in chunks of 256 orders (rows 2 f0 - 1 .. 2 f0 + 510)
  for (int f0 = 0; f0 <= L; f0 += 256) {
    const int f1 = min(f0 + 256, L + 1);
    const int q0 = f0 == 0 ? 0 : 2 * f0 - 1, nrow = 2 * (f1 - 1) - q0 + 1;
    // batches of 8 independent loads per thread, then the LDS stores
    for (int base = 0; base < nrow * W; base += 8 * NT) {
      // unconditional loads (clamped to a valid element, zeroed after) so
      // the eight stay in flight together
      double val[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int idx = base + tid + u * NT;
        const long long line = line0 + (idx % W);
        const bool ok = idx < nrow * W && line < nline;
        const int row = ok ? idx / W : 0;
        val[u] = X[q_offset(xa, q0 + row) + xbase + (ok ? line : line0) * xa.s_line];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int idx = base + tid + u * NT;
        if (idx < nrow * W)
          stage[(idx / W) * SS + (idx % W)] = line0 + (idx % W) < nline ? val[u] : 0.0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int f = j + T * q;  // order of Z_f; Z_{f+N-NP} comes from order NP - f
      if (f >= f0 && f < f1) {
        const double *p = stage + ((f == 0 ? 0 : 2 * f - 1) - q0) * SS + 2 * cl;
        if (f == 0) {
          zlo[q] = make_double2(p[0], p[1]);
        } else {
          const double aa = p[0], ab = p[1], ba = p[SS], bb = p[SS + 1];
          zlo[q] = make_double2(0.5 * (aa + bb), 0.5 * (ab - ba));
        }
      }
      const int g = NP - f;
      if (g >= f0 && g < f1 && g >= 1) {
        const double *p = stage + (2 * g - 1 - q0) * SS + 2 * cl;
        const double aa = p[0], ab = p[1], ba = p[SS], bb = p[SS + 1];
        zhi[q] = make_double2(0.5 * (aa - bb), 0.5 * (ab + ba));
      }
    }
    __syncthreads();
  }
  double2 *ln = lds + cl * (NP + NP / 8);
  const long long nnod = nline * N;
  const bool vec2 = (nline & 1) == 0;
  for (int j1 = 0; j1 < R0; j1++) {
    double2 w[8];
    // e^{+2 pi i m j1 / N} for m = j + T q as e^{+2 pi i j j1 / N} times the
    // q-th power of e^{+2 pi i T j1 / N}, and the alias m + N - NP one factor
    // e^{-2 pi i NP j1 / N} further
    auto root_n = [&](int k) {  // exp(-2 pi i k / N)
      if constexpr (TL)
        return tw_n(tl, twr, k, R0, r0m);
      else
        return tw_n(static_cast<const double2 *>(twp), twr, k, R0, r0m);
    };
    double2 t1 = root_n(j * j1);
    double2 st = root_n(T * j1);
    double2 hf = j1 == 0 ? make_double2(1.0, 0.0) : root_n(N - NP * j1);
    t1.y = -t1.y;  // e^{+...}
    st.y = -st.y;
    hf.y = -hf.y;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      w[q] = cfma(zhi[q], cmul(t1, hf), cmul(zlo[q], t1));
      if (q < 7) t1 = cmul(t1, st);
    }
    if constexpr (TL)
      f3_passes_2l<NP, 1>(w, ln, j, tl, 1.0);
    else
      f3_passes<NP, 1>(w, ln, j, twp, 1, 1.0);
    __syncthreads();  // every wave done with its pass area (under the staging)
    // rows phi = R0 (j + T q) + j1, staged as [j2][NC complex lines] (stride NC + 1)
#pragma unroll
    for (int q = 0; q < 8; q++) lds[(j + T * q) * (NC + 1) + cl] = w[f3_pos<NP>(q)];
    __syncthreads();
    for (int idx = tid; idx < NP * NC; idx += NT) {
      const int j2 = idx / NC, c = idx % NC;
      const long long la = line0 + 2 * c;
      const double2 z = lds[j2 * (NC + 1) + c];
      double *o = v + (long long)nd * nnod + (long long)(R0 * j2 + j1) * nline + la;
      if (la + 1 < nline) {
        if (vec2)
          *reinterpret_cast<double2 *>(o) = z;
        else {
          o[0] = z.x;
          o[1] = z.y;
        }
      } else if (la < nline) {
        o[0] = z.x;
      }
    }
    __syncthreads();
  }
}

// Forward accumulation of decimation j1: e^{-2 pi i k j1 / N} for k = j + T q
// as e^{-2 pi i j j1 / N} times the q-th power of e^{-2 pi i T j1 / N} (three
// table reads per thread instead of sixteen), the alias k + N - NP one factor
// e^{-2 pi i (N - NP j1) / N} further, each from the NP roots and the R0
// roots (tw_n).
template <int NP>
__device__ __forceinline__ void f3_fwd_alias(const double2 (&w)[8], double2 (&zlo)[8],
                                             double2 (&zhi)[8], const double2 *twp,
                                             const double2 *twr, int R0, unsigned r0m, int j,
                                             int j1, int N, int L) {
  constexpr int T = NP / 8;
  double2 t = tw_n(twp, twr, j * j1, R0, r0m);
  const double2 st = tw_n(twp, twr, T * j1, R0, r0m);
  const double2 hf = j1 == 0 ? make_double2(1.0, 0.0) : tw_n(twp, twr, N - NP * j1, R0, r0m);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int k = j + T * q;
    const double2 y = w[f3_pos<NP>(q)];
    if (k <= L) zlo[q] = cfma(y, t, zlo[q]);
    if (k >= NP - L) zhi[q] = cfma(y, cmul(t, hf), zhi[q]);
    if (q < 7) t = cmul(t, st);
  }
}

template <int NP, int NC>
__global__ void __launch_bounds__(NC * NP / 8, NC * NP / 8 <= 256 ? 2 : 1)
    k_fft3_fwd(int N, int R0, unsigned r0m, int L, long long nline, const double *__restrict__ v,
               double *__restrict__ X, FourierAddr xa, const double2 *__restrict__ twn) {
  extern __shared__ double2 lds[];
  constexpr int T = NP / 8, NT = NC * T, W = 2 * NC, SS = W + 1;
  const int tid = threadIdx.x, cl = tid / T, j = tid % T;
  const int nd = blockIdx.y;
  const long long line0 = (long long)xcd_block(blockIdx.x, gridDim.x) * W;
  const long long nnod = nline * N;
  const bool vec2 = (nline & 1) == 0;
  double2 zlo[8], zhi[8];
#pragma unroll
  for (int q = 0; q < 8; q++) zlo[q] = zhi[q] = make_double2(0.0, 0.0);
  double2 *ln = lds + cl * (NP + NP / 8);
  // NP roots of order NP, then the R0 roots exp(-2 pi i b / N), b < R0, as
  // in the backward; read after the first barrier below
  double2 *twp = lds + fft3_tw_off_fwd<NP, NC>(), *twr = twp + NP;
  for (int i = tid; i < NP + R0; i += NT) twp[i] = twn[i < NP ? R0 * i : i - NP];
  // DMA form (lines contiguous in pairs, aligned base, per-component byte
  // offsets within 32 bits): the rows of the next phi decimation j1 + 1 are
  // copied global -> LDS while the FFT of j1 runs, into a stage past the
  // pass areas; row r's slot u holds lines 2 (u ^ (r & (NC - 1))) and +1
  // (XOR swizzle: consecutive threads of a complex line read distinct slots).
  // the copies address each wave instruction's 64 / NC rows from a base of
  // their own, so a lane's byte offset stays below R0 (64 / NC) nline 8
  const bool dma = fft_fwd_dma<NP>() && vec2 && (((unsigned long long)v & 15) == 0) &&
                   (long long)R0 * (64 / NC) * nline * 8 < (1ll << 32);
  if (dma) {
    double2 *st2 = lds + NC * (NP + NP / 8);
    const unsigned st2_byte =
        (unsigned)(size_t)(const __attribute__((address_space(3))) void *)lds +
        16u * (NC * (NP + NP / 8));
    auto issue = [&](int jj) {
      const double *base = v + (long long)nd * nnod + (long long)jj * nline + line0;
#pragma unroll
      for (int i0 = tid & ~63; i0 < NC * NP; i0 += NT) {
        const int i = i0 + (tid & 63), r = i / NC, u = i % NC, c = u ^ (r & (NC - 1));
        const long long cc = line0 + 2 * c + 1 < nline ? 2 * c : 0;  // past the end: line0
        const int r0w = i0 / NC;  // first row of this wave instruction (uniform)
        const unsigned off = (unsigned)(((long long)R0 * (r - r0w) * nline + cc) * 8);
        fft_dma16(base + (long long)R0 * r0w * nline, off, st2_byte + 16u * (unsigned)i0);
      }
    };
    issue(0);
    for (int j1 = 0; j1 < R0; j1++) {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      double2 w[8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int r = j + T * q;
        w[q] = st2[r * NC + (cl ^ (r & (NC - 1)))];
      }
      __syncthreads();
      if (j1 + 1 < R0) issue(j1 + 1);
      f3_passes<NP, 1>(w, ln, j, twp, 1, -1.0);
      f3_fwd_alias<NP>(w, zlo, zhi, twp, twr, R0, r0m, j, j1, N, L);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  } else
  for (int j1 = 0; j1 < R0; j1++) {
    __syncthreads();  // every wave done with its pass area (under the stage)
    {
      // eight independent loads per thread, then the LDS stores; the loads
      // are unconditional (lines clamped into range, zeroed after) so they
      // stay in flight together
      double2 z[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int idx = tid + u * NT, j2 = idx / NC, c = idx % NC;
        const long long la = line0 + 2 * c;
        const double *row = v + (long long)nd * nnod + (long long)(R0 * j2 + j1) * nline;
        if (vec2) {
          // nline even: la and nline - 2 are even, so the pair is aligned
          z[u] = *reinterpret_cast<const double2 *>(row + (la + 1 < nline ? la : nline - 2));
        } else {
          z[u] = make_double2(row[la < nline ? la : nline - 1],
                              row[la + 1 < nline ? la + 1 : nline - 1]);
        }
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int idx = tid + u * NT;
        const long long la = line0 + 2 * (idx % NC);
        lds[(idx / NC) * (NC + 1) + (idx % NC)] =
            make_double2(la < nline ? z[u].x : 0.0, la + 1 < nline ? z[u].y : 0.0);
      }
    }
    __syncthreads();
    double2 w[8];
#pragma unroll
    for (int q = 0; q < 8; q++) w[q] = lds[(j + T * q) * (NC + 1) + cl];
    __syncthreads();
    f3_passes<NP, 1>(w, ln, j, twp, 1, -1.0);
    f3_fwd_alias<NP>(w, zlo, zhi, twp, twr, R0, r0m, j, j1, N, L);
  }
  // Z_{N-m} for m = j + T q sits at index NP - m of the lane (T - j) % T of
  // this complex line, in register 7 - q (or (8 - q) % 8 when j = 0)
  double2 zn[8];
  if constexpr (T <= 64) {
    const int lane = threadIdx.x & 63;
    const int partner = lane - j + (T - j) % T;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const double2 send = (j == 0) ? zhi[(8 - q) % 8] : zhi[7 - q];
      zn[q].x = __shfl(send.x, partner, 64);
      zn[q].y = __shfl(send.y, partner, 64);
    }
  } else {
    // the line spans waves: through LDS (index k = j + T q of the line's
    // pass area holds zhi[q])
    __syncthreads();  // every wave done with the pass areas
#pragma unroll
    for (int q = 0; q < 8; q++) ln[j + T * q] = zhi[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) zn[q] = ln[(NP - (j + T * q)) & (NP - 1)];
  }
  const double sn = 1.0 / (double)N, tsn = 2.0 / (double)N;
  double *stage = reinterpret_cast<double *>(lds);
  const long long xbase = (long long)nd * xa.s_nd;
  for (int f0 = 0; f0 <= L; f0 += 256) {
    const int f1 = min(f0 + 256, L + 1);
    const int q0 = f0 == 0 ? 0 : 2 * f0 - 1, nrow = 2 * (f1 - 1) - q0 + 1;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int m = j + T * q;
      if (m >= f0 && m < f1) {
        const double2 zm = zlo[q], zc = (m == 0) ? zlo[q] : zn[q];
        // X_a = (Z_m + conj Z_{N-m})/2, X_b = -i (Z_m - conj Z_{N-m})/2
        const double xar = 0.5 * (zm.x + zc.x), xai = 0.5 * (zm.y - zc.y);
        const double xbr = 0.5 * (zm.y + zc.y), xbi = -0.5 * (zm.x - zc.x);
        if (m == 0) {
          double *p = stage + 2 * cl;
          p[0] = sn * xar;
          p[1] = sn * xbr;
        } else {
          double *p = stage + (2 * m - 1 - q0) * SS + 2 * cl;
          p[0] = tsn * xar;
          p[1] = tsn * xbr;
          p[SS] = -tsn * xai;
          p[SS + 1] = -tsn * xbi;
        }
      }
    }
    __syncthreads();
    for (int idx = tid; idx < nrow * W; idx += NT) {
      const int row = idx / W, lw = idx % W;
      const long long line = line0 + lw;
      if (line < nline)
        X[q_offset(xa, q0 + row) + xbase + line * xa.s_line] = stage[row * SS + lw];
    }
  }
}

// ---------------------------------------------------------------------------
// Chirp-z form (plan variant fft_form 3): the whole line as one Bluestein
// convolution of length MP, a power of two >= N + 2L, so that any Nphi, with
// any prime factors, costs two MP-point FFTs on the register / LDS passes of
// the split form.  With w_k = exp(i pi k^2 / N) and 2 m j = m^2 + j^2 -
// (j - m)^2:
//   backward x_j = sum_{|m|<=L} Z_m e^{+2 pi i m j/N}
//              = w_j sum_m (Z_m w_m) conj(w_{j-m}),            j < N,
//   forward  Z_q = sum_{j<N} x_j e^{-2 pi i q j/N}
//              = conj(w_q) sum_j (x_j conj(w_j)) w_{q-j},      |q| <= L.
// Each sum is a linear convolution whose kernel spans N + 2L consecutive
// indices (j - m in [-L, N-1+L], q - j in [-(N-1)-L, L]); with every index
// taken mod MP >= N + 2L no two of them collide, so the cyclic convolution
//   y = IDFT_MP(DFT_MP(u) . DFT_MP(h))
// reproduces it at the wanted outputs.  DFT_MP(h) / MP is precomputed per
// direction in long double.  Table (device): the roots exp(-2 pi i k / MP)
// as MP/64 + 64 two-level factors, the backward and the forward filter (MP
// entries each), then w_k, k < N.  The two real lines of a complex line pack
// as in the other forms.
//
// A work-group owns NC complex lines (NC = 2, one at MP = 2048: measured
// faster than 4 and 2, the work-groups being barrier-bound), T = MP/8 threads
// per line, the root factors in LDS (the filters are read from L2, once per
// element and line), and walks a range of line blocks (persistent).
template <int MP>
constexpr int czt_nc() { return MP >= 2048 ? 1 : 2; }
template <int MP>
constexpr int czt_hb() { return MP / 64 + 64; }  // table offsets: backward filter,
template <int MP>
constexpr int czt_hf() { return czt_hb<MP>() + MP; }  // forward filter,
template <int MP>
constexpr int czt_ch() { return czt_hf<MP>() + MP; }  // chirp w_k
// LDS, in double2 units: the pass areas NC (MP + MP/8), which also stage the
// grid rows and Fourier rows read and the grid rows written; in the forward
// then the staging of a 256-order chunk of Fourier rows written (2 NC + 1
// doubles per row: the pass areas still hold the line's Z values); then the
// root factors.
template <int MP, int NC>
constexpr int czt_pa() { return NC * (MP + MP / 8); }
template <int MP, int NC>
constexpr int czt_st() { return 256 * (2 * NC + 1); }
template <int MP>
constexpr int czt_lds_bytes(bool fwd) {
  return (int)sizeof(double2) *
         (czt_pa<MP, czt_nc<MP>()>() + (fwd ? czt_st<MP, czt_nc<MP>()>() : 0) + czt_hb<MP>());
}

// The 8 XCDs each take a contiguous range of line blocks; the G = gridDim.x/8
// work-groups of an XCD (dispatch deals them round-robin to the XCDs) take
// consecutive blocks of its range, so the blocks in flight on one XCD are
// neighbours and share the 128-byte rows of the line-contiguous layouts in
// its L2.  gridDim.x is a multiple of 8.
__device__ __forceinline__ void czt_range(long long nblk, long long &b, long long &bend,
                                          long long &step) {
  const long long x = blockIdx.x % 8, i = blockIdx.x / 8;
  step = gridDim.x / 8;
  b = x * nblk / 8 + i;
  bend = (x + 1) * nblk / 8;
}

__device__ __forceinline__ double2 conjd(double2 a) { return make_double2(a.x, -a.y); }

template <int MP, int NC>
__global__ void __launch_bounds__(NC * MP / 8, 2) k_fftz_bwd(int N, int L, long long nline,
                                                          long long nblk,
                                                          const double *__restrict__ X,
                                                          FourierAddr xa, double *__restrict__ v,
                                                          const double2 *__restrict__ zt) {
  extern __shared__ double2 lds[];
  constexpr int T = MP / 8, NT = NC * T, W = 2 * NC, SS = W + 1;
  // (2L+1)(2NC+1) doubles of Fourier rows fit the pass areas: 2L + 1 <=
  // (MP + 1) / 2 since Nphi >= 2L + 2
  constexpr int OFF = czt_pa<MP, NC>(), CR = OFF / (NC + 1);
  const int tid = threadIdx.x, cl = tid / T, j = tid % T;
  const int nd = blockIdx.y;
  double2 *rts = lds + OFF;
  for (int i = tid; i < czt_hb<MP>(); i += NT) rts[i] = zt[i];
  const TwoLevelRoots twp{rts, rts + MP / 64};
  const double2 *hat = zt + czt_hb<MP>(), *chirp = zt + czt_ch<MP>();
  double *stage = reinterpret_cast<double *>(lds);
  double2 *ln = lds + cl * (MP + MP / 8);
  const long long nnod = nline * N;
  const bool vec2 = (nline & 1) == 0;
  const long long xbase = (long long)nd * xa.s_nd;
  const int nrow = 2 * L + 1;
  long long b, bend, step;
  czt_range(nblk, b, bend, step);
  for (; b < bend; b += step) {
    const long long line0 = b * W;
    __syncthreads();  // root factors stored; the previous block's staging read
    // rows q = 0 .. 2L of the W lines: batches of 8 independent loads per
    // thread (clamped into range, zeroed after), then the LDS stores
    for (int base = 0; base < nrow * W; base += 8 * NT) {
      double val[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int idx = base + tid + u * NT;
        const long long line = line0 + (idx % W);
        const bool ok = idx < nrow * W && line < nline;
        const int row = ok ? idx / W : 0;
        val[u] = X[q_offset(xa, row) + xbase + (ok ? line : line0) * xa.s_line];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int idx = base + tid + u * NT;
        if (idx < nrow * W)
          stage[(idx / W) * SS + (idx % W)] = line0 + (idx % W) < nline ? val[u] : 0.0;
      }
    }
    __syncthreads();
    // u at index k = j + T q: Z_k w_k (k <= L), Z_{N-m} w_m at k = MP - m
    // (1 <= m <= L), zero between
    double2 w[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = j + T * q;
      double2 z = make_double2(0.0, 0.0);
      if (k <= L) {
        const double *p = stage + (k == 0 ? 0 : 2 * k - 1) * SS + 2 * cl;
        if (k == 0) {
          z = make_double2(p[0], p[1]);
        } else {
          const double aa = p[0], ab = p[1], ba = p[SS], bb = p[SS + 1];
          z = make_double2(0.5 * (aa + bb), 0.5 * (ab - ba));
        }
        z = cmul(z, chirp[k]);
      } else if (k >= MP - L) {
        const int m = MP - k;
        const double *p = stage + (2 * m - 1) * SS + 2 * cl;
        const double aa = p[0], ab = p[1], ba = p[SS], bb = p[SS + 1];
        z = cmul(make_double2(0.5 * (aa - bb), 0.5 * (ab + ba)), chirp[m]);
      }
      w[q] = z;
    }
    __syncthreads();  // staging read before the passes overwrite it
    f3_passes_2l<MP, 1>(w, ln, j, twp, -1.0);
    double2 y[8];
#pragma unroll
    for (int q = 0; q < 8; q++) y[q] = cmul(w[f3_pos<MP>(q)], hat[j + T * q]);
    f3_passes_2l<MP, 1>(y, ln, j, twp, 1.0);
    // x_i = w_i y_i for i = j + T q < N (register f3_pos(q)), staged in
    // chunks of CR rows as [i][NC complex lines] (stride NC + 1)
    for (int i0 = 0; i0 < N; i0 += CR) {
      __syncthreads();  // pass areas / the previous chunk free
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int i = j + T * q;
        if (i >= i0 && i < i0 + CR && i < N)
          lds[(i - i0) * (NC + 1) + cl] = cmul(y[f3_pos<MP>(q)], chirp[i]);
      }
      __syncthreads();
      const int nr = min(CR, N - i0);
      for (int idx = tid; idx < nr * NC; idx += NT) {
        const int i = idx / NC, c = idx % NC;
        const long long la = line0 + 2 * c;
        const double2 z = lds[i * (NC + 1) + c];
        double *o = v + (long long)nd * nnod + (long long)(i0 + i) * nline + la;
        if (la + 1 < nline) {
          if (vec2)
            *reinterpret_cast<double2 *>(o) = z;
          else {
            o[0] = z.x;
            o[1] = z.y;
          }
        } else if (la < nline) {
          o[0] = z.x;
        }
      }
    }
  }
}

template <int MP, int NC>
__global__ void __launch_bounds__(NC * MP / 8, 2) k_fftz_fwd(int N, int L, long long nline,
                                                          long long nblk,
                                                          const double *__restrict__ v,
                                                          double *__restrict__ X, FourierAddr xa,
                                                          const double2 *__restrict__ zt) {
  extern __shared__ double2 lds[];
  constexpr int T = MP / 8, NT = NC * T, W = 2 * NC, SS = W + 1;
  constexpr int OFF = czt_pa<MP, NC>(), CR = OFF / (NC + 1);
  const int tid = threadIdx.x, cl = tid / T, j = tid % T;
  const int nd = blockIdx.y;
  double2 *rts = lds + OFF + czt_st<MP, NC>();
  for (int i = tid; i < czt_hb<MP>(); i += NT) rts[i] = zt[i];
  const TwoLevelRoots twp{rts, rts + MP / 64};
  const double2 *hat = zt + czt_hf<MP>(), *chirp = zt + czt_ch<MP>();
  double *stage = reinterpret_cast<double *>(lds + OFF);
  double2 *ln = lds + cl * (MP + MP / 8);
  const long long nnod = nline * N;
  const bool vec2 = (nline & 1) == 0;
  const long long xbase = (long long)nd * xa.s_nd;
  const double sn = 1.0 / (double)N, tsn = 2.0 / (double)N;
  long long b, bend, step;
  czt_range(nblk, b, bend, step);
  for (; b < bend; b += step) {
    const long long line0 = b * W;
    // u_i = x_i conj(w_i), i < N, zero up to MP; rows staged in chunks of CR
    double2 w[8];
#pragma unroll
    for (int q = 0; q < 8; q++) w[q] = make_double2(0.0, 0.0);
    for (int i0 = 0; i0 < N; i0 += CR) {
      __syncthreads();  // root factors stored; the previous chunk / block read
      const int nr = min(CR, N - i0);
      for (int base = 0; base < nr * NC; base += 8 * NT) {
        // eight independent loads per thread (lines clamped into range,
        // zeroed after), then the LDS stores
        double2 z[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int idx = base + tid + u * NT, c = idx % NC;
          const int i = idx < nr * NC ? idx / NC : 0;
          const long long la = line0 + 2 * c;
          const double *row = v + (long long)nd * nnod + (long long)(i0 + i) * nline;
          if (vec2)
            z[u] = *reinterpret_cast<const double2 *>(row + (la + 1 < nline ? la : nline - 2));
          else
            z[u] = make_double2(row[la < nline ? la : nline - 1],
                                row[la + 1 < nline ? la + 1 : nline - 1]);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int idx = base + tid + u * NT;
          if (idx < nr * NC) {
            const long long la = line0 + 2 * (idx % NC);
            lds[(idx / NC) * (NC + 1) + (idx % NC)] =
                make_double2(la < nline ? z[u].x : 0.0, la + 1 < nline ? z[u].y : 0.0);
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int i = j + T * q;
        if (i >= i0 && i < i0 + nr) w[q] = cmul(lds[(i - i0) * (NC + 1) + cl], conjd(chirp[i]));
      }
    }
    __syncthreads();  // staging read before the passes overwrite it
    f3_passes_2l<MP, 1>(w, ln, j, twp, -1.0);
    double2 y[8];
#pragma unroll
    for (int q = 0; q < 8; q++) y[q] = cmul(w[f3_pos<MP>(q)], hat[j + T * q]);
    f3_passes_2l<MP, 1>(y, ln, j, twp, 1.0);
    // Z_m = conj(w_m) y_m (m <= L) and Z_{N-m} = conj(w_m) y_{MP-m}
    // (1 <= m <= L), into the line's pass area at indices m and MP - m (the
    // last pass's reads are ordered before these stores)
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = j + T * q;
      const double2 yk = y[f3_pos<MP>(q)];
      if (k <= L)
        ln[k] = cmul(yk, conjd(chirp[k]));
      else if (k >= MP - L)
        ln[k] = cmul(yk, conjd(chirp[MP - k]));
    }
    pass_lds_order<MP>();
    for (int f0 = 0; f0 <= L; f0 += 256) {
      const int f1 = min(f0 + 256, L + 1);
      const int q0 = f0 == 0 ? 0 : 2 * f0 - 1, nrow = 2 * (f1 - 1) - q0 + 1;
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int m = j + T * q;
        if (m >= f0 && m < f1) {
          const double2 zm = ln[m], zc = ln[(MP - m) & (MP - 1)];
          // X_a = (Z_m + conj Z_{N-m})/2, X_b = -i (Z_m - conj Z_{N-m})/2
          const double xar = 0.5 * (zm.x + zc.x), xai = 0.5 * (zm.y - zc.y);
          const double xbr = 0.5 * (zm.y + zc.y), xbi = -0.5 * (zm.x - zc.x);
          if (m == 0) {
            double *p = stage + 2 * cl;
            p[0] = sn * xar;
            p[1] = sn * xbr;
          } else {
            double *p = stage + (2 * m - 1 - q0) * SS + 2 * cl;
            p[0] = tsn * xar;
            p[1] = tsn * xbr;
            p[SS] = -tsn * xai;
            p[SS + 1] = -tsn * xbi;
          }
        }
      }
      __syncthreads();
      for (int idx = tid; idx < nrow * W; idx += NT) {
        const int row = idx / W, lw = idx % W;
        const long long line = line0 + lw;
        if (line < nline)
          X[q_offset(xa, q0 + row) + xbase + line * xa.s_line] = stage[row * SS + lw];
      }
    }
  }
}

// in-place radix-2 FFT in long double, sign -1 (exp(-2 pi i ...)); n a power of two
static void fft_ld(std::vector<long double> &re, std::vector<long double> &im) {
  const int n = (int)re.size();
  for (int i = 1, jj = 0; i < n; i++) {
    int bit = n >> 1;
    for (; jj & bit; bit >>= 1) jj ^= bit;
    jj ^= bit;
    if (i < jj) {
      std::swap(re[i], re[jj]);
      std::swap(im[i], im[jj]);
    }
  }
  const long double kPi = 3.14159265358979323846264338327950288L;
  for (int len = 2; len <= n; len <<= 1) {
    for (int i = 0; i < n; i += len)
      for (int k = 0; k < len / 2; k++) {
        const long double a = -2.0L * kPi * (long double)k / (long double)len;
        const long double c = cosl(a), s = sinl(a);
        const int u = i + k, t = i + k + len / 2;
        const long double xr = re[t] * c - im[t] * s, xi = re[t] * s + im[t] * c;
        re[t] = re[u] - xr;
        im[t] = im[u] - xi;
        re[u] += xr;
        im[u] += xi;
      }
  }
}

int czt_init(sht_plan *p) {
  FftPlan &fp = p->fft;
  if (fp.zm) return SHT_OK;
  const int N = p->nphi, L = p->L;
  int MP = 256;
  while (MP < N + 2 * L) MP *= 2;
  if (MP > 2048) {
    set_error("chirp-z form: Nphi + 2L = %d exceeds 2048", N + 2 * L);
    return SHT_ERR_UNSUPPORTED;
  }
  const long double kPi = 3.14159265358979323846264338327950288L;
  // w_k = exp(i pi k^2 / N), from k^2 mod 2N (exact)
  auto chirp = [&](long long k, long double &c, long double &s) {
    const long long r = (k * k) % (2LL * N);
    const long double a = kPi * (long double)r / (long double)N;
    c = cosl(a);
    s = sinl(a);
  };
  const int hb = MP / 64 + 64, hf = hb + MP, ch = hf + MP;
  std::vector<double2> tab((size_t)ch + N);
  auto root = [&](int k) {  // exp(-2 pi i k / MP)
    const long double a = -2.0L * kPi * (long double)k / (long double)MP;
    return make_double2((double)cosl(a), (double)sinl(a));
  };
  for (int a = 0; a < MP / 64; a++) tab[a] = root(64 * a);
  for (int r = 0; r < 64; r++) tab[MP / 64 + r] = root(r);
  // backward kernel conj(w_k), k in [-L, N-1+L]; forward kernel w_k, k in
  // [-(N-1)-L, L]; each at k mod MP; transformed and scaled by 1/MP
  for (int dir = 0; dir < 2; dir++) {
    std::vector<long double> re(MP, 0.0L), im(MP, 0.0L);
    const int lo = dir == 0 ? -L : -(N - 1) - L, hi = dir == 0 ? N - 1 + L : L;
    for (int k = lo; k <= hi; k++) {
      long double c, s;
      chirp(k < 0 ? -k : k, c, s);
      const int i = ((k % MP) + MP) % MP;
      re[i] = c;
      im[i] = dir == 0 ? -s : s;
    }
    fft_ld(re, im);
    for (int k = 0; k < MP; k++)
      tab[(size_t)(dir == 0 ? hb : hf) + k] = make_double2((double)(re[k] / MP), (double)(im[k] / MP));
  }
  for (int k = 0; k < N; k++) {
    long double c, s;
    chirp(k, c, s);
    tab[(size_t)ch + k] = make_double2((double)c, (double)s);
  }
  double2 *d = nullptr;
  if (hipMalloc(&d, sizeof(double2) * tab.size()) != hipSuccess) {
    set_error("chirp-z form: table of %zu bytes: out of device memory", sizeof(double2) * tab.size());
    return SHT_ERR_OOM;
  }
  p->t.ztab = d;  // freed with the plan
  p->device_bytes += (long long)(sizeof(double2) * tab.size());
  KSHT_HIP(hipMemcpy(d, tab.data(), sizeof(double2) * tab.size(), hipMemcpyHostToDevice));
#define KSHT_CZT_ATTR(M)                                                                     \
  case M:                                                                                  \
    KSHT_HIP(hipFuncSetAttribute((const void *)k_fftz_bwd<M, czt_nc<M>()>,                   \
                                 hipFuncAttributeMaxDynamicSharedMemorySize,               \
                                 czt_lds_bytes<M>(false)));                                \
    KSHT_HIP(hipFuncSetAttribute((const void *)k_fftz_fwd<M, czt_nc<M>()>,                   \
                                 hipFuncAttributeMaxDynamicSharedMemorySize,               \
                                 czt_lds_bytes<M>(true)));                                 \
    break;
  switch (MP) {
    KSHT_CZT_ATTR(256)
    KSHT_CZT_ATTR(512)
    KSHT_CZT_ATTR(1024)
    KSHT_CZT_ATTR(2048)
    default: break;
  }
#undef KSHT_CZT_ATTR
  fp.zm = MP;
  return SHT_OK;
}

// work-groups per launch: 8 x G, about two rounds of what the CUs hold
static unsigned czt_grid_x(long long nblk, int ncomp, int lds_bytes) {
  const int per_cu = lds_bytes > 80 * 1024 ? 1 : lds_bytes > 53 * 1024 ? 2 : 3;
  long long g = (2LL * 256 * per_cu) / (8LL * (ncomp > 0 ? ncomp : 1));
  if (g < 1) g = 1;
  const long long need = (nblk + 7) / 8;
  if (g > need) g = need;
  return (unsigned)(8 * g);
}

static int launch_czt(sht_plan *p, int ncomp, bool bwd, const double *in, double *out,
                      FourierAddr xa) {
  const long long nline = (long long)p->nr * p->nth;
#define KSHT_CZT_CASE(M)                                                                     \
  case M: {                                                                                \
    constexpr int nc_ = czt_nc<M>();                                                       \
    const long long nblk = (nline + 2 * nc_ - 1) / (2 * nc_);                              \
    const int lds_ = czt_lds_bytes<M>(!bwd);                                               \
    const dim3 g((unsigned)czt_grid_x(nblk, ncomp, lds_), (unsigned)ncomp);                  \
    if (bwd)                                                                               \
      hipLaunchKernelGGL((k_fftz_bwd<M, nc_>), g, dim3(M / 8 * nc_), lds_, p->stream,        \
                         p->nphi, p->L, nline, nblk, in, xa, out, p->t.ztab);              \
    else                                                                                   \
      hipLaunchKernelGGL((k_fftz_fwd<M, nc_>), g, dim3(M / 8 * nc_), lds_, p->stream,        \
                         p->nphi, p->L, nline, nblk, in, out, xa, p->t.ztab);              \
  } break;
  switch (p->fft.zm) {
    KSHT_CZT_CASE(256)
    KSHT_CZT_CASE(512)
    KSHT_CZT_CASE(1024)
    KSHT_CZT_CASE(2048)
    default:
      set_error("chirp-z form: not set up for this plan");
      return SHT_ERR_ARG;
  }
#undef KSHT_CZT_CASE
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

int fft_plan_init(sht_plan *p) {
  FftPlan &fp = p->fft;
  const int n = p->nphi;
  fp.n = n;
  fp.npass = 0;
  int rem = n;
  std::vector<int> fac;
  while (rem % 8 == 0) { fac.push_back(8); rem /= 8; }
  while (rem % 4 == 0) { fac.push_back(4); rem /= 4; }
  while (rem % 2 == 0) { fac.push_back(2); rem /= 2; }
  for (int q : {3, 5, 7, 11, 13})
    while (rem % q == 0) { fac.push_back(q); rem /= q; }
  // primes above 13 (direct-sum passes through an LDS scratch copy), first,
  // so that the first of them runs without twiddles
  fp.big = 0;
  std::vector<int> big;
  for (int q = 17; rem > 1; q += 2) {
    if ((long long)q * q > rem) q = rem;
    while (rem % q == 0) {
      big.push_back(q);
      rem /= q;
      fp.big = 1;
    }
  }
  std::sort(big.begin(), big.end(), [](int a, int b) { return a > b; });  // largest first
  fac.insert(fac.begin(), big.begin(), big.end());
  // the largest prime as a Bluestein convolution when it is large enough to
  // pay (R >= kBluesteinMin) and the M-point work arrays fit in LDS
  int blM = 0;
  if (!big.empty() && big[0] >= kBluesteinMin) {
    int M = 1;
    while (M < 2 * big[0] - 1) M *= 2;
    if (n + (n / big[0]) * M <= kFftElemsLarge) blM = M;
  }
  std::vector<int> mfac;
  for (int m = blM; m > 1;) {
    const int r = m % 8 == 0 ? 8 : m % 4 == 0 ? 4 : 2;
    mfac.push_back(r);
    m /= r;
  }
  if ((int)(fac.size() + mfac.size()) > kMaxPass) {
    set_error("FFT length %d has more than %d factors (unsupported)", n, kMaxPass);
    return SHT_ERR_UNSUPPORTED;
  }
  fp.line_elems = blM ? n + (n / big[0]) * blM : fp.big ? 2 * n : n;
  const int E = (blM || n > kFftElemsSmall) ? kFftElemsLarge : kFftElemsSmall;
  int F = E / fp.line_elems;  // complex FFTs per work-group (with their scratch)
  if (F < 1) {
    set_error("FFT length %d exceeds the LDS capacity (%d%s)", n, kFftElemsLarge,
              fp.big ? " with a scratch copy for its prime factors above 13" : "");
    return SHT_ERR_UNSUPPORTED;
  }
  fp.pad = E;
  // every pass costs the same whatever F is (its loops are sized by E), so
  // a work-group takes as many lines as the LDS holds
  if (F > 32) F = 32;
  fp.lines_per_wg = 2 * F;
  std::vector<double2> tw;
  const long double kPi = 3.14159265358979323846264338327950288L;
  int Ns = 1;
  for (int R : fac) {
    FftPass ps;
    ps.R = R;
    ps.Ns = Ns;
    ps.tw_off = (int)tw.size();
    ps.pad = 0;
    ps.M = ps.chirp_off = ps.hat_off = ps.m0 = ps.nm = 0;
    if (Ns == 1 && blM && R == big[0]) {
      // Bluestein: chirps c_k = exp(s i pi k^2 / R) and the transform of the
      // filter B (B_m = conj c_m, B_{M-k} = conj c_k, zero between), 1/M
      // included, for s = +1 then -1; the M-point passes follow the main ones
      ps.M = blM;
      ps.chirp_off = (int)tw.size();
      for (int sg = 1; sg >= -1; sg -= 2)
        for (int k = 0; k < R; k++) {
          const long double a = sg * kPi * (long double)(((long long)k * k) % (2 * R)) / R;
          tw.push_back(make_double2((double)cosl(a), (double)sinl(a)));
        }
      ps.hat_off = (int)tw.size();
      for (int sg = 1; sg >= -1; sg -= 2) {
        std::vector<long double> br(blM, 0.0L), bi(blM, 0.0L);
        for (int k = 0; k < R; k++) {
          const long double a = sg * kPi * (long double)(((long long)k * k) % (2 * R)) / R;
          br[k] = cosl(a);
          bi[k] = -sinl(a);
          if (k) {
            br[blM - k] = cosl(a);
            bi[blM - k] = -sinl(a);
          }
        }
        std::vector<long double> rc(blM), rs(blM);  // exp(-2 pi i t / M)
        for (int t = 0; t < blM; t++) {
          rc[t] = cosl(-2.0L * kPi * t / blM);
          rs[t] = sinl(-2.0L * kPi * t / blM);
        }
        for (int u = 0; u < blM; u++) {
          long double sr = 0.0L, si = 0.0L;
          for (int m = 0; m < blM; m++) {
            const int t = (int)(((long long)u * m) % blM);
            sr += br[m] * rc[t] - bi[m] * rs[t];
            si += br[m] * rs[t] + bi[m] * rc[t];
          }
          tw.push_back(make_double2((double)(sr / blM), (double)(si / blM)));
        }
      }
      fp.pass[fp.npass++] = ps;
      Ns *= R;
      continue;
    }
    if (R > 13) {
      // direct-sum pass: all Ns R roots of order Ns R
      for (int e = 0; e < Ns * R; e++) {
        const long double a =
            -2.0L * 3.14159265358979323846264338327950288L * (long double)e / (long double)(Ns * R);
        tw.push_back(make_double2((double)cosl(a), (double)sinl(a)));
      }
    } else
      for (int k = 0; k < Ns; k++)
        for (int r = 0; r < R; r++) {
          const long double a = -2.0L * 3.14159265358979323846264338327950288L *
                                (long double)(k * r) / (long double)(Ns * R);
          tw.push_back(make_double2((double)cosl(a), (double)sinl(a)));
        }
    fp.pass[fp.npass++] = ps;
    Ns *= R;
  }
  if (blM) {
    // the Bluestein convolution's M-point passes, stored past the main ones
    fp.pass[0].m0 = fp.npass;
    fp.pass[0].nm = (int)mfac.size();
    int ms = 1;
    for (size_t i = 0; i < mfac.size(); i++) {
      FftPass q;
      q.R = mfac[i];
      q.Ns = ms;
      q.tw_off = (int)tw.size();
      q.pad = 0;
      q.M = q.chirp_off = q.hat_off = q.m0 = q.nm = 0;
      for (int k = 0; k < ms; k++)
        for (int r = 0; r < q.R; r++) {
          const long double a = -2.0L * kPi * (long double)(k * r) / (long double)(ms * q.R);
          tw.push_back(make_double2((double)cosl(a), (double)sinl(a)));
        }
      fp.pass[fp.npass + i] = q;
      ms *= q.R;
    }
  }
  const int lds = (int)(sizeof(double2) * (size_t)(fp.lines_per_wg / 2) * fp.line_elems);
  bool gen = false;
  for (int R : fac) gen |= !(R == 2 || R == 3 || R == 4 || R == 8);
  p->fft_generic = gen;
  const void *kb[4] = {(const void *)k_fft_bwd<kFftElemsSmall, false>, (const void *)k_fft_bwd<kFftElemsSmall, true>,
                       (const void *)k_fft_bwd<kFftElemsLarge, false>, (const void *)k_fft_bwd<kFftElemsLarge, true>};
  const void *kf[4] = {(const void *)k_fft_fwd<kFftElemsSmall, false>, (const void *)k_fft_fwd<kFftElemsSmall, true>,
                       (const void *)k_fft_fwd<kFftElemsLarge, false>, (const void *)k_fft_fwd<kFftElemsLarge, true>};
  const int sel = (E == kFftElemsSmall ? 0 : 2) + (gen ? 1 : 0);
  KSHT_HIP(hipFuncSetAttribute(kb[sel], hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  KSHT_HIP(hipFuncSetAttribute(kf[sel], hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  KSHT_HIP(hipMalloc(&p->t.tw, sizeof(double2) * tw.size()));
  KSHT_HIP(hipMemcpy(p->t.tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice));
  p->device_bytes += (long long)(sizeof(double2) * tw.size());
  // split form n = r0 * np: np the smallest power of two above L, 32..1024
  fp.fast = 0;
  int np = 32;
  while (np < p->L + 1) np *= 2;
  if (np <= 1024 && n % np == 0) {
    fp.fast = 1;
    fp.np = np;
    fp.r0 = n / np;
    int tw_b = 0, tw_f = 0;
#define KSHT_TW_OFF(NP)                                                                      \
  case NP:                                                                                   \
    tw_b = fft3_tw_off_bwd<NP, fft_ncb<NP>()>() + fft_bwd_tw_entries<NP>();                   \
    tw_f = fft3_tw_off_fwd<NP, fft_ncf<NP>()>();                                             \
    break;
    switch (np) {
      KSHT_TW_OFF(32)
      KSHT_TW_OFF(64)
      KSHT_TW_OFF(128)
      KSHT_TW_OFF(256)
      KSHT_TW_OFF(512)
      KSHT_TW_OFF(1024)
      default: break;
    }
#undef KSHT_TW_OFF
    fp.lds3b = (int)sizeof(double2) * (tw_b + fp.r0);
    fp.lds3 = (int)sizeof(double2) * (tw_f + np + fp.r0);
    fp.r0m = fp.r0 == 1 ? 0u : (unsigned)(((1ull << 32) + (unsigned)fp.r0 - 1) / (unsigned)fp.r0);
    int dev = 0, lds_max = 0;
    KSHT_HIP(hipGetDevice(&dev));
    KSHT_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (fp.lds3b > lds_max || fp.lds3 > lds_max) fp.fast = 0;  // does not fit: generic form
  }
  if (fp.fast) {
    const int np = fp.np;
    std::vector<double2> twn(n);
    for (int k = 0; k < n; k++) {
      const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
      twn[k] = make_double2((double)cosl(a), (double)sinl(a));
    }
    KSHT_HIP(hipMalloc(&p->t.twn, sizeof(double2) * n));
    KSHT_HIP(hipMemcpy(p->t.twn, twn.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    p->device_bytes += (long long)(sizeof(double2) * n);
#define KSHT_FFT3_ATTR(NP)                                                                   \
  case NP:                                                                                   \
    KSHT_HIP(hipFuncSetAttribute((const void *)k_fft3_bwd<NP, fft_ncb<NP>()>,                \
                                 hipFuncAttributeMaxDynamicSharedMemorySize, fp.lds3b));     \
    KSHT_HIP(hipFuncSetAttribute((const void *)k_fft3_fwd<NP, fft_ncf<NP>()>,                \
                                 hipFuncAttributeMaxDynamicSharedMemorySize, fp.lds3));      \
    break;
    switch (np) {
      KSHT_FFT3_ATTR(32)
      KSHT_FFT3_ATTR(64)
      KSHT_FFT3_ATTR(128)
      KSHT_FFT3_ATTR(256)
      KSHT_FFT3_ATTR(512)
      KSHT_FFT3_ATTR(1024)
      default: break;
    }
#undef KSHT_FFT3_ATTR
  }
  return SHT_OK;
}

#define KSHT_FFT3_CASE(K, NP, NCV, LDS, ...)                                                 \
  case NP: {                                                                               \
    constexpr int nc_ = (NCV);                                                             \
    const dim3 g3((unsigned)((nline + 2 * nc_ - 1) / (2 * nc_)), (unsigned)ncomp);         \
    hipLaunchKernelGGL((K<NP, nc_>), g3, dim3(NP / 8 * nc_), (LDS), p->stream, __VA_ARGS__); \
  } break;

// the DFT form takes the plan's own Fourier workspace layout (rows of Nr
// Ntheta contiguous lines); the reference's rtp layout keeps the FFT kernels
static bool dft_applies(const sht_plan *p, const FourierAddr &xa) {
  return p->var_fft_form == 2 && p->dft && !xa.fftpack && xa.s_line == 1 &&
         xa.q_stride == (long long)p->nr * p->nth;
}

int launch_fft_bwd(sht_plan *p, int ncomp, const double *X, FourierAddr xa, double *v) {
  if (dft_applies(p, xa)) return launch_dft_bwd(p, ncomp, X, xa, v);
  if (p->var_fft_form == 3 && p->fft.zm) return launch_czt(p, ncomp, true, X, v, xa);
  const long long nline = (long long)p->nr * p->nth;
  const int B = p->fft.lines_per_wg;
  const dim3 grid((unsigned)((nline + B - 1) / B), (unsigned)ncomp);
  const size_t lds = sizeof(double2) * (size_t)(B / 2) * p->fft.line_elems;
#define KSHT_FFT_LAUNCH(K, ...)                                                              \
  do {                                                                                     \
    if (p->fft.pad == kFftElemsSmall) {                                                    \
      if (p->fft_generic)                                                                  \
        hipLaunchKernelGGL((K<kFftElemsSmall, true>), grid, dim3(kFftThreads), lds, p->stream, __VA_ARGS__); \
      else                                                                                 \
        hipLaunchKernelGGL((K<kFftElemsSmall, false>), grid, dim3(kFftThreads), lds, p->stream, __VA_ARGS__); \
    } else {                                                                               \
      if (p->fft_generic)                                                                  \
        hipLaunchKernelGGL((K<kFftElemsLarge, true>), grid, dim3(kFftThreads), lds, p->stream, __VA_ARGS__); \
      else                                                                                 \
        hipLaunchKernelGGL((K<kFftElemsLarge, false>), grid, dim3(kFftThreads), lds, p->stream, __VA_ARGS__); \
    }                                                                                      \
  } while (0)
  if (p->fft.fast && p->var_fft_form) {
#define KSHT_B(NP) KSHT_FFT3_CASE(k_fft3_bwd, NP, fft_ncb<NP>(), p->fft.lds3b, p->nphi, p->fft.r0, \
                                  p->fft.r0m, p->L, nline, X, xa, v, p->t.twn)
    switch (p->fft.np) {
      KSHT_B(32)
      KSHT_B(64)
      KSHT_B(128)
      KSHT_B(256)
      KSHT_B(512)
      KSHT_B(1024)
      default: break;
    }
#undef KSHT_B
  } else
    KSHT_FFT_LAUNCH(k_fft_bwd, p->fft, p->L, nline, X, xa, v, p->t.tw);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

int launch_fft_fwd(sht_plan *p, int ncomp, const double *v, double *X, FourierAddr xa) {
  if (dft_applies(p, xa)) return launch_dft_fwd(p, ncomp, v, X, xa);
  if (p->var_fft_form == 3 && p->fft.zm) return launch_czt(p, ncomp, false, v, X, xa);
  const long long nline = (long long)p->nr * p->nth;
  const int B = p->fft.lines_per_wg;
  const dim3 grid((unsigned)((nline + B - 1) / B), (unsigned)ncomp);
  const size_t lds = sizeof(double2) * (size_t)(B / 2) * p->fft.line_elems;
  if (p->fft.fast && p->var_fft_form) {
#define KSHT_F(NP) KSHT_FFT3_CASE(k_fft3_fwd, NP, fft_ncf<NP>(), p->fft.lds3, p->nphi, p->fft.r0, \
                                  p->fft.r0m, p->L, nline, v, X, xa, p->t.twn)
    switch (p->fft.np) {
      KSHT_F(32)
      KSHT_F(64)
      KSHT_F(128)
      KSHT_F(256)
      KSHT_F(512)
      KSHT_F(1024)
      default: break;
    }
#undef KSHT_F
  } else
    KSHT_FFT_LAUNCH(k_fft_fwd, p->fft, p->L, nline, v, X, xa, p->t.tw);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

}  // namespace ksht
