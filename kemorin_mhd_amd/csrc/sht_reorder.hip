// sht_reorder.hip -- bit-exact exchange permutations between the reference's
// communication buffers on one rank: calypso_sph_comm_N
// (SS/spherical_SRs_N.f90:233-248 -> SR/calypso_SR_core.f90:101-109 self
// copy) followed by the receive-side gather through irev_sr with its zero
// dummy slot (SR/set_from_recv_buf_rev.f90:77-87).  rtm <-> rtp pairs
// (colatitude t, radius k, order m) with (radius k, colatitude t, FFTPACK
// slot p(m)) as set up by CG/set_comm_table_rtp_rj.f90:549-563.
// The permutations move whole ncomp-vectors; tiles of 16 x 16 nodes go
// through LDS so that both the t-contiguous rtm side and the k-contiguous
// rtp side are read / written in coalesced rows.
#include <algorithm>

#include "sht_internal.h"

namespace ksht {

constexpr int kNdChunk = 16;

__device__ __forceinline__ int fft_slot(int m) { return m == 0 ? 0 : (m > 0 ? 2 * m : -2 * m + 1); }

// dir 0: rtm -> rtp, dir 1: rtp -> rtm.  grid: (t tiles, k tiles, slots)
template <int DIR>
__global__ void __launch_bounds__(256) k_rtm_rtp(int L, int nr, int nth, int ncomp,
                                                  const double *__restrict__ src,
                                                  double *__restrict__ dst) {
  __shared__ double tile[16 * 16 * kNdChunk];
  const int s = blockIdx.z, m = s - L, p = fft_slot(m);
  const int t0 = blockIdx.x * 16, k0 = blockIdx.y * 16;
  const long long rtm_base = (long long)ncomp * nth * ((long long)nr * s);
  const long long rtp_base = (long long)ncomp * nr * ((long long)nth * p);
  for (int nd0 = 0; nd0 < ncomp; nd0 += kNdChunk) {
    const int ndc = min(kNdChunk, ncomp - nd0);
    __syncthreads();
    // read side
    for (int idx = threadIdx.x; idx < 16 * 16 * ndc; idx += 256) {
      const int nd = idx % ndc, a = (idx / ndc) % 16, b = idx / (ndc * 16);
      int t, k;
      if (DIR == 0) { t = t0 + a; k = k0 + b; } else { k = k0 + a; t = t0 + b; }
      if (t < nth && k < nr) {
        const long long o = DIR == 0
                                ? rtm_base + nd0 + nd + (long long)ncomp * (t + (long long)nth * k)
                                : rtp_base + nd0 + nd + (long long)ncomp * (k + (long long)nr * t);
        tile[((t - t0) * 16 + (k - k0)) * kNdChunk + nd] = src[o];
      }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 16 * 16 * ndc; idx += 256) {
      const int nd = idx % ndc, a = (idx / ndc) % 16, b = idx / (ndc * 16);
      int t, k;
      if (DIR == 0) { k = k0 + a; t = t0 + b; } else { t = t0 + a; k = k0 + b; }
      if (t < nth && k < nr) {
        const long long o = DIR == 0
                                ? rtp_base + nd0 + nd + (long long)ncomp * (k + (long long)nr * t)
                                : rtm_base + nd0 + nd + (long long)ncomp * (t + (long long)nth * k);
        dst[o] = tile[((t - t0) * 16 + (k - k0)) * kNdChunk + nd];
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_gather_rev(int ncomp, long long n_dst, long long n_src,
                                                     const int *__restrict__ irev,
                                                     const double *__restrict__ src,
                                                     double *__restrict__ dst) {
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= n_dst * ncomp) return;
  const long long i = gid / ncomp;
  const int nd = (int)(gid - i * ncomp);
  const long long r = irev[i];
  dst[gid] = (r >= 1 && r <= n_src) ? src[nd + (long long)ncomp * (r - 1)] : 0.0;
}

__global__ void __launch_bounds__(256) k_fill_zero(double *__restrict__ a, long long n) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    a[i] = 0.0;
}

// Zero the FFTPACK slots of an rtp buffer that carry no order m <= L: the
// Nyquist slot p = 1 and p >= 2L+2 (p is the slowest index).  A fill kernel,
// not hipMemsetAsync, so that captured graphs replay it faithfully.
int launch_zero_unused_rtp(sht_plan *p, int ncomp, double *rtp) {
  const long long slot = (long long)ncomp * p->nr * p->nth;
  const long long tail = slot * (p->nphi - 2 * p->L - 2);
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<long long>((slot + 255) / 256, 4096)),
                     dim3(256), 0, p->stream, rtp + slot, slot);
  if (tail > 0)
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<long long>((tail + 255) / 256, 65536)),
                       dim3(256), 0, p->stream, rtp + slot * (2 * p->L + 2), tail);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

int launch_rtm_to_rtp(sht_plan *p, int ncomp, const double *rtm, double *rtp) {
  int rc = launch_zero_unused_rtp(p, ncomp, rtp);
  if (rc) return rc;
  const dim3 grid((p->nth + 15) / 16, (p->nr + 15) / 16, p->nslot);
  hipLaunchKernelGGL(k_rtm_rtp<0>, grid, dim3(256), 0, p->stream, p->L, p->nr, p->nth, ncomp, rtm,
                     rtp);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

int launch_rtp_to_rtm(sht_plan *p, int ncomp, const double *rtp, double *rtm) {
  const dim3 grid((p->nth + 15) / 16, (p->nr + 15) / 16, p->nslot);
  hipLaunchKernelGGL(k_rtm_rtp<1>, grid, dim3(256), 0, p->stream, p->L, p->nr, p->nth, ncomp, rtp,
                     rtm);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

int launch_gather_rev(sht_plan *p, int ncomp, long long n_dst, long long n_src, const int *irev,
                      const double *src, double *dst) {
  const long long n = n_dst * ncomp;
  if (n == 0) return SHT_OK;
  hipLaunchKernelGGL(k_gather_rev, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream,
                     ncomp, n_dst, n_src, irev, src, dst);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

}  // namespace ksht
