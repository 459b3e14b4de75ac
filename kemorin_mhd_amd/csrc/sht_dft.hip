// sht_dft.hip -- the zonal transform as a pruned real DFT on FP64 MFMA: one
// strided-batched DGEMM per direction against a cos/sin basis of the 2L+1
// Fourier rows the Legendre side carries (rocBLAS, a plain library GEMM).
//
// The plan's "fft_form" variant 2.  It is for lengths whose FFT is slow: the
// mixed-radix kernel takes an LDS round trip and a barrier per pass, and a
// length with a prime factor above 13 (Nphi = 3(L+1) has them whenever
// (L+1)/2 does) needs a scratch copy per pass or a Bluestein convolution.
// The basis product costs 2 Nphi (2L+1) flops per line and component
// against roughly 5 Nphi log2(Nphi) for an FFT, but it runs at the FP64
// matrix rate; which form is faster depends on the shape, so the autotuner
// decides.  Convention (FFTPACK's, as the FFT kernels):
//   backward x_j = a_0 + sum_m (a_m cos(2 pi m j / N) + b_m sin(2 pi m j / N)),
//   forward  a_0 = (1/N) sum_j x_j, a_m = (2/N) sum_j x_j cos(.),
//            b_m = (2/N) sum_j x_j sin(.),
// with the plan's Fourier rows q = 0 (a_0), 2m-1 (a_m), 2m (b_m), m <= L.
// Layouts: the Fourier workspace of component nd is a column-major
// (Nr Ntheta) x (2L+1) matrix (lines contiguous, row stride Nr Ntheta), the
// grid field a column-major (Nr Ntheta) x Nphi matrix, so
//   backward: V_nd = X_nd B^T,  B[j, q] the backward basis (Nphi x (2L+1));
//   forward:  X_nd = V_nd F,    F[j, q] the forward basis.
#include <rocblas/rocblas.h>

#include <cmath>
#include <vector>

#include "sht_internal.h"

namespace ksht {

struct DftState {
  rocblas_handle h = nullptr;
  double *B = nullptr, *F = nullptr;  // Nphi x (2L+1) each, column-major
  void *ws = nullptr;                 // rocBLAS workspace, sized at set-up
  size_t ws_bytes = 0;
};

// the plan's stream for rocBLAS (the legacy default stream as handle 0)
static hipStream_t blas_stream(const sht_plan *p) { return api_stream(p->stream); }

static int blas_ok(rocblas_status s, const char *what) {
  if (s == rocblas_status_success) return SHT_OK;
  set_error("DFT form: %s: rocBLAS status %d", what, (int)s);
  return s == rocblas_status_memory_error ? SHT_ERR_OOM : SHT_ERR_HIP;
}

static int gemm_bwd(sht_plan *p, DftState *st, int ncomp, const double *X, long long x_nd,
                    double *v) {
  const int nline = p->nr * p->nth, K = p->nslot, N = p->nphi;
  const double one = 1.0, zero = 0.0;
  return blas_ok(rocblas_dgemm_strided_batched(
                     st->h, rocblas_operation_none, rocblas_operation_transpose, nline, N, K,
                     &one, X, nline, x_nd, st->B, N, 0, &zero, v, nline,
                     (rocblas_stride)nline * N, ncomp),
                 "backward product");
}

static int gemm_fwd(sht_plan *p, DftState *st, int ncomp, const double *v, double *X,
                    long long x_nd) {
  const int nline = p->nr * p->nth, K = p->nslot, N = p->nphi;
  const double one = 1.0, zero = 0.0;
  return blas_ok(rocblas_dgemm_strided_batched(
                     st->h, rocblas_operation_none, rocblas_operation_none, nline, K, N, &one,
                     v, nline, (rocblas_stride)nline * N, st->F, N, 0, &zero, X, nline, x_nd,
                     ncomp),
                 "forward product");
}

void dft_free(sht_plan *p) {
  DftState *st = (DftState *)p->dft;
  if (!st) return;
  if (st->h) (void)rocblas_destroy_handle(st->h);
  if (st->B) (void)hipFree(st->B);
  if (st->F) (void)hipFree(st->F);
  if (st->ws) (void)hipFree(st->ws);
  delete st;
  p->dft = nullptr;
}

// Bases, handle and a workspace sized for the plan's largest calls (so that
// no call allocates: the transforms may run under graph capture).
int dft_init(sht_plan *p) {
  if (p->dft) return SHT_OK;
  const int N = p->nphi, K = p->nslot, L = p->L;
  std::vector<double> B((size_t)N * K), F((size_t)N * K);
  const long double two_pi = 2.0L * 3.14159265358979323846264338327950288L;
  for (int j = 0; j < N; j++) {
    B[j] = 1.0;
    F[j] = (double)(1.0L / N);
    for (int m = 1; m <= L; m++) {
      const long long r = ((long long)m * j) % N;  // exact reduction of the angle
      const long double a = two_pi * (long double)r / (long double)N;
      const long double c = cosl(a), s = sinl(a);
      B[j + (size_t)(2 * m - 1) * N] = (double)c;
      B[j + (size_t)(2 * m) * N] = (double)s;
      F[j + (size_t)(2 * m - 1) * N] = (double)(2.0L * c / N);
      F[j + (size_t)(2 * m) * N] = (double)(2.0L * s / N);
    }
  }
  DftState *st = new DftState();
  p->dft = st;
  int rc = SHT_OK;
  const size_t bytes = sizeof(double) * B.size();
  if (hipMalloc(&st->B, bytes) != hipSuccess || hipMalloc(&st->F, bytes) != hipSuccess) {
    set_error("DFT form: bases of %zu bytes: out of device memory", 2 * bytes);
    dft_free(p);
    return SHT_ERR_OOM;
  }
  if (hipMemcpy(st->B, B.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(st->F, F.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("DFT form: copying the bases failed");
    dft_free(p);
    return SHT_ERR_HIP;
  }
  if ((rc = blas_ok(rocblas_create_handle(&st->h), "handle")) ||
      (rc = blas_ok(rocblas_set_stream(st->h, blas_stream(p)), "stream"))) {
    dft_free(p);
    return rc;
  }
  // workspace the largest calls need (max_ncomp components, both directions)
  size_t need = 0;
  if ((rc = blas_ok(rocblas_start_device_memory_size_query(st->h), "size query")) == SHT_OK) {
    const long long x_nd = (long long)p->nslot * p->nr * p->nth;
    (void)gemm_bwd(p, st, p->max_ncomp, nullptr, x_nd, nullptr);
    (void)gemm_fwd(p, st, p->max_ncomp, nullptr, nullptr, x_nd);
    rc = blas_ok(rocblas_stop_device_memory_size_query(st->h, &need), "size query");
  }
  if (rc) {
    dft_free(p);
    return rc;
  }
  if (need) {
    if (hipMalloc(&st->ws, need) != hipSuccess) {
      set_error("DFT form: rocBLAS workspace of %zu bytes: out of device memory", need);
      dft_free(p);
      return SHT_ERR_OOM;
    }
    st->ws_bytes = need;
  }
  if ((rc = blas_ok(rocblas_set_workspace(st->h, st->ws, st->ws_bytes), "workspace"))) {
    dft_free(p);
    return rc;
  }
  p->device_bytes += (long long)(2 * bytes + need);
  return SHT_OK;
}

int launch_dft_bwd(sht_plan *p, int ncomp, const double *X, FourierAddr xa, double *v) {
  DftState *st = (DftState *)p->dft;
  int rc = blas_ok(rocblas_set_stream(st->h, blas_stream(p)), "stream");
  if (rc) return rc;
  return gemm_bwd(p, st, ncomp, X, xa.s_nd, v);
}

int launch_dft_fwd(sht_plan *p, int ncomp, const double *v, double *X, FourierAddr xa) {
  DftState *st = (DftState *)p->dft;
  int rc = blas_ok(rocblas_set_stream(st->h, blas_stream(p)), "stream");
  if (rc) return rc;
  return gemm_fwd(p, st, ncomp, v, X, xa.s_nd);
}

}  // namespace ksht
