// sht_capi.hip -- extern "C" entry points of libkemorin_sht.so.
// Whole transforms (sht_backward / sht_forward) run two fused kernels per
// direction: the Legendre kernel reads the reference's comp-interleaved
// spectral buffer directly and writes (or reads) the field-major Fourier
// workspace already in rtp order, so the rj->rlm copy and the rtm->rtp
// exchange of SS/t_sph_transforms.f90:74-123,250-297 cost no extra pass.
#include <cstring>
#include <mutex>

#include "sht_internal.h"

using namespace ksht;

static int check_call(const sht_plan *p, int nv, int ns, const void *a, const void *b,
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

static SpecAddr spec_addr(const sht_plan *p, int ncomp) {
  SpecAddr sa;
  sa.s_nd = 1;
  sa.s_j = ncomp;
  sa.s_k = (long long)ncomp * p->jmax;
  return sa;
}
static GridAddr x_grid_addr(const sht_plan *p) {
  GridAddr ga;
  ga.slot_off = p->t.slot_off_X;
  ga.slot_mul = 1;
  ga.s_nd = (long long)p->nslot * p->nr * p->nth;
  ga.s_t = p->nr;
  ga.s_k = 1;
  return ga;
}
static FourierAddr x_fourier_addr(const sht_plan *p) {
  FourierAddr xa;
  xa.fftpack = 0;
  xa.q_stride = (long long)p->nr * p->nth;
  xa.s_nd = (long long)p->nslot * p->nr * p->nth;
  xa.s_line = 1;
  return xa;
}
static GridAddr rtm_grid_addr(const sht_plan *p, int ncomp) {
  GridAddr ga;
  ga.slot_off = p->t.slot_off_rtm;
  ga.slot_mul = ncomp;
  ga.s_nd = 1;
  ga.s_t = ncomp;
  ga.s_k = (long long)ncomp * p->nth;
  return ga;
}
static FourierAddr rtp_fourier_addr(const sht_plan *p, int ncomp) {
  FourierAddr xa;
  xa.fftpack = 1;
  xa.q_stride = (long long)ncomp * p->nr * p->nth;
  xa.s_nd = 1;
  xa.s_line = ncomp;
  return xa;
}

static int leg_bwd(sht_plan *p, const Workspace &ws, int nv, int ns, const double *sp,
                   SpecAddr sa, double *out, GridAddr ga) {
  StageTimer tm(p, 0);
  if (p->precision == SHT_PREC_SPLIT_INT8)
    return launch_leg_bwd_split(p, ws, nv, ns, sp, sa, out, ga);
  return launch_leg_bwd(p, nv, ns, sp, sa, out, ga);
}
static int leg_fwd(sht_plan *p, const Workspace &ws, int nv, int ns, const double *in,
                   GridAddr ga, double *sp, SpecAddr sa) {
  StageTimer tm(p, 3);
  if (p->precision == SHT_PREC_SPLIT_INT8)
    return launch_leg_fwd_split(p, ws, nv, ns, in, ga, sp, sa);
  return launch_leg_fwd(p, nv, ns, in, ga, sp, sa);
}

static int backward_ws(sht_plan *p, const Workspace &ws, int nv, int ns, const double *WS_rj,
                       double *v_rtp) {
  const int ncomp = 3 * nv + ns;
  int rc = leg_bwd(p, ws, nv, ns, WS_rj, spec_addr(p, ncomp), ws.X, x_grid_addr(p));
  if (rc) return rc;
  StageTimer tm(p, 1);
  return launch_fft_bwd(p, ncomp, ws.X, x_fourier_addr(p), v_rtp);
}

static int check_finite(sht_plan *p, const double *out, long long n, const char *where);

extern "C" int sht_backward(sht_plan *p, int nv, int ns, const double *WS_rj, double *v_rtp) {
  int rc = check_call(p, nv, ns, WS_rj, v_rtp, "sht_backward");
  if (rc) return rc;
  rc = backward_ws(p, plan_workspace(p), nv, ns, WS_rj, v_rtp);
  if (rc) return rc;
  return check_finite(p, v_rtp, (long long)(3 * nv + ns) * p->nr * p->nth * p->nphi,
                      "sht_backward");
}

extern "C" int sht_forward(sht_plan *p, int nv, int ns, const double *v_rtp, double *WR_rj) {
  int rc = check_call(p, nv, ns, v_rtp, WR_rj, "sht_forward");
  if (rc) return rc;
  const int ncomp = 3 * nv + ns;
  const Workspace ws = plan_workspace(p);
  {
    StageTimer tm(p, 2);
    rc = launch_fft_fwd(p, ncomp, v_rtp, ws.X, x_fourier_addr(p));
  }
  if (rc) return rc;
  rc = leg_fwd(p, ws, nv, ns, ws.X, x_grid_addr(p), WR_rj, spec_addr(p, ncomp));
  if (rc) return rc;
  return check_finite(p, WR_rj, (long long)ncomp * p->jmax * p->nr, "sht_forward");
}

// ---------------------------------------------------------------------------
// Fixed device workspace of sht_backward_noalloc.  Each device that loads the
// library holds its own copy of this array until the process ends.  In host
// code the array's name is only a stand-in for the symbol: its device address
// comes from hipGetSymbolAddress, taken once per device with that device
// current (scratch_attach, run at plan creation) and kept in scratch_dev.
constexpr long long kScratchBytes = SHT_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex scratch_mu;
static void *scratch_dev[kMaxDevices];

namespace ksht {
int scratch_attach(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("scratch: device %d outside 0..%d", device, kMaxDevices - 1);
    return SHT_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_dev[device]) return SHT_OK;
  int cur = -1;
  KSHT_HIP(hipGetDevice(&cur));
  if (cur != device) {
    set_error("scratch: device %d is not current (current %d)", device, cur);
    return SHT_ERR_STATE;
  }
  void *ptr = nullptr;
  KSHT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) {
    set_error("scratch: no device address for the workspace array on device %d", device);
    return SHT_ERR_HIP;
  }
  scratch_dev[device] = ptr;
  return SHT_OK;
}
}  // namespace ksht

static long long align256(long long n) { return (n + 255) / 256 * 256; }

// Number of non-finite values (exponent field all ones: NaN or +-Inf) of
// x[0 .. n), added to *cnt.
__global__ void __launch_bounds__(256) k_count_nonfinite(const double *__restrict__ x, long long n,
                                                         unsigned long long *cnt) {
  unsigned c = 0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    c += ((__double_as_longlong(x[i]) >> 52) & 0x7ff) == 0x7ff ? 1u : 0u;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(cnt, (unsigned long long)c);
}

// SHT_FLAG_CHECK_FINITE: count the output's non-finite values after the call
// (the plan's stream is waited for); skipped while the stream is captured
static int check_finite(sht_plan *p, const double *out, long long n, const char *where) {
  if (!p->nonfinite || n <= 0) return SHT_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  const hipStream_t st = api_stream(p->stream);
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return SHT_OK;
  KSHT_HIP(hipMemsetAsync(p->nonfinite, 0, sizeof(unsigned long long), st));
  const long long nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_count_nonfinite, dim3((unsigned)(nb < 4096 ? nb : 4096)), dim3(256), 0,
                     p->stream, out, n, p->nonfinite);
  KSHT_HIP(hipGetLastError());
  unsigned long long h = 0;
  KSHT_HIP(hipMemcpyAsync(&h, p->nonfinite, sizeof(h), hipMemcpyDeviceToHost, st));
  KSHT_HIP(hipStreamSynchronize(st));
  if (h) {
    set_error("%s: %llu non-finite values in the output (SHT_FLAG_CHECK_FINITE)", where, h);
    return SHT_ERR_NONFINITE;
  }
  return SHT_OK;
}

extern "C" int sht_backward_noalloc(sht_plan *p, int nv, int ns, const double *WS_rj,
                                    double *v_rtp) {
  int rc = check_call(p, nv, ns, WS_rj, v_rtp, "sht_backward_noalloc");
  if (rc) return rc;
  if (p->timing) {
    set_error("sht_backward_noalloc: the plan's stage timers are on (they create events)");
    return SHT_ERR_STATE;
  }
  const int ncomp = 3 * nv + ns;
  const long long xb = align256((long long)sizeof(double) * ncomp * p->nslot * p->nr * p->nth);
  long long db = 0, eb = 0;
  if (p->precision == SHT_PREC_SPLIT_INT8) {
    split_ws_bytes(p, ncomp, true, &db, &eb);
    db = align256(db);
    eb = align256(eb);
  }
  if (xb + db + eb > kScratchBytes) {
    set_error("sht_backward_noalloc: the call needs %lld bytes of workspace, above the fixed "
              "%lld (L=%d Nr=%d Ntheta=%d, %d components); use sht_backward",
              xb + db + eb, kScratchBytes, p->L, p->nr, p->nth, ncomp);
    return SHT_ERR_ARG;
  }
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    base = (p->device >= 0 && p->device < kMaxDevices) ? scratch_dev[p->device] : nullptr;
  }
  if (!base) {
    // plan creation could not attach it: this is the first call on the device
    rc = scratch_attach(p->device);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(scratch_mu);
    base = scratch_dev[p->device];
  }
  unsigned char *b = (unsigned char *)base;
  const Workspace ws{(double *)b, db ? (signed char *)(b + xb) : nullptr,
                     eb ? (int *)(b + xb + db) : nullptr};
  rc = backward_ws(p, ws, nv, ns, WS_rj, v_rtp);
  if (rc) return rc;
  return check_finite(p, v_rtp, (long long)ncomp * p->nr * p->nth * p->nphi,
                      "sht_backward_noalloc");
}

static int host_call(sht_plan *p, int nv, int ns, const double *in, double *out, bool bwd) {
  int rc = check_call(p, nv, ns, in, out, bwd ? "sht_backward_host" : "sht_forward_host");
  if (rc) return rc;
  const int ncomp = 3 * nv + ns;
  const size_t nsp = sizeof(double) * (size_t)ncomp * p->jmax * p->nr;
  const size_t ngr = sizeof(double) * (size_t)ncomp * p->nr * p->nth * p->nphi;
  double *din = nullptr, *dout = nullptr;
  const size_t bin = bwd ? nsp : ngr, bout = bwd ? ngr : nsp;
  if (hipMalloc(&din, bin) != hipSuccess || hipMalloc(&dout, bout) != hipSuccess) {
    if (din) (void)hipFree(din);
    set_error("host transform: device allocation failed");
    return SHT_ERR_OOM;
  }
  rc = SHT_OK;
  if (hipMemcpyAsync(din, in, bin, hipMemcpyHostToDevice, p->stream) != hipSuccess) rc = SHT_ERR_HIP;
  if (!rc) rc = bwd ? sht_backward(p, nv, ns, din, dout) : sht_forward(p, nv, ns, din, dout);
  if (!rc && hipMemcpyAsync(out, dout, bout, hipMemcpyDeviceToHost, p->stream) != hipSuccess)
    rc = SHT_ERR_HIP;
  if (hipStreamSynchronize(api_stream(p->stream)) != hipSuccess && !rc) rc = SHT_ERR_HIP;
  if (rc == SHT_ERR_HIP) set_error("host transform: %s", hipGetErrorString(hipGetLastError()));
  (void)hipFree(din);
  (void)hipFree(dout);
  return rc;
}

extern "C" int sht_backward_host(sht_plan *p, int nv, int ns, const double *WS, double *v) {
  return host_call(p, nv, ns, WS, v, true);
}
extern "C" int sht_forward_host(sht_plan *p, int nv, int ns, const double *v, double *WR) {
  return host_call(p, nv, ns, v, WR, false);
}

extern "C" int sht_legendre_backward(sht_plan *p, int nv, int ns, const double *WR_rlm,
                                     double *WS_rtm) {
  int rc = check_call(p, nv, ns, WR_rlm, WS_rtm, "sht_legendre_backward");
  if (rc) return rc;
  const int ncomp = 3 * nv + ns;
  return leg_bwd(p, plan_workspace(p), nv, ns, WR_rlm, spec_addr(p, ncomp), WS_rtm,
                 rtm_grid_addr(p, ncomp));
}

extern "C" int sht_legendre_forward(sht_plan *p, int nv, int ns, const double *WR_rtm,
                                    double *WS_rlm) {
  int rc = check_call(p, nv, ns, WR_rtm, WS_rlm, "sht_legendre_forward");
  if (rc) return rc;
  const int ncomp = 3 * nv + ns;
  return leg_fwd(p, plan_workspace(p), nv, ns, WR_rtm, rtm_grid_addr(p, ncomp), WS_rlm,
                 spec_addr(p, ncomp));
}

extern "C" int sht_fft_backward(sht_plan *p, int ncomp, const double *WR_rtp, double *v_rtp) {
  int rc = check_call(p, 0, ncomp, WR_rtp, v_rtp, "sht_fft_backward");
  if (rc) return rc;
  StageTimer tm(p, 1);
  return launch_fft_bwd(p, ncomp, WR_rtp, rtp_fourier_addr(p, ncomp), v_rtp);
}

extern "C" int sht_fft_forward(sht_plan *p, int ncomp, const double *v_rtp, double *WS_rtp) {
  int rc = check_call(p, 0, ncomp, v_rtp, WS_rtp, "sht_fft_forward");
  if (rc) return rc;
  StageTimer tm(p, 2);
  rc = launch_zero_unused_rtp(p, ncomp, WS_rtp);
  if (rc) return rc;
  return launch_fft_fwd(p, ncomp, v_rtp, WS_rtp, rtp_fourier_addr(p, ncomp));
}

extern "C" int sht_exchange_rtm_to_rtp(sht_plan *p, int ncomp, const double *rtm, double *rtp) {
  int rc = check_call(p, 0, ncomp, rtm, rtp, "sht_exchange_rtm_to_rtp");
  if (rc) return rc;
  StageTimer tm(p, 4);
  return launch_rtm_to_rtp(p, ncomp, rtm, rtp);
}

extern "C" int sht_exchange_rtp_to_rtm(sht_plan *p, int ncomp, const double *rtp, double *rtm) {
  int rc = check_call(p, 0, ncomp, rtp, rtm, "sht_exchange_rtp_to_rtm");
  if (rc) return rc;
  StageTimer tm(p, 4);
  return launch_rtp_to_rtm(p, ncomp, rtp, rtm);
}

extern "C" int sht_exchange_gather_rev(sht_plan *p, int ncomp, long long n_dst, long long n_src,
                                       const int *irev, const double *src, double *dst) {
  if (!p || ncomp < 1 || n_dst < 0 || n_src < 0 || (n_dst > 0 && (!irev || !dst)) ||
      (n_src > 0 && !src)) {
    set_error("sht_exchange_gather_rev: bad arguments");
    return SHT_ERR_ARG;
  }
  StageTimer tm(p, 4);
  return launch_gather_rev(p, ncomp, n_dst, n_src, irev, src, dst);
}

extern "C" int sht_plan_timer_enable(sht_plan *p, int on) {
  if (!p) {
    set_error("sht_plan_timer_enable: null plan");
    return SHT_ERR_ARG;
  }
  p->timing = on ? 1 : 0;
  return SHT_OK;
}

extern "C" int sht_plan_timer_read(sht_plan *p, double *ms, long long *calls, int n) {
  if (!p || n < 0 || n > kNumStages || (n > 0 && (!ms || !calls))) {
    set_error("sht_plan_timer_read: bad arguments");
    return SHT_ERR_ARG;
  }
  KSHT_HIP(hipStreamSynchronize(api_stream(p->stream)));
  int rc = SHT_OK;
  for (int s = 0; s < kNumStages; s++) {
    double tot = 0.0;
    for (auto &e : p->tev[s]) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, e.first, e.second) != hipSuccess) rc = SHT_ERR_HIP;
      tot += t;
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    if (s < n) {
      ms[s] = tot;
      calls[s] = (long long)p->tev[s].size();
    }
    p->tev[s].clear();
  }
  if (rc) set_error("sht_plan_timer_read: event query failed");
  return rc;
}

// ---------------------------------------------------------------------------
// Self-test of the MFMA operand / result layouts assumed by the kernels.
__global__ void k_selftest(const double *a, const double *b, double *d, const int *ai,
                           const int *bi, int *di) {
  const int lane = threadIdx.x;
  d4 c = {0.0, 0.0, 0.0, 0.0};
  c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[lane], b[lane], c, 0, 0, 0);
  for (int r = 0; r < 4; r++) d[lane * 4 + r] = c[r];
  i4 A = {ai[lane * 4], ai[lane * 4 + 1], ai[lane * 4 + 2], ai[lane * 4 + 3]};
  i4 B = {bi[lane * 4], bi[lane * 4 + 1], bi[lane * 4 + 2], bi[lane * 4 + 3]};
  i4 C = {0, 0, 0, 0};
  C = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B, C, 0, 0, 0);
  for (int r = 0; r < 4; r++) di[lane * 4 + r] = C[r];
}

extern "C" int sht_selftest(void) {
  // A[i][k] = i + 1 + 10k, B[k][j] = (j + 1) * (k + 2); FP64 result row
  // 4r + lane/16 (element r); INT8 result row 4*(lane/16) + r.  For INT8
  // each lane's 16 bytes hold one row i (A) or column j (B) segment of the
  // K = 64 reduction, with the same K assignment for A and B.
  double ha[64], hb[64], hd[256];
  int hai[256], hbi[256], hdi[256];
  for (int lane = 0; lane < 64; lane++) {
    const int i = lane % 16, k = lane / 16;
    ha[lane] = i + 1 + 10 * k;
    hb[lane] = (i + 1) * (k + 2);  // here i plays j
  }
  for (int lane = 0; lane < 64; lane++) {
    const int row = lane % 16;
    signed char av[16], bv[16];
    for (int e = 0; e < 16; e++) {
      av[e] = (signed char)((row + 2 * e + lane / 16) % 7 - 3);
      bv[e] = (signed char)((row * 3 + e + lane / 16) % 5 - 2);
    }
    memcpy(&hai[lane * 4], av, 16);
    memcpy(&hbi[lane * 4], bv, 16);
  }
  double *da, *db, *dd;
  int *dai, *dbi, *ddi;
  KSHT_HIP(hipMalloc(&da, sizeof(ha)));
  KSHT_HIP(hipMalloc(&db, sizeof(hb)));
  KSHT_HIP(hipMalloc(&dd, sizeof(hd)));
  KSHT_HIP(hipMalloc(&dai, sizeof(hai)));
  KSHT_HIP(hipMalloc(&dbi, sizeof(hbi)));
  KSHT_HIP(hipMalloc(&ddi, sizeof(hdi)));
  KSHT_HIP(hipMemcpy(da, ha, sizeof(ha), hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(db, hb, sizeof(hb), hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(dai, hai, sizeof(hai), hipMemcpyHostToDevice));
  KSHT_HIP(hipMemcpy(dbi, hbi, sizeof(hbi), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_selftest, dim3(1), dim3(64), 0, 0, da, db, dd, dai, dbi, ddi);
  KSHT_HIP(hipGetLastError());
  KSHT_HIP(hipMemcpy(hd, dd, sizeof(hd), hipMemcpyDeviceToHost));
  KSHT_HIP(hipMemcpy(hdi, ddi, sizeof(hdi), hipMemcpyDeviceToHost));
  (void)hipFree(da); (void)hipFree(db); (void)hipFree(dd);
  (void)hipFree(dai); (void)hipFree(dbi); (void)hipFree(ddi);
  int bad = 0;
  for (int lane = 0; lane < 64; lane++)
    for (int r = 0; r < 4; r++) {
      const int i = 4 * r + lane / 16, j = lane % 16;  // FP64: row 4r + lane/16
      double ref = 0;
      for (int k = 0; k < 4; k++) ref += (i + 1 + 10 * k) * (double)((j + 1) * (k + 2));
      if (hd[lane * 4 + r] != ref) bad |= 1;
      // int8: row 4*(lane/16) + r; sum over the 4 lane groups g and 16 bytes e
      const int ii = 4 * (lane / 16) + r;
      long long iref = 0;
      for (int g = 0; g < 4; g++)
        for (int e = 0; e < 16; e++) {
          const int av = (ii + 2 * e + g) % 7 - 3, bv = (j * 3 + e + g) % 5 - 2;
          iref += av * bv;
        }
      if (hdi[lane * 4 + r] != iref) bad |= 2;
    }
  if (bad) set_error("sht_selftest: MFMA layout mismatch (mask %d)", bad);
  return bad ? SHT_ERR_STATE : SHT_OK;
}
