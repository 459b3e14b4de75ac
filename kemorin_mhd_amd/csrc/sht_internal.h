// sht_internal.h -- plan state and kernel launch interfaces shared by the
// translation units of libkemorin_sht.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/kemorin_sht.h"

namespace ksht {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 16;   // MFMA tile edge (rows, radii, degrees)
constexpr int kMaxPass = 24;
// Legendre bwd, FFT bwd, FFT fwd, Legendre fwd, exchanges, and the vector-group
// Legendre kernels alone (bwd, fwd): the dominant kernels of a round trip
constexpr int kNumStages = 7;

void set_error(const char *fmt, ...);

#define KSHT_HIP(call)                                                   \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      ::ksht::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,       \
                        hipGetErrorString(e_));                          \
      return SHT_ERR_HIP;                                                \
    }                                                                    \
  } while (0)

// One pass of the Stockham FFT: radix, span of finished sub-transforms and
// offset of its twiddle block (Ns*R entries) in the twiddle table.
struct FftPass {
  int R, Ns, tw_off, pad;
  // Bluestein form of a large prime pass (first pass only, M = 0 otherwise):
  // convolution length M, offsets in the twiddle table of the chirps
  // exp(s i pi k^2 / R), k < R, and of the filter transforms (M entries),
  // each for s = +1 then s = -1, and the M-point passes at pass[m0 .. m0+nm)
  int M, chirp_off, hat_off, m0, nm;
};
struct FftPlan {
  int n, npass, lines_per_wg, pad;  // pad: LDS elements per work-group
  int big;                          // a prime factor above 13 (LDS scratch copy)
  int line_elems;                   // LDS elements per complex line (with scratch)
  FftPass pass[kMaxPass];
  // split form n = r0 * np (np a power of two above L)
  int fast, np, r0, lds3, lds3b;  // LDS bytes: forward, backward (whole-block staging)
  unsigned r0m;                   // ceil(2^32 / r0): division by r0 as a multiply-high
  // chirp-z form (set up on first selection): convolution length, 0 when
  // not set up
  int zm = 0;
};

// Work-group order of the GEMMs.  The hardware deals work-groups round-robin
// to the 8 XCDs, each with its own L2, and dispatches them in order, so an
// XCD with more work than the others holds up the whole grid.  Logical ids
// are grouped in runs of G work-groups that share operands through L2; run r
// goes to XCD r % 8, so each L2 sees whole runs while the 8 XCDs work
// through neighbouring orders at the same time (the work per order falls
// with M, so contiguous M ranges per XCD would be unbalanced; consecutive
// runs must be consecutive orders for the same reason).  Ids past the last
// full set of 8 runs keep their value.
__device__ __forceinline__ long long xcd_group_order(long long b, long long n, long long G) {
  const long long sup = 8 * G, nsup = n / sup;
  if (b >= nsup * sup) return b;
  const long long x = b % 8, i = b / 8;
  return ((i / G) * 8 + x) * G + i % G;
}

// Addressing of an (order m, colatitude t, radius k, component nd) element
// of a Legendre-side buffer:
//   slot_off[m + L]*slot_mul + nd*s_nd + t*s_t + k*s_k.
struct GridAddr {
  const long long *slot_off;
  long long slot_mul, s_nd, s_t, s_k;
};
// Addressing of a spectral element (rlm index j, radius k, component nd):
// nd*s_nd + j*s_j + k*s_k.
struct SpecAddr {
  long long s_nd, s_j, s_k;
};
// Addressing of an FFT-side element: row(q)*q_stride + nd*s_nd + line*s_line,
// q = 0 (a0), 2m-1 (cos m), 2m (sin m), line = k + nr*t.  Rows are in Fourier
// order (row = q: the internal layout) or in FFTPACK slot order (a0 -> 0,
// cos m -> 2m, sin m -> 2m+1: row = q + 1 for q >= 1, the rtp layout), so the
// row offset is arithmetic and FFT kernels need no table lookup per element.
struct FourierAddr {
  int fftpack;
  long long q_stride, s_nd, s_line;
};
__host__ __device__ inline long long q_offset(const FourierAddr &xa, int q) {
  return (long long)((xa.fftpack && q) ? q + 1 : q) * xa.q_stride;
}
// Grid field (v_rtp): line + nline*i + nnod*nd.

struct Tables {
  // per order M = 0..L: offsets (in doubles) of the even/odd blocks of the
  // backward (i-major, theta fastest) and forward (theta-major, i fastest)
  // P tables, and the padded degree counts.
  long long *offb = nullptr, *offf = nullptr;  // device, 2*(L+1) each
  int *ipad = nullptr;                          // device, 2*(L+1)
  int *jst = nullptr;                           // device, 2L+1: lstack[s]
  double *Pb = nullptr, *dPb = nullptr, *Pf = nullptr, *dPf = nullptr;
  double *radius = nullptr, *a_r = nullptr;     // device, Nr
  double *weight = nullptr, *asin_t = nullptr;  // device, Ntheta
  double *cos_t = nullptr, *sin_t = nullptr;    // device, Ntheta
  double2 *tw = nullptr;                        // FFT twiddles
  double2 *twn = nullptr;                       // exp(-2 pi i k / Nphi), k < Nphi
  double2 *ztab = nullptr;                      // chirp-z form: roots, filters, chirp
  long long *slot_off_X = nullptr, *slot_off_rtm = nullptr;  // 2L+1 each
  double *pole = nullptr;  // P_l^0, dP_l^1/dtheta at theta = 0 and pi, 4 (L+1)
};

}  // namespace ksht

struct sht_plan {
  int L, nr, nth, nphi, nslot, jmax, nth_h, nth_hp, nr_p, max_ncomp;
  int precision, n_slices;
  int device;
  hipStream_t stream;
  hipStream_t own_stream;
  // side stream (non-blocking) and fork / join events: a call's scalar-group
  // Legendre launch runs there beside its vector group's (ForkJoin)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::vector<double> radius, gpoint, weight, colat, asin_t;
  std::vector<int> lstack;
  std::vector<long long> h_offb, h_offf;
  std::vector<int> h_ipad;
  long long p_table_elems_b, p_table_elems_f;
  ksht::Tables t;
  ksht::FftPlan fft;
  bool fft_generic = false;  // FFT needs radices other than 2, 3, 4, 8
  double *X = nullptr;        // field-major Fourier workspace
  long long x_elems = 0;
  long long device_bytes = 0;
  // split-precision state (allocated only in SHT_PREC_SPLIT_INT8 mode)
  void *split = nullptr;
  // P tables generated per call, one m-tile at a time, into tile buffers
  // (SHT_FLAG_P_ON_THE_FLY) instead of stored for all orders
  int p_otf = 0;
  std::vector<std::pair<int, int>> mtiles;          // [M0, M1) per tile
  long long *offb_r = nullptr, *offf_r = nullptr;   // device: offsets rebased to the tile
  double *Pt = nullptr, *dPt = nullptr;             // device: one tile of P and dP/dtheta
  long long tile_elems = 0;
  // kernel variants (sht_plan_set_variant): INT8 MFMA form of the forward
  // split GEMM, 16 (16x16x64) or 32 (32x32x32)
  int var_fwd_mfma = 16;
  // split backward: 1 = each vector group's launch also takes one scalar
  // (sharing its P slices), 0 = scalars in launches of their own
  int var_bwd_fused = 1;
  // SHT_FLAG_CHECK_FINITE: device counter of non-finite outputs (else null)
  unsigned long long *nonfinite = nullptr;
  // FFT form: 1 = split form n = r0 * 2^k where the plan has it, 0 = the
  // mixed-radix Stockham form for every length, 2 = the pruned real DFT as an
  // FP64 GEMM (set-up on first selection: sht_dft.hip)
  int var_fft_form = 1;
  void *dft = nullptr;
  // stage timers (sht_plan_timer_*): event pairs recorded per launch
  int timing = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> tev[ksht::kNumStages];
};

namespace ksht {
// Brackets one stage launch with events when the plan's timers are on.
// The plan's stream as HIP's event and stream-wait calls take it: the legacy
// default stream is bound as the sentinel hipStreamLegacy (kernel launches
// accept it), which these calls need as the null stream instead.
inline hipStream_t api_stream(hipStream_t s) { return s == hipStreamLegacy ? (hipStream_t)0 : s; }

// Fork / join of the plan's side stream around launches that may overlap:
// the side stream waits for everything queued on the plan's stream so far,
// and the plan's stream waits for the side stream's launches at the end of
// the scope.  Off (all on the plan's stream) when on is false, the plan has
// no side stream, or the plan's stream is being captured into a graph (a
// shell batch's graph runs its shells on parallel branches instead; forking
// the plan's side stream inside that capture crashed sht_batch_create, a
// cause not established, so the capture keeps the single-stream order).
struct ForkJoin {
  sht_plan *p;
  bool on;
  ForkJoin(sht_plan *p_, bool want) : p(p_), on(want && p_->side) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (on && (hipStreamIsCapturing(api_stream(p->stream), &cs) != hipSuccess ||
               cs != hipStreamCaptureStatusNone))
      on = false;
    if (on && (hipEventRecord(p->ev_fork, api_stream(p->stream)) != hipSuccess ||
               hipStreamWaitEvent(p->side, p->ev_fork, 0) != hipSuccess))
      on = false;  // ordering not set up: stay on the plan's stream
  }
  hipStream_t stream(bool beside) const { return on && beside ? p->side : p->stream; }
  ~ForkJoin() {
    if (on) {
      (void)hipEventRecord(p->ev_join, p->side);
      (void)hipStreamWaitEvent(api_stream(p->stream), p->ev_join, 0);
    }
  }
};

struct StageTimer {
  sht_plan *p;
  int id;
  hipEvent_t a = nullptr, b = nullptr;
  StageTimer(sht_plan *p_, int id_) : p(p_), id(id_) {
    if (p->timing && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      (void)hipEventRecord(a, api_stream(p->stream));
  }
  ~StageTimer() {
    if (b) {
      (void)hipEventRecord(b, api_stream(p->stream));
      p->tev[id].push_back({a, b});
    } else if (a) {
      (void)hipEventDestroy(a);
    }
  }
};
}  // namespace ksht

namespace ksht {
// Per-call workspace of a whole transform: the field-major Fourier workspace
// between the Legendre and FFT kernels and, for split-precision plans, the
// sliced data operand (digits and exponents).  Normally the plan's own
// buffers (plan_workspace); sht_backward_noalloc carves it from the library's
// fixed device array instead.
struct Workspace {
  double *X;
  signed char *dig;
  int *dexp;
};
Workspace plan_workspace(const sht_plan *p);
// bytes of the split engine's data-operand digits and exponents for a call of
// ncomp components in one direction
void split_ws_bytes(const sht_plan *p, int ncomp, bool backward, long long *dig_bytes,
                    long long *dexp_bytes);
// attaches the library's fixed device array on the current device (called at
// plan creation; sht_capi.hip)
int scratch_attach(int device);

// Legendre launchers (sht_legendre.hip)
int launch_leg_bwd(sht_plan *p, int nv, int ns, const double *sp, SpecAddr sa,
                   double *out, GridAddr ga);
int launch_leg_fwd(sht_plan *p, int nv, int ns, const double *in, GridAddr ga,
                   double *sp, SpecAddr sa);
// FFT launchers (sht_fft.hip)
int fft_plan_init(sht_plan *p);
int launch_fft_bwd(sht_plan *p, int ncomp, const double *X, FourierAddr xa,
                   double *v);
int launch_fft_fwd(sht_plan *p, int ncomp, const double *v, double *X,
                   FourierAddr xa);
// zonal transform as a DFT product (sht_dft.hip)
int dft_init(sht_plan *p);
int czt_init(sht_plan *p);
void dft_free(sht_plan *p);
int launch_dft_bwd(sht_plan *p, int ncomp, const double *X, FourierAddr xa, double *v);
int launch_dft_fwd(sht_plan *p, int ncomp, const double *v, double *X, FourierAddr xa);
// reorders (sht_reorder.hip)
int launch_zero_unused_rtp(sht_plan *p, int ncomp, double *rtp);
int launch_rtm_to_rtp(sht_plan *p, int ncomp, const double *rtm, double *rtp);
int launch_rtp_to_rtm(sht_plan *p, int ncomp, const double *rtp, double *rtm);
int launch_gather_rev(sht_plan *p, int ncomp, long long n_dst, long long n_src,
                      const int *irev, const double *src, double *dst);
// P table generation (sht_plan.hip)
int generate_p_tables(sht_plan *p);
// One m-tile of the FP64 P tables: orders [M0, M1), per-(M, parity) block
// offsets (indexed by the absolute order) into the table pointers.  With
// stored tables there is one tile of all orders; with tables per call each
// tile is generated (in the layout of the direction) when prepared.
struct PTile {
  int M0, M1;
  const long long *offb, *offf;
  const double *Pb, *dPb, *Pf, *dPf;
};
int ptile_count(const sht_plan *p);
int ptile_prepare(sht_plan *p, int i, bool fwd, PTile *t);
int pole_tables_init(sht_plan *p);
// split precision (sht_split.hip)
int split_init(sht_plan *p);
void split_free(sht_plan *p);
int launch_leg_bwd_split(sht_plan *p, const Workspace &ws, int nv, int ns, const double *sp,
                         SpecAddr sa, double *out, GridAddr ga);
int launch_leg_fwd_split(sht_plan *p, const Workspace &ws, int nv, int ns, const double *in,
                         GridAddr ga, double *sp, SpecAddr sa);
}  // namespace ksht
