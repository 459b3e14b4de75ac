// sht_batch.hip -- many independent shells replayed as one HIP graph.
//
// The reference transforms one shell per call (SS/t_sph_transforms.f90:74,
// 250); a solver holding many small shells of differing truncation would pay
// one launch sequence per shell and leave most of the device idle on each.
// A batch captures, once, the backward and forward transforms of every shell
// (each shell its own plan and buffers) into one graph: the shells are dealt
// to kBranches parallel branches forked from and joined to one capture
// stream (calls that share a plan stay on one branch, in order), so
// independent shells may run side by side, and the whole set starts with a
// single hipGraphLaunch.  Plans go to branches largest first, each to the
// branch with the least estimated work so far, so that the branches end
// together.  Replays reuse the captured buffer pointers; the caller refills
// the inputs between launches.
#include <algorithm>
#include <vector>

#include "sht_internal.h"

using namespace ksht;

namespace {
constexpr int kBranches = 4;

// Work estimate of one call, for balancing the branches only: the Legendre
// contractions (nr nth (L+1)^2 per component, 5 per vector) and the zonal
// transform (nr nth nphi per component, weighted as measured: about 200
// Legendre multiply-adds per FFT point and component).
double call_cost(const sht_shell_call &c) {
  const sht_plan *p = c.plan;
  const double lines = (double)p->nr * p->nth, l2 = (double)(p->L + 1) * (p->L + 1);
  auto dir = [&](int nv, int ns) {
    return lines * (l2 * (5.0 * nv + ns) + 200.0 * p->nphi * (3.0 * nv + ns));
  };
  return dir(c.nvector_b, c.nscalar_b) + dir(c.nvector_f, c.nscalar_f);
}
}

struct sht_batch {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t own_stream = nullptr;
  int device = 0;
};

static void batch_free(sht_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
  delete b;
}

static int check_shell(const sht_shell_call &c, int i) {
  const sht_plan *p = c.plan;
  if (!p) {
    set_error("sht_batch_create: shell %d has no plan", i);
    return SHT_ERR_ARG;
  }
  if (p->timing) {
    set_error("sht_batch_create: shell %d's plan has its stage timers on (events are not captured)",
              i);
    return SHT_ERR_STATE;
  }
  const int nb = 3 * c.nvector_b + c.nscalar_b, nf = 3 * c.nvector_f + c.nscalar_f;
  if (c.nvector_b < 0 || c.nscalar_b < 0 || c.nvector_f < 0 || c.nscalar_f < 0 ||
      nb > p->max_ncomp || nf > p->max_ncomp || nb + nf == 0) {
    set_error("sht_batch_create: shell %d: field sets (%d,%d) / (%d,%d) outside 1..max_ncomp %d", i,
              c.nvector_b, c.nscalar_b, c.nvector_f, c.nscalar_f, p->max_ncomp);
    return SHT_ERR_ARG;
  }
  if ((nb && (!c.d_WS_rj || !c.d_v_rtp)) || (nf && (!c.d_v_in || !c.d_WR_rj))) {
    set_error("sht_batch_create: shell %d: null buffer", i);
    return SHT_ERR_ARG;
  }
  return SHT_OK;
}

extern "C" int sht_batch_create(const sht_shell_call *calls, int n, sht_batch **out) {
  if (!out || !calls || n < 1) {
    set_error("sht_batch_create: null argument or n = %d < 1", n);
    return SHT_ERR_ARG;
  }
  *out = nullptr;
  for (int i = 0; i < n; i++) {
    const int rc = check_shell(calls[i], i);
    if (rc) return rc;
    if (calls[i].plan->device != calls[0].plan->device) {
      set_error("sht_batch_create: shell %d is on device %d, shell 0 on %d", i,
                calls[i].plan->device, calls[0].plan->device);
      return SHT_ERR_ARG;
    }
  }
  int dev0 = 0;
  KSHT_HIP(hipGetDevice(&dev0));
  KSHT_HIP(hipSetDevice(calls[0].plan->device));
  sht_batch *b = new sht_batch;
  b->device = calls[0].plan->device;
  hipStream_t cap = nullptr, br[kBranches] = {};
  hipEvent_t fork = nullptr, join[kBranches] = {};
  std::vector<hipStream_t> saved(n);
  for (int i = 0; i < n; i++) saved[i] = calls[i].plan->stream;
  int rc = SHT_OK;
  bool capturing = false;
  auto hip_ok = [&](hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    set_error("sht_batch_create: %s: %s", what, hipGetErrorString(e));
    rc = SHT_ERR_HIP;
    return false;
  };
  do {
    if (!hip_ok(hipStreamCreateWithFlags(&b->own_stream, hipStreamDefault), "stream")) break;
    if (!hip_ok(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "capture stream")) break;
    bool ok = true;
    for (int k = 0; k < kBranches && ok; k++)
      ok = hip_ok(hipStreamCreateWithFlags(&br[k], hipStreamNonBlocking), "branch stream") &&
           hip_ok(hipEventCreateWithFlags(&join[k], hipEventDisableTiming), "event");
    if (!ok || !hip_ok(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "event")) break;
    if (!hip_ok(hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed), "begin capture")) break;
    capturing = true;
    if (!hip_ok(hipEventRecord(fork, cap), "fork")) break;
    for (int k = 0; k < kBranches && ok; k++) ok = hip_ok(hipStreamWaitEvent(br[k], fork, 0), "fork");
    if (!ok) break;
    // calls on one plan share its workspace: they go to one branch, in
    // their order.  Plans in decreasing estimated work, each to the least
    // loaded branch; calls captured in that plan order
    std::vector<int> first(n), order(n), branch(n);
    std::vector<double> pcost(n, 0.0);
    for (int i = 0; i < n; i++) {
      first[i] = i;
      for (int k = 0; k < i; k++)
        if (calls[k].plan == calls[i].plan) {
          first[i] = first[k];
          break;
        }
      pcost[first[i]] += call_cost(calls[i]);
    }
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      const double ca = pcost[first[a]], cb = pcost[first[b]];
      return ca != cb ? ca > cb : first[a] < first[b];
    });
    double load[kBranches] = {};
    for (int oi = 0; oi < n; oi++) {
      const int i = order[oi];
      if (first[i] == i) {
        int k0 = 0;
        for (int k = 1; k < kBranches; k++)
          if (load[k] < load[k0]) k0 = k;
        branch[i] = k0;
        load[k0] += pcost[i];
      } else {
        branch[i] = branch[first[i]];
      }
    }
    for (int oi = 0; oi < n && !rc; oi++) {
      const int i = order[oi];
      const sht_shell_call &c = calls[i];
      c.plan->stream = br[branch[i]];
      if (3 * c.nvector_b + c.nscalar_b)
        rc = sht_backward(c.plan, c.nvector_b, c.nscalar_b, c.d_WS_rj, c.d_v_rtp);
      if (!rc && 3 * c.nvector_f + c.nscalar_f)
        rc = sht_forward(c.plan, c.nvector_f, c.nscalar_f, c.d_v_in, c.d_WR_rj);
    }
    if (rc) break;
    for (int k = 0; k < kBranches && ok; k++)
      ok = hip_ok(hipEventRecord(join[k], br[k]), "join") &&
           hip_ok(hipStreamWaitEvent(cap, join[k], 0), "join");
    if (!ok) break;
    capturing = false;
    if (!hip_ok(hipStreamEndCapture(cap, &b->graph), "end capture")) break;
    if (!hip_ok(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "instantiate")) break;
  } while (false);
  if (capturing) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(cap, &g);  // abandon the partial capture
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
  }
  for (int i = 0; i < n; i++) calls[i].plan->stream = saved[i];
  for (int k = 0; k < kBranches; k++) {
    if (join[k]) (void)hipEventDestroy(join[k]);
    if (br[k]) (void)hipStreamDestroy(br[k]);
  }
  if (fork) (void)hipEventDestroy(fork);
  if (cap) (void)hipStreamDestroy(cap);
  (void)hipSetDevice(dev0);
  if (rc) {
    batch_free(b);
    return rc;
  }
  *out = b;
  return SHT_OK;
}

extern "C" int sht_batch_launch(sht_batch *b, void *hip_stream) {
  if (!b) {
    set_error("sht_batch_launch: null batch");
    return SHT_ERR_ARG;
  }
  KSHT_HIP(hipGraphLaunch(b->exec, hip_stream ? (hipStream_t)hip_stream : b->own_stream));
  return SHT_OK;
}

extern "C" int sht_batch_synchronize(sht_batch *b) {
  if (!b) {
    set_error("sht_batch_synchronize: null batch");
    return SHT_ERR_ARG;
  }
  KSHT_HIP(hipStreamSynchronize(b->own_stream));
  return SHT_OK;
}

extern "C" int sht_batch_destroy(sht_batch *b) {
  batch_free(b);
  return SHT_OK;
}
