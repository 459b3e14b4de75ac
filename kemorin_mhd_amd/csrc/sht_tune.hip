// sht_tune.hip -- tile-shape autotuner: times a plan's kernel variants on
// synthetic data of the caller's field sets and keeps the fastest.
//
// The reference's counterpart is select_legendre_transform
// (MHD_src/sph_MHD/init_sphrical_transform_MHD.f90:201-345), which runs every
// Legendre loop variant at start-up, times it (:240-260) and keeps the
// fastest for the run.  Here the variants are the kernel forms of one plan:
//   split_fwd_mfma  16 or 32: INT8 MFMA form of the forward split products
//                   (split-precision plans);
//   fft_form        1, 0, 2 or 3: split-form FFT (n = r0 2^k, where the
//                   length has it), mixed-radix Stockham FFT, the pruned real
//                   DFT as an FP64 GEMM (sht_dft.hip), or the whole line as
//                   one chirp-z (Bluestein) convolution of power-of-two
//                   length (sht_fft.hip).
// Every combination runs one backward and one forward transform untimed,
// then the best of three timed round trips (HIP events on the plan's
// stream) decides.  Results are cached in a JSON file, one entry per line:
//   {"key": "<shape>", "split_fwd_mfma": 16, "fft_form": 1, "ms": 51.2,
//    "default_ms": 52.9}
// keyed by L, Nr, Ntheta, Nphi, precision, slices, tables per call and the
// field sets, so a later plan of the same shape reads its winner instead of
// timing.  The outputs do not depend on the variant (each variant is held to
// the same parity bars), only the time.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sht_internal.h"

namespace ksht {

// deterministic smooth-ish fill, magnitudes O(1), no NaN or denormals
__global__ void k_tune_fill(double *x, long long n, double scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = scale * sin(0.7 * (double)(i % 100003) + 0.3);
}

static std::string tune_key(const sht_plan *p, int nvb, int nsb, int nvf, int nsf) {
  char buf[256];
  snprintf(buf, sizeof buf, "L=%d nr=%d nth=%d nphi=%d prec=%d slices=%d p_otf=%d bwd=%d,%d fwd=%d,%d",
           p->L, p->nr, p->nth, p->nphi, p->precision, p->n_slices, p->p_otf, nvb, nsb, nvf,
           nsf);
  return buf;
}

struct TuneEntry {
  std::string key;
  int mfma, form;
  double ms, default_ms;
};

static bool parse_entry(const char *line, TuneEntry *e) {
  const char *k = strstr(line, "\"key\": \"");
  if (!k) return false;
  k += 8;
  const char *q = strchr(k, '"');
  if (!q) return false;
  e->key.assign(k, q - k);
  const char *a = strstr(q, "\"split_fwd_mfma\": "), *b = strstr(q, "\"fft_form\": ");
  const char *c = strstr(q, "\"ms\": "), *d = strstr(q, "\"default_ms\": ");
  if (!a || !b || !c || !d) return false;
  e->mfma = atoi(a + 18);
  e->form = atoi(b + 12);
  e->ms = atof(c + 6);
  e->default_ms = atof(d + 14);
  return true;
}

static std::vector<TuneEntry> read_cache(const char *path) {
  std::vector<TuneEntry> out;
  FILE *f = fopen(path, "r");
  if (!f) return out;
  char line[1024];
  while (fgets(line, sizeof line, f)) {
    TuneEntry e;
    if (parse_entry(line, &e)) out.push_back(e);
  }
  fclose(f);
  return out;
}

static int write_cache(const char *path, const std::vector<TuneEntry> &es) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) {
    set_error("sht_plan_autotune: cannot write %s", tmp.c_str());
    return SHT_ERR_ARG;
  }
  fprintf(f, "{\"kemorin_sht_autotune\": 1, \"entries\": [\n");
  for (size_t i = 0; i < es.size(); i++)
    fprintf(f,
            "{\"key\": \"%s\", \"split_fwd_mfma\": %d, \"fft_form\": %d, \"ms\": %.4f, "
            "\"default_ms\": %.4f}%s\n",
            es[i].key.c_str(), es[i].mfma, es[i].form, es[i].ms, es[i].default_ms,
            i + 1 < es.size() ? "," : "");
  fprintf(f, "]}\n");
  fclose(f);
  if (rename(tmp.c_str(), path) != 0) {
    set_error("sht_plan_autotune: cannot replace %s", path);
    return SHT_ERR_ARG;
  }
  return SHT_OK;
}

}  // namespace ksht

using namespace ksht;

extern "C" int sht_plan_autotune(sht_plan *p, int nvb, int nsb, int nvf, int nsf,
                                 const char *cache_path, double *ms_out) {
  static const char *who = "sht_plan_autotune";
  if (!p) {
    set_error("%s: null plan", who);
    return SHT_ERR_ARG;
  }
  const int ncb = 3 * nvb + nsb, ncf = 3 * nvf + nsf;
  if (nvb < 0 || nsb < 0 || nvf < 0 || nsf < 0 || ncb + ncf == 0 || ncb > p->max_ncomp ||
      ncf > p->max_ncomp) {
    set_error("%s: field sets (%d, %d) / (%d, %d) outside 1..max_ncomp = %d components", who,
              nvb, nsb, nvf, nsf, p->max_ncomp);
    return SHT_ERR_ARG;
  }
  const std::string key = tune_key(p, nvb, nsb, nvf, nsf);
  std::vector<TuneEntry> cache;
  if (cache_path) {
    cache = read_cache(cache_path);
    for (const TuneEntry &e : cache)
      if (e.key == key) {
        int rc = sht_plan_set_variant(p, "split_fwd_mfma", e.mfma);
        if (!rc) rc = sht_plan_set_variant(p, "fft_form", e.form);
        if (rc) return rc;
        if (ms_out) {
          ms_out[0] = e.ms;
          ms_out[1] = e.default_ms;
        }
        return SHT_OK;
      }
  }
  // candidate values of each knob that applies to this plan
  std::vector<int> mfmas = {16}, forms = {p->fft.fast ? 1 : 0};
  if (p->precision == SHT_PREC_SPLIT_INT8) mfmas.push_back(32);
  if (p->fft.fast) forms.push_back(0);
  if (dft_init(p) == SHT_OK) forms.push_back(2);  // the DFT product, where its set-up fits
  if (czt_init(p) == SHT_OK) forms.push_back(3);  // the chirp-z form, Nphi + 2L <= 2048
  const int keep_mfma = p->var_fwd_mfma, keep_form = p->var_fft_form;
  const long long nnod = (long long)p->nr * p->nth * p->nphi;
  const long long nsp = (long long)p->jmax * p->nr;
  const int nc = ncb > ncf ? ncb : ncf;
  double *sp = nullptr, *grid = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  int rc = SHT_OK;
  double best = 1e300, def = 0.0;
  int best_mfma = keep_mfma, best_form = keep_form;
  auto fail = [&](int code, const char *what) {
    set_error("%s: %s", who, what);
    rc = code;
  };
  if (hipMalloc(&sp, sizeof(double) * nc * nsp) != hipSuccess ||
      hipMalloc(&grid, sizeof(double) * nc * nnod) != hipSuccess) {
    fail(SHT_ERR_OOM, "device buffers for the timing runs");
    (void)hipGetLastError();  // the failed allocation must not fail the next launch check
  } else if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    fail(SHT_ERR_HIP, "events");
  } else {
    const long long ns = (long long)nc * nsp, ng = (long long)nc * nnod;
    hipLaunchKernelGGL(k_tune_fill, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, p->stream,
                       sp, ns, 0.01);
    hipLaunchKernelGGL(k_tune_fill, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, p->stream,
                       grid, ng, 1.0);
    double *out_grid = grid, *out_sp = sp;  // in place: contents do not matter for timing
    for (int m : mfmas) {
      for (int fm : forms) {
        p->var_fwd_mfma = m;
        p->var_fft_form = fm;
        auto trip = [&]() -> int {
          int r = SHT_OK;
          if (ncb) r = sht_backward(p, nvb, nsb, sp, out_grid);
          if (!r && ncf) r = sht_forward(p, nvf, nsf, grid, out_sp);
          return r;
        };
        if ((rc = trip())) break;
        double ms_best = 1e300;
        for (int rep = 0; rep < 3 && !rc; rep++) {
          (void)hipEventRecord(a, api_stream(p->stream));
          rc = trip();
          (void)hipEventRecord(b, api_stream(p->stream));
          if (hipEventSynchronize(b) != hipSuccess) {
            fail(SHT_ERR_HIP, "timing run failed");
            break;
          }
          float ms = 0.0f;
          (void)hipEventElapsedTime(&ms, a, b);
          if (ms < ms_best) ms_best = ms;
        }
        if (rc) break;
        if (m == mfmas[0] && fm == forms[0]) def = ms_best;
        if (ms_best < best) {
          best = ms_best;
          best_mfma = m;
          best_form = fm;
        }
      }
      if (rc) break;
    }
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  (void)hipStreamSynchronize(api_stream(p->stream));
  if (sp) (void)hipFree(sp);
  if (grid) (void)hipFree(grid);
  if (rc) {
    p->var_fwd_mfma = keep_mfma;
    p->var_fft_form = keep_form;
    return rc;
  }
  p->var_fwd_mfma = best_mfma;
  p->var_fft_form = best_form;
  if (ms_out) {
    ms_out[0] = best;
    ms_out[1] = def;
  }
  if (cache_path) {
    TuneEntry e{key, best_mfma, best_form, best, def};
    bool found = false;
    for (TuneEntry &x : cache)
      if (x.key == key) {
        x = e;
        found = true;
      }
    if (!found) cache.push_back(e);
    return write_cache(cache_path, cache);
  }
  return SHT_OK;
}
