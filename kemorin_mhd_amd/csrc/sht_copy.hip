// sht_copy.hip -- device copies between the field-major rj spectra of the
// MHD solver and the interleaved send / receive buffers of the transform.
//
// copy_mhd_spectr_to_send (MHD_src/sph_MHD/copy_sph_MHD_4_send_recv.f90:63-136)
// runs, per field of the transform, set_to_send_buf_{scalar,vector,tensor}
// (SOLVER_src/send_recv/field_to_send_buffer.f90:43-120):
//   WS(NB (k-1) + i_trans + c) = d_fld(inod_export(k), i_field + c);
// copy_mhd_spectr_from_recv (:140-206) runs set_from_recv_buf_rev_*
// (field_to_send_buffer.f90:246-...):
//   d_fld(k, i_field + c) = WR(NB (irev(k)-1) + i_trans + c), with the dummy
//   record irev = nnod_recv + 1 reading zero.
// Here all fields of one call go in one launch: a work-group stages the
// mapped components of kCopyNodes nodes in LDS, so that both the field-major
// side (node fastest) and the interleaved side (component fastest) are
// accessed in address order.  Pure copies: bit-exact.
#include "sht_internal.h"

namespace ksht {

constexpr int kCopyNodes = 128;
constexpr int kCopyThreads = 256;
constexpr int kMaxCopyComps = 64;

struct CopyMap {
  int n;                         // mapped components
  int fld[kMaxCopyComps];        // component in the field-major array
  int trn[kMaxCopyComps];        // component in the interleaved buffer
};

// interleaved <- field-major
__global__ void __launch_bounds__(kCopyThreads) k_to_send(CopyMap cm, long long nnod_fld,
                                                          const double *__restrict__ fld,
                                                          const int *__restrict__ inod,
                                                          long long nnod_send, int nb,
                                                          double *__restrict__ ws) {
  extern __shared__ double st[];  // [component][node]
  const long long k0 = (long long)blockIdx.x * kCopyNodes;
  const int nk = (int)min((long long)kCopyNodes, nnod_send - k0);
  for (int e = threadIdx.x; e < cm.n * kCopyNodes; e += kCopyThreads) {
    const int j = e / kCopyNodes, kk = e - j * kCopyNodes;
    if (kk < nk) {
      // a table entry outside 1..nnod_fld reads zero instead of faulting
      const long long src = inod ? (long long)inod[k0 + kk] - 1 : k0 + kk;
      st[e] = (src >= 0 && src < nnod_fld) ? fld[src + nnod_fld * cm.fld[j]] : 0.0;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < cm.n * nk; e += kCopyThreads) {
    const int kk = e / cm.n, j = e - kk * cm.n;
    ws[(k0 + kk) * nb + cm.trn[j]] = st[j * kCopyNodes + kk];
  }
}

// field-major <- interleaved
__global__ void __launch_bounds__(kCopyThreads) k_from_recv(CopyMap cm, long long nnod_recv,
                                                            int nb, const double *__restrict__ wr,
                                                            const int *__restrict__ irev,
                                                            long long nnod_fld,
                                                            double *__restrict__ fld) {
  extern __shared__ double st[];  // [component][node]
  const long long k0 = (long long)blockIdx.x * kCopyNodes;
  const int nk = (int)min((long long)kCopyNodes, nnod_fld - k0);
  for (int e = threadIdx.x; e < cm.n * nk; e += kCopyThreads) {
    const int kk = e / cm.n, j = e - kk * cm.n;
    const long long r = irev ? (long long)irev[k0 + kk] - 1 : k0 + kk;
    // the dummy record nnod_recv + 1 (and any entry outside the buffer) reads zero
    st[j * kCopyNodes + kk] = (r >= 0 && r < nnod_recv) ? wr[r * nb + cm.trn[j]] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < cm.n * kCopyNodes; e += kCopyThreads) {
    const int j = e / kCopyNodes, kk = e - j * kCopyNodes;
    if (kk < nk) fld[k0 + kk + nnod_fld * cm.fld[j]] = st[e];
  }
}

// Expands the field list into component pairs after checking every field.
static int make_map(const char *who, int nfield, const sht_field_copy *f, int ntot_phys,
                    int nb, CopyMap *cm) {
  if (nfield < 0 || (nfield > 0 && !f)) {
    set_error("%s: bad field list", who);
    return SHT_ERR_ARG;
  }
  cm->n = 0;
  for (int i = 0; i < nfield; i++) {
    const int nc = f[i].ncomp;
    if (nc != 1 && nc != 3 && nc != 6) {
      set_error("%s: field %d: ncomp = %d, need 1, 3 or 6", who, i, nc);
      return SHT_ERR_ARG;
    }
    if (f[i].i_field < 0 || f[i].i_field + nc > ntot_phys || f[i].i_trans < 0 ||
        f[i].i_trans + nc > nb) {
      set_error("%s: field %d: components %d..%d of %d (fields) or %d..%d of %d (buffer)",
                who, i, f[i].i_field, f[i].i_field + nc - 1, ntot_phys, f[i].i_trans,
                f[i].i_trans + nc - 1, nb);
      return SHT_ERR_ARG;
    }
    for (int c = 0; c < nc; c++) {
      if (cm->n == kMaxCopyComps) {
        set_error("%s: more than %d components", who, kMaxCopyComps);
        return SHT_ERR_ARG;
      }
      cm->fld[cm->n] = f[i].i_field + c;
      cm->trn[cm->n] = f[i].i_trans + c;
      cm->n++;
    }
  }
  return SHT_OK;
}

// Named component pairs (fld[i] -> trn[i]) of a field-major array to an
// interleaved send buffer, for callers that copy single components
// (copy_coriolis_terms_rlm, sht_coriolis.hip).
int launch_to_send_pairs(sht_plan *p, int n, const int *fld, const int *trn, long long nnod_fld,
                         const double *d_fld, const int *d_inod, long long nnod_send, int nb,
                         double *d_WS) {
  if (n < 0 || n > kMaxCopyComps) {
    set_error("to_send: %d components (1..%d)", n, kMaxCopyComps);
    return SHT_ERR_ARG;
  }
  CopyMap cm;
  cm.n = n;
  for (int i = 0; i < n; i++) {
    cm.fld[i] = fld[i];
    cm.trn[i] = trn[i];
  }
  if (n == 0 || nnod_send == 0) return SHT_OK;
  StageTimer tm(p, 4);
  const long long nblk = (nnod_send + kCopyNodes - 1) / kCopyNodes;
  hipLaunchKernelGGL(k_to_send, dim3((unsigned)nblk), dim3(kCopyThreads),
                     sizeof(double) * kCopyNodes * cm.n, p->stream, cm, nnod_fld, d_fld, d_inod,
                     nnod_send, nb, d_WS);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

}  // namespace ksht

using namespace ksht;

extern "C" int sht_rj_fields_to_send(sht_plan *p, int nfield, const sht_field_copy *fields,
                                     int ntot_phys, long long nnod_rj, const double *d_fld,
                                     const int *d_inod_export, long long nnod_send,
                                     int ncomp_send, double *d_WS) {
  static const char *who = "sht_rj_fields_to_send";
  if (!p || ntot_phys < 1 || ncomp_send < 1 || nnod_rj < 0 || nnod_send < 0 ||
      (!d_inod_export && nnod_send != nnod_rj)) {
    set_error("%s: bad arguments (identity export needs nnod_send = nnod_rj)", who);
    return SHT_ERR_ARG;
  }
  CopyMap cm;
  int rc = make_map(who, nfield, fields, ntot_phys, ncomp_send, &cm);
  if (rc) return rc;
  if (cm.n == 0 || nnod_send == 0) return SHT_OK;
  if (!d_fld || !d_WS) {
    set_error("%s: null buffer", who);
    return SHT_ERR_ARG;
  }
  StageTimer tm(p, 4);
  const long long nblk = (nnod_send + kCopyNodes - 1) / kCopyNodes;
  hipLaunchKernelGGL(k_to_send, dim3((unsigned)nblk), dim3(kCopyThreads),
                     sizeof(double) * kCopyNodes * cm.n, p->stream, cm, nnod_rj, d_fld,
                     d_inod_export, nnod_send, ncomp_send, d_WS);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}

extern "C" int sht_rj_fields_from_recv(sht_plan *p, int nfield, const sht_field_copy *fields,
                                       int ncomp_recv, long long nnod_recv, const double *d_WR,
                                       const int *d_irev, int ntot_phys, long long nnod_rj,
                                       double *d_fld) {
  static const char *who = "sht_rj_fields_from_recv";
  if (!p || ntot_phys < 1 || ncomp_recv < 1 || nnod_rj < 0 || nnod_recv < 0 ||
      (!d_irev && nnod_recv != nnod_rj)) {
    set_error("%s: bad arguments (identity receive needs nnod_recv = nnod_rj)", who);
    return SHT_ERR_ARG;
  }
  CopyMap cm;
  int rc = make_map(who, nfield, fields, ntot_phys, ncomp_recv, &cm);
  if (rc) return rc;
  if (cm.n == 0 || nnod_rj == 0) return SHT_OK;
  if (!d_fld || (!d_WR && nnod_recv > 0)) {
    set_error("%s: null buffer", who);
    return SHT_ERR_ARG;
  }
  StageTimer tm(p, 4);
  const long long nblk = (nnod_rj + kCopyNodes - 1) / kCopyNodes;
  hipLaunchKernelGGL(k_from_recv, dim3((unsigned)nblk), dim3(kCopyThreads),
                     sizeof(double) * kCopyNodes * cm.n, p->stream, cm, nnod_recv, ncomp_recv,
                     d_WR, d_irev, nnod_rj, d_fld);
  KSHT_HIP(hipGetLastError());
  return SHT_OK;
}
