// sht_files.hip -- plans driven by the reference's spherical grid / mode files.
//
// The reference reads each rank's grids and modes from four ASCII files that
// gen_sph_grids writes, one per space: rtp (grid), rtm, rlm and rj (modes).
// Readers: read_geom_rtp_data, read_spectr_modes_rj_data, read_geom_rtm_data,
// read_spectr_modes_rlm_data (CORE_src/IO/sph_modes_grids_data_IO.f90:53-183)
// and below them read_domain_info / read_import_data (domain_data_IO.f90:
// 25-80), read_gl_resolution_sph / read_rank_4_sph / read_gl_nodes_sph
// (spherical_model_IO.f90:22-79), read_rtp_gl_1d_table / read_rj_gl_1d_table
// (sph_global_1d_idx_IO.f90:22-100), read_arrays_for_stacks
// (field_data_IO.f90:164-200), read_send_recv_item (comm_table_IO.f90:21-36);
// skip_comment (skip_comment_f.f90:39-51) skips blank lines and lines that
// start with '!' or '#'.  Each file holds, in order: the domain (rank,
// neighbour count, neighbours), the global sizes and truncation, the
// subdomain index per direction, the 1-D global index tables (radii with
// their index, theta, and phi or m; or radii and the (j, l, m) mode table),
// the global node list, the import table (stack, items) and groups (not
// read here).
//
// Value reads follow Fortran list-directed input: a read that follows
// skip_comment takes its values from that one line; a plain record read
// starts on the next line, continues over lines until it has its values and
// drops the rest of its last line.
//
// A plan is driven by a set when it is one subdomain per direction of the
// single-rank layout this engine uses (m_folding = 1); the check compares
// every table with the one gen_sph_grids builds for one rank (orderings of
// CONST_SPH_GRID/set_sph_tranform_ordering.f90:43-293, node lists of
// set_local_sphere_param.f90:45-171, 1-D tables of
// set_sph_1d_global_index.f90:45-222) and the communication tables by what
// they pair (rtm node <-> rtp node of the same radius and colatitude whose
// phi index is the rtm slot's FFT slot, set_comm_table_rtp_rj.f90:499-569;
// rlm node <-> rj node of the same radius and mode), and reports the first
// difference.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sht_internal.h"

struct sht_sph_files {
  struct Space {
    std::string path;
    int numdir = 0;
    int rank = -1, num_neib = 0;
    std::vector<int> id_neib;
    int nidx_gl[3] = {0, 0, 0}, ltr = -1, sph_rank[3] = {0, 0, 0};
    int nidx[3] = {0, 0, 0}, ist[3] = {0, 0, 0}, ied[3] = {0, 0, 0};
    std::vector<int> idx1;       // radial global index
    std::vector<double> r1;      // radius
    std::vector<int> idx2;       // theta (1 per entry) or (j, l, m) (3 per entry)
    std::vector<int> idx3;       // (phi, m) or (FFT slot, m): 2 per entry
    long long numnod = 0;
    std::vector<long long> inod_gl;
    std::vector<int> idx_node;   // numdir per node
    std::vector<int> istack;     // import stack (num_neib entries)
    std::vector<int> item;       // import items, 1-based node numbers
  };
  Space s[4];  // rtp, rtm, rlm, rj
};

namespace {

enum { kRtp = 0, kRtm = 1, kRlm = 2, kRj = 3 };
const char *kSpaceName[4] = {"rtp", "rtm", "rlm", "rj"};

struct Reader {
  FILE *f = nullptr;
  std::string path;
  long line = 0;
  std::string buf;
  bool fail = false;

  bool next_line() {
    buf.clear();
    int c;
    bool any = false;
    while ((c = fgetc(f)) != EOF) {
      any = true;
      if (c == '\n') break;
      if (c != '\r') buf.push_back((char)c);
    }
    if (!any) return false;
    line++;
    return true;
  }
  // skip_comment: the next line that is not blank and does not start with
  // '!' or '#' (after leading blanks)
  bool data_line() {
    while (next_line()) {
      size_t i = buf.find_first_not_of(" \t,");
      if (i == std::string::npos) continue;
      if (buf[i] == '!' || buf[i] == '#') continue;
      return true;
    }
    return false;
  }
  static void split(const std::string &s, std::vector<std::string> &out) {
    out.clear();
    size_t i = 0;
    while (i < s.size()) {
      while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == ',')) i++;
      if (i >= s.size()) break;
      size_t j = i;
      while (j < s.size() && s[j] != ' ' && s[j] != '\t' && s[j] != ',') j++;
      out.push_back(s.substr(i, j - i));
      i = j;
    }
  }
  bool err(const char *what) {
    if (!fail) ksht::set_error("%s:%ld: %s", path.c_str(), line, what);
    fail = true;
    return false;
  }
  static bool to_ll(const std::string &t, long long *v) {
    char *e = nullptr;
    errno = 0;
    const long long x = strtoll(t.c_str(), &e, 10);
    if (errno || e == t.c_str() || *e) return false;
    *v = x;
    return true;
  }
  static bool to_d(const std::string &t, double *v) {
    std::string u = t;
    for (char &c : u)
      if (c == 'd' || c == 'D') c = 'e';  // Fortran double-precision exponent
    char *e = nullptr;
    errno = 0;
    const double x = strtod(u.c_str(), &e);
    if (e == u.c_str() || *e) return false;
    *v = x;
    return true;
  }
  // values of the line after skip_comment (at least n of them)
  bool line_vals(int n, std::vector<std::string> &tok, const char *what) {
    if (!data_line()) return err(what);
    split(buf, tok);
    if ((int)tok.size() < n) return err(what);
    return true;
  }
  // one list-directed record read of n values
  bool record(int n, std::vector<std::string> &tok, const char *what) {
    tok.clear();
    std::vector<std::string> t;
    while ((int)tok.size() < n) {
      if (!next_line()) return err(what);
      split(buf, t);
      for (auto &x : t) tok.push_back(x);
    }
    tok.resize(n);
    return true;
  }
  bool ints(const std::vector<std::string> &tok, int n, long long *v, const char *what) {
    for (int i = 0; i < n; i++)
      if (!to_ll(tok[i], &v[i])) return err(what);
    return true;
  }
};

bool read_space(sht_sph_files::Space &sp, int kind) {
  Reader r;
  r.path = sp.path;
  r.f = fopen(sp.path.c_str(), "r");
  if (!r.f) {
    ksht::set_error("%s: cannot open (%s file)", sp.path.c_str(), kSpaceName[kind]);
    return false;
  }
  const bool grid = kind == kRtp || kind == kRtm;
  sp.numdir = grid ? 3 : 2;
  std::vector<std::string> t;
  long long v[8];
  bool ok = true;
  // domain: rank, neighbour count, neighbours
  ok = ok && r.line_vals(1, t, "domain rank") && r.ints(t, 1, v, "domain rank");
  if (ok) sp.rank = (int)v[0];
  ok = ok && r.record(1, t, "neighbour count") && r.ints(t, 1, v, "neighbour count");
  if (ok) sp.num_neib = (int)v[0];
  if (ok && (sp.num_neib < 0 || sp.num_neib > 1 << 20)) ok = r.err("neighbour count out of range");
  if (ok && sp.num_neib > 0) {
    ok = r.record(sp.num_neib, t, "neighbour ranks");
    for (int i = 0; ok && i < sp.num_neib; i++) {
      ok = r.ints({t[i]}, 1, v, "neighbour rank");
      if (ok) sp.id_neib.push_back((int)v[0]);
    }
  }
  // global sizes, truncation, subdomain index
  ok = ok && r.line_vals(sp.numdir, t, "global sizes") && r.ints(t, sp.numdir, v, "global sizes");
  for (int i = 0; ok && i < sp.numdir; i++) sp.nidx_gl[i] = (int)v[i];
  ok = ok && r.line_vals(1, t, "truncation") && r.ints(t, 1, v, "truncation");
  if (ok) sp.ltr = (int)v[0];
  ok = ok && r.line_vals(sp.numdir, t, "subdomain index") && r.ints(t, sp.numdir, v, "subdomain index");
  for (int i = 0; ok && i < sp.numdir; i++) sp.sph_rank[i] = (int)v[i];
  // 1-D tables
  const int ncomp[3] = {1, grid ? 1 : 3, 2};
  for (int d = 0; ok && d < sp.numdir; d++) {
    ok = r.line_vals(3, t, "1-D table size") && r.ints(t, 3, v, "1-D table size");
    if (!ok) break;
    sp.nidx[d] = (int)v[0];
    sp.ist[d] = (int)v[1];
    sp.ied[d] = (int)v[2];
    if (sp.nidx[d] < 0 || sp.nidx[d] > (1 << 26)) {
      ok = r.err("1-D table size out of range");
      break;
    }
    for (int i = 0; ok && i < sp.nidx[d]; i++) {
      if (d == 0) {
        ok = r.record(2, t, "radial table entry") && r.ints(t, 1, v, "radial index");
        double rr = 0.0;
        if (ok && !Reader::to_d(t[1], &rr)) ok = r.err("radius");
        if (ok) {
          sp.idx1.push_back((int)v[0]);
          sp.r1.push_back(rr);
        }
      } else {
        ok = r.record(ncomp[d], t, "1-D table entry") && r.ints(t, ncomp[d], v, "1-D table entry");
        for (int c = 0; ok && c < ncomp[d]; c++) (d == 1 ? sp.idx2 : sp.idx3).push_back((int)v[c]);
      }
    }
  }
  // global node list
  ok = ok && r.line_vals(1, t, "node count") && r.ints(t, 1, v, "node count");
  if (ok) sp.numnod = v[0];
  if (ok && (sp.numnod < 0 || sp.numnod > (1ll << 34))) ok = r.err("node count out of range");
  for (long long i = 0; ok && i < sp.numnod; i++) {
    ok = i == 0 ? r.line_vals(1 + sp.numdir, t, "node entry") : r.record(1 + sp.numdir, t, "node entry");
    ok = ok && r.ints(t, 1 + sp.numdir, v, "node entry");
    if (!ok) break;
    sp.inod_gl.push_back(v[0]);
    for (int c = 0; c < sp.numdir; c++) sp.idx_node.push_back((int)v[1 + c]);
  }
  // import table: stack (num_neib values; the first from the skip_comment
  // line, the rest from records), then items
  if (ok && sp.num_neib > 0) {
    ok = r.data_line();
    if (!ok) r.err("import stack");
    std::vector<std::string> first;
    if (ok) Reader::split(r.buf, first);
    for (size_t i = 0; ok && i < first.size() && (int)sp.istack.size() < sp.num_neib; i++) {
      ok = r.ints({first[i]}, 1, v, "import stack");
      if (ok) sp.istack.push_back((int)v[0]);
    }
    while (ok && (int)sp.istack.size() < sp.num_neib) {
      ok = r.record(1, t, "import stack") && r.ints(t, 1, v, "import stack");
      if (ok) sp.istack.push_back((int)v[0]);
    }
    const long long ntot = ok ? sp.istack.back() : 0;
    if (ok && (ntot < 0 || ntot > sp.numnod)) ok = r.err("import item count out of range");
    for (long long i = 0; ok && i < ntot; i++) {
      ok = i == 0 ? r.line_vals(1, t, "import item") : r.record(1, t, "import item");
      ok = ok && r.ints(t, 1, v, "import item");
      if (ok) sp.item.push_back((int)v[0]);
    }
  }
  fclose(r.f);
  return ok;
}

// FFTPACK slot (1-based) of order m for the ispack ordering with half
// length nth (set_wavenumber_4_ispack_fft): 0 -> 1, nth -> 2, m -> 2m+1,
// -m -> 2m+2
int fft_slot(int m, int nth) {
  if (m == 0) return 1;
  if (m == nth) return 2;
  return m > 0 ? 2 * m + 1 : -2 * m + 2;
}

bool bad(const sht_sph_files::Space &sp, const char *fmt, long long a = 0, long long b = 0,
         long long c = 0) {
  char msg[256];
  snprintf(msg, sizeof msg, fmt, a, b, c);
  ksht::set_error("%s (%s): %s", sp.path.c_str(), "sht_sph_files_check", msg);
  return false;
}

bool check_set(const sht_sph_files *f) {
  const auto &rtp = f->s[kRtp], &rtm = f->s[kRtm], &rlm = f->s[kRlm], &rj = f->s[kRj];
  const int L = rj.ltr, nr = rtp.nidx_gl[0], nth = rtp.nidx_gl[1], nphi = rtp.nidx_gl[2];
  const long long jmax = (long long)(L + 1) * (L + 1);
  // one rank: every space is the whole domain, its neighbour is itself
  for (int k = 0; k < 4; k++) {
    const auto &sp = f->s[k];
    if (sp.ltr != L) return bad(sp, "truncation %lld differs from the rj file's %lld", sp.ltr, L);
    if (sp.rank != rj.rank) return bad(sp, "rank %lld differs from the rj file's %lld", sp.rank, rj.rank);
    for (int d = 0; d < sp.numdir; d++) {
      if (sp.sph_rank[d] != 0)
        return bad(sp, "subdomain index %lld in direction %lld: not a one-rank set", sp.sph_rank[d], d + 1);
      if (sp.nidx[d] != sp.nidx_gl[d] || sp.ist[d] != 1 || sp.ied[d] != sp.nidx_gl[d])
        return bad(sp, "direction %lld holds %lld of %lld global entries: not a one-rank set", d + 1,
                   sp.nidx[d], sp.nidx_gl[d]);
    }
    if (sp.num_neib > 1 || (sp.num_neib == 1 && sp.id_neib[0] != sp.rank))
      return bad(sp, "%lld neighbours (first %lld): not a one-rank set", sp.num_neib,
                 sp.num_neib ? sp.id_neib[0] : -1);
  }
  if (L < 1 || nr < 1) return bad(rj, "truncation %lld, %lld radii", L, nr);
  if (nphi != 2 * nth) return bad(rtp, "Nphi = %lld is not 2 Ntheta = %lld", nphi, 2 * nth);
  if (rtm.nidx_gl[0] != nr || rtm.nidx_gl[1] != nth || rtm.nidx_gl[2] != 2 * L + 1)
    return bad(rtm, "global sizes (%lld, %lld, %lld) do not match the rtp file and 2L+1 slots",
               rtm.nidx_gl[0], rtm.nidx_gl[1], rtm.nidx_gl[2]);
  for (const auto *sp : {&rlm, &rj})
    if (sp->nidx_gl[0] != nr || sp->nidx_gl[1] != jmax)
      return bad(*sp, "global sizes (%lld, %lld): need (Nr, (L+1)^2 = %lld)", sp->nidx_gl[0],
                 sp->nidx_gl[1], jmax);
  // radial tables: index k, the same radii in every file
  for (int s = 0; s < 4; s++)
    for (int k = 0; k < nr; k++) {
      const auto &sp = f->s[s];
      if (sp.idx1[k] != k + 1) return bad(sp, "radial entry %lld has index %lld", k + 1, sp.idx1[k]);
      if (sp.r1[k] != rj.r1[k] || !(sp.r1[k] > 0.0))
        return bad(sp, "radius %lld differs from the rj file's (or is not positive)", k + 1);
    }
  for (const auto *sp : {&rtp, &rtm})
    for (int l = 0; l < nth; l++)
      if (sp->idx2[l] != l + 1) return bad(*sp, "colatitude entry %lld has index %lld", l + 1, sp->idx2[l]);
  // phi table: (i, m of FFTPACK slot i)
  for (int i = 1; i <= nphi; i++) {
    int m = 0;
    for (int mm = -nth + 1; mm <= nth; mm++)
      if (fft_slot(mm, nth) == i) m = mm;
    if (rtp.idx3[2 * (i - 1)] != i || rtp.idx3[2 * (i - 1) + 1] != m)
      return bad(rtp, "phi entry %lld is (%lld, ...): need (i, m of slot i); m_folding = 1 only", i,
                 rtp.idx3[2 * (i - 1)]);
  }
  // rtm slots: order m = -L..L, each with its FFT slot
  for (int s = 0; s <= 2 * L; s++) {
    const int m = s - L;
    if (rtm.idx3[2 * s] != fft_slot(m, nth) || rtm.idx3[2 * s + 1] != m)
      return bad(rtm, "m-slot %lld is (%lld, %lld): need slots in order m = -L..L", s + 1,
                 rtm.idx3[2 * s], rtm.idx3[2 * s + 1]);
  }
  // rlm and rj mode tables: m = -L..L, l = |m|..L, j = l(l+1)+m
  for (const auto *sp : {&rlm, &rj}) {
    long long j = 0;
    for (int m = -L; m <= L; m++)
      for (int l = (m < 0 ? -m : m); l <= L; l++, j++) {
        const int *e = &sp->idx2[3 * j];
        if (e[0] != l * (l + 1) + m || e[1] != l || e[2] != m)
          return bad(*sp, "mode %lld is (%lld, %lld, ...): need the one-rank order m = -L..L, "
                     "l = |m|..L",
                     j + 1, e[0], e[1]);
      }
  }
  // node lists
  {
    long long n = 0;
    for (int m = 1; m <= nphi; m++)
      for (int l = 1; l <= nth; l++)
        for (int k = 1; k <= nr; k++, n++) {
          if (n >= rtp.numnod) return bad(rtp, "%lld nodes, need more", rtp.numnod);
          const int *e = &rtp.idx_node[3 * n];
          if (e[0] != k || e[1] != l || e[2] != m)
            return bad(rtp, "node %lld: need (k, l, phi) with radius fastest", n + 1);
        }
    if (rtp.numnod != n) return bad(rtp, "%lld nodes, need %lld", rtp.numnod, n);
    n = 0;
    for (int s = 0; s <= 2 * L; s++)
      for (int k = 1; k <= nr; k++)
        for (int l = 1; l <= nth; l++, n++) {
          if (n >= rtm.numnod) return bad(rtm, "%lld nodes, need more", rtm.numnod);
          const int *e = &rtm.idx_node[3 * n];
          if (e[0] != k || e[1] != l || e[2] != rtm.idx3[2 * s])
            return bad(rtm, "node %lld: need (k, l, slot) with colatitude fastest", n + 1);
        }
    if (rtm.numnod != n) return bad(rtm, "%lld nodes, need %lld", rtm.numnod, n);
    for (const auto *sp : {&rlm, &rj}) {
      n = 0;
      for (int k = 1; k <= nr; k++)
        for (long long j = 0; j < jmax; j++, n++) {
          if (n >= sp->numnod) return bad(*sp, "%lld nodes, need more", sp->numnod);
          const int *e = &sp->idx_node[2 * n];
          if (e[0] != k || e[1] != sp->idx2[3 * j])
            return bad(*sp, "node %lld: need (k, j) with the mode fastest", n + 1);
        }
      // the rj space may carry one centre node (0, 0) at the end
      const bool centre = sp == &rj && sp->numnod == n + 1 && sp->idx_node[2 * n] == 0 &&
                          sp->idx_node[2 * n + 1] == 0;
      if (sp->numnod != n && !centre) return bad(*sp, "%lld nodes, need %lld", sp->numnod, n);
    }
  }
  // communication tables: the engine's buffers are in rlm / rtm node order
  const long long nrlm = jmax * nr, nrtm = (long long)(2 * L + 1) * nr * nth;
  for (const auto *sp : {&rlm, &rtm}) {
    const long long nn = sp == &rlm ? nrlm : nrtm;
    if ((long long)sp->item.size() != nn) return bad(*sp, "%lld import items, need %lld", sp->item.size(), nn);
    for (long long i = 0; i < nn; i++)
      if (sp->item[i] != i + 1)
        return bad(*sp, "import item %lld is node %lld: the buffer must be in node order", i + 1,
                   sp->item[i]);
  }
  if ((long long)rj.item.size() != nrlm) return bad(rj, "%lld import items, need %lld", rj.item.size(), nrlm);
  for (long long i = 0; i < nrlm; i++) {
    const long long q = rj.item[i] - 1;  // rj node paired with rlm node i
    if (q < 0 || q >= rj.numnod || rj.idx_node[2 * q] != rlm.idx_node[2 * i] ||
        rj.idx_node[2 * q + 1] != rlm.idx_node[2 * i + 1])
      return bad(rj, "import item %lld (node %lld) is not the rlm node's radius and mode", i + 1,
                 q + 1);
  }
  if ((long long)rtp.item.size() != nrtm) return bad(rtp, "%lld import items, need %lld", rtp.item.size(), nrtm);
  for (long long i = 0; i < nrtm; i++) {
    const long long q = rtp.item[i] - 1;
    const int *a = &rtm.idx_node[3 * i];
    if (q < 0 || q >= rtp.numnod || rtp.idx_node[3 * q] != a[0] || rtp.idx_node[3 * q + 1] != a[1] ||
        rtp.idx_node[3 * q + 2] != a[2])
      return bad(rtp, "import item %lld (node %lld) is not the rtm node's radius, colatitude "
                 "and FFT slot", i + 1, q + 1);
  }
  return true;
}

}  // namespace

using namespace ksht;

extern "C" int sht_sph_files_read(const char *rtp, const char *rtm, const char *rlm,
                                  const char *rj, sht_sph_files **out) {
  if (!out || !rtp || !rtm || !rlm || !rj) {
    set_error("sht_sph_files_read: null argument");
    return SHT_ERR_ARG;
  }
  *out = nullptr;
  sht_sph_files *f = new sht_sph_files();
  const char *paths[4] = {rtp, rtm, rlm, rj};
  for (int k = 0; k < 4; k++) {
    f->s[k].path = paths[k];
    if (!read_space(f->s[k], k)) {
      delete f;
      return SHT_ERR_ARG;
    }
  }
  *out = f;
  return SHT_OK;
}

extern "C" void sht_sph_files_destroy(sht_sph_files *f) { delete f; }

extern "C" int sht_sph_files_check(const sht_sph_files *f) {
  if (!f) {
    set_error("sht_sph_files_check: null argument");
    return SHT_ERR_ARG;
  }
  return check_set(f) ? SHT_OK : SHT_ERR_UNSUPPORTED;
}

extern "C" int sht_sph_files_shape(const sht_sph_files *f, int dims[5], double *radius) {
  if (!f || !dims) {
    set_error("sht_sph_files_shape: null argument");
    return SHT_ERR_ARG;
  }
  const auto &rtp = f->s[kRtp], &rj = f->s[kRj];
  dims[0] = rj.ltr;
  dims[1] = rtp.nidx_gl[0];
  dims[2] = rtp.nidx_gl[1];
  dims[3] = rtp.nidx_gl[2];
  dims[4] = (int)rj.numnod;
  if (radius)
    for (size_t k = 0; k < rj.r1.size(); k++) radius[k] = rj.r1[k];
  return SHT_OK;
}

extern "C" long long sht_sph_files_table(const sht_sph_files *f, const char *name, int *out,
                                         long long n) {
  if (!f || !name) {
    set_error("sht_sph_files_table: null argument");
    return -SHT_ERR_ARG;
  }
  static const char *spaces[4] = {"rtp", "rtm", "rlm", "rj"};
  for (int k = 0; k < 4; k++) {
    const auto &sp = f->s[k];
    const std::string pre = spaces[k];
    const std::vector<int> *v = nullptr;
    std::vector<int> irev;
    if (name == pre + "_import")
      v = &sp.item;
    else if (name == pre + "_import_rev") {
      // reverse table (set_reverse_import_table): node -> 1-based buffer
      // record, ntot + 1 (the zero dummy record) for nodes not imported
      irev.assign((size_t)sp.numnod, (int)sp.item.size() + 1);
      for (size_t i = 0; i < sp.item.size(); i++)
        if (sp.item[i] >= 1 && sp.item[i] <= sp.numnod) irev[(size_t)sp.item[i] - 1] = (int)i + 1;
      v = &irev;
    } else if (name == pre + "_nodes")
      v = &sp.idx_node;
    else if (name == pre + "_table2")
      v = &sp.idx2;
    else if (name == pre + "_table3")
      v = &sp.idx3;
    if (!v) continue;
    const long long len = (long long)v->size();
    if (out) {
      if (n < len) {
        set_error("sht_sph_files_table: %s has %lld entries, buffer %lld", name, len, n);
        return -SHT_ERR_ARG;
      }
      for (long long i = 0; i < len; i++) out[i] = (*v)[(size_t)i];
    }
    return len;
  }
  set_error("sht_sph_files_table: unknown table '%s'", name);
  return -SHT_ERR_ARG;
}

extern "C" int sht_plan_create_from_files(const sht_sph_files *f, int max_ncomp, int precision,
                                          int n_slices, unsigned flags, sht_plan **plan) {
  if (!f || !plan) {
    set_error("sht_plan_create_from_files: null argument");
    return SHT_ERR_ARG;
  }
  *plan = nullptr;
  if (!check_set(f)) return SHT_ERR_UNSUPPORTED;
  const auto &rtp = f->s[kRtp], &rj = f->s[kRj];
  sht_desc d;
  memset(&d, 0, sizeof d);
  d.l_truncation = rj.ltr;
  d.nidx_r = rtp.nidx_gl[0];
  d.nidx_theta = rtp.nidx_gl[1];
  d.nidx_phi = rtp.nidx_gl[2];
  d.radius = rj.r1.data();
  d.max_ncomp = max_ncomp;
  d.precision = precision;
  d.n_slices = n_slices;
  d.flags = flags;
  return sht_plan_create(&d, plan);
}
