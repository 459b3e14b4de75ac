/*
 * sht_oracle.c -- CPU FP64 restatement of the reference spherical-harmonic
 * transform.  TEST INFRASTRUCTURE ONLY: it is the checker for tests/, the
 * oracle leg of __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py.  The product library (kemorin_mhd_amd) never links or calls it.
 *
 * Each function cites the reference file:line it restates (abbreviations in
 * sht_oracle.h).  Known, documented deviation: the Schmidt P_l^m product
 * df(l)*sin^m is formed as df(l)*(sin^m) with sin^m built by repeated
 * multiplication once per (theta,m), where SH/schmidt_fix_m.f90:133-138
 * multiplies df(l) by sin m times; both orders round to within m ulp.
 *
 * Pins (tests/test_oracle_*.py): compiled reference routines (oracle/_ref),
 * numpy leggauss, closed forms of P_l^m, scipy lpmv, O(N^2) DFT and the
 * known-answer signals of MHD/programs/TESTS/FFT_test/test_FFT.f90:42-50,
 * single-mode closed forms and round-trip identities of the transform.
 */
#include "sht_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

struct orc_plan {
  int L, nr, nth, nphi, nslot, jmax, nth_h, flags;
  double *radius, *a_r;
  double *gpoint, *colat, *weight, *asin_t;
  double *g3, *g6, *g7;
  int *lstack, *lstack_even, *idx_j, *mdx_p, *mdx_n;
  double *P, *dP;   /* nth x jmax, optional */
  double *Ps, *dPs; /* nth_h x jmax */
};

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* GI/gauss_integration.f90:60-133 (const_gauss_points_coefs).  Only the
 * last pass n = n_point of the reference's n = 2..n_point loop survives
 * (every earlier pass is overwritten), so only that pass is computed. */
void orc_gauss_points(int n, int exact, double *x, double *w) {
  const double one = 1.0, two = 2.0, half = 0.5, quad = 0.25;
  const double pi = 4.0 * atan(1.0);
  const double tol = (double)3.0e-12f; /* REAL(4) literal in the source */
  const double x1 = 0.0, x2 = 1.0;
  const double xm = half * (x2 + x1), xl = half * (x2 - x1);
  int i;
  if (n == 1) {
    x[0] = half;
    w[0] = one;
  } else {
    int m = (n + 1 - (n + 1) % 2) / 2;
    for (i = 1; i <= m; i++) {
      double z = cos(pi * ((double)i - quad) / ((double)n + half));
      double z1, pp, p1, p2, p3;
      if (!exact) {
        for (;;) {
          int j;
          p1 = one;
          p2 = 0.0;
          for (j = 1; j <= n; j++) {
            p3 = p2;
            p2 = p1;
            p1 = ((two * (double)j - one) * z * p2 - ((double)j - one) * p3) /
                 (double)j;
          }
          pp = (double)n * (z * p1 - p2) / (z * z - one);
          z1 = z;
          z = z1 - p1 / pp;
          if (fabs(z - z1) < tol) break;
        }
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = two * xl / ((one - z * z) * pp * pp);
        w[n - i] = w[i - 1];
      } else {
        long double zl = z, zl1, ppl = 1, p1l, p2l, p3l;
        int it;
        for (it = 0; it < 100; it++) {
          int j;
          p1l = 1;
          p2l = 0;
          for (j = 1; j <= n; j++) {
            p3l = p2l;
            p2l = p1l;
            p1l = ((2.0L * j - 1) * zl * p2l - (j - 1.0L) * p3l) / j;
          }
          ppl = n * (zl * p1l - p2l) / (zl * zl - 1);
          zl1 = zl;
          zl = zl1 - p1l / ppl;
          if (fabsl(zl - zl1) < 1e-19L) break;
        }
        {
          int j;
          p1l = 1;
          p2l = 0;
          for (j = 1; j <= n; j++) {
            p3l = p2l;
            p2l = p1l;
            p1l = ((2.0L * j - 1) * zl * p2l - (j - 1.0L) * p3l) / j;
          }
          ppl = n * (zl * p1l - p2l) / (zl * zl - 1);
        }
        x[i - 1] = (double)(0.5L - 0.5L * zl);
        x[n - i] = (double)(0.5L + 0.5L * zl);
        w[i - 1] = (double)(1.0L / ((1 - zl * zl) * ppl * ppl));
        w[n - i] = w[i - 1];
      }
    }
  }
  for (i = 0; i < n; i++) {
    x[i] = x[i] * two - one;
    w[i] = w[i] * two;
  }
}

/* ------------------------------------------------------------------ */
/* SH/schmidt_fix_m.f90:89-141 (schmidt_legendre_m).  df has ltr+3 slots. */
static void schmidt_legendre_m(int ltr, int m, double theta, double *p,
                               double *df) {
  int l, k;
  const double c = cos(theta);
  if (m < 0) {
    for (l = 0; l <= ltr; l++) p[l] = 0.0;
  } else if (m == 0) {
    p[0] = 1.0;
    if (ltr >= 1) p[1] = c;
    for (l = 2; l <= ltr; l++)
      p[l] = p[l - 1] * (double)(2 * l - 1) / (double)l * c -
             p[l - 2] * (double)(l - 1) / (double)l;
  } else {
    double sm = 1.0;
    const double s = sin(theta);
    df[m] = 1.0;
    for (k = 1; k <= m; k++) df[m] = df[m] * (double)(2 * k - 1) / (double)(2 * k);
    df[m + 1] = sqrt(2.0 * df[m] * (double)(2 * m + 1)) * c;
    df[m] = sqrt(2.0 * df[m]);
    if (m < ltr - 1) {
      for (l = m + 2; l <= ltr; l++)
        df[l] = (c * (double)(2 * l - 1) * df[l - 1] -
                 sqrt((double)((l - 1) * (l - 1) - m * m)) * df[l - 2]) /
                sqrt((double)(l * l - m * m));
    }
    for (k = 1; k <= m; k++) sm = sm * s;
    for (l = m; l <= ltr; l++) p[l] = df[l] * sm;
  }
}

/* SH/schmidt_fix_m.f90:145-175 (diff_schmidt_legendre_m) */
static void diff_schmidt_legendre_m(int ltr, int m, const double *pmn1,
                                    const double *pmp1, double *dp) {
  int l;
  if (m == 0) {
    dp[0] = 0.0;
    for (l = 1; l <= ltr; l++)
      dp[l] = -sqrt((double)(l * (l + 1) / 2)) * pmp1[l];
  } else if (m == 1) {
    dp[1] = pmn1[1];
    for (l = 2; l <= ltr; l++)
      dp[l] = 0.5 * (sqrt((double)(2 * l * (l + 1))) * pmn1[l] -
                     sqrt((double)((l - 1) * (l + 2))) * pmp1[l]);
  } else {
    dp[m] = 0.5 * sqrt((double)(2 * m)) * pmn1[m];
    for (l = m + 1; l <= ltr; l++)
      dp[l] = 0.5 * (sqrt((double)((l + m) * (l - m + 1))) * pmn1[l] -
                     sqrt((double)((l - m) * (l + m + 1))) * pmp1[l]);
  }
}

/* SH/schmidt_fix_m.f90:62-84 (schmidt_legendres_m) */
void orc_schmidt_legendres_m(int ltr, int m, double theta, double *p,
                             double *dp) {
  double *w = (double *)calloc((size_t)(2 * (ltr + 3)), sizeof(double));
  double *pmn1 = w, *pmp1 = w + (ltr + 3);
  double *dfw = (double *)calloc((size_t)(ltr + 3), sizeof(double));
  schmidt_legendre_m(ltr, m - 1, theta, pmn1, dfw);
  schmidt_legendre_m(ltr, m, theta, p, dfw);
  schmidt_legendre_m(ltr, m + 1, theta, pmp1, dfw);
  diff_schmidt_legendre_m(ltr, m, pmn1, pmp1, dp);
  free(dfw);
  free(w);
}

/* SH/spherical_harmonics.f90:91-140 (sph_normalizations) */
void orc_sph_normalizations(int l, int m, double g[17]) {
  const double pi = 4.0 * atan(1.0);
  g[0] = (double)l;
  g[1] = (double)m;
  g[2] = (double)(l * (l + 1));
  g[3] = (double)((l - 1) * l);
  g[4] = (double)((l + 1) * (l + 2));
  g[7] = 4.0 * pi * (double)(l * (l + 1)) / (double)(2 * l + 1);
  g[9] = 4.0 * pi / (double)(2 * l + 1);
  g[10] = 1.0 / (double)(2 * l + 1);
  g[11] = (double)(l * (l + 1)) / (double)(2 * l + 1);
  g[13] = 1.0;
  g[14] = 1.0;
  g[15] = (double)(2 * l + 1) / 4.0;
  if (l == 0 && m == 0) {
    g[5] = 1.0 / 2.0;
    g[6] = 1.0;
    g[8] = 1.0 / (4.0 * pi);
    g[12] = 1.0;
    g[16] = 0.0;
  } else {
    g[8] = (double)(2 * l + 1) / (4.0 * pi * (double)(l * (l + 1)));
    g[12] = 1.0 / (double)(l * (l + 1));
    g[16] = (double)(2 * l + 1) / (double)(4 * l * (l + 1));
    if (m == 0) {
      g[5] = (double)(2 * l + 1) / 2.0;
      g[6] = (double)(2 * l + 1) / (double)(2 * l * (l + 1));
    } else {
      g[5] = (double)(2 * l + 1) / 4.0;
      g[6] = (double)(2 * l + 1) / (double)(4 * l * (l + 1));
    }
  }
}

void orc_chebyshev_radius(int nr, double ri, double ro, double *r) {
  const double pi = 4.0 * atan(1.0);
  const double shell = ro - ri;
  int k;
  for (k = 0; k < nr; k++)
    r[k] = ri + 0.5 * shell * (1.0 - cos(pi * (double)k / (double)(nr - 1)));
}

/* ------------------------------------------------------------------ */
/* Plan: orderings CG/set_sph_tranform_ordering.f90:69-206 (single rank:
 * m-slot s <-> m = s - L, rlm j by slot then l ascending), stacks
 * SS/set_params_sph_trans.f90:64-137 and SS/set_legendre_matrices.f90:64-82,
 * normalisation SP/schmidt_poly_on_rtm_grid.f90:89-113 (g3(0,0)=1/2),
 * P tables SP/schmidt_poly_on_rtm_grid.f90:145-197, hemisphere tables
 * SS/set_legendre_matrices.f90:86-130, 1/sin SS/set_params_sph_trans.f90:141. */
orc_plan *orc_plan_create(int L, int nr, int nth, int nphi,
                          const double *radius, int flags) {
  orc_plan *p;
  int s, j, k, lt;
  if (L < 1 || nr < 1 || nth < 2 || nphi < 2 * L + 2 || (nphi & 1)) return NULL;
  p = (orc_plan *)calloc(1, sizeof(orc_plan));
  p->L = L;
  p->nr = nr;
  p->nth = nth;
  p->nphi = nphi;
  p->nslot = 2 * L + 1;
  p->jmax = (L + 1) * (L + 1);
  p->nth_h = (nth + 1) / 2;
  p->flags = flags;
  p->radius = (double *)malloc(sizeof(double) * nr);
  p->a_r = (double *)malloc(sizeof(double) * nr);
  for (k = 0; k < nr; k++) {
    p->radius[k] = radius[k];
    p->a_r[k] = 1.0 / radius[k];
  }
  p->gpoint = (double *)malloc(sizeof(double) * nth);
  p->colat = (double *)malloc(sizeof(double) * nth);
  p->weight = (double *)malloc(sizeof(double) * nth);
  p->asin_t = (double *)malloc(sizeof(double) * nth);
  orc_gauss_points(nth, flags & 1, p->gpoint, p->weight);
  for (lt = 0; lt < nth; lt++) {
    p->colat[lt] = acos(p->gpoint[lt]);
    p->asin_t[lt] = 1.0 / sin(p->colat[lt]);
  }
  p->lstack = (int *)malloc(sizeof(int) * (p->nslot + 1));
  p->lstack_even = (int *)malloc(sizeof(int) * (p->nslot + 1));
  p->idx_j = (int *)malloc(sizeof(int) * 3 * p->jmax);
  p->mdx_p = (int *)malloc(sizeof(int) * p->jmax);
  p->mdx_n = (int *)malloc(sizeof(int) * p->jmax);
  p->g3 = (double *)malloc(sizeof(double) * p->jmax);
  p->g6 = (double *)malloc(sizeof(double) * p->jmax);
  p->g7 = (double *)malloc(sizeof(double) * p->jmax);
  p->lstack[0] = 0;
  p->lstack_even[0] = 0;
  j = 0;
  for (s = 0; s < p->nslot; s++) {
    int m = s - L, am = abs(m), l, n = L - am + 1;
    p->lstack[s + 1] = p->lstack[s] + n;
    p->lstack_even[s + 1] = p->lstack[s] + (n + 1) / 2;
    for (l = am; l <= L; l++, j++) {
      double g[17];
      p->idx_j[j] = l * (l + 1) + m;
      p->idx_j[j + p->jmax] = l;
      p->idx_j[j + 2 * p->jmax] = m;
      p->mdx_p[j] = m + L;
      p->mdx_n[j] = -m + L;
      orc_sph_normalizations(l, m, g);
      p->g3[j] = (l == 0 && m == 0) ? 0.5 : g[2];
      p->g6[j] = g[5];
      p->g7[j] = g[6];
    }
  }
  if (flags & 2) {
    p->P = (double *)malloc(sizeof(double) * (size_t)nth * p->jmax);
    p->dP = (double *)malloc(sizeof(double) * (size_t)nth * p->jmax);
  }
  p->Ps = (double *)malloc(sizeof(double) * (size_t)p->nth_h * p->jmax);
  p->dPs = (double *)malloc(sizeof(double) * (size_t)p->nth_h * p->jmax);
#pragma omp parallel for schedule(dynamic)
  for (lt = 0; lt < nth; lt++) {
    double *pm = (double *)calloc((size_t)(L + 3), sizeof(double));
    double *dpm = (double *)calloc((size_t)(L + 3), sizeof(double));
    int am;
    if (!(p->P) && lt >= p->nth_h) {
      free(pm);
      free(dpm);
      continue;
    }
    for (am = 0; am <= L; am++) {
      int sg;
      orc_schmidt_legendres_m(L, am, p->colat[lt], pm, dpm);
      for (sg = 0; sg < (am == 0 ? 1 : 2); sg++) {
        int ss = sg == 0 ? am + L : L - am;
        int jst = p->lstack[ss], n = L - am + 1, ne = (n + 1) / 2, l;
        for (l = am; l <= L; l++) {
          int jj = jst + (l - am);
          if (p->P) {
            p->P[(size_t)jj * nth + lt] = pm[l];
            p->dP[(size_t)jj * nth + lt] = dpm[l];
          }
          if (lt < p->nth_h) {
            int col = ((l - am) & 1) ? jst + ne + (l - am) / 2 : jst + (l - am) / 2;
            p->Ps[(size_t)col * p->nth_h + lt] = pm[l];
            p->dPs[(size_t)col * p->nth_h + lt] = dpm[l];
          }
        }
      }
    }
    free(pm);
    free(dpm);
  }
  return p;
}

void orc_plan_destroy(orc_plan *p) {
  if (!p) return;
  free(p->radius); free(p->a_r); free(p->gpoint); free(p->colat);
  free(p->weight); free(p->asin_t); free(p->g3); free(p->g6); free(p->g7);
  free(p->lstack); free(p->lstack_even); free(p->idx_j); free(p->mdx_p);
  free(p->mdx_n); free(p->P); free(p->dP); free(p->Ps); free(p->dPs);
  free(p);
}

const void *orc_plan_table(const orc_plan *p, const char *name) {
#define T(nm, f) if (!strcmp(name, nm)) return (const void *)p->f
  T("radius", radius); T("a_r", a_r); T("gpoint", gpoint); T("colat", colat);
  T("weight", weight); T("asin_t", asin_t); T("g3", g3); T("g6", g6);
  T("g7", g7); T("P", P); T("dP", dP); T("Ps", Ps); T("dPs", dPs);
  T("lstack", lstack); T("lstack_even", lstack_even); T("idx_j", idx_j);
  T("mdx_p", mdx_p); T("mdx_n", mdx_n);
#undef T
  return NULL;
}

void orc_plan_dims(const orc_plan *p, int d[8]) {
  d[0] = p->L; d[1] = p->nr; d[2] = p->nth; d[3] = p->nphi;
  d[4] = p->nslot; d[5] = p->jmax; d[6] = p->nth_h; d[7] = p->flags;
}

/* ------------------------------------------------------------------ */
/* FFT slot of zonal order m in the rtp spectral layout,
 * CG/set_sph_tranform_ordering.f90:43-65 (0-based: a0 -> 0, Nyquist -> 1,
 * cos(m phi) -> 2m, sin(m phi) -> 2m+1). */
static inline int fft_slot(int m) { return m == 0 ? 0 : (m > 0 ? 2 * m : -2 * m + 1); }

void orc_rtm_to_rtp(const orc_plan *p, int ncomp, const double *rtm,
                    double *rtp) {
  const size_t nline = (size_t)p->nr * p->nth;
  int s;
  memset(rtp, 0, sizeof(double) * ncomp * nline * p->nphi);
#pragma omp parallel for
  for (s = 0; s < p->nslot; s++) {
    int q = fft_slot(s - p->L), k, lt, nd;
    for (k = 0; k < p->nr; k++)
      for (lt = 0; lt < p->nth; lt++)
        for (nd = 0; nd < ncomp; nd++)
          rtp[nd + ncomp * (k + (size_t)p->nr * (lt + (size_t)p->nth * q))] =
              rtm[nd + ncomp * (lt + (size_t)p->nth * (k + (size_t)p->nr * s))];
  }
}

void orc_rtp_to_rtm(const orc_plan *p, int ncomp, const double *rtp,
                    double *rtm) {
  int s;
#pragma omp parallel for
  for (s = 0; s < p->nslot; s++) {
    int q = fft_slot(s - p->L), k, lt, nd;
    for (k = 0; k < p->nr; k++)
      for (lt = 0; lt < p->nth; lt++)
        for (nd = 0; nd < ncomp; nd++)
          rtm[nd + ncomp * (lt + (size_t)p->nth * (k + (size_t)p->nr * s))] =
              rtp[nd + ncomp * (k + (size_t)p->nr * (lt + (size_t)p->nth * q))];
  }
}

/* ------------------------------------------------------------------ */
/* SS/legendre_bwd_trans_org.f90:51-184 (vector) and :188-248 (scalar) */
void orc_leg_bwd_org(const orc_plan *p, int nv, int ns, const double *sp_in,
                     double *vr) {
  const int ncomp = 3 * nv + ns, nth = p->nth, nr = p->nr, jmax = p->jmax;
  const size_t nrlm = (size_t)jmax * nr, nrtm = (size_t)nth * nr * p->nslot;
  double *sp = (double *)malloc(sizeof(double) * ncomp * nrlm);
  int k;
  memcpy(sp, sp_in, sizeof(double) * ncomp * nrlm);
  memset(vr, 0, sizeof(double) * ncomp * nrtm);
  for (k = 0; k < nr; k++) {
    double a1 = p->a_r[k], a2 = a1 * a1;
    int j, nd;
    for (j = 0; j < jmax; j++)
      for (nd = 0; nd < nv; nd++) {
        size_t i = 3 * nd + ncomp * (j + (size_t)jmax * k);
        sp[i] = sp[i] * a2;
        sp[i + 1] = sp[i + 1] * a1;
        sp[i + 2] = sp[i + 2] * a1;
      }
  }
#pragma omp parallel for schedule(dynamic)
  for (k = 0; k < nr; k++) {
    double *Pg3 = (double *)malloc(sizeof(double) * jmax);
    double *dPdt = (double *)malloc(sizeof(double) * jmax);
    int s, lt, j, nd;
    if (nv > 0) {
      for (s = 0; s < p->nslot; s++) {
        int jst = p->lstack[s], jed = p->lstack[s + 1];
        for (lt = 0; lt < nth; lt++) {
          for (j = jst; j < jed; j++) {
            Pg3[j] = p->P[(size_t)j * nth + lt] * p->g3[j];
            dPdt[j] = p->dP[(size_t)j * nth + lt];
          }
          for (nd = 0; nd < nv; nd++) {
            double vr1 = 0, vr2 = 0, vr3 = 0;
            size_t ip;
            for (j = jst; j < jed; j++) {
              size_t i = 3 * nd + ncomp * (j + (size_t)jmax * k);
              vr1 = vr1 + sp[i] * Pg3[j];
              vr2 = vr2 + sp[i + 1] * dPdt[j];
              vr3 = vr3 - sp[i + 2] * dPdt[j];
            }
            ip = 3 * nd + ncomp * (lt + (size_t)nth * (k + (size_t)nr * s));
            vr[ip] = vr[ip] + vr1;
            vr[ip + 1] = vr[ip + 1] + vr2;
            vr[ip + 2] = vr[ip + 2] + vr3;
          }
        }
      }
      for (s = 0; s < p->nslot; s++) {
        int sn = p->nslot - 1 - s;
        int jst = p->lstack[s], jed = p->lstack[s + 1];
        for (lt = 0; lt < nth; lt++) {
          for (j = jst; j < jed; j++)
            Pg3[j] = -p->P[(size_t)j * nth + lt] * (double)p->idx_j[j + 2 * jmax] *
                     p->asin_t[lt];
          for (nd = 0; nd < nv; nd++) {
            double vr2 = 0, vr3 = 0;
            size_t in;
            for (j = jst; j < jed; j++) {
              size_t i = 3 * nd + ncomp * (j + (size_t)jmax * k);
              vr2 = vr2 + sp[i + 2] * Pg3[j];
              vr3 = vr3 + sp[i + 1] * Pg3[j];
            }
            in = 3 * nd + ncomp * (lt + (size_t)nth * (k + (size_t)nr * sn));
            vr[in + 1] = vr[in + 1] + vr2;
            vr[in + 2] = vr[in + 2] + vr3;
          }
        }
      }
    }
    for (s = 0; s < p->nslot; s++) {
      int jst = p->lstack[s], jed = p->lstack[s + 1];
      for (lt = 0; lt < nth; lt++)
        for (nd = 0; nd < ns; nd++) {
          double vr1 = 0;
          for (j = jst; j < jed; j++)
            vr1 = vr1 + sp[nd + 3 * nv + ncomp * (j + (size_t)jmax * k)] *
                            p->P[(size_t)j * nth + lt];
          vr[nd + 3 * nv + ncomp * (lt + (size_t)nth * (k + (size_t)nr * s))] = vr1;
        }
    }
    free(Pg3);
    free(dPdt);
  }
  free(sp);
}

/* SS/legendre_fwd_trans_org.f90:51-145 (vector) and :149-211 (scalar) */
void orc_leg_fwd_org(const orc_plan *p, int nv, int ns, const double *vr,
                     double *sp) {
  const int ncomp = 3 * nv + ns, nth = p->nth, nr = p->nr, jmax = p->jmax;
  int k;
  memset(sp, 0, sizeof(double) * ncomp * (size_t)jmax * nr);
#pragma omp parallel for schedule(dynamic)
  for (k = 0; k < nr; k++) {
    double *Pvw = (double *)malloc(sizeof(double) * nth);
    double *dPvw = (double *)malloc(sizeof(double) * nth);
    double *Pgvw = (double *)malloc(sizeof(double) * nth);
    const double r1 = p->radius[k], r2 = r1 * r1;
    int j, lt, nd;
    for (j = 0; j < jmax; j++) {
      const size_t bp = (size_t)nth * (k + (size_t)nr * p->mdx_p[j]);
      const size_t bn = (size_t)nth * (k + (size_t)nr * p->mdx_n[j]);
      if (nv > 0) {
        for (lt = 0; lt < nth; lt++) {
          Pvw[lt] = p->P[(size_t)j * nth + lt] * p->g7[j] * p->weight[lt];
          dPvw[lt] = p->dP[(size_t)j * nth + lt] * p->g7[j] * p->weight[lt];
          Pgvw[lt] = p->P[(size_t)j * nth + lt] * (double)p->idx_j[j + 2 * jmax] *
                     p->asin_t[lt] * p->g7[j] * p->weight[lt];
        }
        for (nd = 0; nd < nv; nd++) {
          double sp1 = 0, sp2 = 0, sp3 = 0;
          size_t i;
          for (lt = 0; lt < nth; lt++) {
            size_t ip = 3 * nd + ncomp * (lt + bp);
            size_t in = 3 * nd + ncomp * (lt + bn);
            sp1 = sp1 + vr[ip] * Pvw[lt];
            sp2 = sp2 + (vr[ip + 1] * dPvw[lt] - vr[in + 2] * Pgvw[lt]);
            sp3 = sp3 - (vr[in + 1] * Pgvw[lt] + vr[ip + 2] * dPvw[lt]);
          }
          i = 3 * nd + ncomp * (j + (size_t)jmax * k);
          sp[i] = sp[i] + sp1 * r2;
          sp[i + 1] = sp[i + 1] + sp2 * r1;
          sp[i + 2] = sp[i + 2] + sp3 * r1;
        }
      }
      if (ns > 0) {
        for (lt = 0; lt < nth; lt++)
          Pvw[lt] = p->P[(size_t)j * nth + lt] * p->g6[j] * p->weight[lt];
        for (nd = 0; nd < ns; nd++) {
          double sp1 = 0;
          size_t i;
          for (lt = 0; lt < nth; lt++)
            sp1 = sp1 + vr[nd + 3 * nv + ncomp * (lt + bp)] * Pvw[lt];
          i = nd + 3 * nv + ncomp * (j + (size_t)jmax * k);
          sp[i] = sp[i] + sp1;
        }
      }
    }
    free(Pvw);
    free(dPvw);
    free(Pgvw);
  }
}

/* ------------------------------------------------------------------ */
/* Column-major C(MxN) = A(MxK) B(KxN): the restated DGEMM('N','N',...) of
 * SS/matmul_for_legendre_trans.F90:144-169,265-294 (beta = 0). */
static void dgemm_nn(int M, int N, int K, const double *A, int lda,
                     const double *B, int ldb, double *C, int ldc) {
  int i, j, q;
  for (j = 0; j + 4 <= N; j += 4) {
    double *c0 = C + (size_t)ldc * j, *c1 = c0 + ldc, *c2 = c1 + ldc, *c3 = c2 + ldc;
    for (i = 0; i < M; i++) c0[i] = c1[i] = c2[i] = c3[i] = 0.0;
    for (q = 0; q < K; q++) {
      const double *a = A + (size_t)lda * q;
      const double b0 = B[q + (size_t)ldb * j], b1 = B[q + (size_t)ldb * (j + 1)];
      const double b2 = B[q + (size_t)ldb * (j + 2)], b3 = B[q + (size_t)ldb * (j + 3)];
      for (i = 0; i < M; i++) {
        const double av = a[i];
        c0[i] += av * b0;
        c1[i] += av * b1;
        c2[i] += av * b2;
        c3[i] += av * b3;
      }
    }
  }
  for (; j < N; j++) {
    double *c0 = C + (size_t)ldc * j;
    for (i = 0; i < M; i++) c0[i] = 0.0;
    for (q = 0; q < K; q++) {
      const double *a = A + (size_t)lda * q;
      const double b0 = B[q + (size_t)ldb * j];
      for (i = 0; i < M; i++) c0[i] += a[i] * b0;
    }
  }
}

static void radial_chunks(int nr, int nthreads, int t, int *k0, int *nk) {
  int base = nr / nthreads, rem = nr % nthreads;
  *k0 = t * base + (t < rem ? t : rem);
  *nk = base + (t < rem ? 1 : 0);
}

/* symmetric_BLAS_big backward: SS/leg_b_trans_sym_matmul_big.f90:183-293 with
 * pack SS/set_sp_rlm_for_leg_matmul.f90:281-398 and unpack
 * SS/cal_vr_rtm_by_matmul.f90:357-540.  Threads own radial chunks. */
void orc_leg_bwd_sym(const orc_plan *p, int nv, int ns, const double *WR,
                     double *WS) {
  const int ncomp = 3 * nv + ns, nth = p->nth, nr = p->nr, jmax = p->jmax;
  const int nth_h = p->nth_h, L = p->L;
  const size_t nrtm = (size_t)nth * nr * p->nslot;
  memset(WS, 0, sizeof(double) * ncomp * nrtm);
#pragma omp parallel
  {
    int nt = 1, t = 0, k0, nkr;
#ifdef _OPENMP
    nt = omp_get_num_threads();
    t = omp_get_thread_num();
#endif
    radial_chunks(nr, nt, t, &k0, &nkr);
    if (nkr > 0) {
      const int nkrv = nkr * nv, nkrs = ncomp * nkr, nkrt = 2 * nv * nkr;
      const int kmax = (L + 2) / 2;
      double *pol_e = (double *)malloc(sizeof(double) * kmax * nkrs);
      double *pol_o = (double *)malloc(sizeof(double) * kmax * nkrs);
      double *tor_e = (double *)malloc(sizeof(double) * kmax * (nkrt + 1));
      double *tor_o = (double *)malloc(sizeof(double) * kmax * (nkrt + 1));
      double *symp_r = (double *)malloc(sizeof(double) * nth_h * nkrs);
      double *asmp_r = (double *)malloc(sizeof(double) * nth_h * nkrs);
      double *symp_p = (double *)malloc(sizeof(double) * nth_h * (nkrt + 1));
      double *asmp_p = (double *)malloc(sizeof(double) * nth_h * (nkrt + 1));
      int s;
      for (s = 0; s < p->nslot; s++) {
        const int sn = p->nslot - 1 - s, jst = p->lstack[s];
        const int ne = p->lstack_even[s + 1] - jst;
        const int no = p->lstack[s + 1] - p->lstack_even[s + 1];
        int kk, jj, lp;
        /* pack */
        for (kk = 0; kk < nkrv; kk++) {
          const int kl = kk % nkr, nd = kk / nkr, k = k0 + kl;
          const double a1 = p->a_r[k], a2 = a1 * a1;
          for (jj = 0; jj < ne + no; jj++) {
            const int odd = jj >= ne;
            const int j = odd ? jst + 2 * (jj - ne) + 1 : jst + 2 * jj;
            const int c = odd ? jj - ne : jj;
            double *pe = odd ? pol_o : pol_e, *te = odd ? tor_o : tor_e;
            const int ldk = odd ? no : ne;
            const double g3 = p->g3[j], gm = (double)p->idx_j[j + 2 * jmax];
            const size_t i = 3 * nd + ncomp * (j + (size_t)jmax * k);
            pe[c + ldk * kk] = WR[i] * a2 * g3;
            te[c + ldk * (kk + nkrv)] = WR[i + 1] * a1;
            pe[c + ldk * (kk + 2 * nkrv)] = WR[i + 1] * a1 * gm;
            te[c + ldk * kk] = WR[i + 2] * a1;
            pe[c + ldk * (kk + nkrv)] = WR[i + 2] * a1 * gm;
          }
        }
        for (kk = 0; kk < nkr * ns; kk++) {
          const int kl = kk % nkr, nd = kk / nkr, k = k0 + kl;
          for (jj = 0; jj < ne; jj++)
            pol_e[jj + ne * (kk + 3 * nkrv)] =
                WR[nd + 3 * nv + ncomp * (jst + 2 * jj + (size_t)jmax * k)];
          for (jj = 0; jj < no; jj++)
            pol_o[jj + no * (kk + 3 * nkrv)] =
                WR[nd + 3 * nv + ncomp * (jst + 2 * jj + 1 + (size_t)jmax * k)];
        }
        /* 4 GEMMs */
        dgemm_nn(nth_h, nkrs, ne, p->Ps + (size_t)jst * nth_h, nth_h, pol_e, ne, symp_r, nth_h);
        dgemm_nn(nth_h, nkrt, ne, p->dPs + (size_t)jst * nth_h, nth_h, tor_e, ne, asmp_p, nth_h);
        dgemm_nn(nth_h, nkrs, no, p->Ps + (size_t)(jst + ne) * nth_h, nth_h, pol_o, no, asmp_r, nth_h);
        dgemm_nn(nth_h, nkrt, no, p->dPs + (size_t)(jst + ne) * nth_h, nth_h, tor_o, no, symp_p, nth_h);
        /* unpack */
        for (kk = 0; kk < nkrv; kk++)
          for (lp = 0; lp < nth_h; lp++) {
            const double ai = p->asin_t[lp];
            symp_r[lp + nth_h * (kk + nkrv)] = -symp_r[lp + nth_h * (kk + nkrv)] * ai;
            symp_r[lp + nth_h * (kk + 2 * nkrv)] = -symp_r[lp + nth_h * (kk + 2 * nkrv)] * ai;
            asmp_r[lp + nth_h * (kk + nkrv)] = -asmp_r[lp + nth_h * (kk + nkrv)] * ai;
            asmp_r[lp + nth_h * (kk + 2 * nkrv)] = -asmp_r[lp + nth_h * (kk + 2 * nkrv)] * ai;
          }
        for (kk = 0; kk < nkrv; kk++) {
          const int kl = kk % nkr, nd = kk / nkr, k = k0 + kl;
#define SR(c) symp_r[lp + nth_h * (kk + (c) * nkrv)]
#define AR(c) asmp_r[lp + nth_h * (kk + (c) * nkrv)]
#define SP(c) symp_p[lp + nth_h * (kk + (c) * nkrv)]
#define AP(c) asmp_p[lp + nth_h * (kk + (c) * nkrv)]
          for (lp = 0; lp < nth / 2; lp++) {
            const int ln = nth - 1 - lp;
            double *pp = WS + 3 * nd + ncomp * (lp + (size_t)nth * (k + (size_t)nr * s));
            double *np = WS + 3 * nd + ncomp * (lp + (size_t)nth * (k + (size_t)nr * sn));
            double *pn = WS + 3 * nd + ncomp * (ln + (size_t)nth * (k + (size_t)nr * s));
            double *nn = WS + 3 * nd + ncomp * (ln + (size_t)nth * (k + (size_t)nr * sn));
            pp[0] = pp[0] + SR(0) + AR(0);
            pp[1] = pp[1] + AP(1) + SP(1);
            pp[2] = pp[2] - AP(0) - SP(0);
            np[1] = np[1] + SR(1) + AR(1);
            np[2] = np[2] + SR(2) + AR(2);
            pn[0] = pn[0] + SR(0) - AR(0);
            pn[1] = pn[1] - AP(1) + SP(1);
            pn[2] = pn[2] + AP(0) - SP(0);
            nn[1] = nn[1] + SR(1) - AR(1);
            nn[2] = nn[2] + SR(2) - AR(2);
          }
          for (lp = nth / 2; lp < nth_h; lp++) {
            double *pp = WS + 3 * nd + ncomp * (lp + (size_t)nth * (k + (size_t)nr * s));
            double *np = WS + 3 * nd + ncomp * (lp + (size_t)nth * (k + (size_t)nr * sn));
            pp[0] = pp[0] + SR(0);
            pp[1] = pp[1] + SP(1);
            pp[2] = pp[2] - SP(0);
            np[1] = np[1] + SR(1);
            np[2] = np[2] + SR(2);
          }
#undef SR
#undef AR
#undef SP
#undef AP
        }
        for (kk = 0; kk < nkr * ns; kk++) {
          const int kl = kk % nkr, nd = kk / nkr, k = k0 + kl;
          for (lp = 0; lp < nth / 2; lp++) {
            const int ln = nth - 1 - lp;
            const double sy = symp_r[lp + nth_h * (kk + 3 * nkrv)];
            const double as = asmp_r[lp + nth_h * (kk + 3 * nkrv)];
            double *pp = WS + nd + 3 * nv + ncomp * (lp + (size_t)nth * (k + (size_t)nr * s));
            double *pn = WS + nd + 3 * nv + ncomp * (ln + (size_t)nth * (k + (size_t)nr * s));
            *pp = *pp + sy + as;
            *pn = *pn + sy - as;
          }
          for (lp = nth / 2; lp < nth_h; lp++) {
            double *pp = WS + nd + 3 * nv + ncomp * (lp + (size_t)nth * (k + (size_t)nr * s));
            *pp = *pp + symp_r[lp + nth_h * (kk + 3 * nkrv)];
          }
        }
      }
      free(pol_e); free(pol_o); free(tor_e); free(tor_o);
      free(symp_r); free(asmp_r); free(symp_p); free(asmp_p);
    }
  }
}

/* symmetric_BLAS_big forward: SS/leg_f_trans_sym_matmul_big.f90:187-305 with
 * pack SS/set_vr_rtm_for_leg_matmul.f90:369-565 and unpack
 * SS/cal_sp_rlm_by_matmul.f90:365-545. */
void orc_leg_fwd_sym(const orc_plan *p, int nv, int ns, const double *WR,
                     double *WS) {
  const int ncomp = 3 * nv + ns, nth = p->nth, nr = p->nr, jmax = p->jmax;
  const int nth_h = p->nth_h, L = p->L, nlo = nth / 2;
  memset(WS, 0, sizeof(double) * ncomp * (size_t)jmax * nr);
#pragma omp parallel
  {
    int nt = 1, t = 0, k0, nkr;
#ifdef _OPENMP
    nt = omp_get_num_threads();
    t = omp_get_thread_num();
#endif
    radial_chunks(nr, nt, t, &k0, &nkr);
    if (nkr > 0) {
      const int nkrv = nkr * nv, nkrs = ncomp * nkr, nkrt = 2 * nv * nkr;
      const int kmax = (L + 2) / 2;
      double *symp_r = (double *)malloc(sizeof(double) * nkrs * nth_h);
      double *asmp_r = (double *)malloc(sizeof(double) * nkrs * nth_h);
      double *symp_p = (double *)malloc(sizeof(double) * (nkrt + 1) * nth_h);
      double *asmp_p = (double *)malloc(sizeof(double) * (nkrt + 1) * nth_h);
      double *pol_e = (double *)malloc(sizeof(double) * nkrs * kmax);
      double *pol_o = (double *)malloc(sizeof(double) * nkrs * kmax);
      double *tor_e = (double *)malloc(sizeof(double) * (nkrt + 1) * kmax);
      double *tor_o = (double *)malloc(sizeof(double) * (nkrt + 1) * kmax);
      int s;
      for (s = 0; s < p->nslot; s++) {
        const int sn = p->nslot - 1 - s, jst = p->lstack[s];
        const int ne = p->lstack_even[s + 1] - jst;
        const int no = p->lstack[s + 1] - p->lstack_even[s + 1];
        int kk, jj, lp;
        /* pack (rows = columns kk, cols = theta) */
        for (lp = 0; lp < nth_h; lp++) {
          const int eq = lp >= nlo, ln = nth - 1 - lp;
          const double wp = p->weight[lp], as = p->asin_t[lp];
          for (kk = 0; kk < nkrv; kk++) {
            const int kl = kk % nkr, nd = kk / nkr, k = k0 + kl;
            const double *pp = WR + 3 * nd + ncomp * (lp + (size_t)nth * (k + (size_t)nr * s));
            const double *np = WR + 3 * nd + ncomp * (lp + (size_t)nth * (k + (size_t)nr * sn));
            double *sr = symp_r + kk + (size_t)nkrs * lp;
            double *ar = asmp_r + kk + (size_t)nkrs * lp;
            double *sq = symp_p + kk + (size_t)nkrt * lp;
            double *aq = asmp_p + kk + (size_t)nkrt * lp;
            if (!eq) {
              const double *pn = WR + 3 * nd + ncomp * (ln + (size_t)nth * (k + (size_t)nr * s));
              const double *nn = WR + 3 * nd + ncomp * (ln + (size_t)nth * (k + (size_t)nr * sn));
              sr[0] = (pp[0] + pn[0]) * wp;
              sq[nkrv] = (pp[1] + pn[1]) * wp;
              sq[0] = (pp[2] + pn[2]) * wp;
              ar[0] = (pp[0] - pn[0]) * wp;
              aq[nkrv] = (pp[1] - pn[1]) * wp;
              aq[0] = (pp[2] - pn[2]) * wp;
              sr[nkrv] = (np[1] + nn[1]) * wp * as;
              sr[2 * nkrv] = (np[2] + nn[2]) * wp * as;
              ar[nkrv] = (np[1] - nn[1]) * wp * as;
              ar[2 * nkrv] = (np[2] - nn[2]) * wp * as;
            } else {
              sr[0] = pp[0] * wp;
              sq[nkrv] = pp[1] * wp;
              sq[0] = pp[2] * wp;
              ar[0] = 0.0;
              aq[nkrv] = 0.0;
              aq[0] = 0.0;
              sr[nkrv] = np[1] * wp * as;
              sr[2 * nkrv] = np[2] * wp * as;
              ar[nkrv] = 0.0;
              ar[2 * nkrv] = 0.0;
            }
          }
          for (kk = 0; kk < nkr * ns; kk++) {
            const int kl = kk % nkr, nd = kk / nkr, k = k0 + kl;
            const double *pp = WR + nd + 3 * nv + ncomp * (lp + (size_t)nth * (k + (size_t)nr * s));
            if (!eq) {
              const double *pn = WR + nd + 3 * nv + ncomp * (ln + (size_t)nth * (k + (size_t)nr * s));
              symp_r[kk + 3 * nkrv + (size_t)nkrs * lp] = (*pp + *pn) * wp;
              asmp_r[kk + 3 * nkrv + (size_t)nkrs * lp] = (*pp - *pn) * wp;
            } else {
              symp_r[kk + 3 * nkrv + (size_t)nkrs * lp] = *pp * wp;
              asmp_r[kk + 3 * nkrv + (size_t)nkrs * lp] = 0.0;
            }
          }
        }
        dgemm_nn(nkrs, ne, nth_h, symp_r, nkrs, p->Ps + (size_t)jst * nth_h, nth_h, pol_e, nkrs);
        dgemm_nn(nkrt, ne, nth_h, asmp_p, nkrt, p->dPs + (size_t)jst * nth_h, nth_h, tor_e, nkrt);
        dgemm_nn(nkrs, no, nth_h, asmp_r, nkrs, p->Ps + (size_t)(jst + ne) * nth_h, nth_h, pol_o, nkrs);
        dgemm_nn(nkrt, no, nth_h, symp_p, nkrt, p->dPs + (size_t)(jst + ne) * nth_h, nth_h, tor_o, nkrt);
        /* unpack */
        for (jj = 0; jj < ne + no; jj++) {
          const int odd = jj >= ne, c = odd ? jj - ne : jj;
          const int j = odd ? jst + 2 * c + 1 : jst + 2 * c;
          double *pe = (odd ? pol_o : pol_e) + (size_t)nkrs * c;
          double *te = (odd ? tor_o : tor_e) + (size_t)nkrt * c;
          const double g7 = p->g7[j], gm = (double)p->idx_j[j + 2 * jmax];
          for (kk = 0; kk < nkrv; kk++) {
            const int kl = kk % nkr, nd = kk / nkr, k = k0 + kl;
            const double r1 = p->radius[k], r2 = r1 * r1;
            double *w = WS + 3 * nd + ncomp * (j + (size_t)jmax * k);
            pe[kk] = pe[kk] * r2 * g7;
            te[kk + nkrv] = te[kk + nkrv] * r1 * g7;
            pe[kk + 2 * nkrv] = pe[kk + 2 * nkrv] * r1 * g7 * gm;
            te[kk] = te[kk] * r1 * g7;
            pe[kk + nkrv] = pe[kk + nkrv] * r1 * g7 * gm;
            w[0] = w[0] + pe[kk];
            w[1] = w[1] - pe[kk + 2 * nkrv] + te[kk + nkrv];
            w[2] = w[2] - pe[kk + nkrv] - te[kk];
          }
          for (kk = 0; kk < nkr * ns; kk++) {
            const int kl = kk % nkr, nd = kk / nkr, k = k0 + kl;
            double *w = WS + nd + 3 * nv + ncomp * (j + (size_t)jmax * k);
            pe[kk + 3 * nkrv] = pe[kk + 3 * nkrv] * p->g6[j];
            w[0] = w[0] + pe[kk + 3 * nkrv];
          }
        }
      }
      free(symp_r); free(asmp_r); free(symp_p); free(asmp_p);
      free(pol_e); free(pol_o); free(tor_e); free(tor_o);
    }
  }
}

/* ------------------------------------------------------------------ */
/* Mixed-radix Stockham complex FFT (radix 4, 2, 3, 5, generic), used for
 * pairs of real lines packed as real + i*imag. */
typedef struct {
  int n, nf, fac[40];
  cplx *tw[40]; /* per stage: Ns*R twiddles */
  cplx *root[40];
} cfft_plan;

static void cfft_plan_init(cfft_plan *f, int n, int sign) {
  int nn = n, ns = 1, st;
  const long double pi = 3.14159265358979323846264338327950288L;
  f->n = n;
  f->nf = 0;
  while (nn % 4 == 0) { f->fac[f->nf++] = 4; nn /= 4; }
  while (nn % 2 == 0) { f->fac[f->nf++] = 2; nn /= 2; }
  {
    int q = 3;
    while (nn > 1) {
      while (nn % q == 0) { f->fac[f->nf++] = q; nn /= q; }
      q += 2;
    }
  }
  for (st = 0; st < f->nf; st++) {
    int R = f->fac[st], k, r;
    f->tw[st] = (cplx *)malloc(sizeof(cplx) * ns * R);
    f->root[st] = (cplx *)malloc(sizeof(cplx) * R);
    for (k = 0; k < ns; k++)
      for (r = 0; r < R; r++) {
        long double a = sign * 2.0L * pi * (long double)(k * r) / (long double)(ns * R);
        f->tw[st][k * R + r] = (double)cosl(a) + I * (double)sinl(a);
      }
    for (r = 0; r < R; r++) {
      long double a = sign * 2.0L * pi * (long double)r / (long double)R;
      f->root[st][r] = (double)cosl(a) + I * (double)sinl(a);
    }
    ns *= R;
  }
}

static void cfft_plan_free(cfft_plan *f) {
  int st;
  for (st = 0; st < f->nf; st++) {
    free(f->tw[st]);
    free(f->root[st]);
  }
}

/* x: in/out, y: work, both n.  Govindaraju et al. (SC'08) Stockham form. */
static void cfft_exec(const cfft_plan *f, cplx *x, cplx *y, int sign) {
  const int n = f->n;
  int st, ns = 1;
  cplx *a = x, *b = y;
  cplx v[64];
  for (st = 0; st < f->nf; st++) {
    const int R = f->fac[st], nR = n / R;
    int j, r, q;
    for (j = 0; j < nR; j++) {
      const int k = j % ns;
      const cplx *tw = f->tw[st] + k * R;
      const int idx = (j / ns) * ns * R + k;
      cplx *vv = v;
      cplx *big = NULL;
      if (R > 64) big = vv = (cplx *)malloc(sizeof(cplx) * R);
      for (r = 0; r < R; r++) vv[r] = a[j + r * nR] * tw[r];
      if (R == 2) {
        cplx t = vv[0];
        vv[0] = t + vv[1];
        vv[1] = t - vv[1];
      } else if (R == 4) {
        cplx t0 = vv[0] + vv[2], t1 = vv[0] - vv[2];
        cplx t2 = vv[1] + vv[3], t3 = (vv[1] - vv[3]) * (sign > 0 ? I : -I);
        vv[0] = t0 + t2;
        vv[2] = t0 - t2;
        vv[1] = t1 + t3;
        vv[3] = t1 - t3;
      } else {
        cplx o[64], *out = o;
        cplx *bigo = NULL;
        if (R > 64) bigo = out = (cplx *)malloc(sizeof(cplx) * R);
        for (q = 0; q < R; q++) {
          cplx acc = 0;
          for (r = 0; r < R; r++) acc += vv[r] * f->root[st][(r * q) % R];
          out[q] = acc;
        }
        for (q = 0; q < R; q++) vv[q] = out[q];
        free(bigo);
      }
      for (r = 0; r < R; r++) b[idx + r * ns] = vv[r];
      free(big);
    }
    {
      cplx *t = a;
      a = b;
      b = t;
    }
    ns *= R;
  }
  if (a != x) memcpy(x, a, sizeof(cplx) * n);
}

void orc_rfft_fwd(int n, int nlines, const double *x, double *c) {
  cfft_plan f;
  int pr;
  const double sn = 1.0 / (double)n, tsn = 2.0 / (double)n;
  cfft_plan_init(&f, n, -1);
#pragma omp parallel for schedule(static)
  for (pr = 0; pr < (nlines + 1) / 2; pr++) {
    cplx *z = (cplx *)malloc(sizeof(cplx) * n), *y = (cplx *)malloc(sizeof(cplx) * n);
    const int la = 2 * pr, lb = 2 * pr + 1;
    int i, k;
    for (i = 0; i < n; i++)
      z[i] = x[(size_t)la * n + i] + I * (lb < nlines ? x[(size_t)lb * n + i] : 0.0);
    cfft_exec(&f, z, y, -1);
    for (k = 0; 2 * k <= n; k++) {
      const cplx zk = z[k], zn = conj(z[(n - k) % n]);
      const cplx xa = 0.5 * (zk + zn), xb = -0.5 * I * (zk - zn);
      const cplx xs[2] = {xa, xb};
      int t;
      for (t = 0; t < 2; t++) {
        double *cc;
        if (la + t >= nlines) continue;
        cc = c + (size_t)(la + t) * n;
        if (k == 0)
          cc[0] = sn * creal(xs[t]);
        else if (2 * k == n)
          cc[n - 1] = sn * creal(xs[t]);
        else {
          cc[2 * k - 1] = tsn * creal(xs[t]);
          cc[2 * k] = -tsn * cimag(xs[t]);
        }
      }
    }
    free(z);
    free(y);
  }
  cfft_plan_free(&f);
}

void orc_rfft_bwd(int n, int nlines, const double *c, double *x) {
  cfft_plan f;
  int pr;
  cfft_plan_init(&f, n, +1);
#pragma omp parallel for schedule(static)
  for (pr = 0; pr < (nlines + 1) / 2; pr++) {
    cplx *z = (cplx *)malloc(sizeof(cplx) * n), *y = (cplx *)malloc(sizeof(cplx) * n);
    const int la = 2 * pr, lb = 2 * pr + 1;
    int i, k, t;
    for (k = 0; k < n; k++) z[k] = 0;
    for (t = 0; t < 2; t++) {
      const double *cc;
      const cplx fac = t == 0 ? 1.0 : I;
      if (la + t >= nlines) continue;
      cc = c + (size_t)(la + t) * n;
      z[0] += fac * cc[0];
      for (k = 1; 2 * k < n; k++) {
        const cplx ck = 0.5 * (cc[2 * k - 1] - I * cc[2 * k]);
        z[k] += fac * ck;
        z[n - k] += fac * conj(ck);
      }
      if ((n & 1) == 0) z[n / 2] += fac * cc[n - 1];
    }
    cfft_exec(&f, z, y, +1);
    for (i = 0; i < n; i++) {
      x[(size_t)la * n + i] = creal(z[i]);
      if (lb < nlines) x[(size_t)lb * n + i] = cimag(z[i]);
    }
    free(z);
    free(y);
  }
  cfft_plan_free(&f);
}

/* Backward FFT stage, SS/t_sph_FFTPACK5.f90:259-331 (sph_RFFTMB_from_recv):
 * gather the rtp-spectral slots of each (r,theta,comp) line into FFTPACK
 * order, transform, scatter to the field-major grid. */
void orc_fft_bwd_from_rtp(const orc_plan *p, int ncomp, const double *rtp,
                          double *v) {
  const int n = p->nphi;
  const size_t nline = (size_t)p->nr * p->nth, nnod = nline * n;
  const size_t nl = nline * ncomp;
  double *c = (double *)malloc(sizeof(double) * nl * n);
  double *x = (double *)malloc(sizeof(double) * nl * n);
  long long li;
#pragma omp parallel for
  for (li = 0; li < (long long)nl; li++) {
    const int nd = (int)(li % ncomp);
    const size_t line = (size_t)li / ncomp;
    double *cc = c + (size_t)li * n;
    int m;
    cc[0] = rtp[nd + ncomp * (line + nline * 0)];
    cc[n - 1] = rtp[nd + ncomp * (line + nline * 1)];
    for (m = 1; 2 * m < n; m++) {
      cc[2 * m - 1] = rtp[nd + ncomp * (line + nline * (2 * m))];
      cc[2 * m] = rtp[nd + ncomp * (line + nline * (2 * m + 1))];
    }
  }
  orc_rfft_bwd(n, (int)nl, c, x);
#pragma omp parallel for
  for (li = 0; li < (long long)nl; li++) {
    const int nd = (int)(li % ncomp);
    const size_t line = (size_t)li / ncomp;
    int i;
    for (i = 0; i < n; i++) v[line + nline * i + nnod * nd] = x[(size_t)li * n + i];
  }
  free(c);
  free(x);
}

/* Forward FFT stage, SS/t_sph_FFTPACK5.f90:183-255 (sph_RFFTMF_to_send). */
void orc_fft_fwd_to_rtp(const orc_plan *p, int ncomp, const double *v,
                        double *rtp) {
  const int n = p->nphi;
  const size_t nline = (size_t)p->nr * p->nth, nnod = nline * n;
  const size_t nl = nline * ncomp;
  double *c = (double *)malloc(sizeof(double) * nl * n);
  double *x = (double *)malloc(sizeof(double) * nl * n);
  long long li;
#pragma omp parallel for
  for (li = 0; li < (long long)nl; li++) {
    const int nd = (int)(li % ncomp);
    const size_t line = (size_t)li / ncomp;
    int i;
    for (i = 0; i < n; i++) x[(size_t)li * n + i] = v[line + nline * i + nnod * nd];
  }
  orc_rfft_fwd(n, (int)nl, x, c);
#pragma omp parallel for
  for (li = 0; li < (long long)nl; li++) {
    const int nd = (int)(li % ncomp);
    const size_t line = (size_t)li / ncomp;
    const double *cc = c + (size_t)li * n;
    int m;
    rtp[nd + ncomp * (line + nline * 0)] = cc[0];
    rtp[nd + ncomp * (line + nline * 1)] = cc[n - 1];
    for (m = 1; 2 * m < n; m++) {
      rtp[nd + ncomp * (line + nline * (2 * m))] = cc[2 * m - 1];
      rtp[nd + ncomp * (line + nline * (2 * m + 1))] = cc[2 * m];
    }
  }
  free(c);
  free(x);
}

/* SS/t_sph_transforms.f90:74-123.  Single rank: the rj->rlm exchange is the
 * self-copy of SR/calypso_SR_core.f90:101-109 with identity tables. */
void orc_sht_backward(const orc_plan *p, int nv, int ns, const double *WS_rj,
                      double *v_rtp, int variant) {
  const int ncomp = 3 * nv + ns;
  const size_t nrtm = (size_t)p->nth * p->nr * p->nslot;
  const size_t nrtp = (size_t)p->nr * p->nth * p->nphi;
  double *rtm = (double *)malloc(sizeof(double) * ncomp * nrtm);
  double *rtp = (double *)malloc(sizeof(double) * ncomp * nrtp);
  if (variant == 0)
    orc_leg_bwd_org(p, nv, ns, WS_rj, rtm);
  else
    orc_leg_bwd_sym(p, nv, ns, WS_rj, rtm);
  orc_rtm_to_rtp(p, ncomp, rtm, rtp);
  orc_fft_bwd_from_rtp(p, ncomp, rtp, v_rtp);
  free(rtm);
  free(rtp);
}

/* SS/t_sph_transforms.f90:250-297 */
void orc_sht_forward(const orc_plan *p, int nv, int ns, const double *v_rtp,
                     double *WR_rj, int variant) {
  const int ncomp = 3 * nv + ns;
  const size_t nrtm = (size_t)p->nth * p->nr * p->nslot;
  const size_t nrtp = (size_t)p->nr * p->nth * p->nphi;
  double *rtm = (double *)malloc(sizeof(double) * ncomp * nrtm);
  double *rtp = (double *)malloc(sizeof(double) * ncomp * nrtp);
  orc_fft_fwd_to_rtp(p, ncomp, v_rtp, rtp);
  orc_rtp_to_rtm(p, ncomp, rtp, rtm);
  if (variant == 0)
    orc_leg_fwd_org(p, nv, ns, rtm, WR_rj);
  else
    orc_leg_fwd_sym(p, nv, ns, rtm, WR_rj);
  free(rtm);
  free(rtp);
}

/* ------------------------------------------------------------------ */
/* Backward transform at the poles and the centre.
 * Pole tables: SP/schmidt_poly_on_rtm_grid.f90:201-266
 * (set_lagender_pole_rlm): the Schmidt recurrence at theta = 0 and at
 * theta = 4 atan(1); only P of order 0 and dP/dtheta of order 1 are read.
 * Poles: SS/schmidt_b_trans_at_pole.f90:88-141 (scalar) and :145-289
 * (vector); centre: SS/schmidt_b_trans_at_center.f90:45-62 (scalar, leaves
 * the row as it is) and :66-115 (vector).  The vector centre follows that
 * code (x from order -1, y from order +1), not its header comment (which
 * names the cos mode for x); tests pin the centre row bit for bit against
 * that routine compiled from its source.  Rows of v_pole: 0..nr-1 north
 * pole, nr..2nr-1 south pole, 2nr the centre (nnod_pole = 2nr + 1,
 * SP/t_spheric_rtp_data.f90:187); v_pole[i + nnod_pole*nd].  The centre row
 * is written for vector components only and only when center != 0
 * (iflag_MESH_w_center); every other centre entry keeps its value. */
void orc_pole_backward(const orc_plan *p, int nv, int ns, const double *WS,
                       double *v_pole, int center) {
  const int L = p->L, nr = p->nr, ncomp = 3 * nv + ns, npole = 2 * nr + 1;
  const size_t jmax = (size_t)p->jmax;
  const int j0 = p->lstack[L], jm = p->lstack[L - 1], jp = p->lstack[L + 1];
  double *pn = (double *)calloc((size_t)(L + 3), sizeof(double));
  double *ps = (double *)calloc((size_t)(L + 3), sizeof(double));
  double *dn = (double *)calloc((size_t)(L + 3), sizeof(double));
  double *ds = (double *)calloc((size_t)(L + 3), sizeof(double));
  double *tmp = (double *)calloc((size_t)(L + 3), sizeof(double));
  const double pi = 4.0 * atan(1.0);
  int k, nd, l;
  orc_schmidt_legendres_m(L, 0, 0.0, pn, tmp);
  orc_schmidt_legendres_m(L, 1, 0.0, tmp, dn);
  orc_schmidt_legendres_m(L, 0, pi, ps, tmp);
  orc_schmidt_legendres_m(L, 1, pi, tmp, ds);
  for (nd = 0; nd < ncomp; nd++)
    for (k = 0; k < 2 * nr; k++) v_pole[k + (size_t)npole * nd] = 0.0;
#define SP(c, j, kk) WS[(c) + (size_t)ncomp * ((size_t)(j) + jmax * (size_t)(kk))]
#define VP(i, c) v_pole[(size_t)(i) + (size_t)npole * (size_t)(c)]
  for (k = 0; k < nr; k++) {
    const int kn = k, ks = k + nr;
    const double a = p->a_r[k];
    for (nd = 0; nd < nv; nd++) {
      const int cx = 3 * nd, cy = 3 * nd + 1, cz = 3 * nd + 2;
      for (l = 0; l <= L; l++) { /* order 0: pol (c = 3nd) */
        const int j = j0 + l;
        VP(kn, cz) = VP(kn, cz) + SP(3 * nd, j, k) * pn[l] * p->g3[j];
        VP(ks, cz) = VP(ks, cz) - SP(3 * nd, j, k) * ps[l] * p->g3[j];
      }
      for (l = 1; l <= L; l++) { /* order -1 (1s) */
        const int j = jm + l - 1;
        VP(kn, cx) = VP(kn, cx) + SP(3 * nd + 2, j, k) * dn[l];
        VP(kn, cy) = VP(kn, cy) + SP(3 * nd + 1, j, k) * dn[l];
        VP(ks, cx) = VP(ks, cx) + SP(3 * nd + 2, j, k) * ds[l];
        VP(ks, cy) = VP(ks, cy) - SP(3 * nd + 1, j, k) * ds[l];
      }
      for (l = 1; l <= L; l++) { /* order +1 (1c) */
        const int j = jp + l - 1;
        VP(kn, cx) = VP(kn, cx) + SP(3 * nd + 1, j, k) * dn[l];
        VP(kn, cy) = VP(kn, cy) - SP(3 * nd + 2, j, k) * dn[l];
        VP(ks, cx) = VP(ks, cx) - SP(3 * nd + 1, j, k) * ds[l];
        VP(ks, cy) = VP(ks, cy) - SP(3 * nd + 2, j, k) * ds[l];
      }
      VP(kn, cx) = VP(kn, cx) * a;
      VP(kn, cy) = VP(kn, cy) * a;
      VP(kn, cz) = VP(kn, cz) * a * a;
      VP(ks, cx) = VP(ks, cx) * a;
      VP(ks, cy) = VP(ks, cy) * a;
      VP(ks, cz) = VP(ks, cz) * a * a;
    }
    for (nd = 3 * nv; nd < ncomp; nd++)
      for (l = 0; l <= L; l++) {
        VP(kn, nd) = VP(kn, nd) + SP(nd, j0 + l, k) * pn[l];
        VP(ks, nd) = VP(ks, nd) + SP(nd, j0 + l, k) * ps[l];
      }
  }
  if (center) {
    const double a = p->a_r[0];
    for (nd = 0; nd < nv; nd++) {
      VP(2 * nr, 3 * nd + 2) = 2.0 * SP(3 * nd, j0 + 1, 0) * a * a;
      VP(2 * nr, 3 * nd) = -(2.0 * SP(3 * nd, jm, 0) * a * a);
      VP(2 * nr, 3 * nd + 1) = -(2.0 * SP(3 * nd, jp, 0) * a * a);
    }
  }
#undef SP
#undef VP
  free(pn);
  free(ps);
  free(dn);
  free(ds);
  free(tmp);
}

/* ------------------------------------------------------------------ */
/* Grid-space products of the MHD step, MS/cal_nonlinear_sph_MHD.f90:64-128
 * (nonlinear_terms_in_rtp).  Cross products with a coefficient
 * (CORE_src/BASE/cal_products_smp.f90:153-170), vector times scalar
 * (:218-233) and the Coriolis term -coef (omega x u), omega = (cos theta,
 * -sin theta, 0) in (r, theta, phi) (MHD_src/sph_MHD/const_wz_coriolis_rtp.f90:
 * 117-156).  Fields are field-major on the rtp grid, f[inod + nnod*nd],
 * inod = k + nr*(l_theta + nth*i_phi).
 * idx: first component (0-based, -1 = absent / off) of velo, vort, magne,
 * current, temp, light (inputs) and m_advect, lorentz, vp_induct, h_flux,
 * c_flux, coriolis (outputs).  coef: velo, lorentz, induct, h_advect,
 * c_advect, coriolis. */
static void cross_coef(size_t n, double c, const double *a, const double *b,
                       double *o) {
  size_t i;
  for (i = 0; i < n; i++) {
    const double a1 = a[i], a2 = a[i + n], a3 = a[i + 2 * n];
    const double b1 = b[i], b2 = b[i + n], b3 = b[i + 2 * n];
    o[i] = (a2 * b3 - a3 * b2) * c;
    o[i + n] = (a3 * b1 - a1 * b3) * c;
    o[i + 2 * n] = (a1 * b2 - a2 * b1) * c;
  }
}
void orc_nonlinear_rtp(const orc_plan *p, const int idx[12],
                       const double coef[6], int ncomp_fld, const double *fld,
                       int ncomp_frc, double *frc) {
  const size_t n = (size_t)p->nr * p->nth * p->nphi;
  const double *u = idx[0] >= 0 ? fld + n * idx[0] : NULL;
  const double *w = idx[1] >= 0 ? fld + n * idx[1] : NULL;
  const double *b = idx[2] >= 0 ? fld + n * idx[2] : NULL;
  const double *j = idx[3] >= 0 ? fld + n * idx[3] : NULL;
  const double *tt = idx[4] >= 0 ? fld + n * idx[4] : NULL;
  const double *cc = idx[5] >= 0 ? fld + n * idx[5] : NULL;
  (void)ncomp_fld;
  (void)ncomp_frc;
  if (idx[6] >= 0) cross_coef(n, coef[0], w, u, frc + n * idx[6]);
  if (idx[7] >= 0) cross_coef(n, coef[1], j, b, frc + n * idx[7]);
  if (idx[8] >= 0) cross_coef(n, coef[2], u, b, frc + n * idx[8]);
  for (int f = 0; f < 2; f++) {
    const double *s = f == 0 ? tt : cc;
    const int o = idx[9 + f];
    const double c = coef[3 + f];
    size_t i;
    if (o < 0) continue;
    for (i = 0; i < n; i++) {
      frc[i + n * o] = u[i] * s[i] * c;
      frc[i + n * (o + 1)] = u[i + n] * s[i] * c;
      frc[i + n * (o + 2)] = u[i + 2 * n] * s[i] * c;
    }
  }
  if (idx[11] >= 0) {
    double *o = frc + n * idx[11];
    size_t i;
    for (i = 0; i < n; i++) {
      const int lt = (int)((i / (size_t)p->nr) % (size_t)p->nth);
      const double o1 = cos(p->colat[lt]), o2 = -sin(p->colat[lt]), o3 = 0.0;
      const double c = coef[5];
      o[i] = -(c * (o2 * u[i + 2 * n] - o3 * u[i + n]));
      o[i + n] = -(c * (o3 * u[i] - o1 * u[i + 2 * n]));
      o[i + 2 * n] = -(c * (o1 * u[i + n] - o2 * u[i]));
    }
  }
}

/* ------------------------------------------------------------------ */
/* rj fields <-> interleaved send / receive buffers,
 * SR/field_to_send_buffer.f90:43-120 (to send) and :246-330 (from the
 * receive buffer by the reverse table); per field the reference routine of
 * its component count, in the order of the field list. */
void orc_rj_to_send(int nfield, const int *fields, int ntot_phys, long long nnod_rj,
                    const double *fld, const int *inod_export, long long nnod_send, int nb,
                    double *WS) {
  int f, c;
  long long k;
  (void)ntot_phys;
  for (f = 0; f < nfield; f++) {
    const int i_fld = fields[3 * f], i_ws = fields[3 * f + 1], nc = fields[3 * f + 2];
    for (k = 0; k < nnod_send; k++) {
      const long long kk = (long long)nb * k + i_ws;
      const long long jj = inod_export ? (long long)inod_export[k] - 1 : k;
      for (c = 0; c < nc; c++) WS[kk + c] = fld[jj + nnod_rj * (long long)(i_fld + c)];
    }
  }
}

void orc_rj_from_recv(int nfield, const int *fields, int nb, long long nnod_recv,
                      const double *WR, const int *irev, int ntot_phys, long long nnod_rj,
                      double *fld) {
  int f, c;
  long long k;
  (void)ntot_phys;
  for (f = 0; f < nfield; f++) {
    const int i_fld = fields[3 * f], i_wr = fields[3 * f + 1], nc = fields[3 * f + 2];
    for (k = 0; k < nnod_rj; k++) {
      const long long r = irev ? (long long)irev[k] - 1 : k;
      for (c = 0; c < nc; c++)
        fld[k + nnod_rj * (long long)(i_fld + c)] =
            r == nnod_recv ? 0.0 : WR[(long long)nb * r + i_wr + c];
    }
  }
}

/* ------------------------------------------------------------------ */
/* Coriolis term in rlm space (MS/ = MHD/Fortran_src/MHD_src/sph_MHD/).
 *
 * Tables of init_sum_coriolis_rlm (MS/t_coriolis_terms_rlm.f90:133-183):
 * cal_gaunt_coriolis_rlm (MS/interact_coriolis_rlm.f90:42-118) takes, for
 * each rlm j = (l, m), the Gaunt integrals
 *   gi(j,1) = K(1,0; l-1,m; l,m)  (0 unless 1 <= l-1 and |m| <= l-1),
 *   gi(j,2) = K(1,0; l+1,m; l,m)  (0 unless l+1 <= L),
 * and the Elsasser integral ei(j) = L(1,0; l,-m; l,m), with the rlm
 * address of the partner mode (the mode itself where the integral is 0;
 * all partners (0,0) for j = 0).  The reference evaluates them from prime
 * factorisations (SH/normgi_by_primes.f90, SH/normei.f90,
 * SH/cal_gaunt_itgs.f90:84-164); for these index sets they reduce, with the
 * Schmidt recurrence x P_l^m = (sqrt((l+1)^2-m^2) P_(l+1)^m +
 * sqrt(l^2-m^2) P_(l-1)^m)/(2l+1), int (P_l^m)^2 dx = 2(2-d_m0)/(2l+1) and
 * normki's factor 2 for m = 0, to
 *   gi(j,1) = 4 sqrt(l^2 - m^2) / ((2l-1)(2l+1)),
 *   gi(j,2) = 4 sqrt((l+1)^2 - m^2) / ((2l+1)(2l+3)),
 *   ei(j)   = -4 m / (2l+1)  (l >= 1; m = 0 gives 0),
 * the sign being that of setei's combination of Gaunt integrals
 * (SH/normei.f90:76-150) with normli's for m2 = -m3 (tests pin all three
 * against the routines compiled from source: addresses equal, values within
 * a few ulp, since the reference forms them from products of square roots).
 * interact_rot_coriolis_rlm (:122-205) then forms sw(2,3,j) and tw(2,4,j)
 * from these and g3, g4, g5, g17 (g3(0,0) = 1/2), in the reference's
 * operation order.  jgi[j + jmax*c], jei[j] are 1-based as in the
 * reference; sw[a + 2*(b + 3*j)], tw[a + 2*(b + 4*j)]. */
void orc_coriolis_tables(const orc_plan *p, int *jgi, int *jei, double *gi, double *ei,
                         double *sw, double *tw) {
  const int jmax = p->jmax, L = p->L;
  const double two = 2.0, half = 0.5, zero = 0.0;
  int j;
  /* local rlm address of (l, m) in the single-rank order: slot s = m + L
   * holds l = |m|..L from lstack[s] */
#define ORC_JADDR(l, m) (p->lstack[(m) + L] + ((l) - abs(m)) + 1)
  for (j = 0; j < jmax; j++) {
    const int l = p->idx_j[j + jmax], m = p->idx_j[j + 2 * jmax];
    const int j3 = p->idx_j[j];
    double g[17], g3, g4, g5, g17, gi1 = zero, gi2 = zero, e1 = zero;
    int a1 = ORC_JADDR(l, m), a2 = ORC_JADDR(l, m), ae = ORC_JADDR(l, m);
    if (l - 1 >= 1 && abs(m) <= l - 1) {
      gi1 = 4.0 * sqrt((double)(l * l - m * m)) / ((double)(2 * l - 1) * (double)(2 * l + 1));
      a1 = ORC_JADDR(l - 1, m);
    }
    if (l + 1 <= L) {
      gi2 = 4.0 * sqrt((double)((l + 1) * (l + 1) - m * m)) /
            ((double)(2 * l + 1) * (double)(2 * l + 3));
      a2 = ORC_JADDR(l + 1, m);
    }
    if (l >= 1) {
      if (m != 0) e1 = -4.0 * (double)m / (double)(2 * l + 1);
      ae = ORC_JADDR(l, -m);
    }
    if (j3 == 0) {
      a1 = a2 = ae = ORC_JADDR(0, 0);
      gi2 = zero;
      e1 = zero;
    }
    jgi[j] = a1;
    jgi[j + jmax] = a2;
    jei[j] = ae;
    gi[j] = gi1;
    gi[j + jmax] = gi2;
    ei[j] = e1;
    orc_sph_normalizations(l, m, g);
    g3 = (l == 0 && m == 0) ? half : g[2];
    g4 = g[3];
    g5 = g[4];
    g17 = g[16];
    {
      double *s = sw + 6 * (long long)j, *t = tw + 8 * (long long)j;
      s[0] = (two - g4 - g3) * gi1 * g17;
      s[1] = (two - g5 - g3) * gi2 * g17;
      s[2] = g4 * (two - g4 + g3) * gi1 * g17 * half;
      s[3] = g5 * (two - g5 + g3) * gi2 * g17 * half;
      s[4] = two * e1 * g17;
      s[5] = zero;
      t[4] = -(g3 * (two + g4 - g3) + two * (-two + g4 + g3)) * gi1 * g17 * half;
      t[5] = -(g3 * (two + g5 - g3) + two * (-two + g5 + g3)) * gi2 * g17 * half;
      t[6] = -(-two + g4 + g3) * gi1 * g17;
      t[7] = -(-two + g5 + g3) * gi2 * g17;
      t[0] = -two * e1 * g17;
      t[1] = zero;
      t[2] = -g3 * e1 * g17;
      t[3] = zero;
    }
  }
#undef ORC_JADDR
}

/* sum_rot_coriolis_rlm_10 (MS/sum_coriolis_terms_rlm.f90:96-186): curl of
 * the Coriolis force of a rotation along z, poloidal (d_cor[i]) and toroidal
 * (d_cor[i + nnod_rlm]) parts per rlm node i = j + jmax*k, from velocity and
 * vorticity spectra (pol, dpol/dr, tor) at components i_velo, i_vort
 * (0-based) of the comp-interleaved receive buffer WR (record irev[i]-1;
 * irev NULL = identity; the dummy record irev = nnod_rlm + 1 reads zero).  The rotation spectrum is that of
 * set_zdir_rot_earth_4_sph with rotate = 1 (MS/t_poloidal_rotation.f90:
 * 159-250): omega0 = r^2/2, omega1 = r, omega2 = 1.  kr_icb > 0 (1-based
 * radius) adds inner_core_rot_z_coriolis_rlm (:378-418) for a rotatable
 * inner core.  Operation order as in the reference. */
void orc_sum_coriolis_rlm(const orc_plan *p, int ncomp, int i_velo, int i_vort, double coef_cor,
                          const double *WR, const int *irev, int kr_icb, double *d_cor) {
  const int jmax = p->jmax, nr = p->nr;
  const long long nnod = (long long)jmax * nr;
  int *jgi = (int *)malloc(sizeof(int) * 3 * jmax);
  int *jei = jgi + 2 * jmax;
  double *gi = (double *)malloc(sizeof(double) * 17 * jmax);
  double *ei = gi + 2 * jmax, *sw = ei + jmax, *tw = sw + 6 * jmax;
  const double two = 2.0, half = 0.5;
  int k;
  /* record r of the receive buffer; the dummy record r = nnod (irev =
   * nnod_rlm + 1) reads zero, as the reference zeroes it
   * (SR/set_from_recv_buf_rev.f90:77-87) */
#define ORC_WR(c, r) ((r) < 0 || (r) >= nnod ? 0.0 : WR[(c) + (long long)ncomp * (r)])
  orc_coriolis_tables(p, jgi, jei, gi, ei, sw, tw);
#pragma omp parallel for schedule(static)
  for (k = 0; k < nr; k++) {
    const double ar = p->a_r[k], r = p->radius[k];
    const double om0 = half * 1.0 * (r * r), om1 = 1.0 * r, om2 = 1.0;
    int j;
    for (j = 0; j < jmax; j++) {
      const long long i = j + (long long)jmax * k;
      const long long i11 = jgi[j] - 1 + (long long)jmax * k;
      const long long i21 = jgi[j + jmax] - 1 + (long long)jmax * k;
      const long long i12 = jei[j] - 1 + (long long)jmax * k;
      const long long r11 = irev ? irev[i11] - 1 : i11;
      const long long r21 = irev ? irev[i21] - 1 : i21;
      const long long r12 = irev ? irev[i12] - 1 : i12;
      const double *s = sw + 6 * (long long)j, *t = tw + 8 * (long long)j;
      const double dvp_k1 = ORC_WR(i_velo + 1, r11);
      const double dvp_k2 = ORC_WR(i_velo + 1, r21);
      const double wp_l1 = ORC_WR(i_vort, r12);
      const double vp_k1 = ORC_WR(i_velo, r11);
      const double vp_k2 = ORC_WR(i_velo, r21);
      const double d2vp_l1 = ar * ar * p->g3[j] * ORC_WR(i_velo, r12) -
                             ORC_WR(i_vort + 2, r12);
      const double vp_l1 = ORC_WR(i_velo, r12);
      const double wp_k1 = ORC_WR(i_vort, r11);
      const double wp_k2 = ORC_WR(i_vort, r21);
      const double dwp_k1 = ORC_WR(i_vort + 1, r11) - two * ar * ORC_WR(i_vort, r11);
      const double dwp_k2 = ORC_WR(i_vort + 1, r21) - two * ar * ORC_WR(i_vort, r21);
      double rp = s[0] * om0 * dvp_k1 + s[1] * om0 * dvp_k2 + s[4] * om0 * wp_l1 +
                  s[2] * om1 * vp_k1 + s[3] * om1 * vp_k2;
      double rt = t[0] * om0 * d2vp_l1 - t[2] * om2 * vp_l1 + t[4] * om1 * wp_k1 +
                  t[5] * om1 * wp_k2 + t[6] * om0 * dwp_k1 + t[7] * om0 * dwp_k2;
      d_cor[i] = -coef_cor * rp * ar * ar;
      d_cor[i + nnod] = -coef_cor * rt * ar * ar;
    }
  }
  if (kr_icb > 0 && kr_icb <= nr && p->L >= 1) {
    /* inner_core_rot_z_coriolis_rlm: degree-one poloidal terms at the ICB */
    const int L = p->L;
    const long long ko = (long long)jmax * (kr_icb - 1);
    const long long i11s = p->lstack[-1 + L] + (1 - 1) + ko; /* (1, -1) */
    const long long i10c = p->lstack[0 + L] + (1 - 0) + ko;  /* (1, 0) */
    const long long i11c = p->lstack[1 + L] + (1 - 1) + ko;  /* (1, 1) */
    const long long r11s = irev ? irev[i11s] - 1 : i11s;
    const long long r11c = irev ? irev[i11c] - 1 : i11c;
    const double rc = p->radius[kr_icb - 1];
    const double om0 = half * 1.0 * (rc * rc);
    const double wp11s = ORC_WR(i_vort, r11s);
    const double wp11c = ORC_WR(i_vort, r11c);
    d_cor[i10c] = 0.0;
    d_cor[i11s] = -two * coef_cor * rc * om0 * wp11c;
    d_cor[i11c] = two * coef_cor * rc * om0 * wp11s;
  }
#undef ORC_WR
  free(jgi);
  free(gi);
}
