/*
 * sht_oracle.h -- CPU FP64 restatement of the reference spherical-harmonic
 * transform (Kemorin/Calypso sph_trans), used ONLY as test infrastructure:
 * the checker in tests/, the oracle leg of __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product library links it.
 *
 * All arrays are 0-based C arrays laid out exactly like the reference's
 * Fortran arrays (first index fastest).  Citations are to files of the
 * reference tree; abbreviations:
 *   SS/ = MHD/Fortran_src/PARALLEL_src/SPH_SHELL_src/
 *   SH/ = MHD/Fortran_src/CORE_src/spherical_harmonics/
 *   GI/ = MHD/Fortran_src/CORE_src/gauss_integration/
 *   SP/ = MHD/Fortran_src/CORE_src/SPH_SPECTR_src/
 *   CG/ = MHD/Fortran_src/PARALLEL_src/CONST_SPH_GRID/
 *   FP5/= MHD/External_libs/FFTPACK5.1D/
 */
#ifndef KEMORIN_SHT_ORACLE_H
#define KEMORIN_SHT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_plan orc_plan;

/* Gauss-Legendre points on [-1,1] (ascending) and weights.
 * exact==0: the reference's Newton loop, GI/gauss_integration.f90:60-133
 *           (tolerance 3.0e-12 as a REAL(4) literal, weight from P' taken
 *           at the pre-update iterate).
 * exact==1: Newton run to convergence, weight from the converged node. */
void orc_gauss_points(int n, int exact, double *x, double *w);

/* Schmidt quasi-normalised P_l^m(theta) and dP_l^m/dtheta for l=0..ltr,
 * SH/schmidt_fix_m.f90:62-175.  p[l], dp[l] valid for l>=m. */
void orc_schmidt_legendres_m(int ltr, int m, double theta, double *p,
                             double *dp);

/* g(1:17) normalisation table, SH/spherical_harmonics.f90:91-140. */
void orc_sph_normalizations(int l, int m, double g[17]);

/* Chebyshev radial grid r_k = ri + (ro-ri)/2 (1-cos(pi k/(nr-1))),
 * SP/chebyshev_radial_grid.f90 (set_chebyshev_distance_shell). */
void orc_chebyshev_radius(int nr, double ri, double ro, double *r);

/* Plan: truncation L, nr radii, nth Gauss colatitudes, nphi longitudes.
 * flags bit0: exact Gauss weights; bit1: keep full-sphere P/dP tables
 * (needed by the *_org Legendre loops). */
orc_plan *orc_plan_create(int L, int nr, int nth, int nphi,
                          const double *radius, int flags);
void orc_plan_destroy(orc_plan *p);
/* Named table access: "radius","a_r","gpoint","colat","weight","asin_t",
 * "g3","g6","g7" (jmax), "P","dP" (nth x jmax, full tables only),
 * "Ps","dPs" (nth_h x jmax, per-slot even-first order),
 * int tables: "lstack","lstack_even" (nslot+1), "idx_j" (jmax x 3),
 * "mdx_p","mdx_n" (jmax). */
const void *orc_plan_table(const orc_plan *p, const char *name);
void orc_plan_dims(const orc_plan *p, int dims[8]);

/* rtm -> rtp exchange (single rank): comp-interleaved rtm buffer
 * (theta fastest, then r, then m-slot) to the comp-interleaved
 * rtp-spectral buffer (r fastest, then theta, then FFT slot p; slots with
 * no Legendre order are the reference's zero dummy slot). */
void orc_rtm_to_rtp(const orc_plan *p, int ncomp, const double *rtm,
                    double *rtp);
void orc_rtp_to_rtm(const orc_plan *p, int ncomp, const double *rtp,
                    double *rtm);

/* Legendre transforms on comp-interleaved buffers (rlm: j fastest then r;
 * rtm: theta fastest, r, slot).  Vectors first as (pol,dpol,tor) triples
 * then scalars; ncomp = 3*nv + ns.
 * *_org: SS/legendre_bwd_trans_org.f90, SS/legendre_fwd_trans_org.f90.
 * *_sym: symmetric_BLAS_big, SS/leg_b_trans_sym_matmul_big.f90,
 *        SS/leg_f_trans_sym_matmul_big.f90 (OpenMP over radial chunks). */
void orc_leg_bwd_org(const orc_plan *p, int nv, int ns, const double *sp_rlm,
                     double *vr_rtm);
void orc_leg_fwd_org(const orc_plan *p, int nv, int ns, const double *vr_rtm,
                     double *sp_rlm);
void orc_leg_bwd_sym(const orc_plan *p, int nv, int ns, const double *sp_rlm,
                     double *vr_rtm);
void orc_leg_fwd_sym(const orc_plan *p, int nv, int ns, const double *vr_rtm,
                     double *sp_rlm);

/* Real FFT in the FFTPACK5.1 convention (FP5/fftpack5.1d.f90:9288-9335):
 * c[0]=a0, c[2k-1]=a_k, c[2k]=b_k, c[n-1]=a_{n/2} (n even).
 * Lines are contiguous (x[line*n + i]).  Mixed-radix Stockham FFT. */
void orc_rfft_fwd(int n, int nlines, const double *x, double *c);
void orc_rfft_bwd(int n, int nlines, const double *c, double *x);

/* Whole transforms with the reference boundary layouts:
 * WS_rj / WR_rj: comp-interleaved rj buffer, sp[nd + ncomp*(j + jmax*k)]
 * v_rtp: field-major grid, v[inod + nnod_rtp*nd],
 *        inod = k + nr*(l_theta + nth*i_phi).
 * variant 0 = org loops (needs full tables), 1 = sym big. */
void orc_sht_backward(const orc_plan *p, int nv, int ns, const double *WS_rj,
                      double *v_rtp, int variant);
void orc_sht_forward(const orc_plan *p, int nv, int ns, const double *v_rtp,
                     double *WR_rj, int variant);

/* FFT stages alone (between rtp-spectral buffer and grid). */
void orc_fft_bwd_from_rtp(const orc_plan *p, int ncomp, const double *rtp,
                          double *v_rtp);
void orc_fft_fwd_to_rtp(const orc_plan *p, int ncomp, const double *v_rtp,
                        double *rtp);

/* Backward transform at the poles (rows 0..nr-1 north, nr..2nr-1 south)
 * and, when center != 0, the centre (row 2nr) of v_pole[i + (2nr+1)*nd],
 * from the same rj buffer as orc_sht_backward.  Vector components come out
 * as (x, y, z) = (v_theta, v_phi, v_r) at phi = 0 in the north and
 * (-v_theta, v_phi, -v_r) in the south; the centre row of scalars (and of
 * everything when center == 0) is left as it is.
 * SS/pole_sph_transform.f90:107-162, SS/schmidt_b_trans_at_pole.f90,
 * SS/schmidt_b_trans_at_center.f90, SP/schmidt_poly_on_rtm_grid.f90:201-266.
 * Pinned: the pole tables by closed forms, the pole rows by the limit of the
 * full-grid backward at phi = 0 and by uniform fields; the centre row bit
 * for bit against SS/schmidt_b_trans_at_center.f90 compiled from source. */
void orc_pole_backward(const orc_plan *p, int nv, int ns, const double *WS_rj,
                       double *v_pole, int center);

/* Grid-space products of the MHD step (nonlinear_terms_in_rtp,
 * MS/cal_nonlinear_sph_MHD.f90:64-128) on field-major rtp buffers.
 * idx[12]: first component (0-based, -1 = absent/off) of the inputs velo,
 * vort, magne, current, temp, light and the outputs m_advect = c w x u,
 * lorentz = c J x B, vp_induct = c u x B, h_flux = c u T, c_flux = c u C,
 * coriolis = -c (omega x u), omega = (cos theta, -sin theta, 0).
 * coef[6] in the same order as the outputs.  Pinned against Cartesian cross
 * products (numpy) of the same vectors; see tests/test_oracle_nonlinear.py. */
void orc_nonlinear_rtp(const orc_plan *p, const int idx[12],
                       const double coef[6], int ncomp_fld, const double *fld,
                       int ncomp_frc, double *frc);

/* Copies between field-major rj spectra fld[inod + nnod_rj*c] and the
 * interleaved send / receive buffers, per field (fields[3f..3f+2] =
 * i_field, i_trans (0-based first components), ncomp 1, 3 or 6).
 * to_send: SR/field_to_send_buffer.f90:43-120 (set_to_send_buf_*),
 *   WS[i_trans + c + nb*k] = fld[inod_export[k]-1 + nnod_rj*(i_field + c)];
 * from_recv: :246-330 (set_from_recv_buf_rev_*), the dummy record nnod_recv+1
 *   reading zero:
 *   fld[k + nnod_rj*(i_field + c)] = WR[i_trans + c + nb*(irev[k]-1)].
 * Tables are 1-based; NULL = identity.  Called by
 * MS/copy_sph_MHD_4_send_recv.f90:63-206. */
void orc_rj_to_send(int nfield, const int *fields, int ntot_phys, long long nnod_rj,
                    const double *fld, const int *inod_export, long long nnod_send, int nb,
                    double *WS);
void orc_rj_from_recv(int nfield, const int *fields, int nb, long long nnod_recv,
                      const double *WR, const int *irev, int ntot_phys, long long nnod_rj,
                      double *fld);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
