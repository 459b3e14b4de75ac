/*
 * kemorin_sht.h -- C ABI of the MI355X (gfx950) spherical-harmonic transform
 * engine that replaces the hot path of Kemorin/Calypso's sph_trans.
 *
 * Plain C: opaque plan handle, plain pointers and sizes, int status codes.
 * Reference files are cited relative to the reference tree, with
 *   SS/ = MHD/Fortran_src/PARALLEL_src/SPH_SHELL_src/
 *   SR/ = MHD/Fortran_src/SOLVER_src/send_recv/
 *
 * Data layouts (single rank; identical to the reference's buffers):
 *   spectral buffer (rj send/recv, rlm):   sp[nd + ncomp*(j + jmax*k)]
 *       j: single-rank rlm order (m from -L to L, then l = |m|..L),
 *       k: radius, nd: component (vectors first as (pol, dpol/dr, tor)
 *       triples, then scalars).  jmax = (L+1)^2.
 *   rtm buffer (Legendre output/input):    v[nd + ncomp*(t + nth*(k + nr*s))]
 *       t: Gauss colatitude (ascending x = cos(theta)), s = m + L.
 *   rtp spectral buffer (FFT input/output): v[nd + ncomp*(k + nr*(t + nth*p))]
 *       p: FFTPACK slot, a0 -> 0, Nyquist -> 1, cos(m phi) -> 2m,
 *       sin(m phi) -> 2m+1.
 *   grid field (v_rtp):                    v[(k + nr*(t + nth*i)) + nnod*nd]
 *       i: longitude phi_i = 2 pi i / nphi, nnod = nr*nth*nphi.
 * All pointers passed to the transform entry points are DEVICE pointers
 * (hipMalloc'd or torch CUDA tensors) unless the name ends in _host.
 * Calls on one plan are stream-ordered on the plan's stream and not
 * re-entrant; different plans may run concurrently.  A plan's own stream is
 * a blocking stream; a caller whose inputs are written by work queued on the
 * legacy default stream (asynchronous copies included) binds the plan to
 * hipStreamLegacy with sht_plan_set_stream, so that both share one queue.
 */
#ifndef KEMORIN_SHT_H
#define KEMORIN_SHT_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SHT_OK = 0,
  SHT_ERR_ARG = 1,         /* bad descriptor, size or null pointer   */
  SHT_ERR_UNSUPPORTED = 2, /* e.g. FFT length beyond the LDS capacity   */
  SHT_ERR_OOM = 3,         /* device allocation failed               */
  SHT_ERR_HIP = 4,         /* HIP runtime error                      */
  SHT_ERR_STATE = 5,       /* plan not ready / wrong precision mode  */
  SHT_ERR_NONFINITE = 6    /* non-finite output (SHT_FLAG_CHECK_FINITE) */
};

enum { SHT_PREC_FP64 = 0, SHT_PREC_SPLIT_INT8 = 1 };

typedef struct sht_plan sht_plan;

typedef struct {
  int l_truncation; /* L: truncation degree (triangular)                */
  int nidx_r;       /* Nr: number of radial points                       */
  int nidx_theta;   /* Ntheta: Gauss-Legendre colatitudes                */
  int nidx_phi;     /* Nphi: longitudes (even, > 2L+1)                   */
  const double *radius; /* host array of Nr radii                        */
  int max_ncomp;    /* largest ncomp a call will pass (workspace size)    */
  int precision;    /* SHT_PREC_FP64 or SHT_PREC_SPLIT_INT8              */
  int n_slices;     /* split mode: INT8 slices per operand (0 = default) */
  int flags;        /* SHT_FLAG_* bits, 0 for none                      */
} sht_desc;

/* Descriptor flags.
 * SHT_FLAG_P_ON_THE_FLY: do not store P_l^m and dP_l^m/dtheta (nor, in split
 * mode, their INT8 slices) for all orders; every transform generates them
 * in FP64 for one tile of orders m at a time into a tile buffer (about an
 * eighth of the tables) just before that tile's Legendre products.  For
 * large truncations (L = 1023: 12.9 GB of FP64 tables and 12.1 GB of
 * slices), at the cost of the recurrence on every call.  Results are bit for
 * bit those of stored tables. */
#define SHT_FLAG_P_ON_THE_FLY 1
/* SHT_FLAG_CHECK_FINITE: after each sht_backward / sht_forward (and their
 * _host and _noalloc forms) count the NaN and infinite values of the output
 * on the device and return SHT_ERR_NONFINITE, the count in the message, if
 * there are any (the output is still written).  Each call then waits for
 * the plan's stream.  A call being captured into a graph skips the check.
 * For debugging runs; the reference has no such check (its routines return
 * nothing). */
#define SHT_FLAG_CHECK_FINITE 2

/* Replaces initialize_sph_trans (SS/init_sph_trans.f90:46-74) and the
 * Legendre/FFT set-up it runs: Gauss points with the reference's Newton
 * loop (MHD/Fortran_src/CORE_src/gauss_integration/gauss_integration.f90:60),
 * Schmidt P_l^m and dP/dtheta generated on the GPU in FP64, 1/sin(theta),
 * index stacks and FFT twiddles.  The plan owns all device memory. */
int sht_plan_create(const sht_desc *desc, sht_plan **plan);
int sht_plan_destroy(sht_plan *plan);
/* Run subsequent work on this hipStream_t (NULL = plan's own stream;
 * hipStreamLegacy = the legacy default stream, for callers whose other work
 * is queued there). */
int sht_plan_set_stream(sht_plan *plan, void *hip_stream);
/* Copy a plan table to host: "gpoint","weight","colat","asin_t","radius"
 * (double) or "lstack" (int, 2L+2 entries).  n = capacity in elements. */
int sht_plan_get_table(const sht_plan *plan, const char *name, void *out, long long n);
/* Device bytes held by the plan (tables + workspace). */
long long sht_plan_device_bytes(const sht_plan *plan);
/* Kernel variants (tile shapes) of a plan, for comparisons and tuning.  The
 * reference's closest counterpart is its choice of Legendre loop by timing
 * at start-up (MS/init_sphrical_transform_MHD.f90:201-345,
 * MS/ = MHD/Fortran_src/MHD_src/sph_MHD/).  Names and values:
 *   "split_fwd_mfma": 16 or 32, the INT8 MFMA form of the forward split
 *                     Legendre products (v_mfma_i32_16x16x64_i8 or
 *                     v_mfma_i32_32x32x32_i8);
 *   "split_bwd_fused": 1 (default) or 0: in the backward split products,
 *                     each vector's launch also takes one scalar field,
 *                     whose product shares the vector's P slices (1, for
 *                     5 or 6 slices), or the scalars run in launches of
 *                     their own (0); the results are the same bit for bit;
 *   "fft_form":       1, 0, 2 or 3: the split-form FFT (Nphi = r0 2^k,
 *                     2^k > L; default where the length has it), the
 *                     mixed-radix Stockham FFT, the pruned real DFT as one
 *                     FP64 GEMM per direction against a cos/sin basis, or
 *                     the whole line as one chirp-z (Bluestein) convolution
 *                     of power-of-two length MP >= Nphi + 2L (MP <= 2048; any
 *                     prime factors).  Forms 2 and 3 set up their tables (and
 *                     count them in the plan's device bytes) when first
 *                     selected; SHT_ERR_UNSUPPORTED where they do not apply.
 * Results do not depend on the variant beyond round-off.  SHT_ERR_ARG for an
 * unknown name or value. */
int sht_plan_set_variant(sht_plan *plan, const char *name, int value);
int sht_plan_get_variant(const sht_plan *plan, const char *name, int *value);
/* Tile-shape autotuner: times every applicable combination of the plan's
 * variants ("split_fwd_mfma" for split-precision plans, "fft_form" among the
 * forms that apply to the length) on synthetic device data of the given field
 * sets (one backward of nvector_b/nscalar_b and one forward of
 * nvector_f/nscalar_f per round trip, best of three) and keeps the fastest.
 * Replaces select_legendre_transform (MS/init_sphrical_transform_MHD.f90:
 * 201-345), which times the reference's Legendre loop variants at start-up.
 * cache_path (NULL = none): a JSON file of earlier results, one entry per
 * shape (L, Nr, Ntheta, Nphi, precision, slices, flags, field sets); a hit
 * sets the cached winner without timing, a miss appends the new winner.
 * ms_out (NULL = none): [round-trip ms of the winner, of the default].
 * Allocates temporary device buffers of the field sets' size; runs on the
 * plan's stream and returns after it drained. */
int sht_plan_autotune(sht_plan *plan, int nvector_b, int nscalar_b, int nvector_f,
                      int nscalar_f, const char *cache_path, double *ms_out);

/* Stage timers, the counterpart of the reference's elapsed-time ids for the
 * transform (MHD/Fortran_src/MHD_src/IO/init_sph_MHD_elapsed_label.f90:59-65:
 * 22 Legendre backward, 23 Legendre forward, 24 FFT, 18-21 exchanges).
 * When enabled, every kernel launch of the plan is bracketed by HIP events
 * on the plan's stream.  Stage ids: 0 Legendre backward, 1 FFT backward,
 * 2 FFT forward, 3 Legendre forward, 4 exchanges; 5 and 6 time the
 * vector-group Legendre kernel alone (backward, forward; inside 0 and 3,
 * without the operand slicing of the split engine).
 * sht_plan_timer_read synchronises, returns per-stage total milliseconds and
 * launch counts (n entries each, n <= 7) and resets the counters. */
int sht_plan_timer_enable(sht_plan *plan, int on);
int sht_plan_timer_read(sht_plan *plan, double *ms, long long *calls, int n);

/* Whole transforms on device-resident buffers.
 * Replaces sph_backward_transforms (SS/t_sph_transforms.f90:74-123):
 * d_WS_rj is the rj->rlm send buffer (spectral layout above, ncomp =
 * 3*nvector + nscalar), d_v_rtp receives the grid fields. */
int sht_backward(sht_plan *plan, int nvector, int nscalar,
                 const double *d_WS_rj, double *d_v_rtp);
/* Replaces sph_forward_transforms (SS/t_sph_transforms.f90:250-297):
 * d_v_rtp grid fields in, d_WR_rj (rlm->rj receive buffer) out. */
int sht_forward(sht_plan *plan, int nvector, int nscalar,
                const double *d_v_rtp, double *d_WR_rj);
/* Same, with host buffers (one H2D and one D2H copy; PCIe-bound). */
int sht_backward_host(sht_plan *plan, int nvector, int nscalar,
                      const double *WS_rj, double *v_rtp);
int sht_forward_host(sht_plan *plan, int nvector, int nscalar,
                     const double *v_rtp, double *WR_rj);

/* No counterpart in the reference.  sht_backward for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The call allocates nothing; its workspace
 * (the Fourier workspace between the Legendre and FFT kernels and, for
 * split-precision plans, the sliced data operand) is a fixed device array of
 * the library, SHT_NOALLOC_SCRATCH_BYTES on each device, in place of the
 * plan's.  Returns SHT_ERR_ARG when the call needs more than that (8 bytes x
 * ncomp x (2L+1) x Nr x Ntheta, plus the data slices in split mode: e.g. up
 * to L = 63, Nr = 65, Ntheta = 96 with 9 components in FP64, L = 31, Nr = 33,
 * Ntheta = 48 in split mode) and SHT_ERR_STATE when the plan's stage timers
 * are on.  Every call on a device shares that device's
 * array and leaves its contents behind: calls that overlap on one device
 * (other streams, plans or threads) race. */
#define SHT_NOALLOC_SCRATCH_BYTES (128LL << 20)
int sht_backward_noalloc(sht_plan *plan, int nvector, int nscalar,
                         const double *d_WS_rj, double *d_v_rtp);

/* Backward transform at the poles and the centre, device buffers.
 * Replaces pole_b_transform (SS/t_sph_transforms.f90:203-245), i.e.
 * pole_backward_transforms (SS/pole_sph_transform.f90:107-162) on one rank.
 * d_WS_rj as for sht_backward; d_v_pole[i + (2*Nr + 1)*nd] receives rows
 * i = 0..Nr-1 (north pole), Nr..2Nr-1 (south pole) and, for vector
 * components when with_center != 0 (iflag_MESH_w_center), 2Nr (the centre,
 * from the innermost radius).  Vectors come out Cartesian: (x, y, z) =
 * (v_theta, v_phi, v_r) at phi = 0 in the north, (-v_theta, v_phi, -v_r)
 * in the south.  Centre entries not written keep their values, as in the
 * reference. */
int sht_pole_backward(sht_plan *plan, int nvector, int nscalar,
                      const double *d_WS_rj, double *d_v_pole, int with_center);
/* sht_pole_backward, then sht_backward on the same spectra.
 * Replaces sph_b_trans_w_poles (SS/t_sph_transforms.f90:127-199). */
int sht_backward_with_poles(sht_plan *plan, int nvector, int nscalar,
                            const double *d_WS_rj, double *d_v_rtp,
                            double *d_v_pole, int with_center);

/* Grid-space products of one MHD step, on the grid layout of sht_backward's
 * output and sht_forward's input.  Replaces nonlinear_terms_in_rtp
 * (MS/cal_nonlinear_sph_MHD.f90:64-128): each output is a vector written at
 * its first component (0-based) of d_frc_rtp, -1 = term off:
 *   m_advect  = coef_velo     (vort x velo)
 *   lorentz   = coef_lor      (current x magne)
 *   vp_induct = coef_induct   (velo x magne)
 *   h_flux    = coef_h_advect  velo * temp
 *   c_flux    = coef_c_advect  velo * light
 *   coriolis  = -coef_cor (omega x velo), omega = (cos theta, -sin theta, 0)
 * Inputs are first components in d_fld_rtp (vectors in r, theta, phi); only
 * those an enabled term reads must be set (-1 otherwise).  Components of
 * d_frc_rtp no term writes are left alone; d_fld_rtp and d_frc_rtp must not
 * overlap. */
typedef struct {
  int i_velo, i_vort, i_magne, i_current, i_temp, i_light;
  int i_m_advect, i_lorentz, i_vp_induct, i_h_flux, i_c_flux, i_coriolis;
  double coef_velo, coef_lor, coef_induct, coef_h_advect, coef_c_advect, coef_cor;
} sht_nonlinear_desc;
int sht_nonlinear_rtp(sht_plan *plan, const sht_nonlinear_desc *desc,
                      int ncomp_fld, const double *d_fld_rtp,
                      int ncomp_frc, double *d_frc_rtp);

/* Copies between the solver's field-major rj spectra d_fld[inod + nnod_rj*c]
 * (c < ntot_phys) and the interleaved send / receive buffers of the
 * transform.  A field of ncomp components (1 scalar, 3 vector, 6 tensor)
 * sits at component i_field (0-based) of d_fld and i_trans of the buffer.
 * All tables are device int arrays of 1-based node numbers, as the
 * reference's comm_rj tables; NULL = the single-rank identity this engine
 * uses (the buffer's node order is the rj order).  Up to 64 components per
 * call; bit-exact copies, on the plan's stream. */
typedef struct {
  int i_field, i_trans, ncomp;
} sht_field_copy;
/* Replaces copy_mhd_spectr_to_send (MS/copy_sph_MHD_4_send_recv.f90:63-136)
 * and its set_to_send_buf_* (SR/field_to_send_buffer.f90:43-120):
 * WS[i_trans + c + ncomp_send*k] = d_fld[inod_export[k]-1 + nnod_rj*(i_field + c)]
 * for k < nnod_send; components of WS no field names are left alone. */
int sht_rj_fields_to_send(sht_plan *plan, int nfield, const sht_field_copy *fields,
                          int ntot_phys, long long nnod_rj, const double *d_fld,
                          const int *d_inod_export, long long nnod_send, int ncomp_send,
                          double *d_WS);
/* Replaces copy_mhd_spectr_from_recv (MS/copy_sph_MHD_4_send_recv.f90:140-206)
 * with the reverse-table copies set_from_recv_buf_rev_*
 * (SR/field_to_send_buffer.f90:246-330):
 * d_fld[k + nnod_rj*(i_field + c)] = WR[i_trans + c + ncomp_recv*(irev[k]-1)]
 * for k < nnod_rj, zero where irev[k] = nnod_recv + 1 (the dummy record; the
 * reference zeroes it inside WR, here WR is not written). */
int sht_rj_fields_from_recv(sht_plan *plan, int nfield, const sht_field_copy *fields,
                            int ncomp_recv, long long nnod_recv, const double *d_WR,
                            const int *d_irev, int ntot_phys, long long nnod_rj,
                            double *d_fld);

/* Coriolis term in rlm space, for rotating runs whose Coriolis force is
 * taken by Gaunt integrals between the stages of the MHD step
 * (MS/sph_trans_w_coriols.f90:146-149, 233-235;
 * MS/ = MHD/Fortran_src/MHD_src/sph_MHD/).  All buffers are device buffers
 * on the plan's single-rank rlm layout; calls run on the plan's stream.
 * Replaces init_sum_coriolis_rlm (MS/t_coriolis_terms_rlm.f90:133-183):
 * the Gaunt and Elsasser integrals of cal_gaunt_coriolis_rlm
 * (MS/interact_coriolis_rlm.f90:42-118), the coefficients of
 * interact_rot_coriolis_rlm (:122-205) and the rotation spectrum of
 * set_zdir_rot_earth_4_sph (MS/t_poloidal_rotation.f90:155-250: r^2/2, r,
 * 1).  kr_icb: 1-based radius of a rotatable inner core boundary
 * (sph_bc_U%iflag_icb = iflag_rotatable_ic), 0 = none. */
typedef struct sht_coriolis sht_coriolis;
int sht_coriolis_create(sht_plan *plan, int kr_icb, sht_coriolis **cor);
int sht_coriolis_destroy(sht_coriolis *cor);
/* Replaces sum_coriolis_rlm (MS/t_coriolis_terms_rlm.f90:187-258), i.e.
 * sum_rot_coriolis_rlm_10 (MS/sum_coriolis_terms_rlm.f90:96-186) and, with
 * kr_icb, inner_core_rot_z_coriolis_rlm (:378-418): the curl of
 * -coef_cor (z x u), poloidal part to d_cor_rlm[i] and toroidal part to
 * d_cor_rlm[i + nnod_rlm], i = j + jmax*k.  d_WR: receive buffer of
 * ncomp_trans interleaved components with velocity and vorticity
 * (pol, dpol/dr, tor) at 0-based components i_velo and i_vort; record of
 * node i = d_irev_sr_rlm[i] - 1 (1-based device int table; NULL = the
 * single-rank identity); records at or past nnod_recv (the dummy record)
 * read zero. */
int sht_sum_coriolis_rlm(sht_coriolis *cor, int ncomp_trans, int i_velo, int i_vort,
                         double coef_cor, const double *d_WR, long long nnod_recv,
                         const int *d_irev_sr_rlm, double *d_cor_rlm);
/* Replaces copy_coriolis_terms_rlm (MS/t_coriolis_terms_rlm.f90:262-296):
 * the poloidal and toroidal parts of d_cor_rlm go to components
 * i_rot_coriolis and i_rot_coriolis + 2 (0-based) of the forward transform's
 * send buffer, WS[c + ncomp_trans*k] = d_cor_rlm[item[k]-1 (+ nnod_rlm)]
 * for k < nnod_send (item: 1-based device int table of the rlm send
 * order, NULL = identity with nnod_send = nnod_rlm).  Other components of
 * WS are left alone.  Bit-exact copies. */
int sht_copy_coriolis_terms_rlm(sht_coriolis *cor, int ncomp_trans, int i_rot_coriolis,
                                const double *d_cor_rlm, const int *d_item_sr_rlm,
                                long long nnod_send, double *d_WS);

/* Stage entry points on the reference's buffers, for callers that run work
 * between stages (the Coriolis hooks of MS/sph_trans_w_coriols.f90:146,233).
 * Legendre: replaces sel_backward_legendre_trans / sel_forward_legendre_trans
 * (SS/t_legendre_trans_select.f90:184,273) in symmetric_BLAS_big mode. */
int sht_legendre_backward(sht_plan *plan, int nvector, int nscalar,
                          const double *d_WR_rlm, double *d_WS_rtm);
int sht_legendre_forward(sht_plan *plan, int nvector, int nscalar,
                         const double *d_WR_rtm, double *d_WS_rlm);
/* FFT: replaces back_FFT_select_from_recv / fwd_FFT_select_to_send
 * (SS/t_sph_FFT_selector.F90:258,214) for FFTPACK (SS/t_sph_FFTPACK5.f90:
 * 183-331).  The forward writes slots of orders m <= L and zeroes the rest. */
int sht_fft_backward(sht_plan *plan, int ncomp, const double *d_WR_rtp,
                     double *d_v_rtp);
int sht_fft_forward(sht_plan *plan, int ncomp, const double *d_v_rtp,
                    double *d_WS_rtp);
/* Exchanges: single-rank calypso_sph_comm_N (SS/spherical_SRs_N.f90:233)
 * followed by the receive-side gather; bit-exact permutations.
 * rtm -> rtp fills slots with no Legendre order with zero (the dummy slot
 * of SR/set_from_recv_buf_rev.f90:77-87). */
int sht_exchange_rtm_to_rtp(sht_plan *plan, int ncomp, const double *d_rtm,
                            double *d_rtp);
int sht_exchange_rtp_to_rtm(sht_plan *plan, int ncomp, const double *d_rtp,
                            double *d_rtm);
/* Generic receive gather (SR/set_from_recv_buf_rev.f90:77-87):
 * dst[nd + ncomp*i] = irev[i] == n_src+1 ? 0 : src[nd + ncomp*(irev[i]-1)],
 * i < n_dst, irev 1-based (device int array). */
int sht_exchange_gather_rev(sht_plan *plan, int ncomp, long long n_dst,
                            long long n_src, const int *d_irev,
                            const double *d_src, double *d_dst);

/* Shell batch: many independent shells (plans of differing L, Nr and grid,
 * each with its own device buffers) captured once into one HIP graph, one
 * backward and one forward transform per shell, and replayed with a single
 * launch.  No counterpart in the reference, which transforms one shell per
 * call (SS/t_sph_transforms.f90:74,250); it is for callers that carry many
 * uneven shells and would otherwise pay one launch sequence per shell with
 * most of the device idle.  Shells are dealt to parallel graph branches;
 * calls that share a plan run in order on one branch.
 * A field set of (0, 0) skips that direction.  The plans' stage timers must
 * be off; the graph keeps the plans' workspaces and the buffer pointers, so
 * neither may be freed, and the plans may not run other work concurrently
 * with a launch.  Refill the inputs in place between launches. */
typedef struct sht_batch sht_batch;
typedef struct {
  sht_plan *plan;
  int nvector_b, nscalar_b;      /* backward field set                  */
  const double *d_WS_rj;         /* backward input (spectral layout)    */
  double *d_v_rtp;               /* backward output (grid fields)       */
  int nvector_f, nscalar_f;      /* forward field set                   */
  const double *d_v_in;          /* forward input (grid fields)         */
  double *d_WR_rj;               /* forward output (spectral layout)    */
} sht_shell_call;
int sht_batch_create(const sht_shell_call *calls, int n, sht_batch **batch);
/* Launch on hip_stream (NULL = the batch's own blocking stream). */
int sht_batch_launch(sht_batch *batch, void *hip_stream);
/* Wait for the launches on the batch's own stream. */
int sht_batch_synchronize(sht_batch *batch);
int sht_batch_destroy(sht_batch *batch);

/* ---- Plans from the reference's grid / mode files ---------------------
 * The reference reads a rank's spherical grids and modes from the four
 * ASCII files gen_sph_grids writes (rtp, rtm, rlm, rj): read_geom_rtp_data,
 * read_geom_rtm_data, read_spectr_modes_rlm_data, read_spectr_modes_rj_data
 * (MHD/Fortran_src/CORE_src/IO/sph_modes_grids_data_IO.f90:53-183), copied
 * into the grid types by copy_sph_node_4_*_from_IO
 * (MHD/Fortran_src/CORE_src/SPH_SPECTR_src/copy_sph_node_4_IO.f90:119-281).
 * sht_sph_files_read parses such a set (domain, sizes, truncation,
 * subdomain indices, 1-D index tables with the radii, node lists, import
 * tables; groups are not read); SHT_ERR_ARG with file:line on a malformed
 * file. */
typedef struct sht_sph_files sht_sph_files;
int sht_sph_files_read(const char *rtp_path, const char *rtm_path, const char *rlm_path,
                       const char *rj_path, sht_sph_files **files);
void sht_sph_files_destroy(sht_sph_files *files);
/* SHT_OK when the set is one subdomain per direction of the single-rank
 * layout this engine computes in (m_folding = 1): every index table and
 * node list as gen_sph_grids builds them for one rank, rlm and rtm import
 * tables in node order, and rj / rtp import tables that pair each rlm / rtm
 * node with the rj / rtp node of the same radius and mode / colatitude and
 * FFT slot.  Otherwise SHT_ERR_UNSUPPORTED, the message naming the file and
 * the first entry that differs. */
int sht_sph_files_check(const sht_sph_files *files);
/* dims = {L, Nr, Ntheta, Nphi, nnod_rj}; radius (NULL = skip) gets Nr radii. */
int sht_sph_files_shape(const sht_sph_files *files, int dims[5], double *radius);
/* A table of the set by name: "<space>_import" (1-based node numbers in
 * buffer order; the reference's item_sr), "<space>_import_rev" (node ->
 * 1-based buffer record, ntot + 1 for a node not in the table; irev_sr),
 * "<space>_nodes" (global indices per node), "<space>_table2" /
 * "<space>_table3" (the 1-D tables); space = rtp, rtm, rlm or rj.  Copies
 * into out (NULL: only the length) and returns the number of ints, or minus
 * an error code.  The rj tables drive sht_rj_fields_to_send /
 * sht_rj_fields_from_recv. */
long long sht_sph_files_table(const sht_sph_files *files, const char *name, int *out,
                              long long n);
/* sht_plan_create for the shape and radii of a set that passes
 * sht_sph_files_check (its error otherwise). */
int sht_plan_create_from_files(const sht_sph_files *files, int max_ncomp, int precision,
                               int n_slices, unsigned flags, sht_plan **plan);

/* Diagnostics: checks the MFMA lane layouts this library assumes on the
 * current device (FP64 16x16x4 and INT8 16x16x64).  0 = as assumed. */
int sht_selftest(void);
/* Last error message of the calling thread. */
int sht_get_last_error(char *buf, int n);
const char *sht_version(void);

#ifdef __cplusplus
}
#endif
#endif
