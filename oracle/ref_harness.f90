!> ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!! A driver of this project's own that calls the reference's numerical
!! routines, compiled from their sources in the reference tree (see
!! oracle/Makefile.ref), so that tests can compare the C oracle with the
!! reference's own arithmetic.  Nothing here is shipped or used by the
!! product library.
!!
!! The driver loops restate the single-thread (np_smp=1, kst=0, nkr=nr)
!! control flow of SS/leg_b_trans_sym_matmul_big.f90:80-178 and
!! SS/leg_f_trans_sym_matmul_big.f90:80-178 (the matmul flavour of
!! symmetric_BLAS_big); that file itself needs MPI (calypso_mpi), which the
!! image lacks.  Strides follow the generator's node ordering:
!! istep_rtm = (nth, 1, nth*nr), istep_rlm = (jmax, 1); exchange tables are
!! identities (single rank).
      module kemo_ref_harness
!
      use iso_c_binding
      use m_precision
      use m_constants
      use gauss_integration
      use schmidt_fix_m
      use spherical_harmonics
      use set_legendre_matrices
      use matmul_for_legendre_trans
      use set_sp_rlm_for_leg_matmul
      use cal_vr_rtm_by_matmul
      use set_vr_rtm_for_leg_matmul
      use cal_sp_rlm_by_matmul
      use schmidt_b_trans_at_center
      use field_to_send_buffer
      use t_phys_address
      use t_gaunt_coriolis_rlm
      use interact_coriolis_rlm
      use sum_coriolis_terms_rlm
!
      implicit none
!
      integer(kind = kint), save :: ltr_h = 0, nr_h = 0, nth_h = 0
      integer(kind = kint), save :: nslot_h = 0, jmax_h = 0, nhemi_h = 0
      integer(kind = kint), save :: nidx_rtm_h(3), nidx_rlm_h(2)
      integer(kind = kint), save :: istep_rtm_h(3), istep_rlm_h(2)
      integer(kind = kint), allocatable, save :: idx_j_h(:,:)
      integer(kind = kint), allocatable, save :: lstack_h(:)
      integer(kind = kint), allocatable, save :: lstack_e_h(:)
      real(kind = kreal), allocatable, save :: radius_h(:), a_r_h(:)
      real(kind = kreal), allocatable, save :: gpt_h(:), wgt_h(:)
      real(kind = kreal), allocatable, save :: colat_h(:), asin_h(:)
      real(kind = kreal), allocatable, save :: g_h(:,:)
      real(kind = kreal), allocatable, save :: P_h(:,:), dP_h(:,:)
      real(kind = kreal), allocatable, save :: Ps_h(:,:), dPs_h(:,:)
!
      contains
!
! ----------------------------------------------------------------------
!
      subroutine ref_gauss_points(n, x, w) bind(C, name='ref_gauss_points')
      integer(c_int), value :: n
      real(c_double), intent(inout) :: x(n), w(n)
      integer(kind = kint) :: nk
      nk = n
      call const_gauss_points_coefs(nk, x, w)
      end subroutine ref_gauss_points
!
      subroutine ref_schmidt_legendres_m(ltr, m, theta, p, dp)          &
     &          bind(C, name='ref_schmidt_legendres_m')
      integer(c_int), value :: ltr, m
      real(c_double), value :: theta
      real(c_double), intent(inout) :: p(0:ltr), dp(0:ltr)
      real(kind = kreal), allocatable :: pmn1(:), pmp1(:), df(:)
      integer(kind = kint) :: lk, mk
      allocate(pmn1(0:ltr), pmp1(0:ltr), df(0:ltr+2))
      pmn1 = 0.0d0
      pmp1 = 0.0d0
      df = 0.0d0
      lk = ltr
      mk = m
      call schmidt_legendres_m(lk, mk, theta, p, dp, pmn1, pmp1, df)
      deallocate(pmn1, pmp1, df)
      end subroutine ref_schmidt_legendres_m
!
      subroutine ref_sph_normalizations(l, m, g)                        &
     &          bind(C, name='ref_sph_normalizations')
      integer(c_int), value :: l, m
      real(c_double), intent(inout) :: g(17)
      integer(kind = kint) :: idx(2), lk, mk
      lk = l
      mk = m
      call sph_normalizations(lk, mk, idx, g)
      end subroutine ref_sph_normalizations
!
!> FFTPACK5.1 multiple real FFT.  x(lot, n): the line index is fastest, as
!! in SS/t_sph_FFTPACK5.f90:206-214 (jump = 1, inc = lot).
      subroutine ref_rfftm(idir, n, lot, x) bind(C, name='ref_rfftm')
      integer(c_int), value :: idir, n, lot
      real(c_double), intent(inout) :: x(lot*n)
      real(kind = 8), allocatable :: wsave(:), work(:)
      integer(kind = 4) :: lensav, ier, n4, lot4
      n4 = n
      lot4 = lot
      lensav = n4 + int(log(real(n4, kind = 8))/log(2.0d0)) + 4
      allocate(wsave(lensav), work(lot4*n4))
      call rfftmi(n4, wsave, lensav, ier)
      if(idir .gt. 0) then
        call rfftmf(lot4, 1, n4, lot4, x, lot4*n4, wsave, lensav,      &
     &      work, lot4*n4, ier)
      else
        call rfftmb(lot4, 1, n4, lot4, x, lot4*n4, wsave, lensav,      &
     &      work, lot4*n4, ier)
      end if
      deallocate(wsave, work)
      end subroutine ref_rfftm
!
! ----------------------------------------------------------------------
!
!> Set up single-rank tables with the reference routines: Gauss points
!! (GI/gauss_integration.f90:60), g_sph_rlm with g3(0,0)=1/2
!! (SP/schmidt_poly_on_rtm_grid.f90:89-113), P and dP/dtheta per rlm j
!! (SP/schmidt_poly_on_rtm_grid.f90:145-197) and the hemisphere tables
!! (SS/set_legendre_matrices.f90:86-130).
      subroutine ref_setup(ltr, nr, nth, radius)                        &
     &          bind(C, name='ref_setup')
      integer(c_int), value :: ltr, nr, nth
      real(c_double), intent(in) :: radius(nr)
      integer(kind = kint) :: s, m, l, j, i, idx(2), n, lk
      real(kind = kreal), allocatable :: p(:), dp(:), pmn1(:), pmp1(:)
      real(kind = kreal), allocatable :: df(:)
      real(kind = kreal) :: g(17)
!
      call ref_release()
      ltr_h = ltr
      nr_h = nr
      nth_h = nth
      nslot_h = 2*ltr + 1
      jmax_h = (ltr+1)**2
      nhemi_h = (nth + 1) / 2
      nidx_rtm_h = (/nr_h, nth_h, nslot_h/)
      nidx_rlm_h = (/nr_h, jmax_h/)
      istep_rtm_h = (/nth_h, 1, nth_h*nr_h/)
      istep_rlm_h = (/jmax_h, 1/)
!
      allocate(radius_h(nr), a_r_h(nr))
      radius_h(1:nr) = radius(1:nr)
      a_r_h(1:nr) = one / radius(1:nr)
      allocate(gpt_h(nth), wgt_h(nth), colat_h(nth), asin_h(nth))
      call const_gauss_points_coefs(nth_h, gpt_h, wgt_h)
      do i = 1, nth_h
        colat_h(i) = acos(gpt_h(i))
        asin_h(i) = one / sin(colat_h(i))
      end do
!
      allocate(idx_j_h(jmax_h,3), lstack_h(0:nslot_h))
      allocate(lstack_e_h(0:nslot_h), g_h(jmax_h,17))
      lstack_h(0) = 0
      lstack_e_h(0) = 0
      j = 0
      do s = 1, nslot_h
        m = s - ltr_h - 1
        n = ltr_h - abs(m) + 1
        lstack_h(s) = lstack_h(s-1) + n
        lstack_e_h(s) = lstack_h(s-1) + (n+1)/2
        do l = abs(m), ltr_h
          j = j + 1
          idx_j_h(j,1) = l*(l+1) + m
          idx_j_h(j,2) = l
          idx_j_h(j,3) = m
          call sph_normalizations(l, m, idx, g)
          g_h(j,1:17) = g(1:17)
          if(l.eq.0 .and. m.eq.0) g_h(j,3) = half
        end do
      end do
!
      allocate(P_h(nth_h,jmax_h), dP_h(nth_h,jmax_h))
      allocate(Ps_h(nhemi_h,jmax_h), dPs_h(nhemi_h,jmax_h))
      allocate(p(0:ltr_h), dp(0:ltr_h), pmn1(0:ltr_h), pmp1(0:ltr_h))
      allocate(df(0:ltr_h+2))
      lk = ltr_h
      do i = 1, nth_h
        do s = 1, nslot_h
          m = abs(s - ltr_h - 1)
          call schmidt_legendres_m(lk, m, colat_h(i), p, dp,            &
     &        pmn1, pmp1, df)
          do j = lstack_h(s-1)+1, lstack_h(s)
            l = idx_j_h(j,2)
            P_h(i,j) = p(l)
            dP_h(i,j) = dp(l)
          end do
        end do
      end do
      call set_symmetric_legendre_lj(nth_h, nslot_h, jmax_h, nhemi_h,   &
     &    lstack_h, lstack_e_h, P_h, dP_h, Ps_h, dPs_h)
      deallocate(p, dp, pmn1, pmp1, df)
      end subroutine ref_setup
!
!> Centre row of the backward transform at the poles and the centre
!! (SS/schmidt_b_trans_at_center.f90:45-115) on the tables of ref_setup:
!! ist_rtm_order_zero / _1s / _1c are the first rlm j of orders 0, -1, +1
!! (CORE_src/SPH_SPECTR_src/set_special_sph_lm_flags.f90:79-101),
!! irev_sr_rlm is the identity, the innermost radius has global index 1.
!! WR: rlm receive buffer ncomp*(jmax*nr); v_pole(2nr+1, ncomp): only the
!! centre row 2nr+1 is touched (vector components set, scalars kept).
      subroutine ref_center(nvector, nscalar, WR, v_pole)               &
     &          bind(C, name='ref_center')
      integer(c_int), value :: nvector, nscalar
      real(c_double), intent(in) :: WR(*)
      real(c_double), intent(inout) :: v_pole(*)
      integer(kind = kint) :: ncomp, nnod_rlm, nnod_pole, n_WR, j
      integer(kind = kint) :: ist0, ist1s, ist1c, nvk, nsk
      integer(kind = kint), allocatable :: irev(:)
!
      nvk = nvector
      nsk = nscalar
      ncomp = 3*nvk + nsk
      nnod_rlm = jmax_h * nr_h
      nnod_pole = 2*nr_h + 1
      n_WR = ncomp * nnod_rlm
      ist0 = 0
      ist1s = 0
      ist1c = 0
      do j = jmax_h, 1, -1
        if(idx_j_h(j,3) .eq.  0) ist0 = j
        if(idx_j_h(j,3) .eq. -1) ist1s = j
        if(idx_j_h(j,3) .eq.  1) ist1c = j
      end do
      allocate(irev(nnod_rlm))
      do j = 1, nnod_rlm
        irev(j) = j
      end do
      call schmidt_b_trans_center_vect(ncomp, nvk, ist0, ist1s, ist1c,  &
     &    nnod_rlm, nnod_pole, 1, a_r_h(1), irev, n_WR, WR, v_pole)
      call schmidt_b_trans_center_scalar(ncomp, nvk, nsk, nnod_pole,    &
     &    v_pole)
      deallocate(irev)
      end subroutine ref_center
!
!> One field to the interleaved send buffer with the reference's
!! set_to_send_buf_scalar / _vector / _tensor (SR/field_to_send_buffer.f90);
!! i_fld_X, i_fld_WS and inod_export 1-based as in the reference.
      subroutine ref_to_send_buf(ncomp, NB, nnod_org, nnod_send,         &
     &          inod_export, ncomp_X, i_fld_X, i_fld_WS, d_org, WS)      &
     &          bind(C, name='ref_to_send_buf')
      integer(c_int), value :: ncomp, NB, nnod_org, nnod_send
      integer(c_int), value :: ncomp_X, i_fld_X, i_fld_WS
      integer(c_int), intent(in) :: inod_export(nnod_send)
      real(c_double), intent(in) :: d_org(nnod_org, ncomp_X)
      real(c_double), intent(inout) :: WS(NB*nnod_send)
      integer(kind = kint) :: nbk, nok, nsk, nxk, ifx, ifw
      integer(kind = kint), allocatable :: inod(:)
      nbk = NB
      nok = nnod_org
      nsk = nnod_send
      nxk = ncomp_X
      ifx = i_fld_X
      ifw = i_fld_WS
      allocate(inod(nsk))
      inod(1:nsk) = inod_export(1:nsk)
      select case(ncomp)
      case(1)
        call set_to_send_buf_scalar(nbk, nok, nsk, inod, nxk, ifx,      &
     &      ifw, d_org, WS)
      case(3)
        call set_to_send_buf_vector(nbk, nok, nsk, inod, nxk, ifx,      &
     &      ifw, d_org, WS)
      case(6)
        call set_to_send_buf_tensor(nbk, nok, nsk, inod, nxk, ifx,      &
     &      ifw, d_org, WS)
      end select
      deallocate(inod)
      end subroutine ref_to_send_buf
!
!> One field from the receive buffer by the reverse table with the
!! reference's set_from_recv_buf_rev_scalar / _vector / _tensor; WR holds
!! NB*(nnod_recv+1) values (the routine zeroes the dummy record).
      subroutine ref_from_recv_rev(ncomp, NB, nnod_new, nnod_recv,      &
     &          irev_import, ncomp_X, i_fld_X, i_fld_WR, WR, d_new)     &
     &          bind(C, name='ref_from_recv_rev')
      integer(c_int), value :: ncomp, NB, nnod_new, nnod_recv
      integer(c_int), value :: ncomp_X, i_fld_X, i_fld_WR
      integer(c_int), intent(in) :: irev_import(nnod_new)
      real(c_double), intent(inout) :: WR(NB*(nnod_recv+1))
      real(c_double), intent(inout) :: d_new(nnod_new, ncomp_X)
      integer(kind = kint) :: nbk, nnk, nrk, nxk, ifx, ifw
      integer(kind = kint), allocatable :: irev(:)
      nbk = NB
      nnk = nnod_new
      nrk = nnod_recv
      nxk = ncomp_X
      ifx = i_fld_X
      ifw = i_fld_WR
      allocate(irev(nnk))
      irev(1:nnk) = irev_import(1:nnk)
      select case(ncomp)
      case(1)
        call set_from_recv_buf_rev_scalar(nbk, nnk, nrk, irev, nxk,     &
     &      ifx, ifw, WR, d_new)
      case(3)
        call set_from_recv_buf_rev_vector(nbk, nnk, nrk, irev, nxk,     &
     &      ifx, ifw, WR, d_new)
      case(6)
        call set_from_recv_buf_rev_tensor(nbk, nnk, nrk, irev, nxk,     &
     &      ifx, ifw, WR, d_new)
      end select
      deallocate(irev)
      end subroutine ref_from_recv_rev
!
!> Coriolis tables of init_sum_coriolis_rlm
!! (MS/t_coriolis_terms_rlm.f90:133-183) on the tables of ref_setup:
!! cal_gaunt_coriolis_rlm and interact_rot_coriolis_rlm
!! (MS/interact_coriolis_rlm.f90), MS/ = MHD/Fortran_src/MHD_src/sph_MHD/.
!! Output arrays in the reference's shapes (first index fastest).
      subroutine ref_coriolis_tables(jgi, jei, gi, ei, sw, tw, sd, td,  &
     &          sr, tr) bind(C, name='ref_coriolis_tables')
      integer(c_int), intent(inout) :: jgi(jmax_h,2), jei(jmax_h,1)
      real(c_double), intent(inout) :: gi(jmax_h,2), ei(jmax_h,1)
      real(c_double), intent(inout) :: sw(2,3,jmax_h), tw(2,4,jmax_h)
      real(c_double), intent(inout) :: sd(2,2,jmax_h), td(2,jmax_h)
      real(c_double), intent(inout) :: sr(2,jmax_h), tr(2,jmax_h)
      integer(kind = kint), allocatable :: jg(:,:), je(:,:)
!
      allocate(jg(jmax_h,2), je(jmax_h,1))
      jg = 0
      je = 0
      gi = 0.0d0
      ei = 0.0d0
      sw = 0.0d0
      tw = 0.0d0
      sd = 0.0d0
      td = 0.0d0
      sr = 0.0d0
      tr = 0.0d0
      call cal_gaunt_coriolis_rlm(ltr_h, jmax_h, idx_j_h, jg, je,       &
     &    gi, ei)
      call interact_rot_coriolis_rlm(jmax_h, g_h, gi, ei,               &
     &    sw, tw, sd, td, sr, tr)
      jgi(1:jmax_h,1:2) = jg(1:jmax_h,1:2)
      jei(1:jmax_h,1) = je(1:jmax_h,1)
      deallocate(jg, je)
      end subroutine ref_coriolis_tables
!
!> sum_rot_coriolis_rlm_10 and, for kr_icb > 0, inner_core_rot_z_coriolis_rlm
!! (MS/sum_coriolis_terms_rlm.f90) on the tables of ref_setup, as
!! sum_coriolis_rlm (MS/t_coriolis_terms_rlm.f90:187-258) calls them:
!! i_velo / i_vort 1-based components of WR (NB = ncomp), irev_sr_rlm the
!! identity, omega_rlm as set_zdir_rot_earth_4_sph with rotate = 1
!! (MS/t_poloidal_rotation.f90:155-250: r^2/2, r, 1).  d_cor(nnod_rlm,2).
      subroutine ref_sum_coriolis(ncomp, i_velo, i_vort, coef_cor,      &
     &          kr_icb, WR, d_cor) bind(C, name='ref_sum_coriolis')
      integer(c_int), value :: ncomp, i_velo, i_vort, kr_icb
      real(c_double), value :: coef_cor
      real(c_double), intent(in) :: WR(*)
      real(c_double), intent(inout) :: d_cor(jmax_h*nr_h,2)
      type(phys_address) :: b_trns
      integer(kind = kint) :: nnod, n_WR, k, j, nbk, m
      integer(kind = kint) :: jgi(jmax_h,2), jei(jmax_h,1)
      integer(kind = kint) :: idx1(-1:1)
      integer(kind = kint), allocatable :: irev(:)
      real(kind = kreal) :: gi(jmax_h,2), ei(jmax_h,1)
      real(kind = kreal), allocatable :: sw(:,:,:), tw(:,:,:)
      real(kind = kreal), allocatable :: sd(:,:,:), td(:,:)
      real(kind = kreal), allocatable :: sr(:,:), tr(:,:)
      real(kind = kreal), allocatable :: omega(:,:), rj(:)
!
      nnod = jmax_h * nr_h
      nbk = ncomp
      n_WR = nbk * nnod
      allocate(irev(nnod), omega(nr_h,0:2), rj(max(jmax_h, nr_h)))
      allocate(sw(2,3,jmax_h), tw(2,4,jmax_h), sd(2,2,jmax_h))
      allocate(td(2,jmax_h), sr(2,jmax_h), tr(2,jmax_h))
      do j = 1, nnod
        irev(j) = j
      end do
      do k = 1, nr_h
        omega(k,0) = half * one * radius_h(k)**2
        omega(k,1) =        one * radius_h(k)
        omega(k,2) =        one
      end do
      call ref_coriolis_tables(jgi, jei, gi, ei, sw, tw, sd, td, sr, tr)
      b_trns%i_velo = i_velo
      b_trns%i_vort = i_vort
      d_cor = 0.0d0
      call sum_rot_coriolis_rlm_10(b_trns, nnod, nidx_rlm_h, a_r_h,     &
     &    g_h, omega, coef_cor, jgi, jei, sw, tw, nbk, n_WR, irev, WR,  &
     &    d_cor(1,1), d_cor(1,2))
      if(kr_icb .gt. 0) then
        do m = -1, 1
          idx1(m) = find_local_sph_rlm_address(jmax_h, idx_j_h, ione, m)
        end do
        rj = 0.0d0
        rj(1:nr_h) = radius_h(1:nr_h)
        call inner_core_rot_z_coriolis_rlm(b_trns, nnod, nidx_rlm_h,    &
     &      rj, omega, coef_cor, nbk, n_WR, irev, WR,                   &
     &      int(kr_icb, kind = kint), idx1, d_cor(1,1))
      end if
      deallocate(irev, omega, rj, sw, tw, sd, td, sr, tr)
      end subroutine ref_sum_coriolis
!
      subroutine ref_release() bind(C, name='ref_release')
      if(allocated(radius_h)) deallocate(radius_h, a_r_h)
      if(allocated(gpt_h)) deallocate(gpt_h, wgt_h, colat_h, asin_h)
      if(allocated(idx_j_h)) deallocate(idx_j_h, lstack_h, lstack_e_h)
      if(allocated(g_h)) deallocate(g_h)
      if(allocated(P_h)) deallocate(P_h, dP_h, Ps_h, dPs_h)
      end subroutine ref_release
!
!> Copy tables out: which = 1 points, 2 weights, 3 P(nth,jmax),
!! 4 dP(nth,jmax), 5 Ps(nhemi,jmax), 6 dPs(nhemi,jmax), 7 g(jmax,17)
      subroutine ref_get_table(which, out) bind(C, name='ref_get_table')
      integer(c_int), value :: which
      real(c_double), intent(inout) :: out(*)
      integer(kind = kint) :: n
      select case(which)
      case(1)
        out(1:nth_h) = gpt_h(1:nth_h)
      case(2)
        out(1:nth_h) = wgt_h(1:nth_h)
      case(3)
        n = nth_h*jmax_h
        out(1:n) = reshape(P_h, (/n/))
      case(4)
        n = nth_h*jmax_h
        out(1:n) = reshape(dP_h, (/n/))
      case(5)
        n = nhemi_h*jmax_h
        out(1:n) = reshape(Ps_h, (/n/))
      case(6)
        n = nhemi_h*jmax_h
        out(1:n) = reshape(dPs_h, (/n/))
      case(7)
        n = jmax_h*17
        out(1:n) = reshape(g_h, (/n/))
      end select
      end subroutine ref_get_table
!
! ----------------------------------------------------------------------
!
!> Backward Legendre, symmetric_BLAS_big (matmul flavour), one thread.
!! WR: rlm receive buffer ncomp*(jmax*nr), WS: rtm send buffer
!! ncomp*(nth*nr*nslot); identity exchange tables.
      subroutine ref_leg_bwd(nvector, nscalar, WR, WS)                  &
     &          bind(C, name='ref_leg_bwd')
      integer(c_int), value :: nvector, nscalar
      real(c_double), intent(inout) :: WR(*), WS(*)
      integer(kind = kint) :: ncomp, nv, ns, nnod_rlm, nnod_rtm
      integer(kind = kint) :: kst, nkr, nkrs, nkrt, nl_rtm
      integer(kind = kint) :: mp_rlm, mn_rlm, jst, jst_h, n_jk_e, n_jk_o
      integer(kind = kint) :: n_WR, n_WS, i, maxk
      integer(kind = kint), allocatable :: irev_rlm(:), irev_rtm(:)
      real(kind = kreal), allocatable :: pol_e(:), tor_e(:)
      real(kind = kreal), allocatable :: pol_o(:), tor_o(:)
      real(kind = kreal), allocatable :: symp_r(:), asmp_p(:)
      real(kind = kreal), allocatable :: asmp_r(:), symp_p(:)
!
      nv = nvector
      ns = nscalar
      ncomp = 3*nv + ns
      nnod_rlm = jmax_h * nr_h
      nnod_rtm = nth_h * nr_h * nslot_h
      n_WR = ncomp * nnod_rlm
      n_WS = ncomp * nnod_rtm
      allocate(irev_rlm(nnod_rlm), irev_rtm(nnod_rtm))
      do i = 1, nnod_rlm
        irev_rlm(i) = i
      end do
      do i = 1, nnod_rtm
        irev_rtm(i) = i
      end do
      kst = 0
      nkr = nr_h
      nkrs = ncomp*nkr
      nkrt = 2*nv*nkr
      nl_rtm = (nth_h + 1)/2
      maxk = (ltr_h+2)/2
      allocate(pol_e(maxk*nkrs), pol_o(maxk*nkrs))
      allocate(tor_e(maxk*nkrt+1), tor_o(maxk*nkrt+1))
      allocate(symp_r(nl_rtm*nkrs), asmp_r(nl_rtm*nkrs))
      allocate(asmp_p(nl_rtm*nkrt+1), symp_p(nl_rtm*nkrt+1))
!
      WS(1:n_WS) = 0.0d0
      do mp_rlm = 1, nslot_h
        mn_rlm = nslot_h - mp_rlm + 1
        jst = lstack_h(mp_rlm-1)
        jst_h = lstack_e_h(mp_rlm) + 1
        n_jk_e = lstack_e_h(mp_rlm) - lstack_h(mp_rlm-1)
        n_jk_o = lstack_h(mp_rlm) - lstack_e_h(mp_rlm)
        call set_sp_rlm_vec_sym_matmul_big                              &
     &     (nnod_rlm, nidx_rlm_h, istep_rlm_h, idx_j_h, a_r_h, g_h,     &
     &      kst, nkr, jst, n_jk_e, n_jk_o, ncomp, nv, irev_rlm,         &
     &      n_WR, WR, pol_e, tor_e, pol_o, tor_o)
        call set_sp_rlm_scl_sym_matmul_big                              &
     &     (nnod_rlm, nidx_rlm_h, istep_rlm_h,                          &
     &      kst, nkr, jst, n_jk_e, n_jk_o, ncomp, nv, ns, irev_rlm,     &
     &      n_WR, WR, pol_e, pol_o)
        call matmul_bwd_leg_trans(nl_rtm, nkrs, n_jk_e,                 &
     &      Ps_h(1,jst+1), pol_e, symp_r)
        call matmul_bwd_leg_trans(nl_rtm, nkrt, n_jk_e,                 &
     &      dPs_h(1,jst+1), tor_e, asmp_p)
        call matmul_bwd_leg_trans(nl_rtm, nkrs, n_jk_o,                 &
     &      Ps_h(1,jst_h), pol_o, asmp_r)
        call matmul_bwd_leg_trans(nl_rtm, nkrt, n_jk_o,                 &
     &      dPs_h(1,jst_h), tor_o, symp_p)
        call cal_vr_rtm_vec_sym_matmul_big                              &
     &     (nnod_rtm, nidx_rtm_h, istep_rtm_h, nidx_rlm_h, asin_h,      &
     &      kst, nkr, mp_rlm, mn_rlm, nl_rtm, symp_r, asmp_p,           &
     &      asmp_r, symp_p, ncomp, nv, irev_rtm, n_WS, WS)
        call cal_vr_rtm_scl_sym_matmul_big                              &
     &     (nnod_rtm, nidx_rtm_h, istep_rtm_h, nidx_rlm_h,              &
     &      kst, nkr, mp_rlm, nl_rtm, symp_r, asmp_r,                   &
     &      ncomp, nv, ns, irev_rtm, n_WS, WS)
      end do
      deallocate(irev_rlm, irev_rtm, pol_e, pol_o, tor_e, tor_o)
      deallocate(symp_r, asmp_r, asmp_p, symp_p)
      end subroutine ref_leg_bwd
!
!> Forward Legendre, symmetric_BLAS_big (matmul flavour), one thread.
!! WR: rtm receive buffer, WS: rlm send buffer.
      subroutine ref_leg_fwd(nvector, nscalar, WR, WS)                  &
     &          bind(C, name='ref_leg_fwd')
      integer(c_int), value :: nvector, nscalar
      real(c_double), intent(inout) :: WR(*), WS(*)
      integer(kind = kint) :: ncomp, nv, ns, nnod_rlm, nnod_rtm
      integer(kind = kint) :: kst, nkr, nkrs, nkrt, nle_rtm, nlo_rtm
      integer(kind = kint) :: mp_rlm, mn_rlm, jst, jst_h, n_jk_e, n_jk_o
      integer(kind = kint) :: n_WR, n_WS, i, maxk
      integer(kind = kint), allocatable :: irev_rlm(:), irev_rtm(:)
      real(kind = kreal), allocatable :: pol_e(:), tor_e(:)
      real(kind = kreal), allocatable :: pol_o(:), tor_o(:)
      real(kind = kreal), allocatable :: symp_r(:), asmp_p(:)
      real(kind = kreal), allocatable :: asmp_r(:), symp_p(:)
!
      nv = nvector
      ns = nscalar
      ncomp = 3*nv + ns
      nnod_rlm = jmax_h * nr_h
      nnod_rtm = nth_h * nr_h * nslot_h
      n_WR = ncomp * nnod_rtm
      n_WS = ncomp * nnod_rlm
      allocate(irev_rlm(nnod_rlm), irev_rtm(nnod_rtm))
      do i = 1, nnod_rlm
        irev_rlm(i) = i
      end do
      do i = 1, nnod_rtm
        irev_rtm(i) = i
      end do
      kst = 0
      nkr = nr_h
      nkrs = ncomp*nkr
      nkrt = 2*nv*nkr
      nle_rtm = (nth_h + 1)/2
      nlo_rtm = nth_h / 2
      maxk = (ltr_h+2)/2
      allocate(pol_e(maxk*nkrs), pol_o(maxk*nkrs))
      allocate(tor_e(maxk*nkrt+1), tor_o(maxk*nkrt+1))
      allocate(symp_r(nle_rtm*nkrs), asmp_r(nle_rtm*nkrs))
      allocate(asmp_p(nle_rtm*nkrt+1), symp_p(nle_rtm*nkrt+1))
!
      WS(1:n_WS) = 0.0d0
      do mp_rlm = 1, nslot_h
        mn_rlm = nslot_h - mp_rlm + 1
        jst = lstack_h(mp_rlm-1)
        jst_h = lstack_e_h(mp_rlm) + 1
        n_jk_e = lstack_e_h(mp_rlm) - lstack_h(mp_rlm-1)
        n_jk_o = lstack_h(mp_rlm) - lstack_e_h(mp_rlm)
        call set_vr_rtm_vec_sym_matmul_big                              &
     &     (nnod_rtm, nidx_rtm_h, istep_rtm_h, nidx_rlm_h, asin_h,      &
     &      wgt_h, kst, nkr, mp_rlm, mn_rlm, nle_rtm, nlo_rtm,          &
     &      ncomp, nv, irev_rtm, n_WR, WR,                              &
     &      symp_r, asmp_p, asmp_r, symp_p)
        call set_vr_rtm_scl_sym_matmul_big                              &
     &     (nnod_rtm, nidx_rtm_h, istep_rtm_h, nidx_rlm_h, wgt_h,       &
     &      kst, nkr, mp_rlm, nle_rtm, nlo_rtm,                         &
     &      ncomp, nv, ns, irev_rtm, n_WR, WR, symp_r, asmp_r)
        call matmul_fwd_leg_trans(nkrs, n_jk_e, nle_rtm,                &
     &      symp_r, Ps_h(1,jst+1), pol_e)
        call matmul_fwd_leg_trans(nkrt, n_jk_e, nle_rtm,                &
     &      asmp_p, dPs_h(1,jst+1), tor_e)
        call matmul_fwd_leg_trans(nkrs, n_jk_o, nle_rtm,                &
     &      asmp_r, Ps_h(1,jst_h), pol_o)
        call matmul_fwd_leg_trans(nkrt, n_jk_o, nle_rtm,                &
     &      symp_p, dPs_h(1,jst_h), tor_o)
        call cal_sp_rlm_vec_sym_matmul_big                              &
     &     (nnod_rlm, nidx_rlm_h, istep_rlm_h, idx_j_h, radius_h, g_h,  &
     &      kst, nkr, jst, n_jk_o, n_jk_e, pol_e, pol_o, tor_e, tor_o,  &
     &      ncomp, nv, irev_rlm, n_WS, WS)
        call cal_sp_rlm_scl_sym_matmul_big(nnod_rlm, nidx_rlm_h,        &
     &      istep_rlm_h, g_h, kst, nkr, jst, n_jk_o, n_jk_e,            &
     &      pol_e, pol_o, ncomp, nv, ns, irev_rlm, n_WS, WS)
      end do
      deallocate(irev_rlm, irev_rtm, pol_e, pol_o, tor_e, tor_o)
      deallocate(symp_r, asmp_r, asmp_p, symp_p)
      end subroutine ref_leg_fwd
!
      end module kemo_ref_harness
