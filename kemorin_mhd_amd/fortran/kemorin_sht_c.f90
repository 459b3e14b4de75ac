!> @file  kemorin_sht_c.f90
!!       ISO_C_BINDING interface to libkemorin_sht.so (include/kemorin_sht.h)
!!
!! Host-side binding for the Fortran solver: the plan descriptor as an
!! interoperable derived type and one interface per C entry point.  Status
!! codes are returned as integer(c_int) (0 = ok); sht_get_last_error gives
!! the message.  Device buffers are passed as type(c_ptr) (hipMalloc'd or
!! mapped), host buffers of the *_host entry points as assumed-size arrays.
      module kemorin_sht_c
!
      use iso_c_binding
      implicit none
!
      integer(c_int), parameter :: SHT_OK = 0
      integer(c_int), parameter :: SHT_PREC_FP64 = 0
      integer(c_int), parameter :: SHT_PREC_SPLIT_INT8 = 1
!     descriptor flags: P_l^m generated per call, one tile of orders at a time;
!     count the outputs' non-finite values after each whole transform
      integer(c_int), parameter :: SHT_FLAG_P_ON_THE_FLY = 1
      integer(c_int), parameter :: SHT_FLAG_CHECK_FINITE = 2
!     status of a whole transform whose output holds NaN or Inf (with the flag)
      integer(c_int), parameter :: SHT_ERR_NONFINITE = 6
!
!>      Plan descriptor (sht_desc)
      type, bind(C) :: sht_desc
        integer(c_int) :: l_truncation
        integer(c_int) :: nidx_r
        integer(c_int) :: nidx_theta
        integer(c_int) :: nidx_phi
        type(c_ptr) :: radius
        integer(c_int) :: max_ncomp
        integer(c_int) :: precision
        integer(c_int) :: n_slices
        integer(c_int) :: flags
      end type sht_desc
!
!>    Grid-space products (nonlinear_terms_in_rtp): first components,
!!    0-based, -1 = absent / term off
      type, bind(C) :: sht_nonlinear_desc
        integer(c_int) :: i_velo = -1, i_vort = -1, i_magne = -1
        integer(c_int) :: i_current = -1, i_temp = -1, i_light = -1
        integer(c_int) :: i_m_advect = -1, i_lorentz = -1
        integer(c_int) :: i_vp_induct = -1, i_h_flux = -1
        integer(c_int) :: i_c_flux = -1, i_coriolis = -1
        real(c_double) :: coef_velo = 0.0d0, coef_lor = 0.0d0
        real(c_double) :: coef_induct = 0.0d0, coef_h_advect = 0.0d0
        real(c_double) :: coef_c_advect = 0.0d0, coef_cor = 0.0d0
      end type sht_nonlinear_desc
!
      interface
!
        integer(c_int) function sht_plan_create(desc, plan)           &
     &                          bind(C, name = 'sht_plan_create')
          import :: c_int, c_ptr, sht_desc
          type(sht_desc), intent(in) :: desc
          type(c_ptr), intent(out) :: plan
        end function sht_plan_create
!
        integer(c_int) function sht_plan_destroy(plan)                 &
     &                          bind(C, name = 'sht_plan_destroy')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
        end function sht_plan_destroy
!
        integer(c_int) function sht_plan_set_stream(plan, stream)      &
     &                          bind(C, name = 'sht_plan_set_stream')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan, stream
        end function sht_plan_set_stream
!
        integer(c_int) function sht_plan_get_table(plan, name, out, n) &
     &                          bind(C, name = 'sht_plan_get_table')
          import :: c_int, c_ptr, c_char, c_long_long
          type(c_ptr), value :: plan
          character(kind = c_char), intent(in) :: name(*)
          type(c_ptr), value :: out
          integer(c_long_long), value :: n
        end function sht_plan_get_table
!
        integer(c_long_long) function sht_plan_device_bytes(plan)      &
     &                          bind(C, name = 'sht_plan_device_bytes')
          import :: c_long_long, c_ptr
          type(c_ptr), value :: plan
        end function sht_plan_device_bytes
!
        integer(c_int) function sht_plan_timer_enable(plan, on)        &
     &                          bind(C, name = 'sht_plan_timer_enable')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: on
        end function sht_plan_timer_enable
!
        integer(c_int) function sht_plan_timer_read(plan, ms, calls, n) &
     &                          bind(C, name = 'sht_plan_timer_read')
          import :: c_int, c_ptr, c_double, c_long_long
          type(c_ptr), value :: plan
          real(c_double), intent(out) :: ms(*)
          integer(c_long_long), intent(out) :: calls(*)
          integer(c_int), value :: n
        end function sht_plan_timer_read
!
!>      sph_backward_transforms (SS/t_sph_transforms.f90:74-123), device
        integer(c_int) function sht_backward(plan, nvector, nscalar,   &
     &                          d_WS_rj, d_v_rtp)                      &
     &                          bind(C, name = 'sht_backward')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: nvector, nscalar
          type(c_ptr), value :: d_WS_rj, d_v_rtp
        end function sht_backward
!
!>      No reference counterpart: sht_backward with the library's fixed
!!      device workspace (for callers that cannot allocate)
        integer(c_int) function sht_backward_noalloc(plan, nvector,    &
     &                          nscalar, d_WS_rj, d_v_rtp)             &
     &                          bind(C, name = 'sht_backward_noalloc')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: nvector, nscalar
          type(c_ptr), value :: d_WS_rj, d_v_rtp
        end function sht_backward_noalloc
!
!>      sph_forward_transforms (SS/t_sph_transforms.f90:250-297), device
        integer(c_int) function sht_forward(plan, nvector, nscalar,    &
     &                          d_v_rtp, d_WR_rj)                      &
     &                          bind(C, name = 'sht_forward')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: nvector, nscalar
          type(c_ptr), value :: d_v_rtp, d_WR_rj
        end function sht_forward
!
!>      Same with host arrays (one H2D and one D2H copy)
        integer(c_int) function sht_backward_host(plan, nvector,       &
     &                          nscalar, WS_rj, v_rtp)                 &
     &                          bind(C, name = 'sht_backward_host')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: plan
          integer(c_int), value :: nvector, nscalar
          real(c_double), intent(in) :: WS_rj(*)
          real(c_double), intent(out) :: v_rtp(*)
        end function sht_backward_host
!
        integer(c_int) function sht_forward_host(plan, nvector,        &
     &                          nscalar, v_rtp, WR_rj)                 &
     &                          bind(C, name = 'sht_forward_host')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: plan
          integer(c_int), value :: nvector, nscalar
          real(c_double), intent(in) :: v_rtp(*)
          real(c_double), intent(out) :: WR_rj(*)
        end function sht_forward_host
!
!>      pole_b_transform (SS/t_sph_transforms.f90:203-245), device
!!      buffers; v_pole(2*Nr+1, ncomp), with_center = 1 for
!!      iflag_MESH_w_center
        integer(c_int) function sht_pole_backward(plan, nvector,       &
     &                          nscalar, d_WS_rj, d_v_pole,            &
     &                          with_center)                           &
     &                          bind(C, name = 'sht_pole_backward')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: nvector, nscalar, with_center
          type(c_ptr), value :: d_WS_rj, d_v_pole
        end function sht_pole_backward
!
!>      nonlinear_terms_in_rtp (MS/cal_nonlinear_sph_MHD.f90:64-128),
!!      device buffers in the grid layout
        integer(c_int) function sht_nonlinear_rtp(plan, desc,          &
     &                          ncomp_fld, d_fld_rtp, ncomp_frc,       &
     &                          d_frc_rtp)                             &
     &                          bind(C, name = 'sht_nonlinear_rtp')
          import :: c_int, c_ptr, sht_nonlinear_desc
          type(c_ptr), value :: plan
          type(sht_nonlinear_desc), intent(in) :: desc
          integer(c_int), value :: ncomp_fld, ncomp_frc
          type(c_ptr), value :: d_fld_rtp, d_frc_rtp
        end function sht_nonlinear_rtp
!
!>      sph_b_trans_w_poles (SS/t_sph_transforms.f90:127-199), device
        integer(c_int) function sht_backward_with_poles(plan,          &
     &                          nvector, nscalar, d_WS_rj, d_v_rtp,    &
     &                          d_v_pole, with_center)                 &
     &                          bind(C, name = 'sht_backward_with_poles')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: nvector, nscalar, with_center
          type(c_ptr), value :: d_WS_rj, d_v_rtp, d_v_pole
        end function sht_backward_with_poles
!
!>      Stage entry points (sel_[backward|forward]_legendre_trans,
!!      back_FFT_select_from_recv / fwd_FFT_select_to_send,
!!      calypso_sph_comm_N), device buffers in the reference layouts
        integer(c_int) function sht_legendre_backward(plan, nvector,   &
     &                          nscalar, d_WR_rlm, d_WS_rtm)           &
     &                          bind(C, name = 'sht_legendre_backward')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: nvector, nscalar
          type(c_ptr), value :: d_WR_rlm, d_WS_rtm
        end function sht_legendre_backward
!
        integer(c_int) function sht_legendre_forward(plan, nvector,    &
     &                          nscalar, d_WR_rtm, d_WS_rlm)           &
     &                          bind(C, name = 'sht_legendre_forward')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: nvector, nscalar
          type(c_ptr), value :: d_WR_rtm, d_WS_rlm
        end function sht_legendre_forward
!
        integer(c_int) function sht_fft_backward(plan, ncomp,          &
     &                          d_WR_rtp, d_v_rtp)                     &
     &                          bind(C, name = 'sht_fft_backward')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: ncomp
          type(c_ptr), value :: d_WR_rtp, d_v_rtp
        end function sht_fft_backward
!
        integer(c_int) function sht_fft_forward(plan, ncomp,           &
     &                          d_v_rtp, d_WS_rtp)                     &
     &                          bind(C, name = 'sht_fft_forward')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: ncomp
          type(c_ptr), value :: d_v_rtp, d_WS_rtp
        end function sht_fft_forward
!
        integer(c_int) function sht_exchange_rtm_to_rtp(plan, ncomp,   &
     &                          d_rtm, d_rtp)                          &
     &                          bind(C, name = 'sht_exchange_rtm_to_rtp')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: ncomp
          type(c_ptr), value :: d_rtm, d_rtp
        end function sht_exchange_rtm_to_rtp
!
        integer(c_int) function sht_exchange_rtp_to_rtm(plan, ncomp,   &
     &                          d_rtp, d_rtm)                          &
     &                          bind(C, name = 'sht_exchange_rtp_to_rtm')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: ncomp
          type(c_ptr), value :: d_rtp, d_rtm
        end function sht_exchange_rtp_to_rtm
!
        integer(c_int) function sht_exchange_gather_rev(plan, ncomp,   &
     &                          n_dst, n_src, d_irev, d_src, d_dst)    &
     &                          bind(C, name = 'sht_exchange_gather_rev')
          import :: c_int, c_ptr, c_long_long
          type(c_ptr), value :: plan
          integer(c_int), value :: ncomp
          integer(c_long_long), value :: n_dst, n_src
          type(c_ptr), value :: d_irev, d_src, d_dst
        end function sht_exchange_gather_rev
!
!       rj fields <-> interleaved send / receive buffers (sht_copy.hip)
        integer(c_int) function sht_rj_fields_to_send(plan, nfield,    &
     &          fields, ntot_phys, nnod_rj, d_fld, d_inod_export,       &
     &          nnod_send, ncomp_send, d_WS)                            &
     &                          bind(C, name = 'sht_rj_fields_to_send')
          import :: c_int, c_ptr, c_long_long
          type(c_ptr), value :: plan
          integer(c_int), value :: nfield
          integer(c_int), intent(in) :: fields(3, *)
          integer(c_int), value :: ntot_phys, ncomp_send
          integer(c_long_long), value :: nnod_rj, nnod_send
          type(c_ptr), value :: d_fld, d_inod_export, d_WS
        end function sht_rj_fields_to_send
!
        integer(c_int) function sht_rj_fields_from_recv(plan, nfield,  &
     &          fields, ncomp_recv, nnod_recv, d_WR, d_irev, ntot_phys, &
     &          nnod_rj, d_fld)                                         &
     &                          bind(C, name = 'sht_rj_fields_from_recv')
          import :: c_int, c_ptr, c_long_long
          type(c_ptr), value :: plan
          integer(c_int), value :: nfield
          integer(c_int), intent(in) :: fields(3, *)
          integer(c_int), value :: ncomp_recv, ntot_phys
          integer(c_long_long), value :: nnod_recv, nnod_rj
          type(c_ptr), value :: d_WR, d_irev, d_fld
        end function sht_rj_fields_from_recv
!
!       Coriolis term in rlm space (sht_coriolis.hip)
        integer(c_int) function sht_coriolis_create(plan, kr_icb, cor) &
     &                          bind(C, name = 'sht_coriolis_create')
          import :: c_int, c_ptr
          type(c_ptr), value :: plan
          integer(c_int), value :: kr_icb
          type(c_ptr), intent(out) :: cor
        end function sht_coriolis_create
!
        integer(c_int) function sht_coriolis_destroy(cor)              &
     &                          bind(C, name = 'sht_coriolis_destroy')
          import :: c_int, c_ptr
          type(c_ptr), value :: cor
        end function sht_coriolis_destroy
!
        integer(c_int) function sht_sum_coriolis_rlm(cor, ncomp_trans, &
     &          i_velo, i_vort, coef_cor, d_WR, nnod_recv,              &
     &          d_irev_sr_rlm, d_cor_rlm)                               &
     &                          bind(C, name = 'sht_sum_coriolis_rlm')
          import :: c_int, c_ptr, c_double, c_long_long
          type(c_ptr), value :: cor
          integer(c_int), value :: ncomp_trans, i_velo, i_vort
          real(c_double), value :: coef_cor
          type(c_ptr), value :: d_WR
          integer(c_long_long), value :: nnod_recv
          type(c_ptr), value :: d_irev_sr_rlm, d_cor_rlm
        end function sht_sum_coriolis_rlm
!
        integer(c_int) function sht_copy_coriolis_terms_rlm(cor,       &
     &          ncomp_trans, i_rot_coriolis, d_cor_rlm, d_item_sr_rlm,  &
     &          nnod_send, d_WS)                                        &
     &                  bind(C, name = 'sht_copy_coriolis_terms_rlm')
          import :: c_int, c_ptr, c_long_long
          type(c_ptr), value :: cor
          integer(c_int), value :: ncomp_trans, i_rot_coriolis
          type(c_ptr), value :: d_cor_rlm, d_item_sr_rlm
          integer(c_long_long), value :: nnod_send
          type(c_ptr), value :: d_WS
        end function sht_copy_coriolis_terms_rlm
!
        integer(c_int) function sht_selftest()                         &
     &                          bind(C, name = 'sht_selftest')
          import :: c_int
        end function sht_selftest
!
        integer(c_int) function sht_get_last_error(buf, n)             &
     &                          bind(C, name = 'sht_get_last_error')
          import :: c_int, c_char
          character(kind = c_char), intent(out) :: buf(*)
          integer(c_int), value :: n
        end function sht_get_last_error
!
!       Plans from the grid / mode files (sht_files.hip); paths are
!       NUL-terminated (trim(path) // c_null_char)
        integer(c_int) function sht_sph_files_read(rtp, rtm, rlm, rj,  &
     &          files) bind(C, name = 'sht_sph_files_read')
          import :: c_int, c_ptr, c_char
          character(kind = c_char), intent(in) :: rtp(*), rtm(*)
          character(kind = c_char), intent(in) :: rlm(*), rj(*)
          type(c_ptr), intent(out) :: files
        end function sht_sph_files_read
!
        subroutine sht_sph_files_destroy(files)                        &
     &                          bind(C, name = 'sht_sph_files_destroy')
          import :: c_ptr
          type(c_ptr), value :: files
        end subroutine sht_sph_files_destroy
!
        integer(c_int) function sht_sph_files_check(files)             &
     &                          bind(C, name = 'sht_sph_files_check')
          import :: c_int, c_ptr
          type(c_ptr), value :: files
        end function sht_sph_files_check
!
        integer(c_int) function sht_sph_files_shape(files, dims,       &
     &          radius) bind(C, name = 'sht_sph_files_shape')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: files
          integer(c_int), intent(out) :: dims(5)
          real(c_double), intent(out) :: radius(*)
        end function sht_sph_files_shape
!
        integer(c_long_long) function sht_sph_files_table(files, name, &
     &          out, n) bind(C, name = 'sht_sph_files_table')
          import :: c_ptr, c_char, c_int, c_long_long
          type(c_ptr), value :: files
          character(kind = c_char), intent(in) :: name(*)
          integer(c_int), intent(out) :: out(*)
          integer(c_long_long), value :: n
        end function sht_sph_files_table
!
        integer(c_int) function sht_plan_create_from_files(files,      &
     &          max_ncomp, precision, n_slices, flags, plan)            &
     &                    bind(C, name = 'sht_plan_create_from_files')
          import :: c_int, c_ptr
          type(c_ptr), value :: files
          integer(c_int), value :: max_ncomp, precision, n_slices, flags
          type(c_ptr), intent(out) :: plan
        end function sht_plan_create_from_files
!
        type(c_ptr) function sht_version()                             &
     &                          bind(C, name = 'sht_version')
          import :: c_ptr
        end function sht_version
!
      end interface
!
      end module kemorin_sht_c
