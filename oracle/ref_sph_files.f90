!> ref_sph_files.f90 -- TEST INFRASTRUCTURE ONLY.
!! Writes a one-rank set of the reference's spherical grid / mode files (rtp,
!! rtm, rlm, rj) with the reference's own writers (write_geom_rtp_data,
!! write_geom_rtm_data, write_modes_rlm_data, write_spectr_modes_rj_data,
!! CORE_src/IO/sph_modes_grids_data_IO.f90:190-343) and the reference's own
!! ordering routines (set_wavenumber_4_ispack_fft, set_zonal_wavenum_4_legendre,
!! set_trans_table_fft_2_lgd, set_merged_index_4_sph_trans,
!! set_merged_index_4_sph_rj; CONST_SPH_GRID/set_sph_tranform_ordering.f90),
!! compiled from their sources (oracle/Makefile.ref).  The 1-D tables and
!! node lists restate the one-rank loops of
!! CONST_SPH_GRID/set_sph_1d_global_index.f90:45-222 and
!! set_local_sphere_param.f90:45-171; the import tables pair nodes as
!! set_comm_table_rtp_rj.f90:499-569 does (rlm and rtm in node order, rtp the
!! node of each rtm node's radius, colatitude and FFT slot), the rj table is
!! the caller's.  Nothing here is shipped or used by the product library.
      module kemo_ref_sph_files
!
      use iso_c_binding
      use m_precision
      use m_constants
      use t_comm_table
      use t_node_id_spherical_IO
      use t_group_data
      use t_spheric_group
      use sph_modes_grids_data_IO
      use set_sph_tranform_ordering
!
      implicit none
!
      contains
!
! ----------------------------------------------------------------------
!
      subroutine c_path(src, dst)
      character(kind = c_char), intent(in) :: src(512)
      character(len = 512), intent(inout) :: dst
      integer :: i
      dst = ''
      do i = 1, 512
        if (src(i) .eq. c_null_char) exit
        dst(i:i) = src(i)
      end do
      end subroutine c_path
!
      subroutine empty_groups(grp)
      type(group_data), intent(inout) :: grp
      grp%num_grp = 0
      call allocate_grp_type(grp)
      end subroutine empty_groups
!
      subroutine self_comm(comm, items, n)
      type(communication_table), intent(inout) :: comm
      integer(kind = kint), intent(in) :: n
      integer(kind = kint), intent(in) :: items(n)
      comm%num_neib = 1
      call allocate_type_neib_id(comm)
      comm%id_neib(1) = 0
      comm%ntot_import = n
      call allocate_type_import_num(comm)
      comm%num_import(1) = n
      comm%istack_import(1) = n
      call allocate_type_import_item(comm)
      comm%item_import(1:n) = items(1:n)
      end subroutine self_comm
!
! ----------------------------------------------------------------------
!
!> paths: four NUL-terminated paths of up to 512 bytes (rtp, rtm, rlm, rj);
!! item_rj: the rj import items (jmax*nr, 1-based rj node numbers).
      subroutine ref_write_sph_files(ltr, nr, nth, nphi, radius,        &
     &          item_rj, paths) bind(C, name='ref_write_sph_files')
      integer(c_int), value :: ltr, nr, nth, nphi
      real(c_double), intent(in) :: radius(nr)
      integer(c_int), intent(in) :: item_rj(*)
      character(kind = c_char), intent(in) :: paths(512, 4)
!
      integer(kind = kint) :: jmax, nslot, i, k, l, m, s, j, inod, id
      integer(kind = kint), allocatable :: mspec(:), mdx_ispack(:)
      integer(kind = kint), allocatable :: jdx_fsph(:), mdx_4_lgd(:)
      integer(kind = kint), allocatable :: mtbl(:), jdx_lag(:,:)
      integer(kind = kint), allocatable :: jdx_rj(:,:), items(:)
      integer(kind = kint) :: istack_m(0:1), istack_ml(0:1)
      integer(kind = kint) :: nidx_ml(1), nidx_j(1), istack_j(0:1)
      type(sph_IO_data) :: sph_IO
      type(communication_table) :: comm_IO
      type(sph_group_data) :: grps
      character(len = 512) :: path
!
      jmax = (ltr + 1)**2
      nslot = 2*ltr + 1
      allocate(mspec(-nth:nth), mdx_ispack(nphi))
      allocate(jdx_fsph(-nth:nth), mdx_4_lgd(0:nphi), mtbl(0:nphi))
      allocate(jdx_lag(0:jmax,3), jdx_rj(0:jmax,3))
      call set_wavenumber_4_ispack_fft(nth, nphi, ione,                 &
     &    mspec, mdx_ispack)
      call set_zonal_wavenum_4_legendre(ione, ltr, ione, nth, nphi,     &
     &    jdx_fsph, mdx_4_lgd)
      call set_trans_table_fft_2_lgd(ltr, nth, nphi, ione,              &
     &    mspec, jdx_fsph, mtbl)
      istack_m(0) = -1
      istack_m(1) = 2*ltr
      call set_merged_index_4_sph_trans(ione, ltr, jmax, nphi, ione,    &
     &    istack_m, mdx_4_lgd, nidx_ml, istack_ml, jdx_lag)
      call set_merged_index_4_sph_rj(ione, ione, ione, jmax,            &
     &    istack_ml, jdx_lag, nidx_j, istack_j, jdx_rj)
      id = 31
!
!     rtp: (r, theta, phi), phi slowest
      sph_IO%numdir_sph = 3
      sph_IO%nidx_gl_sph(1:3) = (/nr, nth, nphi/)
      sph_IO%ltr_gl = ltr
      sph_IO%sph_rank(1:3) = 0
      sph_IO%ncomp_table_1d(1:3) = (/1, 1, 2/)
      call alloc_num_idx_sph_IO(sph_IO)
      sph_IO%nidx_sph(1:3) = (/nr, nth, nphi/)
      sph_IO%ist_sph(1:3) = 1
      sph_IO%ied_sph(1:3) = sph_IO%nidx_sph(1:3)
      call alloc_idx_sph_1d1_IO(sph_IO)
      call alloc_idx_sph_1d2_IO(sph_IO)
      call alloc_idx_sph_1d3_IO(sph_IO)
      do k = 1, nr
        sph_IO%idx_gl_1(k) = k
        sph_IO%r_gl_1(k) = radius(k)
      end do
      do l = 1, nth
        sph_IO%idx_gl_2(l,1) = l
      end do
      do i = 1, nphi
        sph_IO%idx_gl_3(i,1) = i
        sph_IO%idx_gl_3(i,2) = mdx_ispack(i)
      end do
      sph_IO%numnod_sph = nr*nth*nphi
      call alloc_nod_id_sph_IO(sph_IO)
      inod = 0
      do m = 1, nphi
        do l = 1, nth
          do k = 1, nr
            inod = inod + 1
            sph_IO%inod_gl_sph(inod) = inod
            sph_IO%idx_gl_sph(inod,1:3) = (/k, l, m/)
          end do
        end do
      end do
!     import: the rtp node of each rtm node (radius, colatitude, FFT slot)
      allocate(items(nr*nth*nslot))
      inod = 0
      do s = 0, nslot-1
        do k = 1, nr
          do l = 1, nth
            inod = inod + 1
            items(inod) = k + (l-1)*nr + (mtbl(s)-1)*nr*nth
          end do
        end do
      end do
      call self_comm(comm_IO, items, nr*nth*nslot)
      call empty_groups(grps%bc_rtp_grp)
      call empty_groups(grps%radial_rtp_grp)
      call empty_groups(grps%theta_rtp_grp)
      call empty_groups(grps%zonal_rtp_grp)
      call c_path(paths(:,1), path)
      open(id, file = trim(path), status = 'replace')
      call write_geom_rtp_data(id, izero, comm_IO, sph_IO, grps)
      close(id)
!
!     rtm: (r, theta, m-slot), colatitude fastest
      sph_IO%numdir_sph = 3
      sph_IO%nidx_gl_sph(1:3) = (/nr, nth, nslot/)
      sph_IO%ncomp_table_1d(1:3) = (/1, 1, 2/)
      call alloc_num_idx_sph_IO(sph_IO)
      sph_IO%nidx_sph(1:3) = (/nr, nth, nslot/)
      sph_IO%ist_sph(1:3) = 1
      sph_IO%ied_sph(1:3) = sph_IO%nidx_sph(1:3)
      call alloc_idx_sph_1d1_IO(sph_IO)
      call alloc_idx_sph_1d2_IO(sph_IO)
      call alloc_idx_sph_1d3_IO(sph_IO)
      do k = 1, nr
        sph_IO%idx_gl_1(k) = k
        sph_IO%r_gl_1(k) = radius(k)
      end do
      do l = 1, nth
        sph_IO%idx_gl_2(l,1) = l
      end do
      do s = 0, nslot-1
        sph_IO%idx_gl_3(s+1,1) = mtbl(s)
        sph_IO%idx_gl_3(s+1,2) = mdx_4_lgd(s)
      end do
      sph_IO%numnod_sph = nr*nth*nslot
      call alloc_nod_id_sph_IO(sph_IO)
      inod = 0
      do s = 0, nslot-1
        do k = 1, nr
          do l = 1, nth
            inod = inod + 1
            sph_IO%inod_gl_sph(inod) = inod
            sph_IO%idx_gl_sph(inod,1:3) = (/k, l, mtbl(s)/)
            items(inod) = inod
          end do
        end do
      end do
      call self_comm(comm_IO, items, nr*nth*nslot)
      call c_path(paths(:,2), path)
      open(id, file = trim(path), status = 'replace')
      call write_geom_rtm_data(id, izero, comm_IO, sph_IO)
      close(id)
      deallocate(items)
!
!     rlm and rj: (r, j), the mode fastest
      allocate(items(nr*jmax))
      do s = 3, 4
        sph_IO%numdir_sph = 2
        sph_IO%nidx_gl_sph(1:2) = (/nr, jmax/)
        sph_IO%ncomp_table_1d(1:2) = (/1, 3/)
        call alloc_num_idx_sph_IO(sph_IO)
        sph_IO%nidx_sph(1:2) = (/nr, jmax/)
        sph_IO%ist_sph(1:2) = 1
        sph_IO%ied_sph(1:2) = sph_IO%nidx_sph(1:2)
        call alloc_idx_sph_1d1_IO(sph_IO)
        call alloc_idx_sph_1d2_IO(sph_IO)
        do k = 1, nr
          sph_IO%idx_gl_1(k) = k
          sph_IO%r_gl_1(k) = radius(k)
        end do
        do j = 1, jmax
          if (s .eq. 3) then
            sph_IO%idx_gl_2(j,1:3) = jdx_lag(j-1,1:3)
          else
            sph_IO%idx_gl_2(j,1:3) = jdx_rj(j-1,1:3)
          end if
        end do
        sph_IO%numnod_sph = nr*jmax
        call alloc_nod_id_sph_IO(sph_IO)
        inod = 0
        do k = 1, nr
          do j = 1, jmax
            inod = inod + 1
            sph_IO%inod_gl_sph(inod) = inod
            sph_IO%idx_gl_sph(inod,1) = k
            sph_IO%idx_gl_sph(inod,2) = sph_IO%idx_gl_2(j,1)
            if (s .eq. 3) then
              items(inod) = inod
            else
              items(inod) = item_rj(inod)
            end if
          end do
        end do
        call self_comm(comm_IO, items, nr*jmax)
        call c_path(paths(:,s), path)
        open(id, file = trim(path), status = 'replace')
        if (s .eq. 3) then
          call write_modes_rlm_data(id, izero, comm_IO, sph_IO)
        else
          call empty_groups(grps%radial_rj_grp)
          call empty_groups(grps%sphere_rj_grp)
          call write_spectr_modes_rj_data(id, izero, comm_IO, sph_IO,   &
     &        grps)
        end if
        close(id)
      end do
      deallocate(items, mspec, mdx_ispack, jdx_fsph, mdx_4_lgd, mtbl)
      deallocate(jdx_lag, jdx_rj)
!
      end subroutine ref_write_sph_files
!
! ----------------------------------------------------------------------
!
      end module kemo_ref_sph_files
