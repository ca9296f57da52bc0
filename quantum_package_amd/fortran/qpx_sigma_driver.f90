! qpx_sigma_driver.f90 -- a plain Fortran host calling libqpx through qpx_mod,
! the way a Quantum Package plugin would.  Reads a problem from a binary
! stream file, runs
!   call qpx_H_S2_u_0_nstates(v_0,s_0,u_0,N_st,sze)     (H_S2_u_0_nstates_openmp)
!   call qpx_davidson_diag(...)                          (davidson_diag_hs2)
!   call qpx_s2_completion(...)                          (make_s2_eigenfunction)
!   call qpx_pt2_generators(...)                         (select_connected, PT2 task)
! and writes v_0, s_0, the lowest energy, the determinants completing the
! spin multiplets of the first third of psi, and the PT2 contributions of
! its first four generators (that third as psi, u(:,1) as coefficients).
!
! input  (stream, native endianness): int32 mo_num, n_alpha, n_beta, n_int,
!        int64 n_det, int32 n_st, int64 n8, real64 h(mo_num,mo_num),
!        real64 eri(n8), int64 psi_det(n_int,2,n_det), real64 u(n_det,n_st)
! output (stream): real64 v(n_det,n_st), real64 s(n_det,n_st), real64 e(1),
!        int64 n_new, int64 new_dets(n_int,2,n_new), real64 pt2_detail(4)
program qpx_sigma_driver
  use qpx_mod
  implicit none
  character(len=1024) :: fin, fout
  integer(4) :: mo_num, na, nb, nint, nst
  integer(8) :: ndet, n8
  double precision, allocatable :: h(:,:), eri(:), u(:,:), v(:,:), s(:,:), ud(:,:), e(:), s2(:)
  integer(8), allocatable :: psi(:,:,:), newd(:,:,:)
  integer(8) :: nnew
  integer :: un, nthird, gens(4)
  double precision :: pt2d(1, 4), e0(1), c3(2)
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=un, file=trim(fin), access='stream', form='unformatted', status='old')
  read(un) mo_num, na, nb, nint, ndet, nst, n8
  allocate(h(mo_num,mo_num), eri(n8), psi(nint,2,ndet), u(ndet,nst), v(ndet,nst), s(ndet,nst))
  read(un) h, eri, psi, u
  close(un)
  call qpx_init(0, int(mo_num), int(na), int(nb), h, eri, n8)
  call qpx_touch_psi(psi, int(nint), int(ndet))
  call qpx_H_S2_u_0_nstates(v, s, u, int(nst), int(ndet))
  allocate(ud(ndet,4), e(4), s2(4))
  ud = 0.d0
  ud(:,1) = u(:,1)
  call qpx_davidson_diag(ud, s2, int(ndet), e, 1, 4, 1.d-10, 8, .true., .false.)
  nthird = int(ndet / 3)
  call qpx_s2_completion(psi(:, :, 1:nthird), int(nint), nthird, int(na), newd, nnew)
  ! the first third as psi, u(:,1) as its coefficients, E0 = <u|H|u>/<u|u>
  call qpx_touch_psi(psi(:, :, 1:nthird), int(nint), nthird)
  gens = (/ 1, 2, 3, 4 /)
  block
    double precision, allocatable :: c(:,:), hc(:,:), sc(:,:)
    allocate(c(nthird, 1), hc(nthird, 1), sc(nthird, 1))
    c(:, 1) = u(1:nthird, 1)
    call qpx_H_S2_u_0_nstates(hc, sc, c, 1, nthird)
    c3(1) = dot_product(c(:, 1), hc(:, 1))
    c3(2) = dot_product(c(:, 1), c(:, 1))
    e0(1) = c3(1) / c3(2)
    call qpx_pt2_generators(c, nthird, 1, nthird, e0, 4, gens, pt2d)
  end block
  open(newunit=un, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(un) v, s, e(1)
  write(un) nnew
  if (nnew > 0) write(un) newd(:, :, 1:nnew)
  write(un) pt2d(1, :)
  close(un)
  call qpx_finalize()
end program
