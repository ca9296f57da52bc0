! qpx_mod.f90 -- Fortran 2008 binding of libqpx (iso_c_binding).
!
! Drop-in replacements for the reference's hot-path subroutines.  A plugin
! that today calls
!     call H_S2_u_0_nstates_openmp(v_0,s_0,u_0,N_st,sze)   (src/Davidson/u0Hu0.irp.f:15)
!     call davidson_diag_hs2(...)                           (src/Davidson/diagonalization_hs2.irp.f:1)
!     call ZMQ_selection(N_in, pt2)                         (plugins/Full_CI_ZMQ/zmq_selection.irp.f:1)
! calls qpx_H_S2_u_0_nstates / qpx_davidson_diag / qpx_zmq_selection with
! the same arrays after qpx_init and qpx_touch_psi (the SOFT_TOUCH of psi).
! Errors stop the program with libqpx's message, as the reference stops.
module qpx_mod
  use iso_c_binding
  implicit none

  type(c_ptr), save :: qpx_ctx = c_null_ptr
  integer(c_int), parameter :: QPX_FP64 = 0, QPX_FP32 = 1

  interface
    integer(c_int) function qpx_create(device, ctx) bind(C, name="qpx_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
    end function
    integer(c_int) function qpx_destroy(ctx) bind(C, name="qpx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function qpx_last_error(ctx, buf, n) bind(C, name="qpx_last_error")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: buf(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function qpx_set_verbosity(ctx, level) bind(C, name="qpx_set_verbosity")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
    end function
    integer(c_int) function qpx_set_hamiltonian(ctx, mo_num, na, nb, h, eri, n8) bind(C, name="qpx_set_hamiltonian")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: mo_num, na, nb
      real(c_double) :: h(*), eri(*)
      integer(c_int64_t), value :: n8
    end function
    integer(c_int) function qpx_set_wavefunction(ctx, nint, ndet, psi) bind(C, name="qpx_set_wavefunction")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nint
      integer(c_int64_t), value :: ndet
      integer(c_int64_t) :: psi(*)
    end function
    integer(c_int) function qpx_sigma(ctx, prec, nst, sze, u, v, s) bind(C, name="qpx_sigma")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: prec, nst
      integer(c_int64_t), value :: sze
      real(c_double) :: u(*), v(*), s(*)
    end function
    integer(c_int) function qpx_davidson(ctx, nst, nstd, thr, szemax, s2eig, follow, mixed, swthr, seed, u, dimin, &
        e, s2, nsig) bind(C, name="qpx_davidson")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nst, nstd, szemax, s2eig, follow, mixed
      real(c_double), value :: thr, swthr
      integer(c_int64_t), value :: seed, dimin
      real(c_double) :: u(*), e(*), s2(*)
      integer(c_int) :: nsig(2)
    end function
    integer(c_int) function qpx_select_generators(ctx, nstates, coef, nsel, e0, nlist, gens, genpt2) &
        bind(C, name="qpx_select_generators")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nstates, nsel, nlist
      real(c_double) :: coef(*), e0(*), genpt2(*)
      integer(c_int) :: gens(*)
    end function
    integer(c_int) function qpx_select_tasks(ctx, nstates, coef, nsel, e0, nkeep, ntasks, tgen, tsub, pt2, dets, &
        vals, nout, taskpt2) bind(C, name="qpx_select_tasks")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nstates, nsel, ntasks
      integer(c_int64_t), value :: nkeep
      real(c_double) :: coef(*), e0(*), pt2(*), vals(*), taskpt2(*)
      integer(c_int) :: tgen(*), tsub(*)
      integer(c_int64_t) :: dets(*)
      integer(c_int64_t) :: nout
    end function
    integer(c_int) function qpx_make_s2_eigenfunction(nint, ndet, psi, nalpha, maxout, dets, nout) &
        bind(C, name="qpx_make_s2_eigenfunction")
      import :: c_int, c_int64_t
      integer(c_int), value :: nint, nalpha
      integer(c_int64_t), value :: ndet, maxout
      integer(c_int64_t) :: psi(*), dets(*)
      integer(c_int64_t) :: nout
    end function
    integer(c_int) function qpx_s2_complete(ctx, nint, ndet, psi, nalpha, maxout, dets, nout) &
        bind(C, name="qpx_s2_complete")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nint, nalpha
      integer(c_int64_t), value :: ndet, maxout
      integer(c_int64_t) :: psi(*), dets(*)
      integer(c_int64_t) :: nout
    end function
    integer(c_int) function qpx_select(ctx, nstates, coef, ngen, nsel, e0, nkeep, subset, pt2, dets, vals, nout) &
        bind(C, name="qpx_select")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nstates, ngen, nsel, subset
      real(c_double) :: coef(*), e0(*), pt2(*), vals(*)
      integer(c_int64_t), value :: nkeep
      integer(c_int64_t) :: dets(*)
      integer(c_int64_t) :: nout
    end function
  end interface

contains

  subroutine qpx_check(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, ierr
    if (rc == 0) return
    ierr = qpx_last_error(qpx_ctx, buf, 512)
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
    print *, 'libqpx error: ', trim(msg)
    stop 1
  end subroutine

  subroutine qpx_init(device, mo_num, na, nb, h, eri, n8)
    integer, intent(in) :: device, mo_num, na, nb
    double precision, intent(in) :: h(mo_num, mo_num), eri(*)
    integer(8), intent(in) :: n8
    if (.not. c_associated(qpx_ctx)) call qpx_check(qpx_create(int(device, c_int), qpx_ctx))
    call qpx_check(qpx_set_hamiltonian(qpx_ctx, int(mo_num, c_int), int(na, c_int), int(nb, c_int), h, eri, &
        int(n8, c_int64_t)))
  end subroutine

  ! after psi_det changes (SOFT_TOUCH psi_det psi_coef N_det)
  subroutine qpx_touch_psi(psi_det, nint, ndet)
    integer, intent(in) :: nint, ndet
    integer(8), intent(in) :: psi_det(nint, 2, ndet)
    call qpx_check(qpx_set_wavefunction(qpx_ctx, int(nint, c_int), int(ndet, c_int64_t), psi_det))
  end subroutine

  ! H_S2_u_0_nstates_openmp(v_0,s_0,u_0,N_st,sze)  (u0Hu0.irp.f:15-66)
  subroutine qpx_H_S2_u_0_nstates(v_0, s_0, u_0, N_st, sze)
    integer, intent(in) :: N_st, sze
    double precision, intent(inout) :: v_0(sze, N_st), s_0(sze, N_st), u_0(sze, N_st)
    call qpx_check(qpx_sigma(qpx_ctx, QPX_FP64, int(N_st, c_int), int(sze, c_int64_t), u_0, v_0, s_0))
  end subroutine

  ! davidson_diag_hs2(dets_in,u_in,s2_out,dim_in,energies,sze,N_st,N_st_diag,Nint,iunit)
  ! (diagonalization_hs2.irp.f:1); dets_in must be the psi passed to qpx_touch_psi
  subroutine qpx_davidson_diag(u_in, s2_out, dim_in, energies, N_st, N_st_diag, threshold, sze_max, s2_eig, mixed)
    integer, intent(in) :: dim_in, N_st, N_st_diag, sze_max
    double precision, intent(inout) :: u_in(dim_in, N_st_diag)
    double precision, intent(out) :: s2_out(N_st_diag), energies(N_st_diag)
    double precision, intent(in) :: threshold
    logical, intent(in) :: s2_eig, mixed
    integer(c_int) :: nsig(2), is2, imx
    is2 = 0
    if (s2_eig) is2 = 1
    imx = 0
    if (mixed) imx = 1
    call qpx_check(qpx_davidson(qpx_ctx, int(N_st, c_int), int(N_st_diag, c_int), threshold, int(sze_max, c_int), &
        is2, 0_c_int, imx, 1.d-7, int(z'5EED', c_int64_t), u_in, int(dim_in, c_int64_t), energies, s2_out, nsig))
  end subroutine

  ! ZMQ_selection(N_in, pt2) (zmq_selection.irp.f:1-100) on psi_det_sorted
  subroutine qpx_zmq_selection(N_in, pt2, psi_coef_sorted, ndet, N_states, N_det_generators, N_det_selectors, &
      E0, nint, sel_dets, sel_vals, n_sel_out)
    integer, intent(in) :: N_in, ndet, N_states, N_det_generators, N_det_selectors, nint
    double precision, intent(in) :: psi_coef_sorted(ndet, N_states), E0(N_states)
    double precision, intent(out) :: pt2(N_states), sel_vals(N_in)
    integer(8), intent(out) :: sel_dets(nint, 2, N_in)
    integer(8), intent(out) :: n_sel_out
    call qpx_check(qpx_select(qpx_ctx, int(N_states, c_int), psi_coef_sorted, int(N_det_generators, c_int), &
        int(N_det_selectors, c_int), E0, int(N_in, c_int64_t), 0_c_int, pt2, sel_dets, sel_vals, n_sel_out))
  end subroutine

  ! pt2_detail(:, i) for the generators gens(i), as the reference's PT2 task
  ! (run_pt2_slave.irp.f) gets it from select_connected, on psi_det_sorted
  subroutine qpx_pt2_generators(psi_coef_sorted, ndet, N_states, N_det_selectors, E0, n_list, gens, pt2_detail)
    integer, intent(in) :: ndet, N_states, N_det_selectors, n_list
    double precision, intent(in) :: psi_coef_sorted(ndet, N_states), E0(N_states)
    integer, intent(in) :: gens(n_list)
    double precision, intent(out) :: pt2_detail(N_states, n_list)
    integer(c_int), allocatable :: g(:)
    allocate(g(n_list))
    g = int(gens, c_int)
    call qpx_check(qpx_select_generators(qpx_ctx, int(N_states, c_int), psi_coef_sorted, &
        int(N_det_selectors, c_int), E0, int(n_list, c_int), g, pt2_detail))
  end subroutine

  ! The selection task farm (zmq_selection.irp.f:54-100) over an explicit
  ! list of "i_generator subset" tasks (run_selection_slave.irp.f:46-106) on
  ! psi_det_sorted, with one merged buffer; task_pt2(:, t) is each task's pt2
  subroutine qpx_selection_tasks(N_in, pt2, psi_coef_sorted, ndet, N_states, N_det_selectors, E0, nint, &
      n_tasks, task_generator, task_subset, sel_dets, sel_vals, n_sel_out, task_pt2)
    integer, intent(in) :: N_in, ndet, N_states, N_det_selectors, nint, n_tasks
    integer, intent(in) :: task_generator(n_tasks), task_subset(n_tasks)
    double precision, intent(in) :: psi_coef_sorted(ndet, N_states), E0(N_states)
    double precision, intent(out) :: pt2(N_states), sel_vals(N_in), task_pt2(N_states, n_tasks)
    integer(8), intent(out) :: sel_dets(nint, 2, N_in)
    integer(8), intent(out) :: n_sel_out
    integer(c_int), allocatable :: g(:), sb(:)
    allocate(g(n_tasks), sb(n_tasks))
    g = int(task_generator, c_int)
    sb = int(task_subset, c_int)
    call qpx_check(qpx_select_tasks(qpx_ctx, int(N_states, c_int), psi_coef_sorted, int(N_det_selectors, c_int), &
        E0, int(N_in, c_int64_t), int(n_tasks, c_int), g, sb, pt2, sel_dets, sel_vals, n_sel_out, task_pt2))
  end subroutine

  ! make_s2_eigenfunction (occ_pattern.irp.f:251-309): the determinants that
  ! complete the spin multiplets of psi_det, returned in a buffer allocated
  ! here; the caller appends them (fill_H_apply_buffer_no_selection)
  subroutine qpx_s2_completion(psi_det, nint, ndet, n_alpha, new_dets, n_new)
    integer, intent(in) :: nint, ndet, n_alpha
    integer(8), intent(in) :: psi_det(nint, 2, ndet)
    integer(8), allocatable, intent(out) :: new_dets(:,:,:)
    integer(8), intent(out) :: n_new
    integer(c_int64_t) :: nout, dummy(1)
    integer(c_int) :: rc
    rc = qpx_make_s2_eigenfunction(int(nint, c_int), int(ndet, c_int64_t), psi_det, int(n_alpha, c_int), &
        0_c_int64_t, dummy, nout)
    if (rc /= 0 .and. rc /= 2) stop 'qpx_make_s2_eigenfunction: bad arguments'
    allocate(new_dets(nint, 2, max(nout, 1_c_int64_t)))
    rc = qpx_make_s2_eigenfunction(int(nint, c_int), int(ndet, c_int64_t), psi_det, int(n_alpha, c_int), &
        nout, new_dets, nout)
    if (rc /= 0) stop 'qpx_make_s2_eigenfunction failed'
    n_new = nout
  end subroutine

  ! the same on the GPU of the module's context (qpx_s2_complete), same
  ! determinants in the same order
  subroutine qpx_s2_completion_gpu(psi_det, nint, ndet, n_alpha, new_dets, n_new)
    integer, intent(in) :: nint, ndet, n_alpha
    integer(8), intent(in) :: psi_det(nint, 2, ndet)
    integer(8), allocatable, intent(out) :: new_dets(:,:,:)
    integer(8), intent(out) :: n_new
    integer(c_int64_t) :: nout, dummy(1)
    integer(c_int) :: rc
    rc = qpx_s2_complete(qpx_ctx, int(nint, c_int), int(ndet, c_int64_t), psi_det, int(n_alpha, c_int), &
        0_c_int64_t, dummy, nout)
    if (rc /= 0 .and. rc /= 4) stop 'qpx_s2_complete failed'
    allocate(new_dets(nint, 2, max(nout, 1_c_int64_t)))
    rc = qpx_s2_complete(qpx_ctx, int(nint, c_int), int(ndet, c_int64_t), psi_det, int(n_alpha, c_int), &
        nout, new_dets, nout)
    if (rc /= 0) stop 'qpx_s2_complete failed'
    n_new = nout
  end subroutine

  subroutine qpx_finalize()
    integer(c_int) :: rc
    if (c_associated(qpx_ctx)) rc = qpx_destroy(qpx_ctx)
    qpx_ctx = c_null_ptr
  end subroutine

end module qpx_mod
