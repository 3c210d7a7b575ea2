! fetd_main.f90 -- FETD forward driver on the MI355X library.
!
!   fetd_amd -f FILE_ROOT [-v full|none] [-r]
!
! Same command line, inputs and outputs as the reference program
! (src/FETD_MAIN.f90:23-95): reads FILE_ROOT.mesh and FILE_ROOT.cntl, runs the
! forward time-stepping loop, writes FILE_ROOT.field (one row per step) and
! FILE_ROOT.field2 (one row per time_points gate), and for every vtkout time
! the run reaches FILE_ROOT_i.msh, a copy of FILE_ROOT.msh with the per-tet
! fields (msh_output, src/FETD_FWD.f90:1642-1656,1713-1864).  -r selects the
! reference's small-loop build (src_small): conductivities of domains 2..n
! scaled by 1 / sigma of the last domain, written values by
! pi**log10(sp) / sp**2 (src_small/FETD_FWD.f90:314-328,1526,1540).  The mesh and control
! reading and the output writing stay in Fortran (module fetd_io); topology,
! assembly, factorisation, BDF2 stepping with the doubling controller and the
! receiver readout run on the GPU through the C ABI (module temfetd).
program fetd_amd
  use iso_c_binding
  use fetd_io
  use temfetd
  implicit none

  character(512) :: arg, root, verbose
  character(256) :: msg
  character(1) :: job
  logical :: loud, rescale, have_msh
  integer :: narg, k, ierr, nN, nT, nseg, nrows, cap, i, rc
  real(c_double), allocatable :: xyz(:, :), sigma(:), inv_mu(:), row_t(:), row_dt(:), fields(:, :, :)
  real(c_double), allocatable :: sig9(:, :), nu9(:, :), f9(:, :)
  real(c_double) :: slice_t
  integer(c_int) :: slice_step
  character(32) :: num
  integer(c_int), allocatable :: tet(:, :), tet_phy(:), seg(:, :), rec_tet(:), row_step(:)
  type(cntl_t) :: c
  type(c_ptr) :: ctx
  type(field_writer_t) :: w
  real(c_double) :: tm(8)
  integer(c_int) :: nE, nnzU

  ! ---- command line (src/FETD_MAIN.f90:23-95) ----
  job = ' '
  root = ' '
  verbose = 'full'
  rescale = .false.
  narg = command_argument_count()
  k = 1
  do while (k <= narg)
    call get_command_argument(k, arg)
    if (trim(arg) == '-r') then
      rescale = .true.
      k = k + 1
      cycle
    end if
    if (arg(1:2) == '-v' .and. k < narg) then
      call get_command_argument(k + 1, verbose)
      k = k + 2
      cycle
    end if
    if (arg(1:1) == '-' .and. job == ' ') then
      job = arg(2:2)
    else if (root == ' ') then
      root = arg
    end if
    k = k + 1
  end do
  if (job /= 'f' .or. root == ' ') then
    write(*, '(a)') 'Usage: fetd_amd -[job] FILE_ROOT'
    write(*, '(a)') '[FORWARD]    -f    Calculates the predicted data and saves the EM solution'
    write(*, '(a)') 'Optional     -v    [full|none]'
    write(*, '(a)') 'Optional     -r    small-loop variant (conductivity rescaling)'
    stop 1
  end if
  loud = trim(verbose) /= 'none'
  if (loud) write(*, '(a)') 'fetd_amd: run job -f with root name: '//trim(root)

  ! ---- model_mesh / model_ctl (host) ----
  call read_medit(trim(root)//'.mesh', nN, xyz, nT, tet, tet_phy, nseg, seg, ierr, msg)
  if (ierr /= 0) call die(msg)
  call read_cntl(trim(root)//'.cntl', c, ierr, msg)
  if (ierr /= 0) call die(msg)
  if (c%sp_mthd /= 2) call die('error!')  ! only the adaptive BDF2 method exists (src/FETD_FWD.f90:1210-1219)
  if (c%bound_cond /= 0) call die('absorbing boundary condition is not supported')
  if (rescale) then
    if (c%aniso) call die('the conductivity rescaling applies to isotropic domain tables')
    call rescale_domains(c)
  end if
  allocate(sigma(nT), inv_mu(nT), rec_tet(c%rec_num))
  if (c%aniso) then
    allocate(sig9(9, nT), nu9(9, nT))
    call tet_tensors(c, nT, tet_phy, sig9, nu9, ierr)
  else
    call tet_materials(c, nT, tet_phy, sigma, inv_mu, ierr)
  end if
  if (ierr /= 0) call die('error - physical domain not found!')
  if (loud) write(*, '(a,i0,a,i0,a,i0,a)') ' mesh: ', nN, ' vertices, ', nT, ' tets, ', nseg, ' source segments'

  ! ---- GPU path through the C ABI ----
  call check(tf_create(ctx, 0_c_int), 'tf_create')
  call check(tf_set_mesh(ctx, int(nN, c_int), xyz, int(nT, c_int), tet), 'tf_set_mesh')
  call check(tf_get_sizes(ctx, nE, nnzU), 'tf_get_sizes')
  if (loud) write(*, '(a,i0,a,i0)') ' edges: ', nE, '  upper nonzeros: ', nnzU
  if (c%aniso) then
    call check(tf_set_materials_tensor(ctx, sig9, nu9), 'tf_set_materials_tensor')
  else
    call check(tf_set_materials(ctx, sigma, inv_mu), 'tf_set_materials')
  end if
  call check(tf_set_source(ctx, 0_c_int, int(nseg, c_int), seg, int(c%source_type, c_int), &
                           int(c%pulse_type, c_int), c%pulse_width, c%current), 'tf_set_source')
  call check(tf_set_receivers(ctx, int(c%rec_num, c_int), c%rec, c%rec_charac, rec_tet), 'tf_set_receivers')
  call check(tf_set_schedule(ctx, int(c%sp_division, c_int), int(c%sp_double, c_int), &
                             int(c%sp_dblesize, c_int), c%time_maximum), 'tf_set_schedule')
  if (c%vtk_num > 0) call check(tf_set_slices(ctx, int(c%vtk_num, c_int), c%vtk_times), 'tf_set_slices')
  call check(tf_assemble(ctx), 'tf_assemble')

  ! rows are returned by the loop; a run longer than the buffer is repeated
  ! with a buffer of the counted size (the loop is deterministic)
  cap = 20000
  do
    allocate(row_step(cap), row_t(cap), row_dt(cap), fields(9, c%rec_num, cap))
    call check(tf_run(ctx, 0_c_int, int(cap, c_int), nrows, row_step, row_t, row_dt, fields), 'tf_run')
    if (nrows <= cap) exit
    cap = nrows
    deallocate(row_step, row_t, row_dt, fields)
  end do

  ! ---- solution_output: .field and .field2 ----
  if (rescale) then
    call fw_open(w, trim(root)//'.field', trim(root)//'.field2', c%rec_num, c%o_f, c%time_points, c%sig_parameter)
  else
    call fw_open(w, trim(root)//'.field', trim(root)//'.field2', c%rec_num, c%o_f, c%time_points)
  end if
  do i = 1, nrows
    call fw_row(w, int(row_step(i)), row_t(i), fields(:, :, i))
  end do
  call fw_close(w)

  ! ---- field slices: ROOT_i.msh for every vtkout time the run reached ----
  inquire(file=trim(root)//'.msh', exist=have_msh)
  if (c%vtk_num > 0 .and. have_msh) then
    allocate(f9(9, nT))
    do i = 1, c%vtk_num
      if (tf_get_slice(ctx, int(i - 1, c_int), f9, slice_step, slice_t) /= 0) cycle  ! not reached
      write(num, '(i0)') i
      call write_msh_slice(trim(root)//'.msh', trim(root)//'_'//trim(num)//'.msh', nT, f9, c%o_f, ierr, msg)
      if (ierr /= 0) call die(msg)
      if (loud) write(*, '(a,es12.5,a)') ' field slice at t = ', slice_t, ' -> '//trim(root)//'_'//trim(num)//'.msh'
    end do
  end if

  rc = tf_get_timers(ctx, tm)
  if (loud) then
    write(*, '(a,i0,a,es12.5)') ' steps: ', row_step(nrows), '  final time: ', row_t(nrows)
    write(*, '(a,f10.3,a)') ' time of assembling:    ', tm(1), ' s'
    write(*, '(a,f10.3,a,i0,a)') ' time of factorization: ', tm(2), ' s (', nint(tm(8)), ' factorisations)'
    write(*, '(a,f10.3,a)') ' time of stepping:      ', tm(6) - tm(2), ' s'
  end if
  call tf_destroy(ctx)

contains

  subroutine check(status, what)
    integer(c_int), intent(in) :: status
    character(*), intent(in) :: what
    if (status /= 0) call die(what//': '//tf_error_message(ctx))
  end subroutine check

  subroutine die(text)
    character(*), intent(in) :: text
    write(*, '(a)') 'fetd_amd: '//trim(text)
    stop 2
  end subroutine die
end program fetd_amd
