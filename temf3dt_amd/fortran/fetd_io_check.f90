! fetd_io_check.f90 -- the driver's Fortran readers and writers without the
! GPU, for tests:
!   fetd_io_check read ROOT        print what read_medit/read_cntl found
!   fetd_io_check write IN ROOT [SP]  rows from IN -> ROOT.field, ROOT.field2
!                                  (SP: the small-loop variant's scale)
!   fetd_io_check rescale ROOT     the small-loop variant's domain table
!   fetd_io_check slice MSH IN DST per-tet values from IN -> DST (copy of MSH
!                                  with $ElementData sections)
! IN of write (list-directed): nrec ngates / o_f(9) / gates / nrows / per
! row: step t and 9*nrec values (component-major per receiver).  IN of
! slice: nT / o_f(9) / nT rows of 9 values.
program fetd_io_check
  use iso_c_binding
  use fetd_io
  implicit none
  character(512) :: mode, a1, a2, msg
  integer :: ierr, nN, nT, nseg, i, nrec, ng, nrows, step, u, o_f(9)
  real(8) :: sp
  character(512) :: a3
  real(c_double), allocatable :: xyz(:, :), gates(:), vals(:, :), s9(:, :), n9(:, :)
  integer(c_int), allocatable :: tet(:, :), tet_phy(:), seg(:, :)
  real(8) :: t
  type(cntl_t) :: c
  type(field_writer_t) :: w

  call get_command_argument(1, mode)
  call get_command_argument(2, a1)
  if (trim(mode) == 'read') then
    call read_medit(trim(a1)//'.mesh', nN, xyz, nT, tet, tet_phy, nseg, seg, ierr, msg)
    if (ierr /= 0) then
      write(*, '(a)') 'error '//trim(msg)
      stop 2
    end if
    write(*, '(a,3i12)') 'counts', nN, nT, nseg
    write(*, '(a,es25.17)') 'xyzsum', sum(xyz)
    write(*, '(a,3es25.17)') 'vfirst', xyz(:, 1)
    write(*, '(a,3es25.17)') 'vlast', xyz(:, nN)
    write(*, '(a,i20)') 'tetsum', sum(int(tet(1, :), 8)) + 2 * sum(int(tet(2, :), 8)) + &
                                  3 * sum(int(tet(3, :), 8)) + 4 * sum(int(tet(4, :), 8))
    write(*, '(a,i20)') 'physum', sum(int(tet_phy, 8))
    do i = 1, nseg
      write(*, '(a,2i12)') 'seg', seg(:, i)
    end do
    call read_cntl(trim(a1)//'.cntl', c, ierr, msg)
    if (ierr /= 0) then
      write(*, '(a)') 'error '//trim(msg)
      stop 2
    end if
    do i = 1, c%rec_num
      write(*, '(a,3es25.17)') 'rec', c%rec(:, i)
    end do
    do i = 1, c%dom_num
      write(*, '(a,4es25.17)') 'dom', c%dom(:, i)
    end do
    write(*, '(a,2i4,2es25.17)') 'source', c%source_type, c%pulse_type, c%pulse_width, c%current
    write(*, '(a,4i8,es25.17)') 'schedule', c%sp_mthd, c%sp_division, c%sp_double, c%sp_dblesize, c%time_maximum
    write(*, '(a,9i2)') 'output', c%o_f
    write(*, '(a,es25.17)') 'rec_charac', c%rec_charac
    if (c%aniso) then
      allocate(s9(9, nT), n9(9, nT))
      call tet_tensors(c, nT, tet_phy, s9, n9, ierr)
      write(*, '(a,9es25.17)') 'tsig1', s9(:, 1)
      write(*, '(a,9es25.17)') 'tnu1', n9(:, 1)
      write(*, '(a,9es25.17)') 'tsigN', s9(:, nT)
      write(*, '(a,9es25.17)') 'tnuN', n9(:, nT)
    end if
    write(*, '(a,i6)') 'ngates', size(c%time_points)
  else if (trim(mode) == 'write') then
    call get_command_argument(3, a2)
    open(newunit=u, file=trim(a1), status='old', action='read')
    read(u, *) nrec, ng
    read(u, *) o_f
    allocate(gates(ng), vals(9, nrec))
    do i = 1, ng
      read(u, *) gates(i)
    end do
    read(u, *) nrows
    call get_command_argument(4, a3)
    if (len_trim(a3) > 0) then
      read(a3, *) sp
      call fw_open(w, trim(a2)//'.field', trim(a2)//'.field2', nrec, o_f, gates, sp)
    else
      call fw_open(w, trim(a2)//'.field', trim(a2)//'.field2', nrec, o_f, gates)
    end if
    do i = 1, nrows
      read(u, *) step, t, vals
      call fw_row(w, step, t, vals)
    end do
    call fw_close(w)
    close(u)
  else if (trim(mode) == 'rescale') then
    call read_cntl(trim(a1)//'.cntl', c, ierr, msg)
    if (ierr /= 0) then
      write(*, '(a)') 'error '//trim(msg)
      stop 2
    end if
    call rescale_domains(c)
    write(*, '(a,es25.17)') 'sp', c%sig_parameter
    do i = 1, c%dom_num
      write(*, '(a,4es25.17)') 'dom', c%dom(:, i)
    end do
  else if (trim(mode) == 'slice') then
    call get_command_argument(3, a2)
    call get_command_argument(4, a3)
    open(newunit=u, file=trim(a2), status='old', action='read')
    read(u, *) nT
    read(u, *) o_f
    allocate(vals(9, nT))
    do i = 1, nT
      read(u, *) vals(:, i)
    end do
    close(u)
    call write_msh_slice(trim(a1), trim(a3), nT, vals, o_f, ierr, msg)
    if (ierr /= 0) then
      write(*, '(a)') 'error '//trim(msg)
      stop 2
    end if
  else
    write(*, '(a)') 'usage: fetd_io_check read ROOT | write IN ROOT [SP] | rescale ROOT | slice MSH IN DST'
    stop 1
  end if
end program fetd_io_check
