! temfetd.f90 -- ISO_C_BINDING interface to libtemfetd.so (include/temfetd.h).
!
! A Fortran caller binds the C ABI through these interface blocks; arrays
! are passed by reference in the caller's column-major storage, which for
! (3, nN) coordinates, (4, nT) tets and (2, nseg) segments is exactly the
! row-major [item][component] layout the ABI expects.  Every function
! returns 0 on success and a negative status on error (tf_error_message
! converts tf_last_error to a Fortran string).
module temfetd
  use iso_c_binding
  implicit none

  interface
    integer(c_int) function tf_create(ctx, device) bind(C, name='tf_create')
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function
    subroutine tf_destroy(ctx) bind(C, name='tf_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function tf_last_error(ctx) bind(C, name='tf_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tf_set_mesh(ctx, nN, nod_cor, nT, tet_vex) bind(C, name='tf_set_mesh')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nN, nT
      real(c_double), intent(in) :: nod_cor(*)
      integer(c_int), intent(in) :: tet_vex(*)
    end function
    integer(c_int) function tf_get_sizes(ctx, nE, nnzU) bind(C, name='tf_get_sizes')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: nE, nnzU
    end function
    integer(c_int) function tf_set_materials(ctx, sigma, inv_mu) bind(C, name='tf_set_materials')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: sigma(*), inv_mu(*)
    end function
    integer(c_int) function tf_set_materials_tensor(ctx, sigma9, nu9) bind(C, name='tf_set_materials_tensor')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: sigma9(*), nu9(*)
    end function
    integer(c_int) function tf_set_source(ctx, isrc, nseg, seg_vex, source_type, pulse_type, pulse_width, &
                                          current) bind(C, name='tf_set_source')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: isrc, nseg, source_type, pulse_type
      integer(c_int), intent(in) :: seg_vex(*)
      real(c_double), value :: pulse_width, current
    end function
    integer(c_int) function tf_set_receivers(ctx, nR, xyz, rec_charac, rec_tet_out) bind(C, name='tf_set_receivers')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nR
      real(c_double), intent(in) :: xyz(*)
      real(c_double), value :: rec_charac
      integer(c_int), intent(out) :: rec_tet_out(*)
    end function
    integer(c_int) function tf_set_schedule(ctx, sp_division, sp_double, sp_dblesize, time_maximum) &
        bind(C, name='tf_set_schedule')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: sp_division, sp_double, sp_dblesize
      real(c_double), value :: time_maximum
    end function
    integer(c_int) function tf_assemble(ctx) bind(C, name='tf_assemble')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tf_analyze(ctx, leaf_size) bind(C, name='tf_analyze')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: leaf_size
    end function
    integer(c_int) function tf_factor(ctx, slot, dt) bind(C, name='tf_factor')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      real(c_double), value :: dt
    end function
    integer(c_int) function tf_factor_csr(ctx, slot, n, ia, ja, va) bind(C, name='tf_factor_csr')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot, n
      integer(c_int), intent(in) :: ia(*), ja(*)
      real(c_double), intent(in) :: va(*)
    end function
    integer(c_int) function tf_solve(ctx, slot, nrhs, b, x) bind(C, name='tf_solve')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot, nrhs
      real(c_double), intent(in) :: b(*)
      real(c_double), intent(out) :: x(*)
    end function
    integer(c_int) function tf_release(ctx, slot) bind(C, name='tf_release')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
    end function
    ! pivots of the last factorisation: the counterpart of the SPD test of
    ! pardiso_direct_1 (a negative pivot is where mtype 2 would fail)
    integer(c_int) function tf_factor_pivots(ctx, n_negative, min_abs, max_abs) bind(C, name='tf_factor_pivots')
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: n_negative
      real(c_double), intent(out) :: min_abs, max_abs
    end function
    integer(c_int) function tf_set_tuning(ctx, key, value) bind(C, name='tf_set_tuning')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: key(*)
      integer(c_int), value :: value
    end function
    integer(c_int) function tf_run(ctx, max_steps, max_rows, nrows, row_step, row_t, row_dt, fields) &
        bind(C, name='tf_run')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: max_steps, max_rows
      integer(c_int), intent(out) :: nrows
      integer(c_int), intent(out) :: row_step(*)
      real(c_double), intent(out) :: row_t(*), row_dt(*), fields(*)
    end function
    integer(c_int) function tf_set_replay(ctx, nchecks, keep) bind(C, name='tf_set_replay')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nchecks
      integer(c_int), intent(in) :: keep(*)
    end function
    integer(c_int) function tf_set_slices(ctx, nslice, times) bind(C, name='tf_set_slices')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nslice
      real(c_double), intent(in) :: times(*)
    end function
    integer(c_int) function tf_get_slice(ctx, i, fields9, step, t) bind(C, name='tf_get_slice')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: i
      real(c_double), intent(out) :: fields9(*)
      integer(c_int), intent(out) :: step
      real(c_double), intent(out) :: t
    end function
    integer(c_int) function tf_get_timers(ctx, t8) bind(C, name='tf_get_timers')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: t8(8)
    end function
  end interface

contains

  function tf_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    p = tf_last_error(ctx)
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(n) :: msg)
    msg = transfer(s(1:n), msg)
  end function tf_error_message
end module temfetd
