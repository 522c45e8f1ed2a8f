! iso_c_binding interface to lib2dense_amd.so (include/dense_capi.h).
! A Fortran caller replaces the reference's per-stage hot-path calls
! (general_main.f90:26-31) with one dense_step() per time step, or replays
! the steps between two saves with one dense_run_saving().
module dense_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: DENSE_OK = 0
  integer(c_int), parameter :: DENSE_BC_PERIODIC = 0, DENSE_BC_EXTRAP = 1, DENSE_BC_REFLECT = 2, &
                               DENSE_BC_DIRICHLET = 3, DENSE_BC_DMR_WALL = 4, DENSE_BC_DMR_TOP = 5

  ! struct dense_cfg.  C bc_state[side][var] is Fortran bc_state(var, side).
  type, bind(C) :: dense_cfg
    integer(c_int) :: ni = 0, nj = 0, nghost = 4, neq = 4
    real(c_double) :: dx = 0.d0, dy = 0.d0, gamma = 1.4d0, cfl = 0.5d0, delta = 0.5d0, dt = 0.01d0
    integer(c_int) :: splitting = 101, recon = 255, stages = 3, if_const_dt = 0, if_bodyforce = 0, benchmark = 0
    integer(c_int) :: bc(4) = 0
    real(c_double) :: bc_state(4, 4) = 0.d0
    integer(c_int) :: tile = 0
  end type dense_cfg

  interface
    integer(c_int) function dense_problem_defaults(benchmark, ni, nj, cfg) bind(C, name="dense_problem_defaults")
      import :: c_int, dense_cfg
      integer(c_int), value :: benchmark, ni, nj
      type(dense_cfg), intent(inout) :: cfg
    end function

    integer(c_int) function dense_create(cfg, h) bind(C, name="dense_create")
      import :: c_int, c_ptr, dense_cfg
      type(dense_cfg), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function dense_destroy(h) bind(C, name="dense_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function dense_set_state(h, cons) bind(C, name="dense_set_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: cons(*)
    end function

    integer(c_int) function dense_get_state(h, cons, prim) bind(C, name="dense_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: cons(*)
      type(c_ptr), value :: prim
    end function

    integer(c_int) function dense_get_dump(h, cons) bind(C, name="dense_get_dump")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: cons(*)
    end function

    integer(c_int) function dense_run_saving(h, max_steps, total_time, total_step, dt_save, elasp, t, step, &
                                             dt_hist, steps_taken) bind(C, name="dense_run_saving")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: max_steps, total_step
      real(c_double), value :: total_time, dt_save
      real(c_double), intent(inout) :: elasp, t
      integer(c_int), intent(inout) :: step
      real(c_double), intent(out) :: dt_hist(*)
      integer(c_int), intent(out) :: steps_taken
    end function

    integer(c_int) function dense_step(h, current_time, dt) bind(C, name="dense_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: current_time
      real(c_double), intent(out) :: dt
    end function

    integer(c_int) function dense_update_params(h, cfg) bind(C, name="dense_update_params")
      import :: c_int, c_ptr, dense_cfg
      type(c_ptr), value :: h
      type(dense_cfg), intent(in) :: cfg
    end function

    integer(c_int) function dense_autotune(h, trial_steps, tile, best_ms) bind(C, name="dense_autotune")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: trial_steps
      integer(c_int), intent(out) :: tile
      real(c_double), intent(out) :: best_ms
    end function

    integer(c_int) function dense_get_theta(h, theta) bind(C, name="dense_get_theta")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: theta(*)
    end function

    type(c_ptr) function dense_last_error(h) bind(C, name="dense_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! Stop with the library's message when a call fails (the reference stops on
  ! bad switch codes, flux_splitting.f90:30-32).
  subroutine dense_check(rc, h, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    if (rc == DENSE_OK) return
    p = dense_last_error(h)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    write(*, '(a,a,i0,a,*(a))') what, ' failed (', rc, '): ', chars
    stop 1
  end subroutine dense_check

  ! Print the library's message for a failed call and carry on.
  subroutine dense_warn(rc, h, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    if (rc == DENSE_OK) return
    p = dense_last_error(h)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    write(*, '(a,a,i0,a,*(a))') what, ' failed (', rc, '): ', chars
  end subroutine dense_warn

end module dense_capi
