! Fortran driver of the MI355X explicit update.
!
! Reads the reference's input.in (the same six NAMELIST groups, variable
! names and switch codes, flow_mod_general.f90:51-180, read in the order of
! read_par_namelist, :313-332), builds one of the canned problems
! (benchmark 1 isentropic vortex, 2 Sedov blast, 3 Rayleigh-Taylor,
! 4 Richtmyer-Meshkov, 5 double Mach reflection, 6 shear layer,
! 9 shock/vortex), advances it on the GPU through the C ABI and writes the
! reference's outputs: per-step stdout, 2d_<step>.dat2d binary dumps,
! 2d_<step>.dat Tecplot files and CpuTime.dat (general_main.f90:1-63).
!
! Addition over the reference: START_STEP in TIME_STEP_PARS restarts from
! 2d_<START_STEP>.dat2d (the reference's loader, read_WriteAndLoad.f90:30-70,
! is unreachable because START_STEP is in no NAMELIST).
module driver_state
  use dense_capi
  implicit none

  integer :: omp_thread_num = 4
  namelist /omp_pars/ omp_thread_num

  integer :: neq = 4
  double precision :: mach_inf = 1.d0, reynolds = 500.d0
  logical :: if_source = .false., if_bodyforce = .false., if_viscous = .false.
  integer :: benchmark_index = 0
  namelist /model_pars/ neq, if_source, if_viscous, if_bodyforce, mach_inf, reynolds, benchmark_index

  logical :: if_curve_grid = .false., if_read_from_file = .false.
  integer :: single_ni = 64, single_nj = 64
  namelist /grid_pars/ if_curve_grid, if_read_from_file, single_ni, single_nj

  integer :: stages = 3
  double precision :: total_time = 0.d0, current_time = 0.d0, elasp_time = 0.d0
  double precision :: dt = 0.01d0, cfl = 0.5d0
  logical :: if_const_dt = .false., if_save_tec = .true., if_save_binary = .true.
  integer :: total_step = 1, start_step = 0, current_step = 0, save_step = 100
  double precision :: dt_save = 0.01d0
  namelist /time_step_pars/ stages, total_time, dt_save, dt, cfl, total_step, save_step, if_const_dt, &
                            if_save_binary, if_save_tec, start_step

  integer :: scheme_splitting_index = 101, scheme_inv_index = 201, scheme_vis_index = 301
  logical :: if_save_theta = .false.
  double precision :: delta = 0.5d0
  namelist /scheme_pars/ scheme_splitting_index, scheme_inv_index, scheme_vis_index, if_save_theta, delta

  double precision :: gamma = 1.4d0
  namelist /gas_pars/ gamma

  integer, parameter :: bf = 4
  double precision, parameter :: pie = 3.141592653589793238462643383279502884d0
  integer :: ni, nj
  double precision :: dx, dy
  double precision, allocatable, target :: cons(:, :, :), prim(:, :, :)
  double precision, allocatable :: xc(:, :), yc(:, :), vet_x(:, :), vet_y(:, :)

contains

  ! Whether to pick the stage kernel's tile shape by timing before the run
  ! (dense_autotune: every shape gives the same bits).  Worth its 160 trial
  ! steps on long runs only; DENSE_AUTOTUNE=1 or 0 forces it on or off.
  logical function want_autotune(steps_left)
    integer, intent(in) :: steps_left
    character(len=8) :: v
    integer :: n, st
    call get_environment_variable('DENSE_AUTOTUNE', v, n, st)
    if (st == 0 .and. n > 0) then
      want_autotune = v(1:1) == '1'
    else
      want_autotune = steps_left >= 2000
    end if
  end function want_autotune

  subroutine read_input()
    logical :: there
    inquire(file='input.in', exist=there)
    if (.not. there) then
      print *, 'Input file input.in does not exist!'
      open(10, file='input.in.example')
      write(10, nml=omp_pars)
      write(10, nml=model_pars)
      write(10, nml=grid_pars)
      write(10, nml=time_step_pars)
      write(10, nml=scheme_pars)
      write(10, nml=gas_pars)
      close(10)
      stop 1
    end if
    open(10, file='input.in')
    read(10, nml=omp_pars)
    read(10, nml=model_pars)
    read(10, nml=grid_pars)
    read(10, nml=time_step_pars)
    read(10, nml=scheme_pars)
    read(10, nml=gas_pars)
    close(10)
  end subroutine read_input

  ! Conserved variables of a primitive state (the expression the reference's
  ! initial conditions use for cons(:,:,4)).
  subroutine put_state(i, j, rho, u, v, p)
    integer, intent(in) :: i, j
    double precision, intent(in) :: rho, u, v, p
    cons(i + bf, j + bf, 1) = rho
    cons(i + bf, j + bf, 2) = rho * u
    cons(i + bf, j + bf, 3) = rho * v
    cons(i + bf, j + bf, 4) = p / (gamma - 1.d0) + 0.5 * rho * (u**2 + v**2)
  end subroutine put_state

  ! Grid and initial state of the canned problems (init_general.f90:13-37).
  subroutine init_problem()
    integer :: i, j
    double precision :: x, y, r, rho, u, v, p, c, tau, b
    double precision :: x_c, y_c, epsl, r_c, alp
    allocate(cons(ni + 2 * bf, nj + 2 * bf, 4), prim(ni + 2 * bf, nj + 2 * bf, 6))
    allocate(xc(ni, nj), yc(ni, nj), vet_x(ni + 1, nj + 1), vet_y(ni + 1, nj + 1))
    cons = 0.d0
    prim = 0.d0
    select case (benchmark_index)
    case (1)  ! isentropic vortex: init_IVC_general.f90:13-56
      b = 0.5d0
      dx = 10.d0 / ni
      dy = 10.d0 / nj
      do j = 1, nj
        do i = 1, ni
          xc(i, j) = (real(i) - 0.5) * dx
          yc(i, j) = (real(j) - 0.5) * dy
        end do
      end do
      do j = 1, nj + 1
        do i = 1, ni + 1
          vet_x(i, j) = (real(i) - 1.0) * dx
          vet_y(i, j) = (real(j) - 1.0) * dy
        end do
      end do
      do j = 1, nj
        do i = 1, ni
          r = sqrt((xc(i, j) - 5.d0)**2 + (yc(i, j) - 5.d0)**2)
          rho = (1.d0 - (gamma - 1.d0) * b**2 * exp(1 - r**2) / (8.d0 * gamma * pie**2))**2.5d0
          u = 0.0 - b / (2 * pie) * exp((1 - r**2) / 2) * (yc(i, j) - 5.d0)
          v = b / (2 * pie) * exp((1 - r**2) / 2) * (xc(i, j) - 5.d0)
          p = rho**gamma
          call put_state(i, j, rho, u, v, p)
        end do
      end do
    case (3)  ! Rayleigh-Taylor: init_RT_general.f90:9-65
      dx = 0.25d0 / (ni - 1.d0)
      dy = 1.d0 / (nj - 1.d0)
      do j = 1, nj
        do i = 1, ni
          xc(i, j) = (i - 1.d0) * dx
          yc(i, j) = (j - 1.d0) * dy
        end do
      end do
      do j = 1, nj + 1
        do i = 1, ni + 1
          vet_x(i, j) = (real(i) - 1.5) * dx
          vet_y(i, j) = (real(j) - 1.5) * dy
        end do
      end do
      do j = 1, nj
        do i = 1, ni
          u = 0.d0
          if (yc(i, j) <= 0.5) then
            rho = 2.d0
            p = 2.d0 * yc(i, j) + 1.d0
            c = sqrt(gamma * p / rho)
            v = -0.025d0 * c * cos(8.d0 * pie * xc(i, j))
          else
            rho = 1.d0
            p = yc(i, j) + 1.5d0
            c = sqrt(gamma * p / rho)
            v = -0.025 * c * cos(8.d0 * pie * xc(i, j))
          end if
          call put_state(i, j, rho, u, v, p)
        end do
      end do
    case (5)  ! double Mach reflection: init_doublemach_general.f90:13-59
      dx = 4.d0 / ni
      dy = 1.d0 / nj
      do j = 1, nj
        do i = 1, ni
          xc(i, j) = (i - 0.5d0) * dx
          yc(i, j) = (j - 0.5d0) * dy
        end do
      end do
      do j = 1, nj + 1
        do i = 1, ni + 1
          vet_x(i, j) = (i - 1.0d0) * dx
          vet_y(i, j) = (j - 1.0d0) * dy
        end do
      end do
      do j = 1, nj
        do i = 1, ni
          x = xc(i, j)
          y = yc(i, j)
          if (x < (1.d0 / 6.d0 + y * sqrt(1.d0 / 3.d0))) then
            call put_state(i, j, 8.d0, 8.25 * sqrt(3.d0) * 0.5, -8.25 * 0.5d0, 116.5d0)
          else
            call put_state(i, j, 1.4d0, 0.d0, 0.d0, 1.d0)
          end if
        end do
      end do
    case (9)  ! shock/vortex interaction: init_shockvortex_general.f90:1-64
      x_c = 0.25
      y_c = 0.5
      epsl = 0.3
      r_c = 0.05
      alp = 0.204
      dx = 2.d0 / ni
      dy = 1.d0 / nj
      do j = 1, nj
        do i = 1, ni
          xc(i, j) = (real(i) - 0.5) * dx
          yc(i, j) = (real(j) - 0.5) * dy
        end do
      end do
      do j = 1, nj + 1
        do i = 1, ni + 1
          vet_x(i, j) = (real(i) - 1.0) * dx
          vet_y(i, j) = (real(j) - 1.0) * dy
        end do
      end do
      do j = 1, nj
        do i = 1, ni
          x = xc(i, j)
          y = yc(i, j)
          r = sqrt((x - x_c)**2 + (y - y_c)**2)
          tau = r / r_c
          if (x < 0.5) then
            rho = (1.d0 - (gamma - 1.d0) * epsl**2 * exp(2 * alp * (1 - tau**2)) / (4 * alp * gamma))**(1.0 / (gamma - 1))
            u = 1.1 * sqrt(gamma) + epsl * tau * exp(alp * (1.0 - tau**2)) * (y - y_c) / r
            v = -epsl * tau * exp(alp * (1.0 - tau**2)) * (x - x_c) / r
            p = rho**gamma
          else
            rho = 1.169082d0
            u = 1.113299d0
            v = 0.0d0
            p = 1.245000d0
          end if
          call put_state(i, j, rho, u, v, p)
        end do
      end do
    case (2)  ! Sedov blast: init_sedov_general.f90:13-67
      dx = 1.2d0 / ni
      dy = 1.2d0 / nj
      call cell_centres(0.5d0)
      p = e_blast()
      do j = 1, nj
        do i = 1, ni
          call put_state(i, j, 1.d0, 0.d0, 0.d0, dble(1.0e-6))
        end do
      end do
      call put_state(1, 1, 1.d0, 0.d0, 0.d0, p)
    case (4)  ! Richtmyer-Meshkov: init_RM_general.f90:12-64
      dx = 4.d0 / ni
      dy = 1.d0 / nj
      call cell_centres(0.5d0)
      do j = 1, nj
        do i = 1, ni
          x = xc(i, j)
          y = yc(i, j)
          if (x < dble(0.05)) then
            call put_state(i, j, dble(1.376364), dble(0.394729), 0.d0, dble(1.569800))
          else if (x < (1.d0 - dble(0.05) * sin(2 * pie * (y - 0.25d0))) .and. x > dble(0.05)) then
            call put_state(i, j, 1.d0, 0.d0, 0.d0, 1.d0)
          else
            call put_state(i, j, dble(0.1358), 0.d0, 0.d0, 1.d0)
          end if
        end do
      end do
    case (6)  ! shear layer: init_ssl_general.f90:12-81
      dx = 200.d0 / ni
      dy = 40.d0 / nj
      do j = 1, nj
        do i = 1, ni
          xc(i, j) = (i - 0.5d0) * dx
          yc(i, j) = (j - 0.5d0) * dy - 40.d0 / 2.0
        end do
      end do
      do j = 1, nj + 1
        do i = 1, ni + 1
          vet_x(i, j) = (i - 1.0d0) * dx
          vet_y(i, j) = (j - 1.0d0) * dy - 40.d0 / 2.0
        end do
      end do
      do j = 1, nj
        do i = 1, ni
          if (yc(i, j) >= 0) then
            call put_state(i, j, 1.6374d0, 3.d0, 0.d0, 0.3327d0)
          else
            call put_state(i, j, 0.3626d0, 2.d0, 0.d0, 0.3327d0)
          end if
        end do
      end do
      ! the top row in the state of the top boundary (whose ghost rows the library holds fixed)
      do i = 1, ni
        call put_state(i, nj, 2.1101d0, 2.9709d0, -0.1367d0, dble(0.4754))
      end do
    case default
      print *, 'No such Benchmark index!'
      stop 1
    end select
  contains
    ! cell centres (i - h)*dx, (j - h)*dy and vertices (i - 1)*dx, (j - 1)*dy
    subroutine cell_centres(h)
      double precision, intent(in) :: h
      integer :: ii, jj
      do jj = 1, nj
        do ii = 1, ni
          xc(ii, jj) = (ii - h) * dx
          yc(ii, jj) = (jj - h) * dy
        end do
      end do
      do jj = 1, nj + 1
        do ii = 1, ni + 1
          vet_x(ii, jj) = (ii - 1.d0) * dx
          vet_y(ii, jj) = (jj - 1.d0) * dy
        end do
      end do
    end subroutine cell_centres
    ! p_0 = e (gamma - 1) with e = 0.244816 / (dx dy), init_sedov_general.f90:36-37
    double precision function e_blast()
      e_blast = 0.244816d0 / (dx * dy) * (gamma - 1.d0)
    end function e_blast
  end subroutine init_problem

  ! Switches and metrics handed to the library.
  subroutine fill_cfg(cfg)
    type(dense_cfg), intent(inout) :: cfg
    integer(c_int) :: rc
    rc = dense_problem_defaults(int(benchmark_index, c_int), int(ni, c_int), int(nj, c_int), cfg)
    call dense_check(rc, c_null_ptr, 'dense_problem_defaults')
    cfg%dx = dx
    cfg%dy = dy
    cfg%gamma = gamma
    cfg%cfl = cfl
    cfg%delta = delta
    cfg%dt = dt
    cfg%splitting = scheme_splitting_index
    cfg%recon = scheme_inv_index
    cfg%stages = stages
    cfg%if_const_dt = merge(1, 0, if_const_dt)
    ! only the Rayleigh-Taylor boundary routine refreshes a non-zero body force
    ! (boundary_rt_general.f90:154-155); elsewhere it adds dt*0
    cfg%if_bodyforce = merge(1, 0, if_bodyforce .and. benchmark_index == 3)
  end subroutine fill_cfg

  ! write_date_binary, read_WriteAndLoad.f90:1-28
  subroutine write_binary()
    character(len=10) :: temp
    write(temp, '(i8)') current_step
    open(10, file='2d_' // trim(adjustl(temp)) // '.dat2d', form='unformatted')
    write(10) current_step, current_time
    write(10) 1
    write(10) 0
    write(10) ni, nj
    write(10) bf, bf
    write(10) cons
    write(10) xc
    write(10) yc
    write(10) dx, dy
    close(10)
  end subroutine write_binary

  ! load_date_binary, read_WriteAndLoad.f90:30-70 (driver restart)
  subroutine load_binary()
    character(len=10) :: temp
    integer :: nb, idx, bfi, bfj
    write(temp, '(i8)') start_step
    open(10, file='2d_' // trim(adjustl(temp)) // '.dat2d', form='unformatted', status='old')
    read(10) current_step, current_time
    read(10) nb
    read(10) idx
    read(10) ni, nj
    read(10) bfi, bfj
    close(10)
    ! grid arrays of the problem, then the stored state
    call init_problem()
    open(10, file='2d_' // trim(adjustl(temp)) // '.dat2d', form='unformatted', status='old')
    read(10) current_step, current_time
    read(10) nb
    read(10) idx
    read(10) ni, nj
    read(10) bfi, bfj
    read(10) cons
    close(10)
  end subroutine load_binary

  ! write_tec_at_time_cartisian, output_test_general.f90:16-58
  subroutine write_tec()
    character(len=10) :: temp
    integer :: i, j, k
    write(temp, '(i8)') current_step
    open(10, file='2d_' // trim(adjustl(temp)) // '.dat', form='formatted')
    write(10, *) 'title=', '"flow time ', current_time, '"'
    write(10, *) 'variables="x","y","rho","u","v","p","alpha"'
    write(10, *) 'ZONE I=', ni + 1, ' J=', nj + 1, 'DATAPACKING=BLOCK, VARLOCATION=([3-7]=CELLCENTERED)'
    do j = 1, nj + 1
      do i = 1, ni + 1
        write(10, *) vet_x(i, j)
      end do
    end do
    do j = 1, nj + 1
      do i = 1, ni + 1
        write(10, *) vet_y(i, j)
      end do
    end do
    do k = 1, 6
      do j = bf + 1, nj + bf
        do i = bf + 1, ni + bf
          write(10, *) prim(i, j, k)
        end do
      end do
    end do
    close(10)
  end subroutine write_tec

  ! 2d_theta_<step>.dat (output_test_general.f90:61-119): per cell the larger of
  ! AdaWENO's indicator at its two faces, for each sweep and sign.  th holds
  ! theta_xip(ni+1,nj), theta_xin, theta_etap(nj+1,ni), theta_etan.
  subroutine write_theta(th)
    double precision, intent(in) :: th(*)
    character(len=10) :: temp
    integer :: i, j, k, nx, ny, o
    nx = (ni + 1) * nj
    ny = (nj + 1) * ni
    write(temp, '(i8)') current_step
    open(10, file='2d_theta_' // trim(adjustl(temp)) // '.dat', form='formatted')
    write(10, *) 'title=', '"flow time ', current_time, '"'
    write(10, *) 'variables="x","y","xp","xn","ep","en"'
    write(10, *) 'ZONE I=', ni + 1, ' J=', nj + 1, 'DATAPACKING=BLOCK, VARLOCATION=([3-6]=CELLCENTERED)'
    do j = 1, nj + 1
      do i = 1, ni + 1
        write(10, *) vet_x(i, j)
      end do
    end do
    do j = 1, nj + 1
      do i = 1, ni + 1
        write(10, *) vet_y(i, j)
      end do
    end do
    do k = 0, 1
      o = k * nx
      do j = 1, nj
        do i = 1, ni
          write(10, *) max(th(o + i + (j - 1) * (ni + 1)), th(o + i + 1 + (j - 1) * (ni + 1)))
        end do
      end do
    end do
    do k = 0, 1
      o = 2 * nx + k * ny
      do j = 1, nj
        do i = 1, ni
          write(10, *) max(th(o + j + (i - 1) * (nj + 1)), th(o + j + 1 + (i - 1) * (nj + 1)))
        end do
      end do
    end do
    close(10)
  end subroutine write_theta

end module driver_state


program dense_gpu
  use iso_c_binding
  use dense_capi
  use driver_state
  implicit none
  type(dense_cfg) :: cfg
  type(c_ptr) :: h
  integer(c_int) :: rc
  integer(8) :: c0, c1, rate
  double precision :: cpu_t, chunk_t, t_dev, el_dev
  integer(c_int) :: nrun, step_dev, taken
  integer :: k
  ! at most this many steps per device call (the dt buffer's length)
  integer, parameter :: max_chunk = 4096
  double precision, allocatable :: dts(:)
  integer(c_int) :: tile_choice
  real(c_double) :: trial_ms
  double precision, allocatable :: theta_buf(:)
  integer :: theta_step = -1

  call read_input()
  ni = single_ni
  nj = single_nj
  if (start_step == 0) then
    print *, 'Start with initial conditon...'
    if (benchmark_index == 3) then  ! init_RT_general.f90:9-10
      if_bodyforce = .true.
      gamma = 5.d0 / 3.d0
    end if
    call init_problem()
  else
    print *, 'Start with data from file...'
    call load_binary()
  end if

  allocate(theta_buf(2 * ((ni + 1) * nj + (nj + 1) * ni)))
  call fill_cfg(cfg)
  rc = dense_create(cfg, h)
  call dense_check(rc, c_null_ptr, 'dense_create')
  rc = dense_set_state(h, cons)
  call dense_check(rc, h, 'dense_set_state')
  write(*, *) 'Device: MI355X (HIP), FP64'
  if (want_autotune(total_step - current_step)) then
    rc = dense_autotune(h, 16, tile_choice, trial_ms)
    call dense_check(rc, h, 'dense_autotune')
    write(*, *) 'Tile shape chosen by timing (cfg%tile):', tile_choice
  end if
  open(20, file='CpuTime.dat', access='append')

  cpu_t = 0.d0
  call system_clock(count_rate=rate)
  allocate(dts(max_chunk))
  do while ((current_time < total_time) .and. (current_step < total_step))
    ! The steps up to the next save are replayed on the device in one call
    ! (captured graphs, the clock and the save clock on the device); the
    ! reference's per-step lines are printed afterwards from the step sizes.
    nrun = min(save_step - mod(current_step, save_step), max_chunk)
    t_dev = current_time
    step_dev = current_step
    el_dev = elasp_time
    call system_clock(c0)
    rc = dense_run_saving(h, nrun, total_time, total_step, dt_save, el_dev, t_dev, step_dev, dts, taken)
    call dense_check(rc, h, 'dense_run_saving')
    call system_clock(c1)
    chunk_t = dble(c1 - c0) / dble(rate)
    cpu_t = cpu_t + chunk_t
    do k = 1, taken
      dt = dts(k)
      write(*, *) '=========', 'CPU time:', chunk_t / taken, 'Seconds', '=========='
      current_step = current_step + 1
      current_time = current_time + dt
      elasp_time = elasp_time + dt
      write(*, *) current_step, 'Flow time= ', current_time, 'dt=', dt
    end do
    ! the device advanced the same clocks with the same additions
    if (taken == 0 .or. current_step /= step_dev .or. current_time /= t_dev .or. elasp_time /= el_dev) then
      write(*, *) 'device clock differs from the replayed step sizes'
      stop 1
    end if

    if ((mod(current_step, save_step) == 0) .or. (elasp_time >= dt_save)) then
      print *, 'Saving binary file at time step = ', current_step
      rc = dense_get_state(h, cons, c_loc(prim))
      call dense_check(rc, h, 'dense_get_state')
      if (if_save_binary) then
        ! the dump holds the ghost layers of the last boundary pass
        rc = dense_get_dump(h, cons)
        call dense_check(rc, h, 'dense_get_dump')
        call write_binary()
      end if
      if (if_save_tec) then
        print *, 'Saving TEC data file at time = ', current_time
        call write_tec()
        if (if_save_theta) call save_theta()
        elasp_time = 0.d0
      end if
      print *, 'Saving finished!'
      ! general_main.f90:53: the NAMELIST is re-read at every save
      call read_input()
      call fill_cfg(cfg)
      rc = dense_update_params(h, cfg)
      call dense_check(rc, h, 'dense_update_params')
    end if
  end do
  write(20, *) 'Total time and steps:', cpu_t, current_step
  if (current_step > 0) write(20, *) 'averaged time:', cpu_t / current_step
  close(20)
  print *, 'Writing final data... '
  rc = dense_get_state(h, cons, c_loc(prim))
  call dense_check(rc, h, 'dense_get_state')
  call write_tec()
  if (if_save_theta) call save_theta()
  rc = dense_destroy(h)
  print *, 'Calculation is completed! '
contains

  ! the reference writes the theta file beside every Tecplot file when
  ! IF_SAVE_THETA is set (output_test_general.f90:10-12); its final write after
  ! the loop repeats the file of a save at the same step, which stays as written
  ! (the library's indicator state restarts with the NAMELIST re-read)
  subroutine save_theta()
    integer(c_int) :: r
    if (current_step == theta_step) return
    r = dense_get_theta(h, theta_buf)
    if (r == DENSE_OK) then
      call write_theta(theta_buf)
      theta_step = current_step
    else
      call dense_warn(r, h, 'theta file')
    end if
  end subroutine save_theta
end program dense_gpu
