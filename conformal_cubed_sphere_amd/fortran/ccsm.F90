! ccsm.F90 -- driver of the MI355X shallow-water dynamical core.
!
! Drop-in for the reference driver (dynamic_core/src/ccsm.F90:2-73): same
! namelist.input groups (&time_settings, &case_select, &domain,
! &dynamic_opt of parameters_mod.F90:87-100, same defaults :118-129), same
! initialisation order (initParameters, read_mesh, initStat/initTend,
! initTestCase), the same step loop with history cadence
! (ccsm.F90:50-67) and the same "output index/total, MCR, ECR" line.  The
! time step itself runs on the GPU through libccs.so (ccs_gpu_mod).
!
! &dynamic_opt reconstruct_scheme (declared but unused by the reference, which
! swaps spatial_operators_mod sources in its build instead) selects the
! operator: 'WENO' (spatial_operators_mod.F90, the default), '2nd', '4th',
! '6th' (spatial_operators_mod_2nd/4th/6th.F90) or 'WENO_ext'
! (spatial_operators_mod_WENO.F90, with mesh%matrixIG_ext from the mesh file's
! jab or the generated map).  'WENO_ext' runs with precision = 'fp64' only.
!
! Differences, all in the optional group &gpu_opt that follows &dynamic_opt:
!   nx         cells per panel edge.  When &domain mesh_file names an existing
!              file of the generator's schema (classic or 64-bit-offset
!              netCDF, e.g. written by ccs_mesh_file_write) the grid is read
!              from it as io_mod.F90:37-135 does and nx may be left unset;
!              otherwise the conformal mesh of nx cells is generated in memory
!              (ccs_mesh_geometry).
!   precision  'fp64' (bit-faithful to the reference), 'fp32', 'fp64_fast'.
!   cfl        > 0: dt = largest step <= cfl * min(dx / max Rusanov speed)
!              that divides the run, from the initial state; when the
!              history interval divides the run, the step count is rounded
!              up to a multiple of the number of outputs.
!   seed, n_bumps, mountain, u0, h_mean: case 7, the seeded balanced jet.
!   history_file  history output of io_mod.F90:309-541 written at the
!              initial state and every output (default 'ccsm_output.nc', the
!              reference's name; '' for none).
!   radius, gravity, omega, device, use_graph.
!   final_state_file    after the run, the state as raw stream records
!              (n, halo, steps, dt; phiG, u, v, phi padded), a checkpoint.
!   initial_state_file  start from such a checkpoint instead of the test
!              case (same n and halo; the companions are derived from phiG,
!              u, v as every step after the first does, so on the FP64 path
!              a run resumed from a checkpoint continues bit for bit).
!   run_record_file     after the run, one JSON line: grid, precision,
!              scheme, steps, step-loop seconds, cell-updates/s, and the
!              relative mass and energy change (MCR, ECR) of the final state;
!              the same line is printed after the step-loop summary.
module ccs_params_mod
  use iso_c_binding
  implicit none
  ! &time_settings
  real(8) :: dt
  integer :: run_days, run_hours, run_minutes, run_seconds, history_interval
  character(len=200) :: integral_scheme
  ! &case_select
  integer :: case_num
  ! &domain
  character(len=200) :: mesh_file
  integer :: xhalo, yhalo
  ! &dynamic_opt
  character(len=200) :: reconstruct_scheme = 'WENO'
  ! &gpu_opt
  integer :: nx = 0
  character(len=16) :: precision = 'fp64'
  real(8) :: cfl = 0.d0
  integer :: seed = 1, n_bumps = 8, mountain = 0
  real(8) :: u0 = 20.d0, h_mean = 10000.d0
  real(8) :: radius = 6371229.d0, gravity = 9.80616d0, omega = 7.292d-5
  integer :: device = 0, use_graph = 1
  character(len=200) :: final_state_file = ''
  character(len=200) :: initial_state_file = ''
  character(len=200) :: run_record_file = ''
  character(len=200) :: history_file = 'ccsm_output.nc'
  logical :: mesh_from_file = .false.

  integer :: total_run_time, nsteps
  integer :: steps_per_output = 0

  namelist /time_settings/ dt, run_days, run_hours, run_minutes, run_seconds, history_interval, integral_scheme
  namelist /case_select/ case_num
  namelist /domain/ mesh_file, xhalo, yhalo
  namelist /dynamic_opt/ reconstruct_scheme
  namelist /gpu_opt/ nx, precision, cfl, seed, n_bumps, mountain, u0, h_mean, radius, gravity, omega, device, &
                     use_graph, final_state_file, history_file, initial_state_file, run_record_file

contains

  subroutine initParameters
    integer :: ios
    dt = 300.d0
    xhalo = 1
    yhalo = 1
    run_days = 1
    run_hours = 0
    run_minutes = 0
    run_seconds = 0
    history_interval = 360
    integral_scheme = 'RK4'
    case_num = 2
    mesh_file = 'ccs_mesh'
    open(1, file='namelist.input', status='old')
    read(1, nml=time_settings)
    read(1, nml=case_select)
    read(1, nml=domain)
    read(1, nml=dynamic_opt)
    read(1, nml=gpu_opt, iostat=ios)
    close(1)
    total_run_time = run_days * 86400 + run_hours * 3600 + run_minutes * 60 + run_seconds
    if (trim(adjustl(integral_scheme)) /= 'RK3_TVD' .and. trim(adjustl(integral_scheme)) /= 'RK4') then
      error stop 'Unknown integral scheme, please select from RK3_TVD or RK4 ...'
    end if
    select case (trim(adjustl(reconstruct_scheme)))
    case ('WENO', '2nd', '4th', '6th', 'WENO_ext')
    case default
      error stop 'reconstruct_scheme must be WENO, 2nd, 4th, 6th or WENO_ext'
    end select
    if (xhalo /= yhalo) error stop 'xhalo and yhalo must be equal'
    if (xhalo < 3) error stop 'xhalo >= 3 is required (WENO-5 faces and the 6th-order vorticity)'
    inquire(file=trim(mesh_file), exist=mesh_from_file)
    if (.not. mesh_from_file .and. (nx < 4 .or. mod(nx, 2) /= 0)) &
      error stop 'no mesh file: set nx (even, >= 4) in &gpu_opt to generate the mesh'
    if (history_interval <= 0) error stop 'history_interval must be positive'
    nsteps = int(total_run_time / dt)
  end subroutine initParameters

end module ccs_params_mod

program ccsm
  use iso_c_binding
  use ccs_gpu_mod
  use ccs_params_mod
  implicit none

  integer :: n, h, pw, np, ni, it, nxt, k, output_idx, total_output_num
  integer(8) :: t_start, t_end, t_loop0, t_loop1, rate
  real(8), allocatable, target :: lon_deg(:), lat_deg(:), xyz(:), jab(:)
  real(8), allocatable, target :: lon(:), lat(:), sinlon(:), coslon(:), sinlat(:), coslat(:)
  real(8), allocatable, target :: sqrtG(:), f(:), area(:), G(:), IG(:), A(:), IA(:)
  real(8), allocatable :: phiG(:), u(:), v(:), zonal(:), merid(:), phis(:)
  real(8), allocatable, target :: phi(:), uc(:), vc(:)
  integer :: rn, rh, rsteps
  real(8) :: rdt
  character(len=1024) :: record
  real(8), allocatable, target :: jab_ext(:), ig_ext(:)
  type(ccs_metrics) :: m
  type(ccs_case_params) :: cp
  type(ccs_config) :: cfg
  type(c_ptr) :: hnd, hist
  type(ccs_history_desc) :: hd
  real(8) :: mass0, energy0, mass, energy, dummy, loop_s, area_sum
  character(len=*), parameter :: fmt_out = '(1x,a,i8,a,i8,2es25.16)'

  call system_clock(t_start, rate)
  print *, 'Start the MCV Shallow Water model (MI355X, ', c_to_f_string(ccs_version()), ')'
  call initParameters

  ! ---- read_mesh (io_mod.F90:15-307): mesh file or generated conformal mesh,
  ! then the static tables
  if (mesh_from_file) then
    call ccs_check(ccs_mesh_file_read(trim(mesh_file) // c_null_char, n, c_null_ptr, c_null_ptr, c_null_ptr), &
                   'ccs_mesh_file_read ' // trim(mesh_file))
    if (nx /= 0 .and. nx /= n) error stop 'nx in &gpu_opt differs from the mesh file'
  else
    n = nx
  end if
  h = xhalo
  pw = n + 2 * h
  np = pw * pw * 6
  ni = n * n * 6
  allocate(lon_deg(ni), lat_deg(ni), xyz(3 * ni), jab(6 * ni))
  if (mesh_from_file) then
    call ccs_check(ccs_mesh_file_read(trim(mesh_file) // c_null_char, n, c_loc(lon_deg), c_loc(lat_deg), &
                                      c_loc(jab)), 'ccs_mesh_file_read ' // trim(mesh_file))
    print *, 'Mesh read from ', trim(mesh_file), ': C', n
  else
    call ccs_check(ccs_mesh_geometry(n, lon_deg, lat_deg, xyz, jab), 'ccs_mesh_geometry')
  end if
  allocate(lon(np), lat(np), sinlon(np), coslon(np), sinlat(np), coslat(np), sqrtG(np), f(np), area(np))
  allocate(G(4 * np), IG(4 * np), A(4 * np), IA(4 * np))
  m%n = n
  m%halo = h
  m%radius = radius
  m%omega = omega
  m%lon = c_loc(lon); m%lat = c_loc(lat); m%sinlon = c_loc(sinlon); m%coslon = c_loc(coslon)
  m%sinlat = c_loc(sinlat); m%coslat = c_loc(coslat); m%sqrtG = c_loc(sqrtG); m%f = c_loc(f)
  m%areaCell = c_loc(area); m%G = c_loc(G); m%IG = c_loc(IG); m%A = c_loc(A); m%IA = c_loc(IA)
  call ccs_check(ccs_mesh_metrics(lon_deg, lat_deg, jab, m), 'ccs_mesh_metrics')
  area_sum = interior_sum(area)
  print *, 'Spherical area, numerical, analytical :', area_sum, 4.d0 * acos(-1.d0) * radius**2

  ! ---- initTestCase (test_case_mod.F90:13-58)
  allocate(phiG(np), u(np), v(np), phi(np), uc(np), vc(np), zonal(np), merid(np), phis(np))
  call ccs_case_defaults(cp)
  cp%case_num = case_num
  cp%seed = int(seed, c_long_long)
  cp%n_bumps = n_bumps
  cp%mountain = mountain
  cp%gravity = gravity
  cp%u0 = u0
  cp%h_mean = h_mean
  call ccs_check(ccs_init_case(m, cp, phiG, u, v, phi, uc, vc, zonal, merid, phis), 'ccs_init_case')
  print *, ''
  print *, 'test case ', case_num, ' is selected, C', n, ' halo ', h
  if (cfl > 0.d0) then
    call cfl_dt(cfl, dble(total_run_time), dt, nsteps)
    ! keep the history cadence on whole steps: round the step count up to a
    ! multiple of the number of outputs and test outputs by step index
    if (mod(total_run_time, history_interval) == 0) then
      k = total_run_time / history_interval
      nsteps = ((nsteps + k - 1) / k) * k
      dt = dble(total_run_time) / nsteps
      steps_per_output = nsteps / k
    end if
    print *, 'dt from CFL ', cfl, ': ', dt, ' s, ', nsteps, ' steps'
  end if
  print *, ''
  write(*, '(1x,6a)') 'Temporal integration scheme is ', trim(adjustl(integral_scheme)), ', precision ', &
                      trim(precision), ', reconstruct_scheme ', trim(adjustl(reconstruct_scheme))
  print *, ''

  ! ---- GPU state (initStat/initTend)
  cfg%n = n
  cfg%halo = h
  select case (trim(precision))
  case ('fp64')
    cfg%precision = CCS_FP64_PARITY
  case ('fp32')
    cfg%precision = CCS_FP32_SPLIT
  case ('fp64_fast')
    cfg%precision = CCS_FP64_FAST
  case default
    error stop 'precision must be fp64, fp32 or fp64_fast'
  end select
  cfg%scheme = merge(CCS_RK3_TVD, CCS_RK4, trim(adjustl(integral_scheme)) == 'RK3_TVD')
  cfg%eflux = CCS_EFLUX_NONE
  cfg%device = device
  cfg%use_graph = use_graph
  select case (trim(adjustl(reconstruct_scheme)))
  case ('2nd')
    cfg%spatial_op = CCS_OP_CENTRED2
  case ('4th')
    cfg%spatial_op = CCS_OP_CENTRED4
  case ('6th')
    cfg%spatial_op = CCS_OP_CENTRED6
  case ('WENO_ext')
    cfg%spatial_op = CCS_OP_WENO_EXT
  case default
    cfg%spatial_op = CCS_OP_WENO
  end select
  cfg%dt = dt
  cfg%dx = m%dx
  cfg%radius = radius
  cfg%gravity = gravity
  cfg%omega = omega
  call ccs_check(ccs_create(cfg, hnd), 'ccs_create')
  call ccs_check(ccs_set_mesh(hnd, sqrtG, G, IG, A, IA, f, phis), 'ccs_set_mesh', hnd)
  if (cfg%spatial_op == CCS_OP_WENO_EXT) then
    ! mesh%matrixIG_ext (io_mod.F90:223-250) from the mesh file's jab when
    ! the grid came from a file, else from the generated conformal map
    allocate(ig_ext(4 * (2 * n + 1)**2 * 6))
    if (mesh_from_file) then
      allocate(jab_ext(36 * (2 * n + 1)**2))
      call ccs_check(ccs_mesh_file_read_ext(trim(mesh_file) // c_null_char, k, c_loc(jab_ext)), &
                     'ccs_mesh_file_read_ext ' // trim(mesh_file))
      call ccs_check(ccs_mesh_ext_metric(n, radius, c_loc(jab_ext), ig_ext), 'ccs_mesh_ext_metric')
    else
      call ccs_check(ccs_mesh_ext_metric(n, radius, c_null_ptr, ig_ext), 'ccs_mesh_ext_metric')
    end if
    call ccs_check(ccs_set_mesh_ext(hnd, ig_ext), 'ccs_set_mesh_ext', hnd)
  end if
  if (len_trim(initial_state_file) > 0) then
    ! resume from a checkpoint written by final_state_file
    open(3, file=trim(initial_state_file), access='stream', form='unformatted', status='old')
    read(3) rn, rh, rsteps, rdt
    if (rn /= n .or. rh /= h) error stop 'initial_state_file: its n and halo differ from this run'
    read(3) phiG, u, v, phi
    close(3)
    print *, 'State read from ', trim(initial_state_file), ' (written after ', rsteps, ' steps of dt ', rdt, ')'
    call ccs_check(ccs_set_state(hnd, phiG, u, v, c_null_ptr, c_null_ptr, c_null_ptr), 'ccs_set_state', hnd)
  else
    call ccs_check(ccs_set_state(hnd, phiG, u, v, c_loc(phi), c_loc(uc), c_loc(vc)), 'ccs_set_state', hnd)
  end if

  output_idx = 1
  total_output_num = int(nsteps * dt / history_interval)
  if (steps_per_output > 0) total_output_num = nsteps / steps_per_output
  hist = c_null_ptr
  if (len_trim(history_file) > 0) then
    ! history_init + history_write_stat(stat(old), 1) (ccsm.F90:42-43)
    hd%n = n
    hd%halo = h
    hd%case_num = case_num
    hd%pad_ = 0
    hd%dx_deg = 90.d0 / n
    hd%dt = dt
    call ccs_check(ccs_history_create(trim(history_file) // c_null_char, hd, lon, lat, area, phis, hist), &
                   'ccs_history_create ' // trim(history_file))
    call ccs_check(ccs_history_write(hist, 1, phi, phis, u, v, uc, vc, zonal, merid), 'ccs_history_write')
  end if
  call ccs_check(ccs_diagnostics(hnd, mass0, energy0), 'ccs_diagnostics', hnd)
  write(*, fmt_out) 'output index/total, MCR, ECR :', output_idx - 1, '/', total_output_num, 0.d0, 0.d0

  ! ---- time integration (ccsm.F90:50-67): the steps between two outputs go
  ! to the GPU in one call; at an output the energy is taken on the state
  ! before the last step, as the reference does with stat(old) (:62).
  call system_clock(t_loop0)
  it = 0
  do while (it < nsteps)
    nxt = nsteps
    do k = it + 1, nsteps
      if (is_output(k)) then
        nxt = k
        exit
      end if
    end do
    if (is_output(nxt)) then
      if (nxt - it > 1) call ccs_check(ccs_step(hnd, nxt - it - 1), 'ccs_step', hnd)
      call ccs_check(ccs_diagnostics(hnd, dummy, energy), 'ccs_diagnostics', hnd)
      call ccs_check(ccs_step(hnd, 1), 'ccs_step', hnd)
      call ccs_check(ccs_diagnostics(hnd, mass, dummy), 'ccs_diagnostics', hnd)
      output_idx = output_idx + 1
      if (c_associated(hist)) then
        call ccs_check(ccs_get_state(hnd, phiG, u, v, phi, uc, vc, zonal, merid), 'ccs_get_state', hnd)
        call ccs_check(ccs_history_write(hist, output_idx, phi, phis, u, v, uc, vc, zonal, merid), &
                       'ccs_history_write')
      end if
      write(*, fmt_out) 'output index/total, MCR, ECR :', output_idx - 1, '/', total_output_num, &
               (mass - mass0) / mass0, (energy - energy0) / energy0
    else
      call ccs_check(ccs_step(hnd, nxt - it), 'ccs_step', hnd)
    end if
    it = nxt
  end do
  call ccs_check(ccs_synchronize(hnd), 'ccs_synchronize', hnd)
  call system_clock(t_loop1)
  loop_s = dble(t_loop1 - t_loop0) / dble(rate)
  print *, 'Step loop: ', nsteps, ' steps in ', loop_s, ' s, cell-updates/s ', dble(ni) * nsteps / max(loop_s, 1d-9)
  ! run record: one JSON line (printed, and written to run_record_file)
  call ccs_check(ccs_diagnostics(hnd, mass, energy), 'ccs_diagnostics', hnd)
  write(record, '(a,i0,a,i0,a,a,a,a,a,a,a,i0,a,es24.16e3,a,i0,a,es24.16e3,a,es24.16e3,a,es24.16e3,a,es24.16e3,a)') &
    '{"n": ', n, ', "halo": ', h, ', "precision": "', trim(precision), '", "scheme": "', &
    trim(adjustl(integral_scheme)), '", "operator": "', trim(adjustl(reconstruct_scheme)), '", "case": ', case_num, &
    ', "dt": ', dt, ', "steps": ', nsteps, ', "loop_seconds": ', loop_s, ', "cell_updates_per_s": ', &
    dble(ni) * nsteps / max(loop_s, 1d-9), ', "mcr": ', (mass - mass0) / mass0, ', "ecr": ', &
    (energy - energy0) / energy0, '}'
  print '(a)', trim(record)
  if (len_trim(run_record_file) > 0) then
    open(4, file=trim(run_record_file), status='replace', action='write')
    write(4, '(a)') trim(record)
    close(4)
  end if

  if (len_trim(final_state_file) > 0) then
    call ccs_check(ccs_get_state(hnd, phiG, u, v, phi, uc, vc, zonal, merid), 'ccs_get_state', hnd)
    open(2, file=trim(final_state_file), access='stream', form='unformatted', status='replace')
    write(2) n, h, nsteps, dt
    write(2) phiG, u, v, phi
    close(2)
  end if
  call ccs_destroy(hnd)
  if (c_associated(hist)) call ccs_check(ccs_history_close(hist), 'ccs_history_close')

  call system_clock(t_end)
  print *, 'It took ', dble(t_end - t_start) / dble(rate), ' seconds to run this program'

contains

  logical function is_output(k)
    integer, intent(in) :: k
    if (steps_per_output > 0) then
      is_output = mod(k, steps_per_output) == 0
    else
      is_output = (mod(k * dt, dble(history_interval)) == 0.d0) .and. (k * dt >= history_interval)
    end if
  end function is_output

  real(8) function interior_sum(fld)
    real(8), intent(in) :: fld(:)
    integer :: i, j, p
    interior_sum = 0.d0
    do p = 1, 6
      do j = 1, n
        do i = 1, n
          interior_sum = interior_sum + fld((i + h) + pw * ((j + h - 1) + pw * (p - 1)))
        end do
      end do
    end do
  end function interior_sum

  ! Largest dt <= c * min(dx / max lambda) dividing the run, lambda the
  ! Rusanov face speed of spatial_operators_mod.F90:109-116,134-141.
  subroutine cfl_dt(c, total, dt_out, nst)
    real(8), intent(in) :: c, total
    real(8), intent(out) :: dt_out
    integer, intent(out) :: nst
    integer :: i, j, p, d0, d1
    real(8) :: lx, ly, ub, pb, gb
    lx = 0.d0
    ly = 0.d0
    do p = 1, 6
      do j = 1, n
        do i = 0, n
          d0 = (i + h) + pw * ((j + h - 1) + pw * (p - 1))
          d1 = d0 + 1
          ub = 0.5d0 * (uc(d1) + uc(d0))
          pb = 0.5d0 * (phi(d1) + phis(d1) + phi(d0) + phis(d0))
          gb = 0.5d0 * (IG(4 * (d1 - 1) + 1) + IG(4 * (d0 - 1) + 1))
          lx = max(lx, abs(ub) + sqrt(gb * pb))
        end do
      end do
      do j = 0, n
        do i = 1, n
          d0 = (i + h) + pw * ((j + h - 1) + pw * (p - 1))
          d1 = d0 + pw
          ub = 0.5d0 * (vc(d1) + vc(d0))
          pb = 0.5d0 * (phi(d1) + phis(d1) + phi(d0) + phis(d0))
          gb = 0.5d0 * (IG(4 * (d1 - 1) + 4) + IG(4 * (d0 - 1) + 4))
          ly = max(ly, abs(ub) + sqrt(gb * pb))
        end do
      end do
    end do
    dt_out = c * min(m%dx / lx, m%dx / ly)
    nst = ceiling(total / dt_out - 1d-12)
    dt_out = total / nst
  end subroutine cfl_dt

end program ccsm
