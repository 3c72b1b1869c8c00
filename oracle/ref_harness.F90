! ref_harness.F90 -- C-callable harness around the reference dynamic core's
! own modules, compiled unmodified from the reference sources by
! oracle/Makefile (target `ref`, output oracle/_ref/libccs_ref.so).
!
! TEST INFRASTRUCTURE ONLY: it checks the C oracle (oracle/ccs_oracle.c)
! against the reference's actual spatial_operator, WENO_limiter, halo fill,
! vector projections, test cases and diagnostics.  time_scheme_mod cannot be
! built here (it uses io_mod, which needs netCDF), so the three stage updates
! of RK3_TVD (time_scheme_mod.F90:46-115) are restated below around the
! reference's own spatial_operator and correct_bdy_ghost pieces.
!
! The reference reads two never-written automatic arrays
! (spatial_operators_mod.F90:119,144); ccs_ref_clear_stack zeroes the stack
! region spatial_operator's automatic arrays occupy before every call, which
! is the value fresh stack pages hold.
module ref_harness
  use iso_c_binding
  use parameters_mod
  use mesh_mod
  use stat_mod
  use tend_mod
  use ghost_mod
  use spatial_operators_mod
  use diag_mod
  use test_case_mod
  implicit none

  interface
    subroutine clear_stack(nbytes) bind(C, name="ccs_ref_clear_stack")
      import :: c_size_t
      integer(c_size_t), value :: nbytes
    end subroutine clear_stack
  end interface

  integer(c_size_t) :: stack_bytes = 0

contains

  subroutine ref_init(n, h, dx_rad, clear_bytes) bind(C, name="ref_init")
    integer(c_int), value :: n, h
    real(c_double), value :: dx_rad
    integer(c_size_t), value :: clear_bytes
    ids = 1; ide = 2 * n + 1; jds = 1; jde = 2 * n + 1
    its = 1; ite = n; jts = 1; jte = n
    ifs = 1; ife = 6
    xhalo = h; yhalo = h
    ics = its - xhalo; ice = ite + xhalo; jcs = jts - yhalo; jce = jte + yhalo
    Nx = n; Ny = n
    dx = dx_rad; dy = dx_rad
    nIntegralSubSteps = 3
    stack_bytes = clear_bytes
    call initMesh
    call initStat
    call initTend
  end subroutine ref_init

  subroutine ref_set_mesh(sqrtG, G, IG, A, IA, f, phis, lon, lat, sinlat, coslat) bind(C, name="ref_set_mesh")
    real(c_double), intent(in) :: sqrtG(ics:ice, jcs:jce, 6), f(ics:ice, jcs:jce, 6), phis(ics:ice, jcs:jce, 6)
    real(c_double), intent(in) :: lon(ics:ice, jcs:jce, 6), lat(ics:ice, jcs:jce, 6)
    real(c_double), intent(in) :: sinlat(ics:ice, jcs:jce, 6), coslat(ics:ice, jcs:jce, 6)
    real(c_double), intent(in) :: G(2, 2, ics:ice, jcs:jce, 6), IG(2, 2, ics:ice, jcs:jce, 6)
    real(c_double), intent(in) :: A(2, 2, ics:ice, jcs:jce, 6), IA(2, 2, ics:ice, jcs:jce, 6)
    mesh%sqrtG = sqrtG
    mesh%f = f
    mesh%phis = phis
    mesh%lon = lon
    mesh%lat = lat
    mesh%sinlat = sinlat
    mesh%coslat = coslat
    mesh%matrixG = G
    mesh%matrixIG = IG
    mesh%matrixA = A
    mesh%matrixIA = IA
  end subroutine ref_set_mesh

  subroutine ref_get_phis(phis) bind(C, name="ref_get_phis")
    real(c_double), intent(out) :: phis(ics:ice, jcs:jce, 6)
    phis = mesh%phis
  end subroutine ref_get_phis

  subroutine ref_set_state(k, phiG, u, v, phi, uc, vc) bind(C, name="ref_set_state")
    integer(c_int), value :: k
    real(c_double), intent(in), dimension(ics:ice, jcs:jce, 6) :: phiG, u, v, phi, uc, vc
    stat(k)%phiG = phiG
    stat(k)%u = u
    stat(k)%v = v
    stat(k)%phi = phi
    stat(k)%uc = uc
    stat(k)%vc = vc
  end subroutine ref_set_state

  subroutine ref_get_state(k, phiG, u, v, phi, uc, vc, zonal, merid) bind(C, name="ref_get_state")
    integer(c_int), value :: k
    real(c_double), intent(out), dimension(ics:ice, jcs:jce, 6) :: phiG, u, v, phi, uc, vc, zonal, merid
    phiG = stat(k)%phiG
    u = stat(k)%u
    v = stat(k)%v
    phi = stat(k)%phi
    uc = stat(k)%uc
    vc = stat(k)%vc
    zonal = stat(k)%zonal_wind
    merid = stat(k)%meridional_wind
  end subroutine ref_get_state

#ifndef CCS_REF_CENTRED
  ! the centred-difference operator modules have no WENO_limiter
  real(c_double) function ref_weno(q, dir) bind(C, name="ref_weno")
    real(c_double), intent(in) :: q(5)
    integer(c_int), value :: dir
    real(c_double) :: r
    call WENO_limiter(r, q, dir)
    ref_weno = r
  end function ref_weno
#endif

  ! mesh%matrixIG_ext (io_mod.F90:223-250), read by spatial_operators_mod_WENO.F90
  subroutine ref_set_mesh_ext(IG_ext) bind(C, name="ref_set_mesh_ext")
    real(c_double), intent(in) :: IG_ext(2, 2, ids:ide, jds:jde, 6)
    mesh%matrixIG_ext = IG_ext
  end subroutine ref_set_mesh_ext

  subroutine call_spatial_operator(k, t)
    integer, intent(in) :: k, t
    call clear_stack(stack_bytes)
    call spatial_operator(stat(k), tend(t))
  end subroutine call_spatial_operator

  subroutine ref_spatial_operator(k, tphiG, tu, tv) bind(C, name="ref_spatial_operator")
    integer(c_int), value :: k
    real(c_double), intent(out), dimension(ics:ice, jcs:jce, 6) :: tphiG, tu, tv
    call call_spatial_operator(k, -1)
    tphiG = tend(-1)%phiG
    tu = tend(-1)%u
    tv = tend(-1)%v
  end subroutine ref_spatial_operator

  subroutine ghost_update(k)
    integer, intent(in) :: k
    call convert_wind_P2SP(stat(k))
    call fill_ghost(stat(k))
    call convert_wind_SP2P(stat(k))
    call convert_wind_cov2contrav(stat(k))
    stat(k)%phi = stat(k)%phiG / mesh%sqrtG
  end subroutine ghost_update

  subroutine ref_correct_bdy_ghost(k) bind(C, name="ref_correct_bdy_ghost")
    integer(c_int), value :: k
    call ghost_update(k)
  end subroutine ref_correct_bdy_ghost

  ! RK3_TVD (time_scheme_mod.F90:46-67) on stat(0), nsteps times, each step
  ! followed by copyStat(stat(old), stat(new)) (ccsm.F90:66).
  subroutine ref_rk3(nsteps, dt_s) bind(C, name="ref_rk3")
    integer(c_int), value :: nsteps
    real(c_double), value :: dt_s
    integer :: it
    integer, parameter :: old = 0, new = 1, one = -1, two = -2, three = -3
    do it = 1, nsteps
      call call_spatial_operator(old, one)
      stat(two)%phiG = stat(old)%phiG + dt_s * tend(one)%phiG
      stat(two)%u    = stat(old)%u    + dt_s * tend(one)%u
      stat(two)%v    = stat(old)%v    + dt_s * tend(one)%v
      call ghost_update(two)
      call call_spatial_operator(two, two)
      stat(three)%phiG = 0.75 * stat(old)%phiG + 0.25 * stat(two)%phiG + 0.25 * dt_s * tend(two)%phiG
      stat(three)%u    = 0.75 * stat(old)%u    + 0.25 * stat(two)%u    + 0.25 * dt_s * tend(two)%u
      stat(three)%v    = 0.75 * stat(old)%v    + 0.25 * stat(two)%v    + 0.25 * dt_s * tend(two)%v
      call ghost_update(three)
      call call_spatial_operator(three, three)
      stat(new)%phiG = stat(old)%phiG / 3. + 2. / 3. * stat(three)%phiG + 2. / 3. * dt_s * tend(three)%phiG
      stat(new)%u    = stat(old)%u    / 3. + 2. / 3. * stat(three)%u    + 2. / 3. * dt_s * tend(three)%u
      stat(new)%v    = stat(old)%v    / 3. + 2. / 3. * stat(three)%v    + 2. / 3. * dt_s * tend(three)%v
      call ghost_update(new)
      call copyStat(stat(old), stat(new))
    end do
  end subroutine ref_rk3

  subroutine ref_init_case(cnum) bind(C, name="ref_init_case")
    integer(c_int), value :: cnum
    case_num = cnum
    call initTestCase
  end subroutine ref_init_case

  subroutine ref_diag(k, mass, energy) bind(C, name="ref_diag")
    integer(c_int), value :: k
    real(c_double), intent(out) :: mass, energy
    real :: m, e
    call calc_total_mass(m, stat(k))
    call calc_total_energy(e, stat(k))
    mass = m
    energy = e
  end subroutine ref_diag

end module ref_harness
