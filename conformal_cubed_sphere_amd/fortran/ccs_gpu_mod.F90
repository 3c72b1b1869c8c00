! ccs_gpu_mod.F90 -- iso_c_binding interface to libccs.so (include/ccs.h).
!
! Fortran hosts use this module in place of the reference's time_scheme_mod /
! spatial_operators_mod calls: the derived types mirror the C structs field
! for field and every interface mirrors one extern "C" entry point.
module ccs_gpu_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CCS_OK = 0, CCS_ERR_INVALID = 1, CCS_ERR_DEVICE = 2, &
                               CCS_ERR_ORDER = 3, CCS_ERR_NONFINITE = 4, CCS_ERR_NOMEM = 5, &
                               CCS_ERR_SINGULAR = 6
  integer(c_int), parameter :: CCS_FP64_PARITY = 0, CCS_FP32_SPLIT = 1, CCS_FP64_FAST = 2
  integer(c_int), parameter :: CCS_RK3_TVD = 0, CCS_RK4 = 1
  integer(c_int), parameter :: CCS_EFLUX_NONE = 0
  integer(c_int), parameter :: CCS_OP_WENO = 0, CCS_OP_CENTRED2 = 1, CCS_OP_CENTRED4 = 2, &
                               CCS_OP_CENTRED6 = 3, CCS_OP_WENO_EXT = 4

  type, bind(C) :: ccs_metrics
    integer(c_int) :: n, halo
    real(c_double) :: radius, omega
    real(c_double) :: dx
    type(c_ptr) :: lon, lat, sinlon, coslon, sinlat, coslat
    type(c_ptr) :: sqrtG, f, areaCell
    type(c_ptr) :: G, IG, A, IA
  end type ccs_metrics

  type, bind(C) :: ccs_case_params
    integer(c_int) :: case_num, n_bumps, mountain, pad_
    integer(c_long_long) :: seed
    real(c_double) :: gravity, u0, h_mean, bump_amp, bump_rmin, bump_rmax, mountain_h, mountain_r
  end type ccs_case_params

  type, bind(C) :: ccs_config
    integer(c_int) :: n, halo, precision, scheme, eflux, device, use_graph, spatial_op
    real(c_double) :: dt, dx, radius, gravity, omega
  end type ccs_config

  type, bind(C) :: ccs_history_desc
    integer(c_int) :: n, halo, case_num, pad_
    real(c_double) :: dx_deg, dt
  end type ccs_history_desc

  interface
    integer(c_int) function ccs_mesh_file_write(path, n) bind(C, name="ccs_mesh_file_write")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: n
    end function

    integer(c_int) function ccs_mesh_file_read(path, n, lon_deg, lat_deg, jab) bind(C, name="ccs_mesh_file_read")
      import :: c_int, c_char, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: n
      type(c_ptr), value :: lon_deg, lat_deg, jab
    end function

    integer(c_int) function ccs_mesh_file_read_ext(path, n, jab_ext) bind(C, name="ccs_mesh_file_read_ext")
      import :: c_int, c_char, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: n
      type(c_ptr), value :: jab_ext
    end function

    integer(c_int) function ccs_mesh_ext_metric(n, radius, jab_ext, ig_ext) bind(C, name="ccs_mesh_ext_metric")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: n
      real(c_double), value :: radius
      type(c_ptr), value :: jab_ext
      real(c_double) :: ig_ext(*)
    end function

    integer(c_int) function ccs_set_mesh_ext(h, ig_ext) bind(C, name="ccs_set_mesh_ext")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: ig_ext(*)
    end function

    integer(c_int) function ccs_history_create(path, d, lon, lat, areaCell, phis, h) bind(C, name="ccs_history_create")
      import :: c_int, c_char, c_double, c_ptr, ccs_history_desc
      character(kind=c_char), intent(in) :: path(*)
      type(ccs_history_desc), intent(in) :: d
      real(c_double), intent(in) :: lon(*), lat(*), areaCell(*), phis(*)
      type(c_ptr) :: h
    end function

    integer(c_int) function ccs_history_write(h, slot, phi, phis, u, v, uc, vc, zonal, merid) &
        bind(C, name="ccs_history_write")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: slot
      real(c_double), intent(in) :: phi(*), phis(*), u(*), v(*), uc(*), vc(*), zonal(*), merid(*)
    end function

    integer(c_int) function ccs_history_close(h) bind(C, name="ccs_history_close")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function ccs_mesh_geometry(n, lon_deg, lat_deg, xyz, jab) bind(C, name="ccs_mesh_geometry")
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double) :: lon_deg(*), lat_deg(*), xyz(*), jab(*)
    end function

    integer(c_int) function ccs_mesh_metrics(lon_deg, lat_deg, jab, m) bind(C, name="ccs_mesh_metrics")
      import :: c_int, c_double, ccs_metrics
      real(c_double), intent(in) :: lon_deg(*), lat_deg(*), jab(*)
      type(ccs_metrics) :: m
    end function

    subroutine ccs_case_defaults(c) bind(C, name="ccs_case_defaults")
      import :: ccs_case_params
      type(ccs_case_params) :: c
    end subroutine

    integer(c_int) function ccs_init_case(m, c, phiG, u, v, phi, uc, vc, zonal, merid, phis) &
        bind(C, name="ccs_init_case")
      import :: c_int, c_double, ccs_metrics, ccs_case_params
      type(ccs_metrics), intent(in) :: m
      type(ccs_case_params), intent(in) :: c
      real(c_double) :: phiG(*), u(*), v(*), phi(*), uc(*), vc(*), zonal(*), merid(*), phis(*)
    end function

    integer(c_int) function ccs_create(cfg, h) bind(C, name="ccs_create")
      import :: c_int, c_ptr, ccs_config
      type(ccs_config), intent(in) :: cfg
      type(c_ptr) :: h
    end function

    integer(c_int) function ccs_set_mesh(h, sqrtG, G, IG, A, IA, f, phis) bind(C, name="ccs_set_mesh")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: sqrtG(*), G(*), IG(*), A(*), IA(*), f(*), phis(*)
    end function

    ! phi, uc, vc: c_loc of the companion arrays, or c_null_ptr for none
    ! (the first step then derives them from phiG, u, v)
    integer(c_int) function ccs_set_state(h, phiG, u, v, phi, uc, vc) bind(C, name="ccs_set_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: phiG(*), u(*), v(*)
      type(c_ptr), value :: phi, uc, vc
    end function

    integer(c_int) function ccs_step(h, nsteps) bind(C, name="ccs_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function ccs_get_state(h, phiG, u, v, phi, uc, vc, zonal, merid) bind(C, name="ccs_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: phiG(*), u(*), v(*), phi(*), uc(*), vc(*), zonal(*), merid(*)
    end function

    integer(c_int) function ccs_diagnostics(h, mass, energy) bind(C, name="ccs_diagnostics")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: mass, energy
    end function

    integer(c_int) function ccs_synchronize(h) bind(C, name="ccs_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    subroutine ccs_destroy(h) bind(C, name="ccs_destroy")
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine

    type(c_ptr) function ccs_last_error(h) bind(C, name="ccs_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function

    type(c_ptr) function ccs_version() bind(C, name="ccs_version")
      import :: c_ptr
    end function
  end interface

contains

  ! C string -> Fortran string
  function c_to_f_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    interface
      integer(c_size_t) function strlen(q) bind(C, name="strlen")
        import :: c_ptr, c_size_t
        type(c_ptr), value :: q
      end function
    end interface
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    n = int(strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    s = transfer(chars(1:n), s)
  end function c_to_f_string

  ! Mirrors the reference's handle_err (io_mod.F90:604-612): print and stop.
  subroutine ccs_check(status, where, h)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    type(c_ptr), intent(in), optional :: h
    if (status /= CCS_OK) then
      if (present(h)) then
        print *, 'ccs error in ', trim(where), ': status ', status, ' ', c_to_f_string(ccs_last_error(h))
      else
        print *, 'ccs error in ', trim(where), ': status ', status
      end if
      error stop 2
    end if
  end subroutine ccs_check

end module ccs_gpu_mod
