! ref_mesh_harness.F90 -- C-callable harness around the reference mesh
! generator's conformal-map library (mesh_generation/src/cuco.f90 and the
! Purser modules it uses: pkind, pietc, pietc_s, pmat, pmat4, pfft, pfft1,
! pcoco), compiled unmodified from their sources by oracle/Makefile (target
! `ref`, output oracle/_ref/libccs_ref_mesh.so).
!
! TEST INFRASTRUCTURE ONLY: it checks this library's conformal map
! (ccs_mesh_geometry, ccs_mesh_file_write) against the reference's xmtoxc,
! called as generateConformalCubedSphere calls it for gen_method = 'Rancic'
! (mesh_mod.f90:150-166): map coordinates xm in [-1,1]^2, panel ipan, unit
! sphere position xc and Jacobian dxcdxm per map unit.
module ref_mesh_harness
  use iso_c_binding
  use pkind, only: spi, dp
  use cuco, only: inicuco, xmtoxc
  implicit none
contains
  subroutine ref_mesh_points(np, xm1, xm2, ipan, xc, dxcdxm) bind(C, name="ref_mesh_points")
    integer(c_int), value :: np
    real(c_double), intent(in) :: xm1(np), xm2(np)
    integer(c_int), intent(in) :: ipan(np)
    real(c_double), intent(out) :: xc(3, np), dxcdxm(3, 2, np)
    real(dp) :: xm(2), x(3), d(3, 2)
    integer(spi) :: ip
    integer :: k
    call inicuco
    do k = 1, np
      xm(1) = xm1(k)
      xm(2) = xm2(k)
      ip = int(ipan(k), spi)
      call xmtoxc(xm, x, d, ip)
      xc(:, k) = x
      dxcdxm(:, :, k) = d
    end do
  end subroutine ref_mesh_points
end module ref_mesh_harness
