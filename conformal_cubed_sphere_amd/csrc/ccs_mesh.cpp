// ccs_mesh.cpp — conformal cubed-sphere geometry and the dycore's static
// metric tables, computed once on the host in FP64.
//
// Geometry.  The reference generates its mesh offline with R. J. Purser's
// conformal-cubic series (mesh_generation/src/cuco.f90:89-147, pcoco.f90:
// 752-857, 2370-2447).  This file uses a closed form of the same map instead.
// Put the face centre at w = 0 of the stereographic plane w = (X+iY)/(1+Z)
// (face centre (0,0,1)); the cube's face centres are then 0, inf, +-1, +-i
// and its vertices are the zeros of W(w) = w^8 + 14 w^4 + 1 (the octahedral
// face form).  The map from the face square z = xm + i*ym in [-1,1]^2 onto
// the spherical face is the solution of
//        dw/dz = kappa * W(w)^(1/4),   w(0) = 0,
// kappa = (varpi/2) / 108^(1/4), varpi the lemniscate constant: its inverse
//        z(w) = (1/kappa) * int_0^w W(t)^(-1/4) dt
// sends the edge midpoint tan(pi/8) to 1 and the vertex to 1+i.  (Klein's
// octahedral identity W^3 - 108 w^4 (w^4-1)^4 = chi^2, chi = w^12 - 33 w^8
// - 33 w^4 + 1, turns the integral into a lemniscatic arcsine:
//        z(w) = (2/varpi) * arcsl(s),  s = 108^(1/4) w (1-w^4) / W^(3/4),
//        arcsl(s) = s * R_F(1 - s^2, 1 + s^2, 1),  1 - s^4 = chi^2 / W^3,
// with R_F Carlson's symmetric elliptic integral.)  w(z) is obtained by Newton
// iteration on z(w) over one eighth of one face; the eight-fold symmetry and
// the reference's six panel frames (cuco.f90:100-105) give every cell centre.
// The Jacobian is analytic: dw/dz = kappa W(w)^(1/4), so the grid is
// conformal by construction.
//
// Metrics.  ccs_mesh_metrics restates read_mesh (dynamic_core/src/
// io_mod.F90:110-304) operation for operation: degree->radian conversion,
// sin/cos tables, the halo fills of ghost_mod.f90:18-160, G = J^T J with the
// Gauss-Jordan inverse BRINV (math_mod.f90:451-525), sqrtG = sqrt(det G)
// (math_mod.f90:327-347), A = S(lon,lat) J with its inverse, f and areaCell.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "ccs.h"
#include "ccs_internal.h"

namespace ccs {

// ----------------------------------------------------------- halo table
// Rows: destination panel 1..6; columns: W, E, S, N.  Positions follow
// ghost_mod.f90:34-76, signed column swaps follow ghost_mod.f90:100-153.
static const HaloStrip kHalo[6][4] = {
    // P1: W<-P4(n-k+1,t) E<-P2(k,t) S<-P6(t,n-k+1) N<-P5(t,k)
    {{4, 1, 1, -1, 0, 0, 0, 0, 1, 0, 1, 1, 1},
     {2, 0, 0, 1, 0, 0, 0, 0, 1, 0, 1, 1, 1},
     {6, 0, 0, 0, 1, 1, 1, -1, 0, 0, 1, 1, 1},
     {5, 0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 1, 1}},
    // P2: W<-P1(n-k+1,t) E<-P3(k,t) S<-P6(n-k+1,n+1-t) N<-P5(n-k+1,t)
    {{1, 1, 1, -1, 0, 0, 0, 0, 1, 0, 1, 1, 1},
     {3, 0, 0, 1, 0, 0, 0, 0, 1, 0, 1, 1, 1},
     {6, 1, 1, -1, 0, 1, 1, 0, -1, 1, -1, 0, 1},
     {5, 1, 1, -1, 0, 0, 0, 0, 1, 1, 1, 0, -1}},
    // P3: W<-P2(n-k+1,t) E<-P4(k,t) S<-P6(n+1-t,k) N<-P5(n+1-t,n-k+1)
    {{2, 1, 1, -1, 0, 0, 0, 0, 1, 0, 1, 1, 1},
     {4, 0, 0, 1, 0, 0, 0, 0, 1, 0, 1, 1, 1},
     {6, 1, 1, 0, -1, 0, 0, 1, 0, 0, -1, 1, -1},
     {5, 1, 1, 0, -1, 1, 1, -1, 0, 0, -1, 1, -1}},
    // P4: W<-P3(n-k+1,t) E<-P1(k,t) S<-P6(k,t) N<-P5(k,n+1-t)
    {{3, 1, 1, -1, 0, 0, 0, 0, 1, 0, 1, 1, 1},
     {1, 0, 0, 1, 0, 0, 0, 0, 1, 0, 1, 1, 1},
     {6, 0, 0, 1, 0, 0, 0, 0, 1, 1, 1, 0, -1},
     {5, 0, 0, 1, 0, 1, 1, 0, -1, 1, -1, 0, 1}},
    // P5: W<-P4(n+1-t,n-k+1) E<-P2(t,n-k+1) S<-P1(t,n-k+1) N<-P3(n+1-t,n-k+1)
    {{4, 1, 1, 0, -1, 1, 1, -1, 0, 1, 1, 0, -1},
     {2, 0, 0, 0, 1, 1, 1, -1, 0, 1, -1, 0, 1},
     {1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 1, 1, 1},
     {3, 1, 1, 0, -1, 1, 1, -1, 0, 0, -1, 1, -1}},
    // P6: W<-P4(t,k) E<-P2(n+1-t,k) S<-P3(n+1-t,k) N<-P1(t,k)
    {{4, 0, 0, 0, 1, 0, 0, 1, 0, 1, -1, 0, 1},
     {2, 1, 1, 0, -1, 0, 0, 1, 0, 1, 1, 0, -1},
     {3, 1, 1, 0, -1, 0, 0, 1, 0, 0, -1, 1, -1},
     {1, 0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 1, 1}},
};

const HaloStrip& halo_strip(int p, int side) { return kHalo[p - 1][side]; }

bool build_scatter_rules(int n, int h, ScatterRule rules[6][4]) {
  if (n < 2 || h < 2) return false;
  bool filled[6][4] = {};
  for (int p = 1; p <= 6; ++p) {
    for (int s = 0; s < 4; ++s) {
      const HaloStrip& r = kHalo[p - 1][s];
      // Which side of the source panel holds this strip: inspect layer k=1.
      int is1, js1, is2, js2;
      halo_src(n, r, 1, 1, &is1, &js1);
      halo_src(n, r, 1, n, &is2, &js2);
      int sigma;
      if (is1 == 1 && is2 == 1) sigma = kWest;
      else if (is1 == n && is2 == n) sigma = kEast;
      else if (js1 == 1 && js2 == 1) sigma = kSouth;
      else if (js1 == n && js2 == n) sigma = kNorth;
      else return false;
      const int q = r.q;
      if (filled[q - 1][sigma]) return false;
      // Fit the affine map (i_s, j_s) -> (i_d, j_d) from three points.
      int ia, ja, ib, jb, ic, jc, da_i, da_j, db_i, db_j, dc_i, dc_j;
      halo_src(n, r, 1, 1, &ia, &ja); halo_dst(n, s, 1, 1, &da_i, &da_j);
      halo_src(n, r, 2, 1, &ib, &jb); halo_dst(n, s, 2, 1, &db_i, &db_j);
      halo_src(n, r, 1, 2, &ic, &jc); halo_dst(n, s, 1, 2, &dc_i, &dc_j);
      // Source displacements are unit vectors along i or j (signed).
      const int u1i = ib - ia, u1j = jb - ja, u2i = ic - ia, u2j = jc - ja;
      const int det = u1i * u2j - u1j * u2i;
      if (det != 1 && det != -1) return false;
      // Solve [u1 u2]^T-based system for the linear part L with
      // L*u1 = (db-da), L*u2 = (dc-da).  Inverse of [[u1i,u2i],[u1j,u2j]].
      const int inv00 = u2j * det, inv01 = -u2i * det, inv10 = -u1j * det, inv11 = u1i * det;
      const int d1i = db_i - da_i, d1j = db_j - da_j, d2i = dc_i - da_i, d2j = dc_j - da_j;
      ScatterRule sr;
      sr.p = p;
      sr.eii = d1i * inv00 + d2i * inv10;
      sr.eij = d1i * inv01 + d2i * inv11;
      sr.eji = d1j * inv00 + d2j * inv10;
      sr.ejj = d1j * inv01 + d2j * inv11;
      sr.ei0 = da_i - sr.eii * ia - sr.eij * ja;
      sr.ej0 = da_j - sr.eji * ia - sr.ejj * ja;
      sr.c1 = r.c1; sr.s1 = r.s1; sr.c2 = r.c2; sr.s2 = r.s2;
      // Self-check over the whole strip.
      for (int k = 1; k <= h; ++k)
        for (int t = 1; t <= n; ++t) {
          int si, sj, di, dj;
          halo_src(n, r, k, t, &si, &sj);
          halo_dst(n, s, k, t, &di, &dj);
          if (sr.ei0 + sr.eii * si + sr.eij * sj != di) return false;
          if (sr.ej0 + sr.eji * si + sr.ejj * sj != dj) return false;
        }
      rules[q - 1][sigma] = sr;
      filled[q - 1][sigma] = true;
    }
  }
  for (int q = 0; q < 6; ++q)
    for (int s = 0; s < 4; ++s)
      if (!filled[q][s]) return false;
  return true;
}

// ------------------------------------------------------- conformal map
namespace {

using cd = std::complex<double>;

// varpi/2 = int_0^1 dt / sqrt(1 - t^4) (lemniscate constant over two).
constexpr double kHalfVarpi = 1.3110287771460599052324197949455597068;

inline double root4_108() { return std::sqrt(std::sqrt(108.0)); }
inline double kappa() { return kHalfVarpi / root4_108(); }

inline cd face_form(cd w) {  // W(w) = w^8 + 14 w^4 + 1
  const cd q = (w * w) * (w * w);
  return (q + 14.0) * q + 1.0;
}

// Carlson's R_F for complex arguments off the negative real axis
// (B. C. Carlson, Numer. Algorithms 10 (1995) 13-26, duplication algorithm).
cd carlson_rf(cd x, cd y, cd z) {
  const double tol = 1e-17;
  const cd a0 = (x + y + z) / 3.0;
  const double q = std::pow(3.0 * tol, -1.0 / 6.0) *
                   std::max(std::abs(a0 - x), std::max(std::abs(a0 - y), std::abs(a0 - z)));
  cd xm = x, ym = y, zm = z, am = a0;
  double scale = 1.0;
  for (int m = 0; m < 200 && scale * q >= std::abs(am); ++m) {
    const cd sx = std::sqrt(xm), sy = std::sqrt(ym), sz = std::sqrt(zm);
    const cd lam = sx * sy + sx * sz + sy * sz;
    xm = 0.25 * (xm + lam);
    ym = 0.25 * (ym + lam);
    zm = 0.25 * (zm + lam);
    am = 0.25 * (am + lam);
    scale *= 0.25;
  }
  const cd X = (a0 - x) * scale / am;
  const cd Y = (a0 - y) * scale / am;
  const cd Z = -(X + Y);
  const cd e2 = X * Y - Z * Z;
  const cd e3 = X * Y * Z;
  return (1.0 - e2 / 10.0 + e3 / 14.0 + e2 * e2 / 24.0 - 3.0 * e2 * e3 / 44.0) / std::sqrt(am);
}

// Inverse map z(w) for w in the sector 0 <= arg w <= pi/4 of the face.
cd z_of_w(cd w) {
  const cd q = (w * w) * (w * w);
  const cd W = (q + 14.0) * q + 1.0;
  const cd chi = ((q - 33.0) * q - 33.0) * q + 1.0;
  const cd s = root4_108() * w * (1.0 - q) / std::pow(W, 0.75);
  const cd s2 = s * s;
  const cd y = 1.0 + s2;
  const cd x = chi * chi / (W * W * W * y);  // = 1 - s^2 without cancellation
  return s * carlson_rf(x, y, cd(1.0, 0.0)) / kHalfVarpi;
}

// Newton solve of z(w) = zt for zt in the triangle 0 <= Im z <= Re z <= 1,
// starting from `guess`.  Steps are halved whenever they would leave the
// octant sector (where z_of_w's principal branches are valid).
cd w_of_z(cd zt, cd guess, int* iters) {
  const double k = kappa();
  cd w = guess;
  int it = 0, polish = -1;
  for (; it < 100; ++it) {
    const cd dz = z_of_w(w) - zt;
    cd dw = dz * k * std::pow(face_form(w), 0.25);
    // Stay inside the octant of the face: 0 <= arg w <= pi/4 and inside the
    // stereographic image (Re w + 1)^2 + (Im w)^2 = 2 of the edge great circle.
    for (int damp = 0; damp < 60; ++damp) {
      const cd wn = w - dw;
      const double ang = std::arg(wn);
      const double er = (wn.real() + 1.0) * (wn.real() + 1.0) + wn.imag() * wn.imag();
      if (ang >= -1e-9 && ang <= M_PI / 4 + 1e-9 && er <= 2.0 + 1e-12) break;
      dw *= 0.5;
    }
    w -= dw;
    if (polish < 0 && std::abs(dw) <= 1e-14 * std::abs(w)) polish = 2;  // two more steps
    if (polish >= 0 && polish-- == 0) break;
  }
  if (iters) *iters = it;
  return w;
}

// Edge x = 1 of the face, z = 1 + i yt with 0 <= yt < 1.  The edge is mapped
// onto the stereographic image of the edge great circle, w = sqrt(2) e^{it} - 1,
// on which Re z = 1 holds.  There z_of_w's elliptic integral sits on its branch
// cut, so Im z comes out with either sign (the magnitude is right): solve
// |Im z(w(t))| = yt by Newton in the angle t, with
// dz/dt = dz/dw * dw/dt = i sqrt(2) e^{it} / (kappa W(w)^{1/4}).  (The 2-D
// Newton of w_of_z, which must stay inside the circle, stalls along it.)
cd w_on_edge(double yt, cd guess) {
  const double k = kappa(), r2 = std::sqrt(2.0);
  double t = std::arg(guess + 1.0);
  int polish = -1;
  for (int it = 0; it < 60; ++it) {
    const cd e = std::polar(1.0, t);
    const cd w = r2 * e - 1.0;
    const cd dzdt = cd(0.0, r2) * e / (k * std::pow(face_form(w), 0.25));
    const double dt = (std::fabs(z_of_w(w).imag()) - yt) / std::fabs(dzdt.imag());
    t -= dt;
    if (polish < 0 && std::fabs(dt) <= 1e-15) polish = 2;  // two more steps
    if (polish >= 0 && polish-- == 0) break;
  }
  return r2 * std::polar(1.0, t) - 1.0;
}

// Panel frames: global xc = R_p * local xc, local face centre (0,0,1),
// local x toward +Re w, local y toward +Im w.  Columns as in the reference
// (mesh_generation/src/cuco.f90:100-105), stored here as R[p][row][col].
const double kRot[6][3][3] = {
    {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}},
    {{-1, 0, 0}, {0, 0, 1}, {0, 1, 0}},
    {{0, 0, -1}, {-1, 0, 0}, {0, 1, 0}},
    {{1, 0, 0}, {0, 0, -1}, {0, 1, 0}},
    {{0, -1, 0}, {1, 0, 0}, {0, 0, 1}},
    {{0, 1, 0}, {1, 0, 0}, {0, 0, -1}},
};

}  // namespace

// w and dw/dz at the points (coord[i], coord[j]) of one face (m*m, i
// fastest), coord ascending and symmetric about 0 (it may contain 0 and the
// edges +-1).  Newton solves cover the fundamental triangle
// 0 <= y <= x (continuation along rows from the diagonal); the rest follows
// from the face's symmetries.  The cube vertex, where dz/dw is singular, is
// set from its closed form w = (1 + i)/(sqrt(3) + 1).
void map_points(const std::vector<double>& coord, std::vector<cd>& w, std::vector<cd>& dwdz) {
  const int m = static_cast<int>(coord.size());
  w.assign(static_cast<std::size_t>(m) * m, cd(0, 0));
  dwdz.assign(w.size(), cd(0, 0));
  const double k = kappa();
  const int h0 = m / 2;  // first index with coord >= 0
  const cd vertex = cd(1.0, 1.0) / (std::sqrt(3.0) + 1.0);
  std::vector<cd> tri(static_cast<std::size_t>(m) * m, cd(0, 0));
  cd diag_guess = k * cd(coord[h0], coord[h0]);
  for (int j = h0; j < m; ++j) {
    cd guess = diag_guess;
    for (int i = j; i < m; ++i) {
      const cd zt(coord[i], coord[j]);
      cd wi;
      if (coord[i] == 1.0 && coord[j] == 1.0)
        wi = vertex;
      else if (coord[i] == 1.0)  // panel edge x = 1: a 1-D solve on the edge circle
        wi = w_on_edge(coord[j], guess);
      else
        wi = w_of_z(zt, guess, nullptr);
      tri[static_cast<std::size_t>(j) * m + i] = wi;
      if (i == j) {
        // extrapolate along the diagonal for the next row's start
        diag_guess = wi;
        if (j > h0) {
          const cd prev = tri[static_cast<std::size_t>(j - 1) * m + (j - 1)];
          diag_guess = wi + 0.5 * (wi - prev);
        }
      }
      guess = wi;
    }
  }
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      const double x = coord[i], y = coord[j];
      const int ia = (x >= 0) ? i : m - 1 - i;  // |x| index
      const int ja = (y >= 0) ? j : m - 1 - j;  // |y| index
      cd wa;
      if (ia >= ja) {
        wa = tri[static_cast<std::size_t>(ja) * m + ia];
      } else {
        const cd wt = tri[static_cast<std::size_t>(ia) * m + ja];
        wa = cd(0, 1) * std::conj(wt);  // reflection across the diagonal
      }
      const cd wij((x >= 0 ? 1.0 : -1.0) * wa.real(), (y >= 0 ? 1.0 : -1.0) * wa.imag());
      w[static_cast<std::size_t>(j) * m + i] = wij;
      dwdz[static_cast<std::size_t>(j) * m + i] = k * std::pow(face_form(wij), 0.25);
    }
}

// Position on the unit sphere and Jacobian d(xyz)/d(map angle) (3x2,
// column-major, per radian) of the point w of panel p (0-based).
void panel_point(int p, cd w, cd dwdz, double g[3], double jab[6]) {
  const double(*R)[3] = kRot[p];
  const double quarter_pi = 2.0 * std::asin(1.0) / 4.0;  // d(map angle)/d(xm)
  const double r = w.real(), s = w.imag();
  const double D = 1.0 + r * r + s * s;
  const double v[3] = {2.0 * r / D, 2.0 * s / D, (1.0 - r * r - s * s) / D};
  const double D2 = D * D;
  const double dvr[3] = {2.0 * (1.0 - r * r + s * s) / D2, -4.0 * r * s / D2, -4.0 * r / D2};
  const double dvs[3] = {-4.0 * r * s / D2, 2.0 * (1.0 + r * r - s * s) / D2, -4.0 * s / D2};
  const double pr = dwdz.real(), pi_ = dwdz.imag();
  double dx_loc[3], dy_loc[3], gx[3], gy[3];
  for (int a = 0; a < 3; ++a) {
    dx_loc[a] = dvr[a] * pr + dvs[a] * pi_;   // d v / d xm
    dy_loc[a] = -dvr[a] * pi_ + dvs[a] * pr;  // d v / d ym
  }
  for (int a = 0; a < 3; ++a) {
    g[a] = R[a][0] * v[0] + R[a][1] * v[1] + R[a][2] * v[2];
    gx[a] = R[a][0] * dx_loc[0] + R[a][1] * dx_loc[1] + R[a][2] * dx_loc[2];
    gy[a] = R[a][0] * dy_loc[0] + R[a][1] * dy_loc[1] + R[a][2] * dy_loc[2];
  }
  for (int a = 0; a < 3; ++a) {
    jab[a] = gx[a] / quarter_pi;
    jab[3 + a] = gy[a] / quarter_pi;
  }
}

}  // namespace ccs

using namespace ccs;

extern "C" int ccs_mesh_geometry(int n, double* lon_deg, double* lat_deg, double* xyz,
                                 double* jab) {
  if (n < 4 || (n % 2) != 0) return CCS_ERR_INVALID;
  std::vector<double> coord(n);
  for (int i = 0; i < n; ++i) coord[i] = -1.0 + (2.0 * i + 1.0) / n;  // cell centres
  std::vector<std::complex<double>> w, dwdz;
  map_points(coord, w, dwdz);
  const double pi = 2.0 * std::asin(1.0);
  const double r2d = 180.0 / pi;
  const std::size_t ncell = static_cast<std::size_t>(n) * n;
  for (int p = 0; p < 6; ++p) {
    for (std::size_t c = 0; c < ncell; ++c) {
      double g[3], jb[6];
      panel_point(p, w[c], dwdz[c], g, jb);
      const std::size_t cell = c + ncell * static_cast<std::size_t>(p);
      if (xyz) {
        xyz[3 * cell + 0] = g[0];
        xyz[3 * cell + 1] = g[1];
        xyz[3 * cell + 2] = g[2];
      }
      if (lon_deg) lon_deg[cell] = std::atan2(g[1], g[0]) * r2d;
      if (lat_deg) lat_deg[cell] = std::atan2(g[2], std::sqrt(g[0] * g[0] + g[1] * g[1])) * r2d;
      if (jab)
        for (int a = 0; a < 6; ++a) jab[6 * cell + a] = jb[a];
    }
  }
  return CCS_OK;
}

// ------------------------------------------------------------- metrics
namespace {

constexpr double kFillValue = -9999999999999999.0;  // constants_mod.F90:14

// Gauss-Jordan inverse with full pivoting, math_mod.f90:451-525 (N = 2).
// Returns false when the largest pivot is negligible (D + 1 == 1).
bool brinv2(const double ain[4], double a[4]) {
  // a is column-major: a(r,c) at r + 2c, 0-based.
  auto A = [a](int r, int c) -> double& { return a[(r - 1) + 2 * (c - 1)]; };
  for (int e = 0; e < 4; ++e) a[e] = ain[e];
  int is[3], js[3];
  for (int k = 1; k <= 2; ++k) {
    double d = 0.0;
    for (int i = k; i <= 2; ++i)
      for (int j = k; j <= 2; ++j)
        if (std::fabs(A(i, j)) > d) {
          d = std::fabs(A(i, j));
          is[k] = i;
          js[k] = j;
        }
    if (d + 1.0 == 1.0) return false;
    for (int j = 1; j <= 2; ++j) std::swap(A(k, j), A(is[k], j));
    for (int i = 1; i <= 2; ++i) std::swap(A(i, k), A(i, js[k]));
    A(k, k) = 1.0 / A(k, k);
    for (int j = 1; j <= 2; ++j)
      if (j != k) A(k, j) = A(k, j) * A(k, k);
    for (int i = 1; i <= 2; ++i)
      if (i != k)
        for (int j = 1; j <= 2; ++j)
          if (j != k) A(i, j) = A(i, j) - A(i, k) * A(k, j);
    for (int i = 1; i <= 2; ++i)
      if (i != k) A(i, k) = -A(i, k) * A(k, k);
  }
  for (int k = 2; k >= 1; --k) {
    for (int j = 1; j <= 2; ++j) std::swap(A(k, j), A(js[k], j));
    for (int i = 1; i <= 2; ++i) std::swap(A(i, k), A(i, is[k]));
  }
  return true;
}

// Scalar halo fill (CubedSphereFillHalo, ghost_mod.f90:18-82).
void fill_halo_scalar(int n, int h, double* f) {
  for (int p = 1; p <= 6; ++p)
    for (int s = 0; s < 4; ++s) {
      const HaloStrip& r = halo_strip(p, s);
      for (int k = 1; k <= h; ++k)
        for (int t = 1; t <= n; ++t) {
          int di, dj, si, sj;
          halo_dst(n, s, k, t, &di, &dj);
          halo_src(n, r, k, t, &si, &sj);
          f[pidx(n, h, di, dj, p)] = f[pidx(n, h, si, sj, r.q)];
        }
    }
}

// Jacobian halo fill with signed column swaps (CubedSphereFillJab,
// ghost_mod.f90:84-160).  jab is (3,2,cell).
void fill_halo_jab(int n, int h, double* jab) {
  for (int p = 1; p <= 6; ++p)
    for (int s = 0; s < 4; ++s) {
      const HaloStrip& r = halo_strip(p, s);
      for (int k = 1; k <= h; ++k)
        for (int t = 1; t <= n; ++t) {
          int di, dj, si, sj;
          halo_dst(n, s, k, t, &di, &dj);
          halo_src(n, r, k, t, &si, &sj);
          const double* src = jab + 6 * pidx(n, h, si, sj, r.q);
          double* dst = jab + 6 * pidx(n, h, di, dj, p);
          for (int a = 0; a < 3; ++a) {
            dst[a] = r.s1 * src[3 * r.c1 + a];
            dst[3 + a] = r.s2 * src[3 * r.c2 + a];
          }
        }
    }
}

}  // namespace

extern "C" int ccs_mesh_metrics(const double* lon_deg, const double* lat_deg, const double* jab_in,
                                ccs_metrics* m) {
  if (!m || !lon_deg || !lat_deg || !jab_in) return CCS_ERR_INVALID;
  const int n = m->n, h = m->halo;
  if (n < 2 || h < 1 || h > n) return CCS_ERR_INVALID;
  const double pi = 2.0 * std::asin(1.0);     // constants_mod.F90:9
  const double d2r = pi / 180.0;              // constants_mod.F90:12
  const double r2d = 180.0 / pi;              // constants_mod.F90:13
  const std::size_t np = padded_cells(n, h);
  const double dx_deg = 90.0 / n;
  const double dx = dx_deg * d2r;             // io_mod.F90:54
  m->dx = dx;

  std::vector<double> lon(np, kFillValue), lat(np, kFillValue);
  std::vector<double> sinlon(np), coslon(np), sinlat(np), coslat(np);
  std::vector<double> jab(6 * np, 0.0);
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        const std::size_t c = iidx(n, i, j, p), d = pidx(n, h, i, j, p);
        lon[d] = lon_deg[c] * d2r;            // io_mod.F90:110,126
        lat[d] = lat_deg[c] * d2r;
        for (int a = 0; a < 6; ++a) jab[6 * d + a] = jab_in[6 * c + a] * m->radius;  // :117
      }
  for (std::size_t d = 0; d < np; ++d) {      // io_mod.F90:138-141
    sinlon[d] = std::sin(lon[d]);
    coslon[d] = std::cos(lon[d]);
    sinlat[d] = std::sin(lat[d]);
    coslat[d] = std::cos(lat[d]);
  }
  fill_halo_scalar(n, h, lon.data());         // io_mod.F90:143-153
  fill_halo_scalar(n, h, lat.data());
  fill_halo_scalar(n, h, sinlon.data());
  fill_halo_scalar(n, h, coslon.data());
  fill_halo_scalar(n, h, sinlat.data());
  fill_halo_scalar(n, h, coslat.data());
  fill_halo_jab(n, h, jab.data());

  auto put = [](double* dst, std::size_t d, double v) { if (dst) dst[d] = v; };
  int status = CCS_OK;
  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - h; j <= n + h; ++j)
      for (int i = 1 - h; i <= n + h; ++i) {
        const std::size_t d = pidx(n, h, i, j, p);
        const bool corner = is_corner(n, i, j);
        put(m->lon, d, corner ? 0.0 : lon[d]);
        put(m->lat, d, corner ? 0.0 : lat[d]);
        put(m->sinlon, d, corner ? 0.0 : sinlon[d]);
        put(m->coslon, d, corner ? 0.0 : coslon[d]);
        put(m->sinlat, d, corner ? 0.0 : sinlat[d]);
        put(m->coslat, d, corner ? 0.0 : coslat[d]);
        double G[4] = {0, 0, 0, 0}, IG[4] = {0, 0, 0, 0}, A[4] = {0, 0, 0, 0}, IA[4] = {0, 0, 0, 0};
        double sqrtG = 0.0;
        if (!corner) {
          // io_mod.F90:159-220
          const double lo = lon[d], la = lat[d];
          const double S[2][3] = {{-std::sin(lo), std::cos(lo), 0.0},
                                  {-(std::cos(lo) * std::sin(la)), -(std::sin(lo) * std::sin(la)),
                                   std::cos(la)}};
          const double* J = &jab[6 * d];  // J(r,c) = J[r + 3c]
          for (int c = 0; c < 2; ++c)
            for (int r = 0; r < 2; ++r) {
              // matmul(transpose(J), J)(r,c) = sum_k J(k,r) J(k,c)
              double acc = J[0 + 3 * r] * J[0 + 3 * c];
              acc = acc + J[1 + 3 * r] * J[1 + 3 * c];
              acc = acc + J[2 + 3 * r] * J[2 + 3 * c];
              G[r + 2 * c] = acc;
              // matmul(S, J)(r,c) = sum_k S(r,k) J(k,c)
              double a2 = S[r][0] * J[0 + 3 * c];
              a2 = a2 + S[r][1] * J[1 + 3 * c];
              a2 = a2 + S[r][2] * J[2 + 3 * c];
              A[r + 2 * c] = a2;
            }
          if (!brinv2(G, IG)) status = CCS_ERR_SINGULAR;
          const double det = G[0] * G[3] - G[1] * G[2];  // math_mod.f90:340
          sqrtG = std::sqrt(det);
          const double latdeg = la * r2d;
          if (std::fabs(latdeg) != 90.0) {
            if (!brinv2(A, IA)) status = CCS_ERR_SINGULAR;
          } else {
            // analytic pole substitute, io_mod.F90:189-216; IA2 is the
            // identity, so matmul(IA2, IA1) reproduces IA1 exactly.
            // Column-major: north [cos, sin, -sin, cos], south [-cos, sin, sin, cos].
            const double c = std::cos(lo), sn = std::sin(lo);
            if (latdeg == 90.0) {
              IA[0] = c; IA[1] = sn; IA[2] = -sn; IA[3] = c;
            } else {
              IA[0] = -c; IA[1] = sn; IA[2] = sn; IA[3] = c;
            }
            for (int e = 0; e < 4; ++e) IA[e] = IA[e] * m->radius;
          }
        }
        put(m->sqrtG, d, sqrtG);
        put(m->areaCell, d, sqrtG * (dx * dx));          // io_mod.F90:304
        const bool interior = (i >= 1 && i <= n && j >= 1 && j <= n);
        put(m->f, d, interior ? (2.0 * m->omega) * sinlat[d] : 0.0);  // :298
        for (int e = 0; e < 4; ++e) {
          put(m->G, 4 * d + e, G[e]);
          put(m->IG, 4 * d + e, IG[e]);
          put(m->A, 4 * d + e, A[e]);
          put(m->IA, 4 * d + e, IA[e]);
        }
      }
  return status;
}

// Inverse metric on the extended lattice, io_mod.F90:223-250: for every
// extended point except the panel's four corner points (left unset there, so
// zero here), G_ext = J_ext^T J_ext with J_ext = jab_ext * radius and
// matrixIG_ext = BRINV(G_ext).
extern "C" int ccs_mesh_ext_metric(int n, double radius, const double* jab_ext, double* ig_ext) {
  if (n < 4 || (n % 2) != 0 || !ig_ext || !(radius > 0.0)) return CCS_ERR_INVALID;
  const int m = 2 * n + 1;
  const std::size_t np = static_cast<std::size_t>(m) * m;
  std::vector<double> gen;
  if (!jab_ext) {
    std::vector<double> coord(m);
    for (int k = 0; k < m; ++k) coord[k] = -1.0 + static_cast<double>(k) / n;
    std::vector<std::complex<double>> w, dwdz;
    map_points(coord, w, dwdz);
    gen.resize(36 * np);
    for (int p = 0; p < 6; ++p)
      for (std::size_t c = 0; c < np; ++c) {
        double g[3];
        panel_point(p, w[c], dwdz[c], g, &gen[6 * (c + np * p)]);
      }
    jab_ext = gen.data();
  }
  int status = CCS_OK;
  for (int p = 0; p < 6; ++p)
    for (int j = 1; j <= m; ++j)
      for (int i = 1; i <= m; ++i) {
        const std::size_t q = static_cast<std::size_t>(i - 1) + static_cast<std::size_t>(m) *
                              (static_cast<std::size_t>(j - 1) + static_cast<std::size_t>(m) * p);
        double* IG = ig_ext + 4 * q;
        IG[0] = IG[1] = IG[2] = IG[3] = 0.0;
        if ((i == 1 || i == m) && (j == 1 || j == m)) continue;
        double J[6];
        for (int a = 0; a < 6; ++a) J[a] = jab_ext[6 * q + a] * radius;  // io_mod.F90:115
        double G[4];
        for (int c = 0; c < 2; ++c)
          for (int r = 0; r < 2; ++r) {
            double acc = J[0 + 3 * r] * J[0 + 3 * c];
            acc = acc + J[1 + 3 * r] * J[1 + 3 * c];
            acc = acc + J[2 + 3 * r] * J[2 + 3 * c];
            G[r + 2 * c] = acc;
          }
        if (!brinv2(G, IG)) status = CCS_ERR_SINGULAR;
      }
  return status;
}
