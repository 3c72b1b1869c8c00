// ccs_cases.cpp — initial conditions on the padded grid (host, FP64).
//
// Williamson cases 2, 5 and 6 follow test_case_mod.F90:61-204 expression for
// expression; the state is then completed as initTestCase does
// (test_case_mod.F90:39-48): phiG = phi*sqrtG, covariant (u,v) =
// matmul(G, IA)*(us,vs) (projection_mod.F90:60-77) and contravariant
// (uc,vc) = IA*(us,vs) (projection_mod.F90:98-110).
//
// Case 7 is the seeded benchmark recipe, which has no counterpart in the
// reference: zonal jet u = u0 cos(lat), v = 0, geopotential in gradient-wind
// balance gh = gh0 - (a Omega u0 + u0^2/2) sin^2(lat) (the formula of case 2,
// test_case_mod.F90:82) with gh0 = g*h_mean + (a Omega u0 + u0^2/2)/3 so the
// area-mean depth of the balanced part is h_mean, plus n_bumps Gaussian bumps
// A exp(-(d/r)^2) (d great-circle distance) and an optional Gaussian mountain
// hs = H exp(-(d/R)^2) carried as phis = g hs with phi reduced by g hs (the
// case-5 convention, test_case_mod.F90:134).  Random numbers: SplitMix64
// (Steele, Lea, Flood 2014) of `seed`, 53-bit uniforms u = (x >> 11) 2^-53;
// draw order per bump: sin(lat) ~ U(-1,1), lon ~ U(0,2pi), A ~ U(-amp,amp),
// r ~ U(rmin,rmax); then the mountain centre: sin(lat), lon.

#include <cmath>
#include <cstdint>
#include <vector>

#include "ccs.h"
#include "ccs_internal.h"

namespace {

std::uint64_t splitmix64(std::uint64_t* s) {
  std::uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double uniform53(std::uint64_t* s) {
  return static_cast<double>(splitmix64(s) >> 11) * 0x1.0p-53;
}

double great_circle(double radius, double lon1, double lat1, double lon2, double lat2) {
  double c = std::sin(lat1) * std::sin(lat2) + std::cos(lat1) * std::cos(lat2) * std::cos(lon1 - lon2);
  c = std::fmin(1.0, std::fmax(-1.0, c));
  return radius * std::acos(c);
}

struct Recipe {
  std::vector<double> lon, lat, amp, rad;
  double mlon = 0.0, mlat = 0.0;
};

}  // namespace

extern "C" void ccs_case_defaults(ccs_case_params* c) {
  if (!c) return;
  c->case_num = 7;
  c->n_bumps = 8;
  c->mountain = 0;
  c->pad_ = 0;
  c->seed = 1;
  c->gravity = 9.80616;
  c->u0 = 20.0;
  c->h_mean = 10000.0;
  c->bump_amp = 100.0;
  c->bump_rmin = 5.0e5;
  c->bump_rmax = 1.5e6;
  c->mountain_h = 2000.0;
  c->mountain_r = 1.5e6;
}

extern "C" int ccs_init_case(const ccs_metrics* m, const ccs_case_params* c, double* phiG, double* u,
                             double* v, double* phi, double* uc, double* vc, double* zonal,
                             double* merid, double* phis) {
  if (!m || !c || !phiG || !u || !v || !phi || !uc || !vc || !phis) return CCS_ERR_INVALID;
  if (!m->lon || !m->lat || !m->sinlat || !m->coslat || !m->sqrtG || !m->G || !m->IA)
    return CCS_ERR_INVALID;
  const int n = m->n, h = m->halo;
  const int cn = c->case_num;
  if (cn != 2 && cn != 5 && cn != 6 && cn != 7) return CCS_ERR_INVALID;
  if (cn == 7 && (c->n_bumps < 0 || c->n_bumps > 64)) return CCS_ERR_INVALID;
  const double pi = 2.0 * std::asin(1.0);
  const double radius = m->radius, Omega = m->omega, gravity = c->gravity;

  Recipe rc;
  if (cn == 7) {
    std::uint64_t st = c->seed;
    for (int b = 0; b < c->n_bumps; ++b) {
      const double z = 2.0 * uniform53(&st) - 1.0;
      rc.lon.push_back(2.0 * pi * uniform53(&st));
      rc.lat.push_back(std::asin(z));
      rc.amp.push_back(c->bump_amp * (2.0 * uniform53(&st) - 1.0));
      rc.rad.push_back(c->bump_rmin + (c->bump_rmax - c->bump_rmin) * uniform53(&st));
    }
    const double z = 2.0 * uniform53(&st) - 1.0;
    rc.mlon = 2.0 * pi * uniform53(&st);
    rc.mlat = std::asin(z);
  }

  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - h; j <= n + h; ++j)
      for (int i = 1 - h; i <= n + h; ++i) {
        const std::size_t d = ccs::pidx(n, h, i, j, p);
        if (ccs::is_corner(n, i, j)) {
          phiG[d] = u[d] = v[d] = phi[d] = uc[d] = vc[d] = phis[d] = 0.0;
          if (zonal) zonal[d] = 0.0;
          if (merid) merid[d] = 0.0;
          continue;
        }
        double us = 0.0, vs = 0.0, ph = 0.0, ghs = 0.0;
        if (cn == 2) {  // test_case_mod.F90:61-95
          const double gh0 = 29400.0;
          const double u0 = 2.0 * pi * radius / (12.0 * 86400.0);
          ph = gh0 - (radius * Omega * u0 + u0 * u0 / 2.0) * (m->sinlat[d] * m->sinlat[d]);
          us = u0 * m->coslat[d];
          vs = 0.0;
        } else if (cn == 5) {  // test_case_mod.F90:98-144
          const double hs0 = 2000.0, u0 = 20.0, alpha = 0.0, gh0 = 5960.0 * gravity;
          const double rr = pi / 9.0, lc = 1.5 * pi, tc = pi / 6.0;
          double lo = m->lon[d];
          const double la = m->lat[d];
          if (lo < 0) lo = 2.0 * pi + lo;
          const double dd = (lo - lc) * (lo - lc) + (la - tc) * (la - tc);
          const double r = std::sqrt(rr * rr < dd ? rr * rr : dd);
          ghs = gravity * hs0 * (1.0 - r / rr);
          us = u0 * (std::cos(la) * std::cos(alpha) + std::cos(lo) * std::sin(la) * std::sin(alpha));
          vs = -(u0 * std::sin(lo) * std::sin(alpha));
          const double X = -(std::cos(lo) * std::cos(la) * std::sin(alpha)) + std::sin(la) * std::cos(alpha);
          ph = gh0 - (radius * Omega * u0 + u0 * u0 / 2.0) * (X * X) - ghs;
        } else if (cn == 6) {  // test_case_mod.F90:147-204
          const double omg = 7.848e-6, R = 4.0, h0 = 8000.0;
          const double s = m->sinlat[d], co = m->coslat[d], lo = m->lon[d];
          const double u1 = co;
          const double u2 = R * std::pow(co, R - 1) * (s * s) * std::cos(R * lo);
          const double u3 = std::pow(co, R + 1) * std::cos(R * lo);
          us = radius * omg * (u1 + u2 - u3);
          vs = -(radius * omg * R * std::pow(co, R - 1) * s * std::sin(R * lo));
          const double AA1 = omg * 0.5 * (2.0 * Omega + omg) * (co * co);
          const double Ac = 0.25 * (omg * omg);
          const double A21 = (R + 1.0) * std::pow(co, 2.0 * R + 2.0);
          const double A22 = (2.0 * (R * R) - R - 2.0) * std::pow(co, 2.0 * R);
          const double A23 = 2.0 * (R * R) * std::pow(co, 2.0 * R - 2);
          const double Ah = AA1 + Ac * (A21 + A22 - A23);
          const double Bc = 2.0 * (Omega + omg) * omg / ((R + 1) * (R + 2)) * std::pow(co, R);
          const double BB1 = R * R + 2.0 * R + 2.0;
          const double BB2 = ((R + 1.0) * (R + 1.0)) * (co * co);
          const double Bh = Bc * (BB1 - BB2);
          const double CC = 0.25 * (omg * omg) * std::pow(co, 2.0 * R);
          const double CC1 = (R + 1.0) * (co * co);
          const double CC2 = R + 2.0;
          const double Ch = CC * (CC1 - CC2);
          ph = gravity * h0 + (radius * radius) * (Ah + Bh * std::cos(R * lo) + Ch * std::cos(2.0 * R * lo));
        } else {  // seeded balanced jet
          const double K = radius * Omega * c->u0 + c->u0 * c->u0 / 2.0;
          const double gh0 = gravity * c->h_mean + K / 3.0;
          const double lo = m->lon[d], la = m->lat[d];
          double bump = 0.0;
          for (int b = 0; b < c->n_bumps; ++b) {
            const double q = great_circle(radius, lo, la, rc.lon[b], rc.lat[b]) / rc.rad[b];
            bump = bump + rc.amp[b] * std::exp(-(q * q));
          }
          double hs = 0.0;
          if (c->mountain) {
            const double q = great_circle(radius, lo, la, rc.mlon, rc.mlat) / c->mountain_r;
            hs = c->mountain_h * std::exp(-(q * q));
          }
          us = c->u0 * m->coslat[d];
          vs = 0.0;
          ghs = gravity * hs;
          ph = gh0 - K * (m->sinlat[d] * m->sinlat[d]) + gravity * bump - ghs;
        }
        phi[d] = ph;
        phis[d] = ghs;
        if (zonal) zonal[d] = us;
        if (merid) merid[d] = vs;
        phiG[d] = ph * m->sqrtG[d];
        const double* g = m->G + 4 * d;
        const double* ia = m->IA + 4 * d;
        const double M11 = g[0] * ia[0] + g[2] * ia[1];
        const double M21 = g[1] * ia[0] + g[3] * ia[1];
        const double M12 = g[0] * ia[2] + g[2] * ia[3];
        const double M22 = g[1] * ia[2] + g[3] * ia[3];
        u[d] = M11 * us + M12 * vs;
        v[d] = M21 * us + M22 * vs;
        uc[d] = ia[0] * us + ia[2] * vs;
        vc[d] = ia[1] * us + ia[3] * vs;
      }
  return CCS_OK;
}
