/*
 * ccs_oracle.c — CPU FP64 restatement of the reference dynamic-core time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for libccs; see ccs_oracle.h).
 *
 * Every routine follows the cited reference source operation for operation:
 * the same expressions evaluated in the same order with no FMA contraction
 * (built with -ffp-contract=off), so that its results are the ones the
 * reference produces under IEEE FP64 (the -fdefault-real-8 / -r8 build of
 * dynamic_core/CMakeLists.txt:6,12,14).  Where the reference reads memory it
 * never wrote (the Rusanov term of the Bernoulli flux,
 * spatial_operators_mod.F90:119,144) the oracle uses zero, the value of the
 * freshly allocated automatic arrays.  Halo corner cells are never computed.
 *
 * Parallel loops use OpenMP over (panel, row), the reference's own
 * decomposition (spatial_operators_mod.F90:102,127,152): results do not
 * depend on the thread count.
 */
#include "ccs_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PIDX(n, h, i, j, p) \
  ((size_t)((i) + (h)-1) + (size_t)((n) + 2 * (h)) * ((size_t)((j) + (h)-1) + (size_t)((n) + 2 * (h)) * (size_t)((p)-1)))

static int is_corner(int n, int i, int j) { return (i < 1 || i > n) && (j < 1 || j > n); }

void oracle_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* ------------------------------------------------------------------ WENO */
/* spatial_operators_mod.F90:282-361 */
double oracle_weno_limiter(const double* Q, int dir) {
  const double weno_coef[3] = {0.1, 0.6, 0.3};
  const double eps = 1.E-2;
  const double Qim2 = Q[0], Qim1 = Q[1], Qi = Q[2], Qip1 = Q[3], Qip2 = Q[4];
  double stencil[3], coefA[3], coefB[3], beta[3], omega[3], alpha[3];
  if (dir > 0) {
    stencil[0] = Qim2 / 3. - 7. / 6. * Qim1 + 11. / 6. * Qi;
    stencil[1] = -(Qim1 / 6.) + 5. / 6. * Qi + 1. / 3. * Qip1;
    stencil[2] = Qi / 3. + 5. / 6. * Qip1 - 1. / 6. * Qip2;
    coefA[0] = Qim2 - 2. * Qim1 + Qi;
    coefA[1] = Qim1 - 2. * Qi + Qip1;
    coefA[2] = Qi - 2. * Qip1 + Qip2;
    coefB[0] = Qim2 - 4. * Qim1 + 3. * Qi;
    coefB[1] = Qim1 - Qip1;
    coefB[2] = 3. * Qi - 4. * Qip1 + Qip2;
  } else {
    stencil[0] = Qip2 / 3. - 7. / 6. * Qip1 + 11. / 6. * Qi;
    stencil[1] = -(Qip1 / 6.) + 5. / 6. * Qi + 1. / 3. * Qim1;
    stencil[2] = Qi / 3. + 5. / 6. * Qim1 - 1. / 6. * Qim2;
    coefA[0] = Qip2 - 2. * Qip1 + Qi;
    coefA[1] = Qip1 - 2. * Qi + Qim1;
    coefA[2] = Qi - 2. * Qim1 + Qim2;
    coefB[0] = Qip2 - 4. * Qip1 + 3. * Qi;
    coefB[1] = Qip1 - Qim1;
    coefB[2] = 3. * Qi - 4. * Qim1 + Qim2;
  }
  for (int k = 0; k < 3; ++k) beta[k] = coefA[k] * coefA[k] * 13. / 12. + coefB[k] * coefB[k] * 0.25;
  const double tau40 = fabs(beta[0] - beta[1]);
  const double tau41 = fabs(beta[1] - beta[2]);
  const double tau5 = fabs(beta[2] - beta[0]);
  double mb = beta[0];
  if (beta[1] < mb) mb = beta[1];
  if (beta[2] < mb) mb = beta[2];
  if (tau40 <= mb && tau41 > mb) {
    omega[0] = 1. / 3.; omega[1] = 2. / 3.; omega[2] = 0.;
  } else if (tau40 > mb && tau41 < mb) {
    omega[0] = 0.; omega[1] = 2. / 3.; omega[2] = 1. / 3.;
  } else {
    for (int k = 0; k < 3; ++k) alpha[k] = weno_coef[k] * (1. + tau5 / (eps + beta[k]));
    const double s = alpha[0] + alpha[1] + alpha[2];
    for (int k = 0; k < 3; ++k) omega[k] = alpha[k] / s;
  }
  return stencil[0] * omega[0] + stencil[1] * omega[1] + stencil[2] * omega[2];
}

/* spatial_operators_mod.F90:165,169 */
double oracle_d6(const double* f, double dh) {
  return (-f[0] + 9. * f[1] - 45. * f[2] + 45. * f[4] - 9. * f[5] + f[6]) / (60. * dh);
}

/* ------------------------------------------------------------- halo fill */
/* ghost_mod.f90:34-76; source position for destination panel p (1..6),
 * side s (0 W, 1 E, 2 S, 3 N), layer k, position t. */
static void halo_source(int n, int p, int s, int k, int t, int* q, int* i, int* j) {
  static const int tab[6][4][9] = {
      /* q, i0, in, ik, it, j0, jn, jk, jt   (i = i0 + in*n + ik*k + it*t) */
      {{4, 1, 1, -1, 0, 0, 0, 0, 1}, {2, 0, 0, 1, 0, 0, 0, 0, 1}, {6, 0, 0, 0, 1, 1, 1, -1, 0}, {5, 0, 0, 0, 1, 0, 0, 1, 0}},
      {{1, 1, 1, -1, 0, 0, 0, 0, 1}, {3, 0, 0, 1, 0, 0, 0, 0, 1}, {6, 1, 1, -1, 0, 1, 1, 0, -1}, {5, 1, 1, -1, 0, 0, 0, 0, 1}},
      {{2, 1, 1, -1, 0, 0, 0, 0, 1}, {4, 0, 0, 1, 0, 0, 0, 0, 1}, {6, 1, 1, 0, -1, 0, 0, 1, 0}, {5, 1, 1, 0, -1, 1, 1, -1, 0}},
      {{3, 1, 1, -1, 0, 0, 0, 0, 1}, {1, 0, 0, 1, 0, 0, 0, 0, 1}, {6, 0, 0, 1, 0, 0, 0, 0, 1}, {5, 0, 0, 1, 0, 1, 1, 0, -1}},
      {{4, 1, 1, 0, -1, 1, 1, -1, 0}, {2, 0, 0, 0, 1, 1, 1, -1, 0}, {1, 0, 0, 0, 1, 1, 1, -1, 0}, {3, 1, 1, 0, -1, 1, 1, -1, 0}},
      {{4, 0, 0, 0, 1, 0, 0, 1, 0}, {2, 1, 1, 0, -1, 0, 0, 1, 0}, {3, 1, 1, 0, -1, 0, 0, 1, 0}, {1, 0, 0, 0, 1, 0, 0, 1, 0}},
  };
  const int* r = tab[p - 1][s];
  *q = r[0];
  *i = r[1] + r[2] * n + r[3] * k + r[4] * t;
  *j = r[5] + r[6] * n + r[7] * k + r[8] * t;
}

static void halo_dest(int n, int s, int k, int t, int* i, int* j) {
  if (s == 0) { *i = 1 - k; *j = t; }
  else if (s == 1) { *i = n + k; *j = t; }
  else if (s == 2) { *i = t; *j = 1 - k; }
  else { *i = t; *j = n + k; }
}

void oracle_fill_halo(int n, int h, double* f) {
  for (int p = 1; p <= 6; ++p)
    for (int s = 0; s < 4; ++s)
      for (int k = 1; k <= h; ++k)
        for (int t = 1; t <= n; ++t) {
          int q, si, sj, di, dj;
          halo_source(n, p, s, k, t, &q, &si, &sj);
          halo_dest(n, s, k, t, &di, &dj);
          f[PIDX(n, h, di, dj, p)] = f[PIDX(n, h, si, sj, q)];
        }
}

/* CubedSphereFillJab, ghost_mod.f90:84-160: column signs/swaps per strip. */
static void fill_jab(int n, int h, double* jab) {
  /* dst col1 = s1*src col c1, dst col2 = s2*src col c2 */
  static const int sw[6][4][4] = {
      {{0, 1, 1, 1}, {0, 1, 1, 1}, {0, 1, 1, 1}, {0, 1, 1, 1}},
      {{0, 1, 1, 1}, {0, 1, 1, 1}, {1, -1, 0, 1}, {1, 1, 0, -1}},
      {{0, 1, 1, 1}, {0, 1, 1, 1}, {0, -1, 1, -1}, {0, -1, 1, -1}},
      {{0, 1, 1, 1}, {0, 1, 1, 1}, {1, 1, 0, -1}, {1, -1, 0, 1}},
      {{1, 1, 0, -1}, {1, -1, 0, 1}, {0, 1, 1, 1}, {0, -1, 1, -1}},
      {{1, -1, 0, 1}, {1, 1, 0, -1}, {0, -1, 1, -1}, {0, 1, 1, 1}},
  };
  for (int p = 1; p <= 6; ++p)
    for (int s = 0; s < 4; ++s)
      for (int k = 1; k <= h; ++k)
        for (int t = 1; t <= n; ++t) {
          int q, si, sj, di, dj;
          halo_source(n, p, s, k, t, &q, &si, &sj);
          halo_dest(n, s, k, t, &di, &dj);
          const double* src = jab + 6 * PIDX(n, h, si, sj, q);
          double* dst = jab + 6 * PIDX(n, h, di, dj, p);
          const int* c = sw[p - 1][s];
          for (int a = 0; a < 3; ++a) {
            dst[a] = c[1] * src[3 * c[0] + a];
            dst[3 + a] = c[3] * src[3 * c[2] + a];
          }
        }
}

/* ------------------------------------------------------------- metrics */
/* math_mod.f90:451-525 BRINV for N = 2; returns 0 when singular. */
static int brinv2(const double* ain, double* a) {
#define AA(r, c) a[((r)-1) + 2 * ((c)-1)]
  int is[3] = {0, 0, 0}, js[3] = {0, 0, 0};
  for (int e = 0; e < 4; ++e) a[e] = ain[e];
  for (int k = 1; k <= 2; ++k) {
    double d = 0.0, t;
    for (int i = k; i <= 2; ++i)
      for (int j = k; j <= 2; ++j)
        if (fabs(AA(i, j)) > d) { d = fabs(AA(i, j)); is[k] = i; js[k] = j; }
    if (d + 1.0 == 1.0) return 0;
    for (int j = 1; j <= 2; ++j) { t = AA(k, j); AA(k, j) = AA(is[k], j); AA(is[k], j) = t; }
    for (int i = 1; i <= 2; ++i) { t = AA(i, k); AA(i, k) = AA(i, js[k]); AA(i, js[k]) = t; }
    AA(k, k) = 1 / AA(k, k);
    for (int j = 1; j <= 2; ++j)
      if (j != k) AA(k, j) = AA(k, j) * AA(k, k);
    for (int i = 1; i <= 2; ++i)
      if (i != k)
        for (int j = 1; j <= 2; ++j)
          if (j != k) AA(i, j) = AA(i, j) - AA(i, k) * AA(k, j);
    for (int i = 1; i <= 2; ++i)
      if (i != k) AA(i, k) = -AA(i, k) * AA(k, k);
  }
  for (int k = 2; k >= 1; --k) {
    double t;
    for (int j = 1; j <= 2; ++j) { t = AA(k, j); AA(k, j) = AA(js[k], j); AA(js[k], j) = t; }
    for (int i = 1; i <= 2; ++i) { t = AA(i, k); AA(i, k) = AA(i, is[k]); AA(i, is[k]) = t; }
  }
#undef AA
  return 1;
}

/* io_mod.F90:110-304 */
int oracle_metrics(int n, int h, double radius, double omega, const double* lon_deg,
                   const double* lat_deg, const double* jab_in, double* dx_out, double* lon_o,
                   double* lat_o, double* sinlat_o, double* coslat_o, double* sqrtG_o, double* f_o,
                   double* area_o, double* G_o, double* IG_o, double* A_o, double* IA_o) {
  const double pi = 2. * asin(1.);
  const double D2R = pi / 180.;
  const double R2D = 180. / pi;
  const size_t P = (size_t)(n + 2 * h), np = P * P * 6;
  double dx = 90. / n;
  dx = dx * D2R;
  *dx_out = dx;
  double* lon = calloc(np, sizeof(double));
  double* lat = calloc(np, sizeof(double));
  double* sinlat = calloc(np, sizeof(double));
  double* coslat = calloc(np, sizeof(double));
  double* jab = calloc(6 * np, sizeof(double));
  int status = 0;
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        size_t c = (size_t)(i - 1) + (size_t)n * ((size_t)(j - 1) + (size_t)n * (size_t)(p - 1));
        size_t d = PIDX(n, h, i, j, p);
        lon[d] = lon_deg[c] * D2R;
        lat[d] = lat_deg[c] * D2R;
        for (int a = 0; a < 6; ++a) jab[6 * d + a] = jab_in[6 * c + a] * radius;
      }
  for (size_t d = 0; d < np; ++d) {
    sinlat[d] = sin(lat[d]);
    coslat[d] = cos(lat[d]);
  }
  oracle_fill_halo(n, h, lon);
  oracle_fill_halo(n, h, lat);
  oracle_fill_halo(n, h, sinlat);
  oracle_fill_halo(n, h, coslat);
  fill_jab(n, h, jab);
  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - h; j <= n + h; ++j)
      for (int i = 1 - h; i <= n + h; ++i) {
        size_t d = PIDX(n, h, i, j, p);
        double G[4] = {0, 0, 0, 0}, IG[4] = {0, 0, 0, 0}, A[4] = {0, 0, 0, 0}, IA[4] = {0, 0, 0, 0};
        double sg = 0;
        int corner = is_corner(n, i, j);
        if (!corner) {
          double lo = lon[d], la = lat[d];
          double dlondx = -sin(lo), dlatdx = -(cos(lo) * sin(la));
          double dlondy = cos(lo), dlatdy = -(sin(lo) * sin(la));
          double dlondz = 0., dlatdz = cos(la);
          double S[2][3] = {{dlondx, dlondy, dlondz}, {dlatdx, dlatdy, dlatdz}};
          const double* J = jab + 6 * d;
          for (int c = 0; c < 2; ++c)
            for (int r = 0; r < 2; ++r) {
              double g = 0.;
              g = g + J[0 + 3 * r] * J[0 + 3 * c];
              g = g + J[1 + 3 * r] * J[1 + 3 * c];
              g = g + J[2 + 3 * r] * J[2 + 3 * c];
              G[r + 2 * c] = g;
              double aa = 0.;
              aa = aa + S[r][0] * J[0 + 3 * c];
              aa = aa + S[r][1] * J[1 + 3 * c];
              aa = aa + S[r][2] * J[2 + 3 * c];
              A[r + 2 * c] = aa;
            }
          if (!brinv2(G, IG)) status = 6;
          sg = sqrt(G[0] * G[3] - G[1] * G[2]);
          if (fabs(la * R2D) != 90.) {
            if (!brinv2(A, IA)) status = 6;
          } else if (la * R2D == 90.) {
            IA[0] = cos(lo) * radius; IA[1] = sin(lo) * radius; IA[2] = -sin(lo) * radius; IA[3] = cos(lo) * radius;
          } else {
            IA[0] = -cos(lo) * radius; IA[1] = sin(lo) * radius; IA[2] = sin(lo) * radius; IA[3] = cos(lo) * radius;
          }
        }
        if (lon_o) lon_o[d] = corner ? 0. : lon[d];
        if (lat_o) lat_o[d] = corner ? 0. : lat[d];
        if (sinlat_o) sinlat_o[d] = corner ? 0. : sinlat[d];
        if (coslat_o) coslat_o[d] = corner ? 0. : coslat[d];
        if (sqrtG_o) sqrtG_o[d] = sg;
        if (area_o) area_o[d] = sg * (dx * dx);
        if (f_o) f_o[d] = (i >= 1 && i <= n && j >= 1 && j <= n) ? 2. * omega * sinlat[d] : 0.;
        for (int e = 0; e < 4; ++e) {
          if (G_o) G_o[4 * d + e] = G[e];
          if (IG_o) IG_o[4 * d + e] = IG[e];
          if (A_o) A_o[4 * d + e] = A[e];
          if (IA_o) IA_o[4 * d + e] = IA[e];
        }
      }
  free(lon); free(lat); free(sinlat); free(coslat); free(jab);
  return status;
}

/* ------------------------------------------------------------ test cases */
static uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double uniform53(uint64_t* s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

static double gc_dist(double radius, double lon1, double lat1, double lon2, double lat2) {
  double c = sin(lat1) * sin(lat2) + cos(lat1) * cos(lat2) * cos(lon1 - lon2);
  if (c > 1.) c = 1.;
  if (c < -1.) c = -1.;
  return radius * acos(c);
}

/* test_case_mod.F90:13-204 (cases 2, 5, 6) and the seeded recipe (case 7). */
int oracle_init_case(int n, int h, double radius, double Omega, double gravity, int case_num,
                     unsigned long long seed, int n_bumps, int mountain, double u0j, double h_mean,
                     double bump_amp, double bump_rmin, double bump_rmax, double mountain_h,
                     double mountain_r, const double* lon, const double* lat, const double* sinlat,
                     const double* coslat, const double* sqrtG, const double* G, const double* IA,
                     double* phiG, double* u, double* v, double* phi, double* uc, double* vc,
                     double* zonal, double* merid, double* phis) {
  const double pi = 2. * asin(1.);
  const size_t P = (size_t)(n + 2 * h), np = P * P * 6;
  /* seeded recipe parameters */
  double bl[64], bt[64], ba[64], br[64], mlon = 0, mlat = 0;
  if (case_num == 7) {
    if (n_bumps < 0 || n_bumps > 64) return 1;
    uint64_t st = (uint64_t)seed;
    for (int b = 0; b < n_bumps; ++b) {
      double zz = 2. * uniform53(&st) - 1.;
      bl[b] = 2. * pi * uniform53(&st);
      bt[b] = asin(zz);
      ba[b] = bump_amp * (2. * uniform53(&st) - 1.);
      br[b] = bump_rmin + (bump_rmax - bump_rmin) * uniform53(&st);
    }
    double zz = 2. * uniform53(&st) - 1.;
    mlon = 2. * pi * uniform53(&st);
    mlat = asin(zz);
  } else if (case_num != 2 && case_num != 5 && case_num != 6) {
    return 1;
  }
  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - h; j <= n + h; ++j)
      for (int i = 1 - h; i <= n + h; ++i) {
        size_t d = PIDX(n, h, i, j, p);
        if (is_corner(n, i, j)) {
          phiG[d] = u[d] = v[d] = phi[d] = uc[d] = vc[d] = phis[d] = 0.;
          if (zonal) zonal[d] = 0.;
          if (merid) merid[d] = 0.;
          continue;
        }
        double us = 0, vs = 0, ph = 0, ghs = 0;
        if (case_num == 2) {
          const double gh0 = 29400.;
          double u0 = 2. * pi * radius / (12. * 86400.);
          ph = gh0 - (radius * Omega * u0 + u0 * u0 / 2.) * (sinlat[d] * sinlat[d]);
          us = u0 * coslat[d];
          vs = 0.;
          ghs = 0.;
        } else if (case_num == 5) {
          const double hs0 = 2000., u0 = 20., alpha = 0., gh0 = 5960. * gravity;
          double rr = pi / 9., lc = 1.5 * pi, tc = pi / 6.;
          double lo = lon[d], la = lat[d];
          if (lo < 0) lo = 2. * pi + lo;
          double dd = (lo - lc) * (lo - lc) + (la - tc) * (la - tc);
          double r = sqrt(rr * rr < dd ? rr * rr : dd);
          ghs = gravity * hs0 * (1. - r / rr);
          us = u0 * (cos(la) * cos(alpha) + cos(lo) * sin(la) * sin(alpha));
          vs = -(u0 * sin(lo) * sin(alpha));
          double X = -(cos(lo) * cos(la) * sin(alpha)) + sin(la) * cos(alpha);
          ph = gh0 - (radius * Omega * u0 + u0 * u0 / 2.) * (X * X) - ghs;
        } else if (case_num == 6) {
          const double omg = 7.848e-6, R = 4., h0 = 8000.;
          double s = sinlat[d], c = coslat[d], lo = lon[d];
          double u1 = c;
          double u2 = R * pow(c, R - 1) * (s * s) * cos(R * lo);
          double u3 = pow(c, R + 1) * cos(R * lo);
          us = radius * omg * (u1 + u2 - u3);
          vs = -(radius * omg * R * pow(c, R - 1) * s * sin(R * lo));
          double AA1 = omg * 0.5 * (2. * Omega + omg) * (c * c);
          double Ac = 0.25 * (omg * omg);
          double A21 = (R + 1.) * pow(c, 2. * R + 2.);
          double A22 = (2. * (R * R) - R - 2.) * pow(c, 2. * R);
          double A23 = 2. * (R * R) * pow(c, 2. * R - 2);
          double Ah = AA1 + Ac * (A21 + A22 - A23);
          double Bc = 2. * (Omega + omg) * omg / ((R + 1) * (R + 2)) * pow(c, R);
          double BB1 = R * R + 2. * R + 2.;
          double BB2 = ((R + 1.) * (R + 1.)) * (c * c);
          double Bh = Bc * (BB1 - BB2);
          double CC = 0.25 * (omg * omg) * pow(c, 2. * R);
          double CC1 = (R + 1.) * (c * c);
          double CC2 = R + 2.;
          double Ch = CC * (CC1 - CC2);
          ph = gravity * h0 + (radius * radius) * (Ah + Bh * cos(R * lo) + Ch * cos(2.0 * R * lo));
          ghs = 0.;
        } else {
          double K = radius * Omega * u0j + u0j * u0j / 2.;
          double gh0 = gravity * h_mean + K / 3.;
          double lo = lon[d], la = lat[d];
          double bump = 0.;
          for (int b = 0; b < n_bumps; ++b) {
            double dist = gc_dist(radius, lo, la, bl[b], bt[b]);
            double q = dist / br[b];
            bump = bump + ba[b] * exp(-(q * q));
          }
          double hs = 0.;
          if (mountain) {
            double dist = gc_dist(radius, lo, la, mlon, mlat);
            double q = dist / mountain_r;
            hs = mountain_h * exp(-(q * q));
          }
          us = u0j * coslat[d];
          vs = 0.;
          ghs = gravity * hs;
          ph = gh0 - K * (sinlat[d] * sinlat[d]) + gravity * bump - ghs;
        }
        phi[d] = ph;
        phis[d] = ghs;
        if (zonal) zonal[d] = us;
        if (merid) merid[d] = vs;
        phiG[d] = ph * sqrtG[d];
        /* covProjSphere2Plane with matrix = matmul(G, IA) (projection_mod.F90:60-77) */
        const double* g = G + 4 * d;
        const double* ia = IA + 4 * d;
        double M11 = g[0] * ia[0] + g[2] * ia[1];
        double M21 = g[1] * ia[0] + g[3] * ia[1];
        double M12 = g[0] * ia[2] + g[2] * ia[3];
        double M22 = g[1] * ia[2] + g[3] * ia[3];
        u[d] = M11 * us + M12 * vs;
        v[d] = M21 * us + M22 * vs;
        /* contravProjSphere2Plane (projection_mod.F90:98-110) */
        uc[d] = ia[0] * us + ia[2] * vs;
        vc[d] = ia[1] * us + ia[3] * vs;
      }
  (void)np;
  return 0;
}

/* ------------------------------------------------------- spatial operator */
void oracle_spatial_operator(const oracle_mesh* m, const oracle_stat* s, double* t_phiG, double* t_u,
                             double* t_v) {
  const int n = m->n, h = m->h;
  const size_t P = (size_t)(n + 2 * h), np = P * P * 6;
  const double dx = m->dx, dy = m->dx;
  double* phit = malloc(np * sizeof(double));
  double* phitG = malloc(np * sizeof(double));
  double* E = malloc(np * sizeof(double));
  double* Fx = malloc(np * sizeof(double));
  double* Fy = malloc(np * sizeof(double));
  /* face values: x-faces (i+1/2, j) for i = 0..n; y-faces (i, j+1/2), j = 0..n */
  const size_t nf = (size_t)(n + 1) * (size_t)n * 6;
  double* fxF = malloc(nf * sizeof(double));
  double* fxE = malloc(nf * sizeof(double));
  double* fyF = malloc(nf * sizeof(double));
  double* fyE = malloc(nf * sizeof(double));

  /* pointwise terms, spatial_operators_mod.F90:71-78 */
#pragma omp parallel for schedule(static)
  for (long d = 0; d < (long)np; ++d) {
    phit[d] = s->phi[d] + m->phis[d];
    phitG[d] = phit[d] * m->sqrtG[d];
    double K = 0.5 * (s->u[d] * s->uc[d] + s->v[d] * s->vc[d]);
    E[d] = phit[d] + K;
    Fx[d] = s->phiG[d] * s->uc[d];
    Fy[d] = s->phiG[d] * s->vc[d];
  }

  if (m->op >= 1 && m->op <= 3) {
    /* centred differences of the cell values, spatial_operators_mod_2nd.F90:46-67,
     * _4th.F90:46-71, _6th.F90:46-77 (same pointwise terms as above) */
    const int op = m->op;
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 1; p <= 6; ++p)
      for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= n; ++i) {
          size_t d = PIDX(n, h, i, j, p);
          double fx[7], ex[7], vx[7], fy[7], ey[7], uy[7];
          for (int a = 0; a < 7; ++a) {
            if (a == 3 || (op == 1 && (a == 0 || a == 1 || a == 5 || a == 6)) || (op == 2 && (a == 0 || a == 6)))
              continue;
            size_t dxi = PIDX(n, h, i - 3 + a, j, p), dyj = PIDX(n, h, i, j - 3 + a, p);
            fx[a] = Fx[dxi]; ex[a] = E[dxi]; vx[a] = s->v[dxi];
            fy[a] = Fy[dyj]; ey[a] = E[dyj]; uy[a] = s->u[dyj];
          }
          double dfluxdx, dEdx, dvdx, dfluxdy, dEdy, dudy;
          if (op == 1) {
            dfluxdx = (fx[4] - fx[2]) / (2. * dx);
            dEdx = (ex[4] - ex[2]) / (2. * dx);
            dvdx = (vx[4] - vx[2]) / (2. * dx);
            dfluxdy = (fy[4] - fy[2]) / (2. * dy);
            dEdy = (ey[4] - ey[2]) / (2. * dy);
            dudy = (uy[4] - uy[2]) / (2. * dy);
          } else if (op == 2) {
            dfluxdx = (fx[1] - 8. * fx[2] + 8. * fx[4] - fx[5]) / (12. * dx);
            dEdx = (ex[1] - 8. * ex[2] + 8. * ex[4] - ex[5]) / (12. * dx);
            dvdx = (vx[1] - 8. * vx[2] + 8. * vx[4] - vx[5]) / (12. * dx);
            dfluxdy = (fy[1] - 8. * fy[2] + 8. * fy[4] - fy[5]) / (12. * dy);
            dEdy = (ey[1] - 8. * ey[2] + 8. * ey[4] - ey[5]) / (12. * dy);
            dudy = (uy[1] - 8. * uy[2] + 8. * uy[4] - uy[5]) / (12. * dy);
          } else {
            dfluxdx = oracle_d6(fx, dx);
            dEdx = oracle_d6(ex, dx);
            dvdx = oracle_d6(vx, dx);
            dfluxdy = oracle_d6(fy, dy);
            dEdy = oracle_d6(ey, dy);
            dudy = oracle_d6(uy, dy);
          }
          double vort = dvdx - dudy + m->sqrtG[d] * m->f[d];
          t_phiG[d] = -(dfluxdx + dfluxdy);
          t_u[d] = -dEdx + vort * s->vc[d];
          t_v[d] = -dEdy - vort * s->uc[d];
        }
    free(phit); free(phitG); free(E); free(Fx); free(Fy);
    free(fxF); free(fxE); free(fyF); free(fyE);
    return;
  }
  /* op 4: face-lattice G^-1 and upwind face winds (spatial_operators_mod_WENO.F90) */
  const int ext = m->op == 4;
  const size_t me = (size_t)(2 * n + 1);
  /* mesh%matrixIG_ext(r,c,a,b,p), 1-based indices */
#define IGX(r, c, a, b, p) m->IG_ext[((r)-1) + 2 * ((c)-1) + 4 * (((size_t)(a)-1) + me * (((size_t)(b)-1) + me * ((size_t)(p)-1)))]
  double* fxV = ext ? malloc(nf * sizeof(double)) : NULL;
  double* fyU = ext ? malloc(nf * sizeof(double)) : NULL;

  /* x faces, spatial_operators_mod.F90:80-124 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= n; ++j)
      for (int i = 0; i <= n; ++i) {
        double qF[5], qE[5], qP[5], rF[5], rE[5], rP[5];
        for (int a = 0; a < 5; ++a) {
          size_t dl = PIDX(n, h, i - 2 + a, j, p), dr = PIDX(n, h, i - 1 + a, j, p);
          qF[a] = Fx[dl]; qE[a] = E[dl]; qP[a] = phitG[dl];
          rF[a] = Fx[dr]; rE[a] = E[dr]; rP[a] = phitG[dr];
        }
        double Fp = oracle_weno_limiter(qF, 1), Fn = oracle_weno_limiter(rF, -1);
        double Ep = oracle_weno_limiter(qE, 1), En = oracle_weno_limiter(rE, -1);
        double Pp = oracle_weno_limiter(qP, 1), Pn = oracle_weno_limiter(rP, -1);
        size_t d0 = PIDX(n, h, i, j, p), d1 = PIDX(n, h, i + 1, j, p);
        double uBdy = 0.5 * (s->uc[d1] + s->uc[d0]);
        double phitBdy = 0.5 * (phit[d1] + phit[d0]);
        /* op 4 reads matrixIG_ext(1,1,2i+1,j) -- row j, not 2j, as written at
         * spatial_operators_mod_WENO.F90:111-112 */
        double g = ext ? IGX(1, 1, 2 * i + 1, j, p) * phitBdy
                       : 0.5 * (m->IG[4 * d1 + 0] + m->IG[4 * d0 + 0]) * phitBdy;
        double e1 = uBdy - sqrt(g), e2 = uBdy + sqrt(g);
        double lam = fabs(e1) > fabs(e2) ? fabs(e1) : fabs(e2);
        size_t f = (size_t)i + (size_t)(n + 1) * ((size_t)(j - 1) + (size_t)n * (size_t)(p - 1));
        fxF[f] = 0.5 * (Fp + Fn - lam * (Pn - Pp));
        fxE[f] = 0.5 * (Ep + En - lam * (0. - 0.)); /* u_ext_n - u_ext_p: never written */
        if (ext) { /* v_ext, spatial_operators_mod_WENO.F90:88-89,118 */
          double qV[5], rV[5];
          for (int a = 0; a < 5; ++a) {
            qV[a] = s->v[PIDX(n, h, i - 2 + a, j, p)];
            rV[a] = s->v[PIDX(n, h, i - 1 + a, j, p)];
          }
          double Vp = oracle_weno_limiter(qV, 1), Vn = oracle_weno_limiter(rV, -1);
          fxV[f] = 0.5 * (Vp + Vn - copysign(1., uBdy) * (Vn - Vp));
        }
      }

  /* y faces, spatial_operators_mod.F90:90-99,126-149 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int p = 1; p <= 6; ++p)
    for (int j = 0; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        double qF[5], qE[5], qP[5], rF[5], rE[5], rP[5];
        for (int a = 0; a < 5; ++a) {
          size_t dl = PIDX(n, h, i, j - 2 + a, p), dr = PIDX(n, h, i, j - 1 + a, p);
          qF[a] = Fy[dl]; qE[a] = E[dl]; qP[a] = phitG[dl];
          rF[a] = Fy[dr]; rE[a] = E[dr]; rP[a] = phitG[dr];
        }
        double Fp = oracle_weno_limiter(qF, 1), Fn = oracle_weno_limiter(rF, -1);
        double Ep = oracle_weno_limiter(qE, 1), En = oracle_weno_limiter(rE, -1);
        double Pp = oracle_weno_limiter(qP, 1), Pn = oracle_weno_limiter(rP, -1);
        size_t d0 = PIDX(n, h, i, j, p), d1 = PIDX(n, h, i, j + 1, p);
        double vBdy = 0.5 * (s->vc[d1] + s->vc[d0]);
        double phitBdy = 0.5 * (phit[d1] + phit[d0]);
        /* op 4 reads matrixIG_ext(2,2,i,2j+1) -- column i, not 2i, as written at
         * spatial_operators_mod_WENO.F90:134-135 */
        double g = ext ? IGX(2, 2, i, 2 * j + 1, p) * phitBdy
                       : 0.5 * (m->IG[4 * d1 + 3] + m->IG[4 * d0 + 3]) * phitBdy;
        double e1 = vBdy - sqrt(g), e2 = vBdy + sqrt(g);
        double lam = fabs(e1) > fabs(e2) ? fabs(e1) : fabs(e2);
        size_t f = (size_t)(i - 1) + (size_t)n * ((size_t)j + (size_t)(n + 1) * (size_t)(p - 1));
        fyF[f] = 0.5 * (Fp + Fn - lam * (Pn - Pp));
        fyE[f] = 0.5 * (Ep + En - lam * (0. - 0.)); /* v_ext_n - v_ext_p: never written */
        if (ext) { /* u_ext, spatial_operators_mod_WENO.F90:98-99,141 */
          double qU[5], rU[5];
          for (int a = 0; a < 5; ++a) {
            qU[a] = s->u[PIDX(n, h, i, j - 2 + a, p)];
            rU[a] = s->u[PIDX(n, h, i, j - 1 + a, p)];
          }
          double Up = oracle_weno_limiter(qU, 1), Un = oracle_weno_limiter(rU, -1);
          fyU[f] = 0.5 * (Up + Un - copysign(1., vBdy) * (Un - Up));
        }
      }

  /* divergence, vorticity and tendencies, spatial_operators_mod.F90:151-179 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        size_t d = PIDX(n, h, i, j, p);
        size_t fe = (size_t)i + (size_t)(n + 1) * ((size_t)(j - 1) + (size_t)n * (size_t)(p - 1));
        size_t fw = fe - 1;
        size_t fn = (size_t)(i - 1) + (size_t)n * ((size_t)j + (size_t)(n + 1) * (size_t)(p - 1));
        size_t fs = fn - (size_t)n;
        double vx[7], uy[7];
        for (int a = 0; a < 7; ++a) {
          vx[a] = s->v[PIDX(n, h, i - 3 + a, j, p)];
          uy[a] = s->u[PIDX(n, h, i, j - 3 + a, p)];
        }
        double dfluxdx = (fxF[fe] - fxF[fw]) / dx;
        double dEdx = (fxE[fe] - fxE[fw]) / dx;
        double dvdx = ext ? (fxV[fe] - fxV[fw]) / dx : oracle_d6(vx, dx);   /* _WENO.F90:160 */
        double dfluxdy = (fyF[fn] - fyF[fs]) / dy;
        double dEdy = (fyE[fn] - fyE[fs]) / dy;
        double dudy = ext ? (fyU[fn] - fyU[fs]) / dy : oracle_d6(uy, dy);   /* _WENO.F90:163 */
        double vort = dvdx - dudy + m->sqrtG[d] * m->f[d];
        t_phiG[d] = -(dfluxdx + dfluxdy);
        t_u[d] = -dEdx + vort * s->vc[d];
        t_v[d] = -dEdy - vort * s->uc[d];
      }
  free(phit); free(phitG); free(E); free(Fx); free(Fy);
  free(fxF); free(fxE); free(fyF); free(fyE); free(fxV); free(fyU);
#undef IGX
}

/* --------------------------------------------------------- ghost update */
void oracle_correct_bdy_ghost(const oracle_mesh* m, oracle_stat* s) {
  const int n = m->n, h = m->h;
  /* convert_wind_P2SP on interior (spatial_operators_mod.F90:364-378) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        size_t d = PIDX(n, h, i, j, p);
        const double* A = m->A + 4 * d;
        const double* IG = m->IG + 4 * d;
        /* matrix = matmul(A, IG) (projection_mod.F90:79-96) */
        double M11 = A[0] * IG[0] + A[2] * IG[1];
        double M21 = A[1] * IG[0] + A[3] * IG[1];
        double M12 = A[0] * IG[2] + A[2] * IG[3];
        double M22 = A[1] * IG[2] + A[3] * IG[3];
        s->zonal[d] = M11 * s->u[d] + M12 * s->v[d];
        s->merid[d] = M21 * s->u[d] + M22 * s->v[d];
      }
  /* fill_ghost (ghost_mod.f90:163-170) */
  oracle_fill_halo(n, h, s->phiG);
  oracle_fill_halo(n, h, s->zonal);
  oracle_fill_halo(n, h, s->merid);
  /* convert_wind_SP2P and convert_wind_cov2contrav on the padded array
   * (spatial_operators_mod.F90:380-410), then phi = phiG/sqrtG
   * (time_scheme_mod.F90:81) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - h; j <= n + h; ++j)
      for (int i = 1 - h; i <= n + h; ++i) {
        if (is_corner(n, i, j)) continue;
        size_t d = PIDX(n, h, i, j, p);
        const double* G = m->G + 4 * d;
        const double* IA = m->IA + 4 * d;
        const double* IG = m->IG + 4 * d;
        /* matrix = matmul(G, IA) (projection_mod.F90:60-77) */
        double M11 = G[0] * IA[0] + G[2] * IA[1];
        double M21 = G[1] * IA[0] + G[3] * IA[1];
        double M12 = G[0] * IA[2] + G[2] * IA[3];
        double M22 = G[1] * IA[2] + G[3] * IA[3];
        double uu = M11 * s->zonal[d] + M12 * s->merid[d];
        double vv = M21 * s->zonal[d] + M22 * s->merid[d];
        s->u[d] = uu;
        s->v[d] = vv;
        /* cov2contrav (projection_mod.F90:138-148) */
        s->uc[d] = IG[0] * uu + IG[2] * vv;
        s->vc[d] = IG[1] * uu + IG[3] * vv;
        s->phi[d] = s->phiG[d] / m->sqrtG[d];
      }
}

/* ------------------------------------------------------------ time step */
static void copy_stat(size_t np, oracle_stat* dst, const oracle_stat* src) {
  memcpy(dst->phiG, src->phiG, np * sizeof(double));
  memcpy(dst->u, src->u, np * sizeof(double));
  memcpy(dst->v, src->v, np * sizeof(double));
  memcpy(dst->phi, src->phi, np * sizeof(double));
  memcpy(dst->uc, src->uc, np * sizeof(double));
  memcpy(dst->vc, src->vc, np * sizeof(double));
  memcpy(dst->zonal, src->zonal, np * sizeof(double));
  memcpy(dst->merid, src->merid, np * sizeof(double));
  if (dst->phit && src->phit) memcpy(dst->phit, src->phit, np * sizeof(double));
}

static int alloc_stat(size_t np, oracle_stat* s) {
  double* blk = calloc(8 * np, sizeof(double));
  if (!blk) return 0;
  s->phiG = blk; s->u = blk + np; s->v = blk + 2 * np; s->phi = blk + 3 * np;
  s->uc = blk + 4 * np; s->vc = blk + 5 * np; s->zonal = blk + 6 * np; s->merid = blk + 7 * np;
  s->phit = NULL;
  return 1;
}

/* kind 0: new = old + inc*t (update_stat, :69-83)
 * kind 1: new = 0.75 old + 0.25 s1 + 0.25 dt t (update_stat_RK3_TVD_1, :85-99)
 * kind 2: new = old/3 + 2/3 s2 + 2/3 dt t (update_stat_RK3_TVD_2, :101-115) */
static void update(const oracle_mesh* m, int kind, double inc, oracle_stat* nw, const oracle_stat* old,
                   const oracle_stat* mid, const double* t0, const double* t1, const double* t2) {
  const int n = m->n, h = m->h;
  const double dt = m->dt;
#pragma omp parallel for collapse(2) schedule(static)
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        size_t d = PIDX(n, h, i, j, p);
        if (kind == 0) {
          nw->phiG[d] = old->phiG[d] + inc * t0[d];
          nw->u[d] = old->u[d] + inc * t1[d];
          nw->v[d] = old->v[d] + inc * t2[d];
        } else if (kind == 1) {
          nw->phiG[d] = 0.75 * old->phiG[d] + 0.25 * mid->phiG[d] + 0.25 * dt * t0[d];
          nw->u[d] = 0.75 * old->u[d] + 0.25 * mid->u[d] + 0.25 * dt * t1[d];
          nw->v[d] = 0.75 * old->v[d] + 0.25 * mid->v[d] + 0.25 * dt * t2[d];
        } else {
          nw->phiG[d] = old->phiG[d] / 3. + 2. / 3. * mid->phiG[d] + 2. / 3. * dt * t0[d];
          nw->u[d] = old->u[d] / 3. + 2. / 3. * mid->u[d] + 2. / 3. * dt * t1[d];
          nw->v[d] = old->v[d] / 3. + 2. / 3. * mid->v[d] + 2. / 3. * dt * t2[d];
        }
      }
  oracle_correct_bdy_ghost(m, nw);
}

int oracle_steps(const oracle_mesh* m, oracle_stat* s, int nsteps, int scheme) {
  const int n = m->n, h = m->h;
  const size_t P = (size_t)(n + 2 * h), np = P * P * 6;
  oracle_stat s2, s3, s4, snew;
  if (!alloc_stat(np, &s2) || !alloc_stat(np, &s3) || !alloc_stat(np, &s4) || !alloc_stat(np, &snew)) return 5;
  double* tb = calloc(15 * np, sizeof(double));
  if (!tb) return 5;
  double* T[5][3];
  for (int a = 0; a < 5; ++a)
    for (int b = 0; b < 3; ++b) T[a][b] = tb + (size_t)(3 * a + b) * np;
  copy_stat(np, &snew, s);
  for (int it = 0; it < nsteps; ++it) {
    if (scheme == 0) {
      /* RK3_TVD, time_scheme_mod.F90:46-67 */
      oracle_spatial_operator(m, s, T[0][0], T[0][1], T[0][2]);
      update(m, 0, m->dt, &s2, s, NULL, T[0][0], T[0][1], T[0][2]);
      oracle_spatial_operator(m, &s2, T[1][0], T[1][1], T[1][2]);
      update(m, 1, 0., &s3, s, &s2, T[1][0], T[1][1], T[1][2]);
      oracle_spatial_operator(m, &s3, T[2][0], T[2][1], T[2][2]);
      update(m, 2, 0., &snew, s, &s3, T[2][0], T[2][1], T[2][2]);
    } else {
      /* RK4, time_scheme_mod.F90:13-44 */
      const double half = 0.5 * m->dt;
      oracle_spatial_operator(m, s, T[0][0], T[0][1], T[0][2]);
      update(m, 0, half, &s2, s, NULL, T[0][0], T[0][1], T[0][2]);
      oracle_spatial_operator(m, &s2, T[1][0], T[1][1], T[1][2]);
      update(m, 0, half, &s3, s, NULL, T[1][0], T[1][1], T[1][2]);
      oracle_spatial_operator(m, &s3, T[2][0], T[2][1], T[2][2]);
      update(m, 0, m->dt, &s4, s, NULL, T[2][0], T[2][1], T[2][2]);
      oracle_spatial_operator(m, &s4, T[3][0], T[3][1], T[3][2]);
#pragma omp parallel for schedule(static)
      for (long d = 0; d < (long)np; ++d)
        for (int b = 0; b < 3; ++b)
          T[4][b][d] = (T[0][b][d] + 2. * T[1][b][d] + 2. * T[2][b][d] + T[3][b][d]) / 6.;
      update(m, 0, m->dt, &snew, s, NULL, T[4][0], T[4][1], T[4][2]);
    }
    copy_stat(np, s, &snew); /* copyStat(stat(old), stat(new)), ccsm.F90:66 */
  }
  free(tb);
  free(s2.phiG); free(s3.phiG); free(s4.phiG); free(snew.phiG);
  return 0;
}

/* ----------------------------------------------------------- diagnostics */
void oracle_diagnostics(const oracle_mesh* m, const oracle_stat* s, double* mass_seq, double* energy_seq,
                        double* mass_comp, double* energy_comp) {
  const int n = m->n, h = m->h;
  double ms = 0., es = 0.;
  __float128 mq = 0, eq = 0;
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        size_t d = PIDX(n, h, i, j, p);
        double KE = 0.5 * s->phi[d] * (s->u[d] * s->uc[d] + s->v[d] * s->vc[d]);
        double pt = s->phi[d] + m->phis[d];
        double PE = 0.5 * (pt * pt);
        double e = m->sqrtG[d] * (KE + PE);
        ms = ms + s->phiG[d];
        es = es + e;
        mq += (__float128)s->phiG[d];
        eq += (__float128)e;
      }
  if (mass_seq) *mass_seq = ms;
  if (energy_seq) *energy_seq = es;
  if (mass_comp) *mass_comp = (double)mq;
  if (energy_comp) *energy_comp = (double)eq;
}
