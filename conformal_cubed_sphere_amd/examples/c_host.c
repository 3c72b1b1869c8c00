/*
 * c_host.c — a C host of the time step through the C ABI alone (include/ccs.h):
 * the mesh and static tables, the seeded benchmark initial state, a handle on
 * one GPU, steps, diagnostics and the state copied back.  It mirrors the
 * reference driver's order of calls (ccsm.F90:20-67: read_mesh, initTestCase,
 * then the RK step loop with diagnostics at each output).
 *
 *   cc -O2 -I include c_host.c -L conformal_cubed_sphere_amd -lccs \
 *      -Wl,-rpath,conformal_cubed_sphere_amd -lm -o c_host
 *   ./c_host [n] [fp32|fp64] [steps]
 *
 * Prints one line: n, precision, steps, relative mass and energy change, the
 * range of h, and ok / the library's error text.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccs.h"

static int check(int rc, const char* where, const ccs_handle* h) {
  if (rc != CCS_OK) fprintf(stderr, "%s failed (%d): %s\n", where, rc, h ? ccs_last_error(h) : "");
  return rc;
}

int main(int argc, char** argv) {
  const int n = argc > 1 ? atoi(argv[1]) : 48;
  const int fp32 = argc > 2 ? strcmp(argv[2], "fp64") != 0 : 1;
  const int steps = argc > 3 ? atoi(argv[3]) : 50;
  const int h = 3;
  const size_t ni = (size_t)n * n * 6, P = (size_t)n + 2 * h, np = P * P * 6;
  int rc = 1;

  /* mesh and static tables (io_mod.F90:15-307) */
  double *lon_deg = calloc(ni, 8), *lat_deg = calloc(ni, 8), *xyz = calloc(3 * ni, 8), *jab = calloc(6 * ni, 8);
  double* t[13];
  for (int k = 0; k < 13; ++k) t[k] = calloc(k >= 9 ? 4 * np : np, 8);
  ccs_metrics m;
  memset(&m, 0, sizeof m);
  m.n = n; m.halo = h; m.radius = 6.371e6; m.omega = 7.292e-5;
  m.lon = t[0]; m.lat = t[1]; m.sinlon = t[2]; m.coslon = t[3]; m.sinlat = t[4]; m.coslat = t[5];
  m.sqrtG = t[6]; m.f = t[7]; m.areaCell = t[8]; m.G = t[9]; m.IG = t[10]; m.A = t[11]; m.IA = t[12];
  if (check(ccs_mesh_geometry(n, lon_deg, lat_deg, xyz, jab), "ccs_mesh_geometry", NULL) ||
      check(ccs_mesh_metrics(lon_deg, lat_deg, jab, &m), "ccs_mesh_metrics", NULL))
    return 1;

  /* initTestCase: the seeded balanced jet with 8 bumps and a mountain */
  ccs_case_params c;
  ccs_case_defaults(&c);
  c.case_num = 7; c.seed = 1; c.n_bumps = 8; c.mountain = 1;
  double* s[9];
  for (int k = 0; k < 9; ++k) s[k] = calloc(np, 8);
  if (check(ccs_init_case(&m, &c, s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7], s[8]), "ccs_init_case", NULL))
    return 1;

  ccs_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.n = n; cfg.halo = h; cfg.precision = fp32 ? CCS_FP32_SPLIT : CCS_FP64_PARITY; cfg.scheme = CCS_RK3_TVD;
  cfg.eflux = CCS_EFLUX_NONE; cfg.device = 0; cfg.use_graph = 1; cfg.spatial_op = CCS_OP_WENO;
  cfg.dt = 20.0 * 48.0 / n; /* below the CFL-0.5 step of this state on the conformal grid */
  cfg.dx = m.dx; cfg.radius = m.radius; cfg.gravity = c.gravity; cfg.omega = m.omega;
  ccs_handle* hd = NULL;
  if (check(ccs_create(&cfg, &hd), "ccs_create", NULL)) return 1;
  double m0 = 0, e0 = 0, m1 = 0, e1 = 0;
  if (!check(ccs_set_mesh(hd, m.sqrtG, m.G, m.IG, m.A, m.IA, m.f, s[8]), "ccs_set_mesh", hd) &&
      !check(ccs_set_state(hd, s[0], s[1], s[2], s[3], s[4], s[5]), "ccs_set_state", hd) &&
      !check(ccs_diagnostics(hd, &m0, &e0), "ccs_diagnostics", hd) &&
      !check(ccs_step(hd, steps), "ccs_step", hd) &&
      !check(ccs_diagnostics(hd, &m1, &e1), "ccs_diagnostics", hd) &&
      !check(ccs_get_state(hd, s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]), "ccs_get_state", hd)) {
    double hmin = 1e30, hmax = -1e30;
    for (int p = 0; p < 6; ++p)
      for (int j = h; j < n + h; ++j)
        for (int i = h; i < n + h; ++i) {
          const double hh = s[3][(size_t)p * P * P + (size_t)j * P + i] / c.gravity;
          if (hh < hmin) hmin = hh;
          if (hh > hmax) hmax = hh;
        }
    printf("n=%d precision=%s steps=%d mass_change=%.3e energy_change=%.3e h=[%.1f, %.1f] ok\n", n,
           fp32 ? "fp32" : "fp64", steps, (m1 - m0) / m0, (e1 - e0) / e0, hmin, hmax);
    rc = 0;
  }
  ccs_destroy(hd);
  for (int k = 0; k < 13; ++k) free(t[k]);
  for (int k = 0; k < 9; ++k) free(s[k]);
  free(lon_deg); free(lat_deg); free(xyz); free(jab);
  return rc;
}
