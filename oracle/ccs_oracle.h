/*
 * ccs_oracle.h — CPU FP64 restatement of the reference dynamic-core time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for libccs: only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load it.  It is never linked into or called by the product path.
 *
 * Layout: identical to include/ccs.h (Fortran order, padded arrays of
 * (n+2h)^2*6 values, 2x2 matrices column-major per cell).
 */
#ifndef CCS_ORACLE_H
#define CCS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Spatial operator: the module the reference build compiles in.
 *   0  spatial_operators_mod.F90       WENO-Z, Rusanov speed from the mean of
 *                                      the two cells' G^-1, 6th-order vorticity
 *   1  spatial_operators_mod_2nd.F90   centred 2nd-order differences
 *   2  spatial_operators_mod_4th.F90   centred 4th-order differences
 *   3  spatial_operators_mod_6th.F90   centred 6th-order differences
 *   4  spatial_operators_mod_WENO.F90  WENO-Z, Rusanov speed from the
 *                                      extended-lattice G^-1, upwind WENO face
 *                                      winds in the vorticity */
typedef struct oracle_mesh {
  int n, h;
  double dx, dt;
  const double *sqrtG, *G, *IG, *A, *IA, *f, *phis; /* padded */
  int op, pad_;
  /* op 4: mesh%matrixIG_ext, (2,2,2n+1,2n+1,6) column-major (mesh_mod.F90:45) */
  const double* IG_ext;
} oracle_mesh;

/* The nine fields of stat_field (stat_mod.F90:7-17), padded. */
typedef struct oracle_stat {
  double *phiG, *u, *v, *phi, *uc, *vc, *zonal, *merid, *phit;
} oracle_stat;

/* WENO_limiter (spatial_operators_mod.F90:282-361) on one 5-point window. */
double oracle_weno_limiter(const double* q5, int dir);

/* 6th-order centred first derivative (spatial_operators_mod.F90:165). */
double oracle_d6(const double* f7, double dh);

/* read_mesh metric section (io_mod.F90:110-304); arrays as ccs_metrics. */
int oracle_metrics(int n, int h, double radius, double omega,
                   const double* lon_deg, const double* lat_deg,
                   const double* jab, double* dx_out, double* lon,
                   double* lat, double* sinlat, double* coslat,
                   double* sqrtG, double* f, double* areaCell, double* G,
                   double* IG, double* A, double* IA);

/* initTestCase (test_case_mod.F90:13-58) for cases 2, 5, 6 and the seeded
 * balanced-jet recipe (case 7, same parameters as ccs_case_params). */
int oracle_init_case(int n, int h, double radius, double omega,
                     double gravity, int case_num, unsigned long long seed,
                     int n_bumps, int mountain, double u0, double h_mean,
                     double bump_amp, double bump_rmin, double bump_rmax,
                     double mountain_h, double mountain_r, const double* lon,
                     const double* lat, const double* sinlat,
                     const double* coslat, const double* sqrtG,
                     const double* G, const double* IA, double* phiG,
                     double* u, double* v, double* phi, double* uc,
                     double* vc, double* zonal, double* merid, double* phis);

/* spatial_operator: interior tendencies written into padded arrays. */
void oracle_spatial_operator(const oracle_mesh* m, const oracle_stat* s,
                             double* t_phiG, double* t_u, double* t_v);

/* correct_bdy_ghost followed by phi = phiG/sqrtG (time_scheme_mod.F90:117-125,81). */
void oracle_correct_bdy_ghost(const oracle_mesh* m, oracle_stat* s);

/* CubedSphereFillHalo (ghost_mod.f90:18-82) on one padded scalar field. */
void oracle_fill_halo(int n, int h, double* field);

/* nsteps of RK3_TVD (scheme 0) or RK4 (scheme 1) advancing `s` in place,
 * i.e. the loop body RK*(stat(new), stat(old)); copyStat (ccsm.F90:50-66). */
int oracle_steps(const oracle_mesh* m, oracle_stat* s, int nsteps, int scheme);

/* calc_total_mass / calc_total_energy (diag_mod.F90:10-53): plain
 * sequential sums as the reference, and exactly rounded sums (sum of the
 * same terms evaluated with a compensated accumulator). */
void oracle_diagnostics(const oracle_mesh* m, const oracle_stat* s,
                        double* mass_seq, double* energy_seq,
                        double* mass_comp, double* energy_comp);

void oracle_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif

#endif
