/*
 * ccs.h — C ABI of the MI355X shallow-water dynamical core on the conformal
 * cubed sphere (libccs.so).
 *
 * The reference (DwyaneChou/Conformal_Cubed_Sphere, dynamic_core/) is a
 * Fortran program with no plugin or FFI layer: its operator interface is the
 * module-procedure set called by the driver ccsm.F90.  Every entry point below
 * names the reference procedure it replaces (file:line).  A Fortran host binds
 * these through iso_c_binding (conformal_cubed_sphere_amd/fortran/ccs_gpu_mod.F90),
 * Python through ctypes (conformal_cubed_sphere_amd/_lib.py).
 *
 * Conventions shared by all entry points
 *  - Every int-returning function returns CCS_OK (0) or a ccs_status code.
 *    Handle-based calls leave a message readable with ccs_last_error().
 *  - Arrays are plain host pointers in the reference layout: Fortran order,
 *    i fastest, then j, then panel.  "padded" arrays cover
 *    i, j in 1-halo .. n+halo and panels 1..6, i.e. (n+2*halo)^2*6 values;
 *    "interior" arrays cover i, j in 1..n, i.e. n*n*6 values.  2x2 matrices
 *    are stored column-major per cell (element (r,c) at r + 2*c + 4*cell),
 *    3x2 Jacobians likewise (r + 3*c + 6*cell).
 *  - Halo corner cells (both i and j outside 1..n) are never filled and never
 *    read, exactly as in the reference (ghost_mod.f90:34-76); the library
 *    leaves them zero.
 *  - A handle owns its device buffers and one HIP stream.  A handle is not
 *    thread-safe; calls come from one host thread.
 */
#ifndef CCS_H
#define CCS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum ccs_status {
  CCS_OK = 0,
  CCS_ERR_INVALID = 1,   /* bad argument: size, null pointer, unknown option   */
  CCS_ERR_DEVICE = 2,    /* HIP runtime failure or no usable device            */
  CCS_ERR_ORDER = 3,     /* call order violated (e.g. step before set_state)   */
  CCS_ERR_NONFINITE = 4, /* NaN or Inf found in the prognostic state           */
  CCS_ERR_NOMEM = 5,     /* host or device allocation failed                   */
  CCS_ERR_SINGULAR = 6   /* singular metric matrix (reference io_mod.F90:177)   */
};

/* Arithmetic of the time-step path.
 *  CCS_FP64_PARITY: FP64 with the reference's operation order kept operation
 *    for operation (full 2x2 metric transforms, the covariant->sphere->covariant
 *    round trip of correct_bdy_ghost, no FMA contraction).  This is the
 *    drop-in path that matches the CPU reference to round-off.
 *  CCS_FP32_SPLIT: FP32 state, fluxes and RK stages on a non-dimensional form
 *    of the same equations; metric terms computed in FP64 once and cast; the
 *    geopotential is stored as an FP32 deviation from an FP64 reference field;
 *    diagnostics in FP64 with compensated sums.
 *  CCS_FP64_FAST: the split-precision kernel instantiated in FP64 (same
 *    algebraic shortcuts as FP32_SPLIT); a throughput comparison point. */
enum ccs_precision { CCS_FP64_PARITY = 0, CCS_FP32_SPLIT = 1, CCS_FP64_FAST = 2 };

/* Time scheme: reference namelist integral_scheme (parameters_mod.F90:129). */
enum ccs_scheme { CCS_RK3_TVD = 0, CCS_RK4 = 1 };

/* Rusanov dissipation of the Bernoulli-function flux.  The reference reads an
 * uninitialised array there (spatial_operators_mod.F90:119,144), which is zero
 * on fresh stack pages: CCS_EFLUX_NONE reproduces that reading. */
enum ccs_eflux { CCS_EFLUX_NONE = 0 };

/* Spatial operator: which of the reference's interchangeable
 * spatial_operators_mod sources the build compiles in (the namelist's
 * reconstruct_scheme, parameters_mod.F90:30, selects none of them; the module
 * file is swapped in the build).  The split precisions (CCS_FP32_SPLIT,
 * CCS_FP64_FAST) run CCS_OP_WENO and the centred operators; CCS_OP_WENO_EXT
 * runs on CCS_FP64_PARITY only.
 *  CCS_OP_WENO      spatial_operators_mod.F90: WENO-Z + Rusanov, face speed
 *                   from the mean of the two cells' G^-1, 6th-order vorticity
 *  CCS_OP_CENTRED2  spatial_operators_mod_2nd.F90: centred 2nd-order
 *                   differences of the cell fluxes, E and winds
 *  CCS_OP_CENTRED4  spatial_operators_mod_4th.F90: centred 4th order
 *  CCS_OP_CENTRED6  spatial_operators_mod_6th.F90: centred 6th order
 *  CCS_OP_WENO_EXT  spatial_operators_mod_WENO.F90: WENO-Z + Rusanov with the
 *                   face speed from mesh%matrixIG_ext (read at lattice point
 *                   (2i+1, j) / (i, 2j+1) as written there, :111-112,134-135)
 *                   and the vorticity from upwind WENO face winds (:118,141,
 *                   160,163); needs ccs_set_mesh_ext. */
enum ccs_spatial_op {
  CCS_OP_WENO = 0,
  CCS_OP_CENTRED2 = 1,
  CCS_OP_CENTRED4 = 2,
  CCS_OP_CENTRED6 = 3,
  CCS_OP_WENO_EXT = 4
};

/* ---------------------------------------------------------- mesh producer */
/* Conformal cubed-sphere cell-centre geometry on the reference panel layout
 * (mesh_generation/src/cuco.f90:100-105 frames, mesh_mod.f90:150-166 output):
 * replaces the offline generator ccs.exe and its mesh file, which the dycore
 * reads in io_mod.F90:37-117.  Interior arrays:
 *   lon_deg, lat_deg : n*n*6 (degrees, atan2 ranges as cart2sph)
 *   xyz              : 3*n*n*6 (unit sphere, x,y,z per cell)
 *   jab              : 6*n*n*6 (d xyz / d(map angle), per radian, 3x2 per cell)
 * Any output pointer may be NULL.  n must be even and >= 4. */
int ccs_mesh_geometry(int n, double* lon_deg, double* lat_deg, double* xyz,
                      double* jab);

/* Static tables of the dycore on the padded grid, computed in FP64 exactly as
 * io_mod.F90:110-304 does (halo fills of ghost_mod.f90:18-160, G = J^T J,
 * Gauss-Jordan inverses of math_mod.f90:451-525, sqrtG = sqrt(det G),
 * A = S(lon,lat) J, f = 2 Omega sin(lat), areaCell = sqrtG dx^2).
 * Caller allocates every non-NULL array: scalars (n+2h)^2*6, matrices 4x that. */
typedef struct ccs_metrics {
  int n, halo;
  double radius, omega;
  double dx;                                 /* out: grid spacing, radians   */
  double *lon, *lat;                         /* radians; corners zero         */
  double *sinlon, *coslon, *sinlat, *coslat;
  double *sqrtG, *f, *areaCell;
  double *G, *IG, *A, *IA;                   /* 2x2 per cell                  */
} ccs_metrics;

int ccs_mesh_metrics(const double* lon_deg, const double* lat_deg,
                     const double* jab, ccs_metrics* m);

/* mesh%matrixIG_ext as io_mod.F90:223-250 computes it: BRINV(J^T J) with
 * J = jab_ext * radius at every extended-lattice point except each panel's
 * four corner points, which the reference leaves unset (zero here).
 * jab_ext: the mesh file's jab, (3,2,2n+1,2n+1,6) per radian
 * (ccs_mesh_file_read_ext), or NULL for this library's own conformal map.
 * matrixIG_ext: caller-allocated 4*(2n+1)^2*6. */
int ccs_mesh_ext_metric(int n, double radius, const double* jab_ext,
                        double* matrixIG_ext);

/* ---------------------------------------------------------- initial state */
/* Initial conditions: Williamson cases 2, 5, 6 of test_case_mod.F90:61-204
 * (case_num 2,5,6) and the seeded balanced jet of the benchmark recipe
 * (case_num 7: u = u0 cos(lat), v = 0, depth in gradient-wind balance with the
 * given area-mean depth, n_bumps Gaussian bumps and an optional Gaussian
 * mountain, centres/amplitudes/radii from a SplitMix64 stream of `seed`).
 * Fills padded arrays like initTestCase (test_case_mod.F90:13-58): phiG = phi
 * sqrtG, covariant u,v = (G A^-1)(us,vs), contravariant uc,vc = A^-1 (us,vs).
 * Any output pointer may be NULL except phiG, u, v, phi, uc, vc, phis. */
typedef struct ccs_case_params {
  int case_num;
  int n_bumps;           /* case 7 */
  int mountain;          /* case 7: 0 flat bottom, 1 Gaussian mountain */
  int pad_;
  unsigned long long seed;
  double gravity;
  double u0;             /* case 7 jet speed, m/s            (20)      */
  double h_mean;         /* case 7 area-mean depth, m        (10000)   */
  double bump_amp;       /* case 7 amplitudes U(-amp, amp), m (100)     */
  double bump_rmin;      /* case 7 e-folding radii U(rmin, rmax), m    */
  double bump_rmax;      /*                              (5e5, 1.5e6)  */
  double mountain_h;     /* case 7 mountain height, m        (2000)    */
  double mountain_r;     /* case 7 mountain e-folding radius (1.5e6)   */
} ccs_case_params;

void ccs_case_defaults(ccs_case_params* c);
int ccs_init_case(const ccs_metrics* m, const ccs_case_params* c,
                  double* phiG, double* u, double* v, double* phi, double* uc,
                  double* vc, double* zonal, double* merid, double* phis);

/* ------------------------------------------------------------- dycore handle */
typedef struct ccs_handle ccs_handle;

typedef struct ccs_config {
  int n, halo;           /* halo >= 3 (WENO-5 and the 6th-order vorticity)    */
  int precision;         /* enum ccs_precision                                */
  int scheme;            /* enum ccs_scheme                                   */
  int eflux;             /* enum ccs_eflux                                    */
  int device;            /* HIP device ordinal                                */
  int use_graph;         /* replay each step from a captured HIP graph        */
  int spatial_op;        /* enum ccs_spatial_op (0: the reference default)    */
  double dt, dx;         /* seconds; radians (ccs_metrics.dx)                 */
  double radius, gravity, omega;
} ccs_config;

/* Allocate device state for one run (replaces initStat/initTend,
 * stat_mod.F90:22-49, tend_mod.F90:17-28). */
int ccs_create(const ccs_config* cfg, ccs_handle** out);

/* Upload the static tables (the padded mesh%sqrtG, matrixG, matrixIG, matrixA,
 * matrixIA, f, phis of mesh_mod.F90:9-48).  f is read on interior cells only. */
int ccs_set_mesh(ccs_handle* h, const double* sqrtG, const double* G,
                 const double* IG, const double* A, const double* IA,
                 const double* f, const double* phis);

/* Upload mesh%matrixIG_ext (mesh_mod.F90:45, io_mod.F90:223-250): the inverse
 * metric on the extended lattice, (2,2,2n+1,2n+1,6) column-major.  Needed by
 * CCS_OP_WENO_EXT only (rejected otherwise). */
int ccs_set_mesh_ext(ccs_handle* h, const double* matrixIG_ext);

/* Upload the prognostic state stat(old) (stat_mod.F90:7-17): padded phiG and
 * covariant u, v.  phi, uc, vc are the diagnostic companions the reference
 * carries; when given they are used by the first stage exactly as the
 * reference's first spatial_operator call uses initTestCase's values, when
 * NULL they are derived (phi = phiG/sqrtG, (uc,vc) = G^-1 (u,v)). */
int ccs_set_state(ccs_handle* h, const double* phiG, const double* u,
                  const double* v, const double* phi, const double* uc,
                  const double* vc);

/* Advance nsteps time steps: replaces the loop body
 * `call RK3_TVD(stat(new), stat(old)); call copyStat(...)`
 * (ccsm.F90:50-66, time_scheme_mod.F90:46-67) or RK4 (:13-44).
 * Asynchronous on the handle's stream.  The split paths fold their FP32
 * geopotential deviation into the FP64 reference field after every step
 * whose count since ccs_set_state is a multiple of 16, so a run's rounding
 * does not depend on how its steps are grouped into calls. */
int ccs_step(ccs_handle* h, int nsteps);

/* Copy the state back (padded).  zonal/merid are the sphere components on
 * interior cells, as convert_wind_P2SP (spatial_operators_mod.F90:364-378)
 * leaves them before a history write (ccsm.F90:58).  Any pointer may be NULL.
 * Synchronises the stream. */
int ccs_get_state(ccs_handle* h, double* phiG, double* u, double* v,
                  double* phi, double* uc, double* vc, double* zonal,
                  double* merid);

/* Tendency L(q) of the current state on interior cells (n*n*6 each):
 * spatial_operator (spatial_operators_mod.F90:13-181). */
int ccs_tendency(ccs_handle* h, double* tend_phiG, double* tend_u,
                 double* tend_v);

/* Total mass sum(phiG) and total energy sum(sqrtG (KE+PE)) over interior
 * cells (diag_mod.F90:10-53) in FP64 with compensated (Neumaier) sums. */
int ccs_diagnostics(ccs_handle* h, double* total_mass, double* total_energy);

/* Time nsteps steps with HIP events on the handle's stream.  total_ms: wall
 * time of the whole loop; kernel_ms (optional): mean duration of one stage
 * kernel in a second pass of nsteps steps, each stage launched with a start
 * and an end event attached to its dispatch (the kernel's own timestamps). */
int ccs_time_steps(ccs_handle* h, int nsteps, double* total_ms,
                   double* kernel_ms);

int ccs_synchronize(ccs_handle* h);
void ccs_destroy(ccs_handle* h);
const char* ccs_last_error(const ccs_handle* h);
const char* ccs_version(void);

/* ------------------------------------------------------------ file formats */
/* Mesh file in the schema of the reference mesh generator
 * (mesh_generation/src/output_mod.f90:25-83), read by the dycore's read_mesh
 * (io_mod.F90:37-117): the conformal map at the (2n+1)^2 extended points of
 * each panel (cell centres at even indices), xi/eta, x/y/z (unit sphere),
 * lon/lat (degrees), jab (per radian), attributes dx, dy, ids..ife.  Written
 * as 64-bit-offset netCDF (the reference writes netCDF-4; nf90_open reads
 * both).  Replaces the generator program ccs.exe. */
int ccs_mesh_file_write(const char* path, int n);
/* Reads a mesh file of that schema (classic or 64-bit-offset netCDF; not
 * netCDF-4/HDF5): n from ids/ide (io_mod.F90:49-52) and the cell-centre
 * lon/lat (degrees) and jab (io_mod.F90:122-135), laid out as
 * ccs_mesh_geometry's outputs.  Output arrays may be NULL (n only). */
int ccs_mesh_file_read(const char* path, int* n, double* lon_deg,
                       double* lat_deg, double* jab);
/* The whole jab variable of such a file (io_mod.F90:107, mesh%jab_ext):
 * (3,2,2n+1,2n+1,6) per radian, for ccs_mesh_ext_metric.  jab_ext may be
 * NULL (n only). */
int ccs_mesh_file_read_ext(const char* path, int* n, double* jab_ext);

/* History file ccsm_output.nc of io_mod.F90:309-541 (64-bit-offset netCDF,
 * as there): lon/lat (degrees, corners _FillValue), areaCell and phis on the
 * padded grid; per output slot the time index and phi, phit = phi + phis, u,
 * v, uc, vc, zonal_wind, meridional_wind (padded arrays).  Replaces
 * history_init / history_write_stat. */
typedef struct ccs_history ccs_history;
typedef struct ccs_history_desc {
  int n, halo, case_num, pad_;
  double dx_deg;         /* grid spacing in degrees (90/n)                    */
  double dt;             /* time step, seconds                                 */
} ccs_history_desc;
int ccs_history_create(const char* path, const ccs_history_desc* d,
                       const double* lon, const double* lat,
                       const double* areaCell, const double* phis,
                       ccs_history** out);
int ccs_history_write(ccs_history* h, int time_slot, const double* phi,
                      const double* phis, const double* u, const double* v,
                      const double* uc, const double* vc, const double* zonal,
                      const double* merid);
int ccs_history_close(ccs_history* h);

/* Total mass and total energy, as ccs_diagnostics computes them on the FP64
 * parity path (calc_total_mass / calc_total_energy, diag_mod.F90:10-53, as
 * compensated FP64 sums; the same kernels, so the same bits), of a padded
 * state held in caller-owned device memory, without allocating anything.  No
 * counterpart in the reference: this entry point is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  phiG, u, v and the static tables sqrtG, the G^-1
 * entries ig11, ig12, ig21, ig22 and phis are device arrays of
 * (n+2*halo)^2*6 doubles in the reference layout; result is a device array of
 * 2 doubles (mass, energy) written on `stream` (a hipStream_t, NULL for the
 * null stream); the call returns once the work is enqueued.  The partial sums
 * live in a 16 KiB device array of the library, one per GPU, shared by every
 * call on that GPU: calls that overlap in time on one GPU race.  Grids beyond
 * 2^31 - 1 padded cells are rejected (CCS_ERR_INVALID); err, when not NULL,
 * receives err_len bytes of text on failure. */
int ccs_diagnostics_noalloc(int n, int halo, const double* phiG, const double* u,
                            const double* v, const double* sqrtG,
                            const double* ig11, const double* ig12,
                            const double* ig21, const double* ig22,
                            const double* phis, double* result, void* stream,
                            char* err, size_t err_len);

/* ------------------------------------------------------- batched ensemble */
/* Many independent runs, each with its own resolution, mesh, initial state
 * and time step, advanced together: one kernel launch per RK stage covers the
 * tiles of every member.  There is no reference counterpart (the reference
 * runs one member per process, ccsm.F90:2-73); each member is exactly the
 * run ccs_create/ccs_step would make with the same inputs.  Split-precision
 * paths only (CCS_FP32_SPLIT, CCS_FP64_FAST). */
typedef struct ccs_ensemble ccs_ensemble;

typedef struct ccs_ensemble_config {
  int n_members;
  int halo;              /* >= 3, shared by all members                      */
  int precision;         /* CCS_FP32_SPLIT or CCS_FP64_FAST                   */
  int scheme;            /* CCS_RK3_TVD or CCS_RK4                            */
  int device;
  int use_graph;         /* replay steps from captured HIP graphs             */
  int tile_y;            /* 8, 16, 32; 0 = chosen by the autotuner            */
  int persistent;        /* 0 one workgroup per tile, 1 resident grid that
                            loops over tiles, -1 = chosen by the autotuner    */
  double radius, gravity, omega;
} ccs_ensemble_config;

/* n, dt, dx: per member (cells per panel edge, seconds, radians). */
int ccs_ensemble_create(const ccs_ensemble_config* cfg, const int* n,
                        const double* dt, const double* dx, ccs_ensemble** out);
/* As ccs_set_mesh / ccs_set_state / ccs_get_state / ccs_diagnostics, for one
 * member (0-based).  Arrays have that member's padded size. */
int ccs_ensemble_set_mesh(ccs_ensemble* e, int member, const double* sqrtG,
                          const double* G, const double* IG, const double* A,
                          const double* IA, const double* f, const double* phis);
int ccs_ensemble_set_state(ccs_ensemble* e, int member, const double* phiG,
                           const double* u, const double* v);
int ccs_ensemble_get_state(ccs_ensemble* e, int member, double* phiG, double* u,
                           double* v, double* phi, double* uc, double* vc,
                           double* zonal, double* merid);
int ccs_ensemble_diagnostics(ccs_ensemble* e, int member, double* mass,
                             double* energy);
/* Advance every member nsteps of its own dt. */
int ccs_ensemble_step(ccs_ensemble* e, int nsteps);
/* Advance member m by nsteps[m] >= 0 steps of its own dt.  All members step
 * in the same batched launches; a member that has taken its steps retires
 * from the tile list (members are ordered by decreasing step count, so the
 * tiles still stepping are a prefix of the list).  Each member rebases its
 * geopotential deviation at its own step counts since its own
 * ccs_ensemble_set_state, as a single run does, so a member ends bit for bit
 * where ccs_step(nsteps[m]) would. */
int ccs_ensemble_step_members(ccs_ensemble* e, const int* nsteps);
/* Advance every member to run_time seconds after its ccs_ensemble_set_state:
 * member m reaches int(run_time / dt[m]) steps, the integer conversion by
 * which the reference derives its step count (parameters_mod.F90:149). */
int ccs_ensemble_run_until(ccs_ensemble* e, double run_time);
/* Steps member m has taken since its last ccs_ensemble_set_state. */
int ccs_ensemble_steps_done(const ccs_ensemble* e, int member, long long* steps);
/* Time tile heights 16/8/32 x {one workgroup per tile, resident grid} on an
 * SSP-RK3 stage-2 launch (it writes only the buffer that holds a stage-2
 * result between the stages of a step, so no state changes) and keep the
 * fastest where the config left the choice open (a plan must be 1 % faster
 * than an earlier one to replace it); reports the choice and its mean stage
 * time.  Runs by
 * itself on the first step call if not called.  The choice is cached per
 * device, precision, scheme, halo, multiset of member sizes and config
 * constraints for the life of the process, and in the file named by the
 * environment variable CCS_AUTOTUNE_CACHE when it is set; a cached choice is
 * taken without timing (stage_ms is then the time recorded with it). */
int ccs_ensemble_autotune(ccs_ensemble* e, int* tile_y, int* persistent,
                          double* stage_ms);
/* 1 if the last ccs_ensemble_autotune took its choice from the cache. */
int ccs_ensemble_autotune_source(const ccs_ensemble* e, int* from_cache);
/* Wall time of nsteps steps measured with HIP events on the ensemble stream. */
int ccs_ensemble_time_steps(ccs_ensemble* e, int nsteps, double* total_ms);
int ccs_ensemble_synchronize(ccs_ensemble* e);
void ccs_ensemble_destroy(ccs_ensemble* e);
const char* ccs_ensemble_last_error(const ccs_ensemble* e);

#ifdef __cplusplus
}
#endif

#endif /* CCS_H */
