// ccs_kernels.h — device-side data structures and kernel entry points of the
// MI355X time-step path.  Included by the HIP translation units only.
#ifndef CCS_KERNELS_H
#define CCS_KERNELS_H

#include <hip/hip_runtime.h>

namespace ccs {

// Scatter rules of the edge-halo exchange (see ccs_internal.h ScatterRule),
// flattened for kernel arguments: [q-1][sigma][11] =
// {p, ei0, eii, eij, ej0, eji, ejj, c1, s1, c2, s2}.
struct HaloScatter {
  int r[6][4][11];
};

// Static FP64 tables of the parity path, one padded array each.
//   M = matmul(A, IG)  (covProjPlane2Sphere, projection_mod.F90:90)
//   N = matmul(G, IA)  (covProjSphere2Plane, projection_mod.F90:71)
struct StaticsF64 {
  const double* sqrtG;
  const double* IG11;
  const double* IG12;
  const double* IG21;
  const double* IG22;
  const double* phis;
  const double* f;
  const double* M11;
  const double* M12;
  const double* M21;
  const double* M22;
  const double* N11;
  const double* N12;
  const double* N21;
  const double* N22;
};

template <typename T>
struct Prog {  // prognostic fields (padded)
  T* phiG;
  T* u;
  T* v;
};

// Stage kinds of the fused kernel.
enum StageKind : int {
  kTendOnly = 0,  // write L(q) on interior cells
  kRK3s1 = 1,     // q1 = qn + dt L(qn)                       time_scheme_mod.F90:75-77
  kRK3s2 = 2,     // q2 = 3/4 qn + 1/4 q1 + 1/4 dt L(q1)        :91-93
  kRK3s3 = 3,     // qn+1 = qn/3 + 2/3 q2 + 2/3 dt L(q2)        :107-109
  kRK4a = 4,      // q = qn + inc L(q_in), also stores L        :26,31,34
  kRK4f = 5       // qn+1 = qn + dt (L1 + 2 L2 + 2 L3 + L4)/6   :38-42
};

struct StageArgsF64 {
  Prog<const double> in;    // state the tendency is evaluated on
  Prog<const double> base;  // qn
  Prog<double> out;
  const double* phi0;       // explicit phi/uc/vc for the first stage after upload
  const double* uc0;
  const double* vc0;
  double* tend[3];          // kTendOnly / kRK4a output (interior layout n*n*6)
  const double* tacc[9];    // kRK4f: L1, L2, L3 (phiG,u,v each)
  int n, h;
  int op;                   // enum ccs_spatial_op
  double dx, dt, inc;
  // set by launch_stage_f64: RN(1/dx), RN(1/(60 dx)), 2 dx, RN(1/(2 dx)),
  // RN(12 dx), RN(1/RN(12 dx))
  double rdx, r60dx, dx2, rdx2, dx12, rdx12;
  // CCS_OP_WENO_EXT face speeds: matrixIG_ext(1,1,2i+1,j) at x-face (i+1/2, j)
  // for i = 0..n, j = 1..n: [i + (n+1)(j-1 + n p)]; matrixIG_ext(2,2,i,2j+1) at
  // y-face (i, j+1/2) for i = 1..n, j = 0..n: [i-1 + n(j + (n+1) p)]
  const double* igx;
  const double* igy;
  // 1: each XCD takes its contiguous run of tiles backwards.  Consecutive
  // stage launches alternate, so every XCD stays on its own run (its L2) and
  // starts on the tiles it reached last (the most recent lines in the
  // memory-side cache).  Results do not depend on the order.
  unsigned rev;
};

// tile_y: rows per tile, 8 or 7 (stage_f64_tile_y: 7 where the tiles of an
// n x n panel fit one round of workgroups on `cus` compute units)
int stage_f64_tile_y(int n, int cus);
// ev: NULL, or a start and an end event attached to the dispatch itself
void launch_stage_f64(hipStream_t s, int kind, bool explicit_derived, const StageArgsF64& a,
                      const StaticsF64& st, const HaloScatter& hs, int tile_y, const hipEvent_t* ev = nullptr);

// Derived fields for output on the padded grid (non-corner cells):
// phi = phiG/sqrtG, (uc,vc) = IG (u,v); zonal/merid = M (u,v) on interior.
void launch_derived_f64(hipStream_t s, const Prog<const double>& q, const StaticsF64& st, int n,
                        int h, double* phi, double* uc, double* vc, double* zonal, double* merid);

// Compensated FP64 sums of mass and energy over interior cells.
// partials must hold 2*kDiagBlocks*2 doubles; result[0] mass, result[1] energy.
constexpr int kDiagBlocks = 512;
void launch_diag_f64(hipStream_t s, const Prog<const double>& q, const double* phi,
                     const double* uc, const double* vc, const StaticsF64& st, int n, int h,
                     double* partials, double* result);

}  // namespace ccs

#endif
