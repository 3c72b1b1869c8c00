// ccs_split.hip — split-precision time step (FP32 state, metric tables
// computed in FP64 once and cast) and its FP64 instantiation.
//
// Same equations and the same WENO-Z / Rusanov / 6th-order-vorticity method as
// the reference spatial operator (spatial_operators_mod.F90:13-181), written
// in a non-dimensional form that keeps FP32 in range:
//   P = phiG / a^2   (= phi * sqrtg, sqrtg = sqrtG / a^2 the unit-sphere Jacobian)
//   U = u / a, V = v / a   (covariant winds), tau = t / a.
// On a conformal grid G = sqrtG I, so the contravariant winds are U/sqrtg,
// V/sqrtg, G^-1_11 = G^-1_22 = 1/sqrtG, and the covariant->sphere->covariant
// round trip of correct_bdy_ghost is the identity.  The equations become
//   dP/dtau = -(dFx/dx + dFy/dy),        Fx = P Uc, Fy = P Vc
//   dU/dtau = -dE/dx + zeta Vc,          E = phi + phis + (U Uc + V Vc)/2
//   dV/dtau = -dE/dy - zeta Uc,          zeta = dV/dx - dU/dy + a sqrtg f
// and the Rusanov speed is |Ub| + sqrt(gb pb) with gb the face mean of
// 1/sqrtg.  The WENO-Z epsilon scales with the square of each variable's
// scale so the nonlinear weights equal the reference's: 1e-2/a^2 for Fx, Fy,
// 1e-2 for E, 1e-2/a^4 for phitG/a^2.
//
// The geopotential is carried as an FP32 deviation dP from an FP64 reference
// field Pref; the RK stage combinations have weights summing to one, so they
// act on dP alone and the FP32 increments are not lost against the ~1e5
// magnitude of P.  Pref is the uploaded state, and every 16 steps the host
// folds dP into it (split_rebase): a deviation that kept growing would round
// each increment against an ever larger value and lose mass.
//
// Kernel structure (one launch per RK stage, as the FP64 path):
//   phase 1  pointwise terms on the tile and a 3-cell cross border -> LDS
//   phase 2  WENO windows: the smoothness indicators of each 5-point window
//            are computed once and feed both reconstructions the window
//            gives (the value at its centre cell's right face and at its left
//            face; the reference recomputes them per face,
//            spatial_operators_mod.F90:80-99).  Windows are evaluated on
//            differences from their centre value, so reconstructions come
//            out as small deviations; the Rusanov flux of a face,
//            F = (F_L + F_R)/2 - lambda (P_R - P_L)/2, splits into a centre
//            term (F_A + F_B)/2 - lambda (P_B - P_A)/2 and the share each side's
//            deviations contribute.  Every window writes its own share, no
//            separate flux pass (or barrier) is needed, and no sum or
//            difference of large values enters the divergence: in FP32 the
//            flux differences keep their relative precision (in the absolute
//            form they lose 3-5 digits to cancellation and the round-off
//            drives long runs unstable).
//   phase 3  divergence, vorticity, RK update, halo scatter with the signed
//            permutation of the panel-basis change (ghost_mod.f90:100-153).
// The centred operators of spatial_operators_mod_{2nd,4th,6th}.F90 are the
// OP = 1/2/3 instantiations: no phase 2, and phase 3 takes centred
// differences of the phase-1 fields.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "ccs_internal.h"
#include "ccs_share.h"
#include "ccs_split.h"

namespace ccs {
namespace {

// Tile: TX x TY interior cells per workgroup, one thread per cell.  TX is
// fixed at 32 (half a wavefront per row); TY is a template parameter.
constexpr int TX = 32;
constexpr int EX = TX + 6;

__device__ __forceinline__ long gidx(int i, int j, int p, int h, int P) {
  return static_cast<long>(i + h) + static_cast<long>(P) * (static_cast<long>(j + h) + static_cast<long>(P) * p);
}

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
// a * b + c with one rounding.  The face values and Rusanov speeds are written
// with explicit FMAs and no implicit contraction: the tile-boundary windows
// (weno_side) and the interior ones (weno_faces), the two copies of the
// pointwise terms (a tile's own cells and its border) and every instance of
// lam_x / lam_y then round alike, so a face value does not depend on which
// tile computes it, and results do not depend on the tile height.
__device__ __forceinline__ float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Loads and stores at a 32-bit byte offset from a uniform base pointer: the
// compiler then addresses every array of a cell with one shared VGPR offset
// (global_load ... vOff, sBase) instead of a 64-bit address per array.  The
// host guarantees that every array of the split state is below 4 GiB.
template <typename T>
__device__ __forceinline__ T ldo(const T* p, unsigned off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + off);
}
template <typename T>
__device__ __forceinline__ void sto(T* p, unsigned off, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + off) = v;
}

template <typename T>
struct SplitStatics {
  const T* pref;  // FP32 copy of the reference P (padded)
  const T* sg;    // sqrtG / a^2 (padded)
  const T* phis;  // surface geopotential (padded)
  const T* fs;    // a * sqrtg * f (interior cells of the padded layout)
};

template <typename T>
struct SplitArgs {
  Prog<const T> in, base;
  Prog<T> out;
  T* tend[3];
  const T* tacc[9];
  int n, h;
  int nbx, nby;        // tiles per panel along x and y
  unsigned mbx, mby;   // division by nbx, nby: q = hi32((2 t) mb), exact for t < 2^20 (tile_magic)
  T rdx;               // 1 / dx
  T dts, incs;         // dt / a, RK4 increment / a
  // WENO-Z epsilons of Fx/Fy, E (windows scaled by 1/2: times 1/16) and of
  // phitG/a^2 (scaled by 1/4: times 1/64)
  T epsF, epsE, epsP;
  unsigned rev;        // 1: each XCD takes its run of tiles backwards (xcd_tile_dir)
};

// One 5-point window (q0..q4, centre q2): the WENO-Z value at the centre
// cell's right face (dir > 0 of WENO_limiter) and at its left face (dir < 0),
// sharing the smoothness indicators.  Weight rule of
// spatial_operators_mod.F90:348-355; for the mirrored (left-face) window the
// indicators reverse order, which swaps tau40 and tau41.
//
// With a_k, b_k the second and first differences of the Jiang-Shu
// indicators, every candidate stencil is q2 + a_k/12 +- b_k/4 (q2 dropped:
// values are deviations from the window centre):
//   right face: s1 = q2 + a1/12 + b1/4, s2 = q2 + a2/12 - b2/4, s3 = q2 + a3/12 - b3/4
//   left face : s1 = q2 + a3/12 + b3/4, s2 = q2 + a2/12 + b2/4, s3 = q2 + a1/12 - b1/4
// and the branch weights (1/3,2/3,0), (0,2/3,1/3) are the unnormalised
// weights (1,2,0), (0,2,1) fed through the same normalisation.
//
// Scaled window (SC = 2 or 4): the candidates, hence the returned face
// values, are 1/SC of the true deviations, so a flux share 0.5 (F + lam P)
// becomes one FMA of the scaled values.  The window's inputs come scaled by
// 1/(4 SC) (the phase-1 fields are stored that way): then the first
// differences b_k / (4 SC) are a_k - 2 d_k and d1 - d3 of the scaled inputs,
// one operation each.  The indicators come out multiplied by (1/(2 SC))^2 and
// the caller passes epsilon multiplied by the same power of two, so
// tau5 / (eps + beta_k), and with it every weight, is that of the unscaled
// window.
template <typename T, int SC>
struct WenoWin {
  T a1, a2, a3;     // second differences (relative to q2), of the scaled inputs
  T q1, q2, q3;     // first differences b_k / (4 SC)
  // WENO-Z weights 10 d_k (1 + tau5 / (eps + beta_k)), 10 d = (1, 6, 3): the
  // right face takes (u1, u2, u3), the mirrored left face (v3, u2, v1), with
  // v3 = u3 / 3 and v1 = 3 u1
  T u1, u2, u3, v1, v3;
  T t40, t41, mb;   // indicator differences and min (weight-rule branches)
};

// The window is evaluated relative to its centre value q2: the inputs are
// the differences q_k - q2 (exact in FP32 for neighbouring values of one
// sign, by Sterbenz' lemma) and the reconstructions are returned as
// deviations from q2.  The weights only see differences, so they are those
// of the absolute form; the absolute face value is q2 + deviation.  q0..q4
// are the window's values times 1/(4 SC).
template <typename T, int SC>
__device__ __forceinline__ WenoWin<T, SC> weno_window(T q0, T q1, T q2, T q3, T q4, T eps_scaled) {
#pragma clang fp contract(off)
  WenoWin<T, SC> w;
  const T d0 = q0 - q2, d1 = q1 - q2, d3 = q3 - q2, d4 = q4 - q2;
  w.a1 = fmat(T(-2), d1, d0);
  w.a2 = d1 + d3;
  w.a3 = fmat(T(-2), d3, d4);
  // b1 = a1 - 2 d1, b2 = d1 - d3, b3 = a3 - 2 d3 (times 1/(4 SC) with the inputs)
  w.q1 = fmat(T(-2), d1, w.a1);
  w.q2 = d1 - d3;
  w.q3 = fmat(T(-2), d3, w.a3);
  // beta_k / (2 SC)^2 = 13/12 a_k^2 / (2 SC)^2 + (b_k / (4 SC))^2; with the
  // scaled a_k that is 13/3 a_k^2 + q_k^2
  const T k13 = T(13) / T(3);
  const T be1 = fmat(k13 * w.a1, w.a1, w.q1 * w.q1);
  const T be2 = fmat(k13 * w.a2, w.a2, w.q2 * w.q2);
  const T be3 = fmat(k13 * w.a3, w.a3, w.q3 * w.q3);
  w.t40 = fabs(be1 - be2);
  w.t41 = fabs(be2 - be3);
  const T t5 = fabs(be3 - be1);
  w.mb = fmin(be1, fmin(be2, be3));
  // (d1, d2, d3) = (0.1, 0.6, 0.3) times 10, which the normalisation removes;
  // the factors 3 and 6 multiply, since fma cannot take them as an inline
  // addend
  const T r1 = rcp(eps_scaled + be1), r2 = rcp(eps_scaled + be2), r3 = rcp(eps_scaled + be3);
  w.u1 = fmat(t5, r1, T(1));
  w.v3 = fmat(t5, r3, T(1));
  w.u2 = T(6) * fmat(t5, r2, T(1));
  w.u3 = T(3) * w.v3;
  w.v1 = T(3) * w.u1;
  return w;
}

// The branch cases of the weight rule (spatial_operators_mod.F90:348-355):
// the right face takes weights (1/3, 2/3, 0) when tau40 <= mb < tau41 and
// (0, 2/3, 1/3) when tau41 < mb < tau40; the mirrored left face the same with
// tau40 and tau41 swapped.  In each of the four cases exactly one of
// tau40 <= mb, tau41 <= mb holds, so where both tests agree no case applies
// (also with NaN, every compare then being false).  The cases occur only at
// local extrema of the indicators, so the exact tests sit behind that one
// test, in a branch that whole wavefronts usually skip.  The branch weights
// (1/3, 2/3, 0), (0, 2/3, 1/3) enter as the unnormalised (1, 2, 0), (0, 2, 1).
template <typename T>
__device__ __forceinline__ void weno_branch_weights(T t40, T t41, T mb, T& w1, T& w2, T& w3) {
  const bool b1 = (t40 <= mb) && (t41 > mb);
  const bool b2 = (t40 > mb) && (t41 < mb);
  if (b1 || b2) {
    w1 = b1 ? T(1) : T(0);
    w2 = T(2);
    w3 = b1 ? T(0) : T(1);
  }
}

// normalised combination of the three candidates
template <typename T>
__device__ __forceinline__ T weno_combine(T s1, T s2, T s3, T w1, T w2, T w3) {
#pragma clang fp contract(off)
  return fmat(w3, s3, fmat(w2, s2, w1 * s1)) * rcp(w1 + w2 + w3);
}

// values at the right and the left face of the window's centre cell (the
// left from the mirrored window), times 1/SC
template <typename T, int SC>
__device__ __forceinline__ void weno_faces(const WenoWin<T, SC>& w, T& right, T& left) {
#pragma clang fp contract(off)
  constexpr T kc = T(1) / T(3);  // a_k / (12 SC) of the inputs scaled by 1/(4 SC)
  T r1 = w.u1, r2 = w.u2, r3 = w.u3, l1 = w.v3, l2 = w.u2, l3 = w.v1;
  if ((w.t40 <= w.mb) != (w.t41 <= w.mb)) {
    // an empty volatile asm keeps this a branch (skipped by whole
    // wavefronts) instead of being if-converted into per-lane selects
    asm volatile("");
    weno_branch_weights(w.t40, w.t41, w.mb, r1, r2, r3);
    weno_branch_weights(w.t41, w.t40, w.mb, l1, l2, l3);
  }
  right = weno_combine(fmat(w.a1, kc, w.q1), fmat(w.a2, kc, -w.q2), fmat(w.a3, kc, -w.q3), r1, r2, r3);
  left = weno_combine(fmat(w.a3, kc, w.q3), fmat(w.a2, kc, w.q2), fmat(w.a1, kc, -w.q1), l1, l2, l3);
}

// either face, chosen per lane (left = true: the left face), times 1/SC
template <typename T, int SC>
__device__ __forceinline__ T weno_side(const WenoWin<T, SC>& w, bool left) {
#pragma clang fp contract(off)
  constexpr T kc = T(1) / T(3);
  // the candidates of weno_faces, operand for operand
  const T s1 = left ? fmat(w.a3, kc, w.q3) : fmat(w.a1, kc, w.q1);
  const T s2 = fmat(w.a2, kc, left ? w.q2 : -w.q2);
  const T s3 = left ? fmat(w.a1, kc, -w.q1) : fmat(w.a3, kc, -w.q3);
  T w1 = left ? w.v3 : w.u1, w2 = w.u2, w3 = left ? w.v1 : w.u3;
  if ((w.t40 <= w.mb) != (w.t41 <= w.mb)) {
    asm volatile("");
    weno_branch_weights(left ? w.t41 : w.t40, left ? w.t40 : w.t41, w.mb, w1, w2, w3);
  }
  return weno_combine(s1, s2, s3, w1, w2, w3);
}

template <typename T>
__device__ __forceinline__ T d6f(T m3, T m2, T m1, T p1, T p2, T p3, T rdx) {
#pragma clang fp contract(fast)
  return ((p3 - m3) - T(9) * (p2 - m2) + T(45) * (p1 - m1)) * (rdx * (T(1) / T(60)));
}

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs, each with its own L2, so workgroup b runs on XCD b % 8.  Each XCD is
// given a contiguous run of tiles (x fastest, then y, then panel, then
// member) so the border rows and columns that neighbouring tiles share are
// fetched into one L2 instead of up to three.
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
  const unsigned xq = nb >> 3, xr = nb & 7u, xcd = b & 7u;
  return xcd * xq + min(xcd, xr) + (b >> 3);
}

// The same order with each XCD's run of tiles taken backwards when rev is
// set.  Consecutive stage launches alternate, so every XCD stays on its own
// run (what its L2 may still hold helps small grids) and starts on the tiles
// it reached last, whose outputs and statics are the most recent lines in the
// memory-side cache (that helps where a stage's working set exceeds the
// cache: C1440 and large ensembles, about 3-4 %).  Results do not depend on
// the order: every cell and halo cell is written by one thread of one tile.
// (Reversing the whole list instead put each XCD on another XCD's run and
// cost 10 % at C180 in FP32.)
__device__ __forceinline__ unsigned xcd_tile_dir(unsigned b, unsigned nb, unsigned rev) {
  const unsigned xq = nb >> 3, xr = nb & 7u, xcd = b & 7u;
  const unsigned len = xq + (xcd < xr ? 1u : 0u), j = b >> 3;  // workgroup b is number j of len on its XCD
  return xcd * xq + min(xcd, xr) + (rev ? len - 1u - j : j);
}

// One RK stage on one 32 x TY tile of panel p with lower-left interior cell
// (i0, j0); hr = the panel's four halo-scatter rules.  OP: the spatial
// operator (enum ccs_spatial_op): 0 WENO-Z / Rusanov, 1/2/3 the centred
// 2nd/4th/6th-order differences of the reference's swap-in operator modules,
// which skip phase 2.
template <typename T, int KIND, int TY, int OP>
__device__ __forceinline__ void stage_tile(const SplitArgs<T>& a, const SplitStatics<T>& st, const int (*hr)[11],
                                           const int p, const int i0, const int j0, const int tid) {
#pragma clang fp contract(fast)
  constexpr int NT = TX * TY;
  constexpr int EY = TY + 6;
  constexpr int NE = EX * EY;
  constexpr unsigned W = sizeof(T);
  // cell fields on the extended tile: 0 Fx, 1 Fy, 2 E, 3 phitG, 4 1/sqrtg,
  // 5 U, 6 V (phit = phitG / sqrtg is formed where needed; seven fields keep
  // a 16-row tile within a quarter of the CU's LDS)
  __shared__ T sc[7 * NE];
  T* const sFx = sc;
  T* const sFy = sc + NE;
  T* const sE = sc + 2 * NE;
  T* const sPt = sc + 3 * NE;
  T* const sRg = sc + 4 * NE;
  T* const sU = sc + 5 * NE;
  T* const sV = sc + 6 * NE;
  // face shares: x-faces [TY][TX+1], y-faces [TY+1][TX]; per face the share
  // of the flux of phitG*u (F) and of E from its left/lower side (p) and its
  // right/upper side (m)
  constexpr bool kWeno = (OP == 0);
  constexpr int NFXF = TY * (TX + 1);
  constexpr int NFYF = (TY + 1) * TX;
  __shared__ T sfc[kWeno ? 4 * NFXF + 4 * NFYF : 1];
  T* const xpF = sfc;
  T* const xmF = sfc + NFXF;
  T* const xpE = sfc + 2 * NFXF;
  T* const xmE = sfc + 3 * NFXF;
  T* const ypF = sfc + 4 * NFXF;
  T* const ymF = ypF + NFYF;
  T* const ypE = ypF + 2 * NFYF;
  T* const ymE = ypF + 3 * NFYF;

  const int n = a.n, h = a.h;
  const unsigned P = static_cast<unsigned>(n + 2 * h);
  // scale of the stored E, Fx, Fy (kF) and phitG (kP): 1/(4 SC) of the
  // windows that read them (SC = 2 and 4); the centred operators store them
  // unscaled
  constexpr T kF = kWeno ? T(1) / T(8) : T(1);
  constexpr T kP = kWeno ? T(1) / T(16) : T(1);
  // padded index of the extended tile's (0, 0) cell (i0 - 3, j0 - 3); h >= 3
  const unsigned gt = static_cast<unsigned>(p) * P * P + static_cast<unsigned>(j0 - 3 + h) * P +
                      static_cast<unsigned>(i0 - 3 + h);

  // ---- phase 1: pointwise terms of the extended tile -> LDS.  Every
  // thread takes its own cell; the 6 TX cells of the y-border rows and the
  // 6 TY cells of the x-border columns go to the highest threads, so that
  // they do not fall on the waves that evaluate the boundary windows of
  // phase 2 (the lowest threads).  Both cells' loads are issued before
  // either is used (a cell outside the padded panel, or in a halo corner,
  // which is never read, loads from the tile origin and is not stored), so a
  // thread with two cells waits for memory once.
  struct Raw {
    T pr, dp, u, v, sg, ps;
    int l;       // extended-tile index, -1: not stored
  };
  auto fetch = [&](const int li, const int lj, const bool want) {
    const int i = i0 + li - 3, j = j0 + lj - 3;
    const bool ok = want && i >= -h && i < n + h && j >= -h && j < n + h && !((i < 0 || i >= n) && (j < 0 || j >= n));
    const unsigned go = ok ? (gt + __umul24(static_cast<unsigned>(lj), P) + static_cast<unsigned>(li)) * W
                           : (gt + 3u * P + 3u) * W;
    Raw r;
    r.pr = ldo(st.pref, go);
    r.dp = ldo(a.in.phiG, go);
    r.u = ldo(a.in.u, go);
    r.v = ldo(a.in.v, go);
    r.sg = ldo(st.sg, go);
    r.ps = ldo(st.phis, go);
    r.l = ok ? lj * EX + li : -1;
    return r;
  };
  auto pointwise = [&](const Raw& r) {
#pragma clang fp contract(off)
    if (r.l < 0) return;
    const T Pv = r.pr + r.dp;
    const T U = r.u, V = r.v;
    const T rg = rcp(r.sg);
    // for the WENO windows E, Fx, Fy are stored times kF = 1/8 and phitG
    // times kP = 1/16 (see weno_window; the surface geopotential static is
    // uploaded times kF): the contravariant winds carry the factor kF
    const T rgc = rg * kF;
    const T Uc = U * rgc, Vc = V * rgc;
    const T phitc = fmat(Pv, rgc, r.ps);
    const int l = r.l;
    sE[l] = fmat(T(0.5), fmat(U, Uc, V * Vc), phitc);
    sFx[l] = Pv * Uc;
    sFy[l] = Pv * Vc;
    sPt[l] = phitc * (r.sg * (kP / kF));
    sRg[l] = rg;
    sU[l] = U;
    sV[l] = V;
  };
  {
    constexpr int NYB = 6 * TX, NB = NYB + 6 * TY;
    static_assert(NB <= NT, "one pass of border cells");
    const int k = tid - (NT - NB);
    int bli, blj;
    if (k < NYB) {
      const int r = k / TX;
      bli = 3 + k % TX;
      blj = r < 3 ? r : TY + r;
    } else {
      const int kk = k - NYB, c = kk % 6;
      bli = c < 3 ? c : TX + c;
      blj = 3 + kk / 6;
    }
    const Raw own = fetch((tid % TX) + 3, (tid / TX) + 3, true);
    const Raw border = fetch(bli, blj, k >= 0);
    pointwise(own);
    pointwise(border);
  }
  __syncthreads();

  // Twice the Rusanov speed at the x-face between extended-tile cells l and
  // l + 1, and at the y-face between l and l + EX
  // (spatial_operators_mod.F90:109-116, 134-141, on a conformal grid): the
  // face means are left as sums, 2 |ub| + sqrt(4 gb pb) exactly (the stored
  // phitG carries kP = 1/16, so its square root a factor 1/4).
  constexpr T kRootP = T(4);
  auto lam_x = [&](int l) {
#pragma clang fp contract(off)
    const T r0 = sRg[l], r1 = sRg[l + 1];
    const T ub = fmat(sU[l], r0, sU[l + 1] * r1);
    const T pb = fmat(sPt[l], r0, sPt[l + 1] * r1);
    return fmat(fsqrt((r0 + r1) * pb), kRootP, static_cast<T>(fabs(ub)));
  };
  auto lam_y = [&](int l) {
#pragma clang fp contract(off)
    const T r0 = sRg[l], r1 = sRg[l + EX];
    const T vb = fmat(sV[l], r0, sV[l + EX] * r1);
    const T pb = fmat(sPt[l], r0, sPt[l + EX] * r1);
    return fmat(fsqrt((r0 + r1) * pb), kRootP, static_cast<T>(fabs(vb)));
  };

  // ---- phase 2: WENO windows and face shares.  Thread (tx, ty) owns the
  // windows centred on its own cell along x and along y.  With face values
  // written as centre value + deviation, the flux of the face between cells
  // A (left/below) and B is
  //   F(f) = C(f) + (dF_A + lam dP_A)/2 + (dF_B - lam dP_B)/2,
  //   C(f) = (F_A + F_B)/2 - lam (P_B - P_A)/2,
  // dF_A, dP_A the deviations of A's window at f, dF_B, dP_B those of B's.
  // Each window writes its own share (small numbers); phase 3 adds the
  // centre terms from differences of neighbouring centre values.  Own shares
  // and Rusanov speeds stay in registers.
  const int tx = tid % TX, ty = tid / TX;
  // phase-3 inputs of the thread's own cell: the Coriolis static, the base
  // state and, where the stage combines it, the stage input.  The WENO path
  // fetches them before the barrier that ends phase 2, so their memory
  // latency overlaps the wait (a thread whose cell lies outside the panel
  // fetches the tile origin and stores nothing).
  struct Pre {
    T fs, b0, b1, b2, c0, c1, c2;
  };
  auto preload = [&]() {
    const bool ok = (i0 + tx < n) && (j0 + ty < n);
    const unsigned g = (gt + __umul24(static_cast<unsigned>(ok ? ty + 3 : 3), P) +
                        static_cast<unsigned>(ok ? tx + 3 : 3)) * W;
    Pre q{};
    q.fs = ldo(st.fs, g);
    if (KIND != kTendOnly) {
      q.b0 = ldo(a.base.phiG, g);
      q.b1 = ldo(a.base.u, g);
      q.b2 = ldo(a.base.v, g);
    }
    if (KIND == kRK3s2 || KIND == kRK3s3) {
      q.c0 = ldo(a.in.phiG, g);
      q.c1 = ldo(a.in.u, g);
      q.c2 = ldo(a.in.v, g);
    }
    return q;
  };
  Pre pre{};
  T fr, fl, er, el, fu, fd, eu, ed, lr, ll, lu, ld;
  if constexpr (kWeno) {
    const int l = (ty + 3) * EX + tx + 3;
    const int fx = ty * (TX + 1) + tx;   // left x-face of the cell
    const int fy = ty * TX + tx;         // bottom y-face of the cell
    {
      const auto wp = weno_window<T, 4>(sPt[l - 2], sPt[l - 1], sPt[l], sPt[l + 1], sPt[l + 2], a.epsP);
      const auto wf = weno_window<T, 2>(sFx[l - 2], sFx[l - 1], sFx[l], sFx[l + 1], sFx[l + 2], a.epsF);
      const auto we = weno_window<T, 2>(sE[l - 2], sE[l - 1], sE[l], sE[l + 1], sE[l + 2], a.epsE);
      lr = lam_x(l);
      ll = lam_x(l - 1);
      T fR, fL, pR, pL;
      weno_faces(wf, fR, fL);
      weno_faces(wp, pR, pL);
      weno_faces(we, er, el);
      // (F + lam P)/2 = F/2 + (2 lam) (P/4)
      fr = fmat(lr, pR, fR);
      fl = fmat(-ll, pL, fL);
      xpF[fx + 1] = fr;
      xmF[fx] = fl;
      xpE[fx + 1] = er;
      xmE[fx] = el;
    }
    {
      const auto wp = weno_window<T, 4>(sPt[l - 2 * EX], sPt[l - EX], sPt[l], sPt[l + EX], sPt[l + 2 * EX], a.epsP);
      const auto wf = weno_window<T, 2>(sFy[l - 2 * EX], sFy[l - EX], sFy[l], sFy[l + EX], sFy[l + 2 * EX], a.epsF);
      const auto we = weno_window<T, 2>(sE[l - 2 * EX], sE[l - EX], sE[l], sE[l + EX], sE[l + 2 * EX], a.epsE);
      lu = lam_y(l);
      ld = lam_y(l - EX);
      T fR, fL, pR, pL;
      weno_faces(wf, fR, fL);
      weno_faces(wp, pR, pL);
      weno_faces(we, eu, ed);
      fu = fmat(lu, pR, fR);
      fd = fmat(-ld, pL, fL);
      ypF[fy + TX] = fu;
      ymF[fy] = fd;
      ypE[fy + TX] = eu;
      ymE[fy] = ed;
    }
  }
  if constexpr (kWeno) {
    // windows centred one cell outside the tile, which give only the share
    // of the tile's boundary face: x: centres -1 (right face -> x-face 0) and
    // TX (left face -> x-face TX) of each row; y: centres -1 and TY of each
    // column.  One lane per window centre and role: the lowest NEXTRA
    // threads evaluate the phitG and flux windows and the Rusanov speed (the
    // flux share), the NEXTRA threads from the next wavefront boundary on the
    // E window, so that the two roles fall on different wavefronts.
    constexpr int NXE = 2 * TY, NEXTRA = 2 * TY + 2 * TX;
    constexpr int E0 = (NEXTRA + 63) / 64 * 64;
    static_assert(E0 + NEXTRA <= NT, "one pass of extra windows");
    const bool eRole = tid >= E0 && tid < E0 + NEXTRA;
    if (tid < NEXTRA || eRole) {
      const int q = eRole ? tid - E0 : tid;
      int l0, stride, lf, f;
      bool left;
      if (q < NXE) {
        const int row = q >> 1;
        left = (q & 1) != 0;
        l0 = (row + 3) * EX + (left ? TX : -1) + 3;
        stride = 1;
        f = row * (TX + 1) + (left ? TX : 0);
        lf = left ? l0 - 1 : l0;  // cell left of the face
      } else {
        const int k2 = q - NXE;
        const int col = k2 % TX;
        left = k2 >= TX;
        l0 = ((left ? TY : -1) + 3) * EX + col + 3;
        stride = EX;
        f = (left ? TY * TX : 0) + col;
        lf = left ? l0 - EX : l0;  // cell below the face
      }
      if (eRole) {
        const auto we = weno_window<T, 2>(sE[l0 - 2 * stride], sE[l0 - stride], sE[l0], sE[l0 + stride],
                                          sE[l0 + 2 * stride], a.epsE);
        T* oE = (q < NXE) ? (left ? xmE : xpE) : (left ? ymE : ypE);
        oE[f] = weno_side(we, left);
      } else {
        const T* sF = (q < NXE) ? sFx : sFy;
        const auto wp = weno_window<T, 4>(sPt[l0 - 2 * stride], sPt[l0 - stride], sPt[l0], sPt[l0 + stride],
                                          sPt[l0 + 2 * stride], a.epsP);
        const auto wf = weno_window<T, 2>(sF[l0 - 2 * stride], sF[l0 - stride], sF[l0], sF[l0 + stride],
                                          sF[l0 + 2 * stride], a.epsF);
        const T lam = (q < NXE) ? lam_x(lf) : lam_y(lf);  // twice the speed
        const T vp = weno_side(wp, left), vf = weno_side(wf, left);
        T* oF = (q < NXE) ? (left ? xmF : xpF) : (left ? ymF : ypF);
        oF[f] = fmat(left ? -lam : lam, vp, vf);  // as fr, fl of the interior windows
      }
    }
    pre = preload();
    __syncthreads();
  }

  // ---- phase 3: cells
  if constexpr (!kWeno) pre = preload();
  const int i = i0 + tx, j = j0 + ty;
  if (i >= n || j >= n) return;
  const T rdx = a.rdx;
  const int lc = (ty + 3) * EX + (tx + 3);
  // F(right face) - F(left face): centre terms from differences of
  // neighbouring centre values, plus the face shares.  Both cells beside a
  // face form the same share sum and the same lam (P_B - P_A) product
  // (identical operands, IEEE addition commutes), so the fluxes telescope.
  T dFx, dEx, dFy, dEy, dvdx, dudy;
  if constexpr (kWeno) {
    const int fx = ty * (TX + 1) + tx, fy = ty * TX + tx;
    const T pc = sPt[lc];
    // (lr, ll, lu, ld hold twice the Rusanov speeds; the stored fields carry
    // kF = 1/8 and kP = 1/16, so 0.5 dF and 0.25 lam dP both become 4 times
    // the stored differences)
    const T cx = T(4) * ((sFx[lc + 1] - sFx[lc - 1]) - (lr * (sPt[lc + 1] - pc) - ll * (pc - sPt[lc - 1])));
    const T cy = T(4) * ((sFy[lc + EX] - sFy[lc - EX]) - (lu * (sPt[lc + EX] - pc) - ld * (pc - sPt[lc - EX])));
    dFx = (cx + ((fr + xmF[fx + 1]) - (xpF[fx] + fl))) * rdx;
    dEx = (T(4) * (sE[lc + 1] - sE[lc - 1]) + ((er + xmE[fx + 1]) - (xpE[fx] + el))) * rdx;
    dFy = (cy + ((fu + ymF[fy + TX]) - (ypF[fy] + fd))) * rdx;
    dEy = (T(4) * (sE[lc + EX] - sE[lc - EX]) + ((eu + ymE[fy + TX]) - (ypE[fy] + ed))) * rdx;
    dvdx = d6f(sV[lc - 3], sV[lc - 2], sV[lc - 1], sV[lc + 1], sV[lc + 2], sV[lc + 3], rdx);
    dudy = d6f(sU[lc - 3 * EX], sU[lc - 2 * EX], sU[lc - EX], sU[lc + EX], sU[lc + 2 * EX], sU[lc + 3 * EX], rdx);
  } else {
    // centred differences of the cell values, stride 1 (x) or EX (y), of
    // the same fields: the flux divergence, the gradient of E and the
    // vorticity all take the operator's order
    // (spatial_operators_mod_{2nd,4th,6th}.F90)
    auto cd = [&](const T* f, int s) -> T {
      if (OP == 1) return (f[lc + s] - f[lc - s]) * (T(0.5) * rdx);
      if (OP == 2)
        return ((f[lc - 2 * s] - f[lc + 2 * s]) + T(8) * (f[lc + s] - f[lc - s])) * (rdx * (T(1) / T(12)));
      return d6f(f[lc - 3 * s], f[lc - 2 * s], f[lc - s], f[lc + s], f[lc + 2 * s], f[lc + 3 * s], rdx);
    };
    dFx = cd(sFx, 1);
    dEx = cd(sE, 1);
    dFy = cd(sFy, EX);
    dEy = cd(sE, EX);
    dvdx = cd(sV, 1);
    dudy = cd(sU, EX);
  }
  const unsigned go = (gt + __umul24(static_cast<unsigned>(ty + 3), P) + static_cast<unsigned>(tx + 3)) * W;
  const T zeta = dvdx - dudy + pre.fs;
  const T tp = -(dFx + dFy);
  const T rgc = sRg[lc];
  const T tu = -dEx + zeta * (sV[lc] * rgc);
  const T tv = -dEy - zeta * (sU[lc] * rgc);
  const unsigned un_ = static_cast<unsigned>(n);
  const unsigned io = (static_cast<unsigned>(p) * un_ * un_ + static_cast<unsigned>(j0) * un_ +
                       static_cast<unsigned>(i0) + __umul24(static_cast<unsigned>(ty), un_) +
                       static_cast<unsigned>(tx)) * W;
  if (KIND == kTendOnly) {
    sto(a.tend[0], io, tp);
    sto(a.tend[1], io, tu);
    sto(a.tend[2], io, tv);
    return;
  }
  T pn, un, vn;
  if (KIND == kRK3s1) {
    pn = pre.b0 + a.dts * tp;
    un = pre.b1 + a.dts * tu;
    vn = pre.b2 + a.dts * tv;
  } else if (KIND == kRK3s2 || KIND == kRK3s3) {
    // q2 = 3/4 q + 1/4 (q1 + dt L) and q' = 1/3 q + 2/3 (q2 + dt L)
    // (time_scheme_mod.F90:85-115) as q + c ((q_k - q) + dt L): the base
    // state q enters with the exact coefficient 1 and c rounds only the
    // increment.  (In FP32, fl(1/3) + fl(2/3) = 1 + 2^-25, so the
    // weighted-sum form scaled the winds U, V by that factor every step:
    // total energy grew by 1.0e-6 in 8192 steps at C720 while the FP64 path
    // lost 2.6e-8.)
    const T c = (KIND == kRK3s2) ? T(0.25) : T(2) / T(3);
    pn = pre.b0 + c * ((pre.c0 - pre.b0) + a.dts * tp);
    un = pre.b1 + c * ((pre.c1 - pre.b1) + a.dts * tu);
    vn = pre.b2 + c * ((pre.c2 - pre.b2) + a.dts * tv);
  } else if (KIND == kRK4a) {
    sto(a.tend[0], io, tp);
    sto(a.tend[1], io, tu);
    sto(a.tend[2], io, tv);
    pn = pre.b0 + a.incs * tp;
    un = pre.b1 + a.incs * tu;
    vn = pre.b2 + a.incs * tv;
  } else {
    const T c6 = T(1) / T(6);
    pn = pre.b0 +
         a.dts * ((ldo(a.tacc[0], io) + T(2) * (ldo(a.tacc[3], io) + ldo(a.tacc[6], io)) + tp) * c6);
    un = pre.b1 +
         a.dts * ((ldo(a.tacc[1], io) + T(2) * (ldo(a.tacc[4], io) + ldo(a.tacc[7], io)) + tu) * c6);
    vn = pre.b2 +
         a.dts * ((ldo(a.tacc[2], io) + T(2) * (ldo(a.tacc[5], io) + ldo(a.tacc[8], io)) + tv) * c6);
  }
  sto(a.out.phiG, go, pn);
  sto(a.out.u, go, un);
  sto(a.out.v, go, vn);
  const int i1 = i + 1, j1 = j + 1;
  const bool strip[4] = {i1 <= h, i1 >= n - h + 1, j1 <= h, j1 >= n - h + 1};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (!strip[s]) continue;
    const int* r = hr[s];
    // halo destination (1-based cell id, jd of panel r[0]); the coefficients
    // r[2], r[3], r[5], r[6] are 0 or +-1
    const int id = r[1] + __mul24(r[2], i1) + __mul24(r[3], j1);
    const int jd = r[4] + __mul24(r[5], i1) + __mul24(r[6], j1);
    const unsigned gd = (static_cast<unsigned>(r[0] - 1) * P * P + __umul24(static_cast<unsigned>(jd - 1 + h), P) +
                         static_cast<unsigned>(id - 1 + h)) * W;
    const T c[2] = {un, vn};
    sto(a.out.phiG, gd, pn);
    sto(a.out.u, gd, static_cast<T>(r[8]) * c[r[7]]);
    sto(a.out.v, gd, static_cast<T>(r[10]) * c[r[9]]);
  }
}

template <typename T, int KIND, int TY, int OP>
__global__ void __launch_bounds__(TX * TY) stage_split(SplitArgs<T> a, SplitStatics<T> st, HaloScatter hs) {
  const unsigned t = xcd_tile_dir(blockIdx.x, gridDim.x, a.rev);
  const unsigned tq = __umulhi(t << 1, a.mbx);
  const int p = static_cast<int>(__umulhi(tq << 1, a.mby));
  const int i0 = static_cast<int>(t - tq * a.nbx) * TX;
  const int j0 = static_cast<int>(tq - static_cast<unsigned>(p) * a.nby) * TY;
  stage_tile<T, KIND, TY, OP>(a, st, hs.r[p], p, i0, j0, static_cast<int>(threadIdx.x));
}

// ---- batched ensemble: every member's tiles in one launch per RK stage.
// A member's buffers, constants and halo rules sit in a device table; the
// stage selects which of the three state buffers are input, base and output
// and which tendency slots it writes (the same for every member).
template <typename T>
struct EnsMember {
  SplitArgs<T> a;      // n, h, nbx, nby, rdx, dts, eps (per member)
  SplitStatics<T> st;
  T* buf[3][3];        // [buffer][dP, U, V]
  T* tend[9];
  T inc_half, inc_full;  // RK4 increments dt/(2a), dt/a
  int target;            // steps the member takes in the current run
  int member;            // index of the member in the ensemble (its slot in the step counters)
  int hs[6][4][11];
  double* pref64;        // FP64 reference field and its cast (rebase)
  T* pref_rw;
  unsigned long long np; // padded cells
};

struct EnsStage {
  int in_b, base_b, out_b;
  int tend0;    // first tendency slot written (RK4 stages a)
  int inc_full; // RK4: 1 -> dt/a increment, 0 -> dt/(2a)
  int rev;      // 1: each XCD's run of tiles backwards (xcd_tile_dir; one workgroup per tile only)
};

// Step index within the current run: the index at the start of the graph
// being replayed (device counter, advanced by ens_advance at the end of each
// graph, so a plain uniform load sees it) plus the step's offset in the
// graph.  The autotuner's launches, where every member steps, read a counter
// that holds -1.
__device__ __forceinline__ int run_step(const int* __restrict__ base, int k) { return *base + k; }

// A tile of the global list, with what its first phase needs, so that a
// workgroup issues its tile's loads after one load that depends on nothing
// but its index.  A member's 13 padded arrays Pref, sqrtg, phis, a sqrtg f
// and the three state buffers (dP, U, V) follow each other at a fixed
// stride; the first lies `fbase` bytes from the kernel's `arena` argument (an
// offset from a kernel argument keeps the accesses in the global address
// space).
struct EnsTile {
  unsigned slot;             // the member's slot in the member table
  unsigned pk;               // panel | tile column << 3 | tile row << 15 (tile_pack)
  int target;                // steps the member takes in the current run
  unsigned fstride;          // bytes from one array of the member to the next
  long long fbase;           // the member's first array (Pref), from the arena argument
  int n, h;
};
static_assert(sizeof(EnsTile) == 32, "one scalar load per tile record");

__host__ __device__ constexpr unsigned tile_pack(int p, int bx, int by) {
  return static_cast<unsigned>(p) | (static_cast<unsigned>(bx) << 3) | (static_cast<unsigned>(by) << 15);
}

// One tile of the global list; false when its member has taken its steps
// (members are ordered by decreasing step count, so every later tile of the
// list is finished too).  The member table's fields (constants, halo rules,
// tendency slots) load while phase 1 runs.
template <typename T, int KIND, int TY>
__device__ __forceinline__ bool ens_tile(const EnsMember<T>* __restrict__ md, const EnsTile* __restrict__ tiles,
                                         char* arena, const EnsStage& sel, const unsigned t, const int tid,
                                         const int step) {
  // the whole record in one round trip (loaded as two 16-byte vectors, held
  // until the step test: field by field, the compiler would wait for each)
  const uint4* rv = reinterpret_cast<const uint4*>(tiles + t);
  const uint4 r0 = rv[0], r1 = rv[1];
  asm volatile("" ::"s"(r0.x), "s"(r0.y), "s"(r0.z), "s"(r0.w), "s"(r1.x), "s"(r1.y), "s"(r1.z), "s"(r1.w));
  EnsTile rec;
  rec.slot = r0.x;
  rec.pk = r0.y;
  rec.target = static_cast<int>(r0.z);
  rec.fstride = r0.w;
  rec.fbase = static_cast<long long>(static_cast<unsigned long long>(r1.x) | (static_cast<unsigned long long>(r1.y) << 32));
  rec.n = static_cast<int>(r1.z);
  rec.h = static_cast<int>(r1.w);
  if (step >= rec.target) return false;
  const EnsMember<T>* M = md + rec.slot;
  char* const mbase = arena + rec.fbase;
  auto arr = [&](int k) { return reinterpret_cast<T*>(mbase + static_cast<long long>(k) * rec.fstride); };
  SplitArgs<T> a = M->a;
  a.n = rec.n;
  a.h = rec.h;
  a.in = Prog<const T>{arr(4 + 3 * sel.in_b), arr(5 + 3 * sel.in_b), arr(6 + 3 * sel.in_b)};
  a.base = Prog<const T>{arr(4 + 3 * sel.base_b), arr(5 + 3 * sel.base_b), arr(6 + 3 * sel.base_b)};
  a.out = Prog<T>{arr(4 + 3 * sel.out_b), arr(5 + 3 * sel.out_b), arr(6 + 3 * sel.out_b)};
  for (int k = 0; k < 3; ++k) a.tend[k] = M->tend[sel.tend0 + k];
  for (int k = 0; k < 9; ++k) a.tacc[k] = M->tend[k];
  a.incs = sel.inc_full ? M->inc_full : M->inc_half;
  const SplitStatics<T> st{arr(0), arr(1), arr(2), arr(3)};
  const int p = static_cast<int>(rec.pk & 7u);
  const int i0 = static_cast<int>((rec.pk >> 3) & 0xFFFu) * TX;
  const int j0 = static_cast<int>(rec.pk >> 15) * TY;
  stage_tile<T, KIND, TY, 0>(a, st, M->hs[p], p, i0, j0, tid);
  return true;
}

// PERSIST = false: one workgroup per tile, in the XCD-aware order.
// PERSIST = true: a grid of resident workgroups; each belongs to a group
// (ens_group, one per XCD) owning a contiguous run of the global tile list,
// and claims the group's tiles one at a time from the group's counter in
// `claim`, so the tiles spread over the workgroups as they finish.  Workgroup
// 0 clears the counters of the next launch on the stream (`claim_next`): its
// kernel starts after this one ends, and the launch that used them before
// has ended.
template <typename T, int KIND, int TY, bool PERSIST>
__global__ void __launch_bounds__(TX * TY) stage_ens(const EnsMember<T>* __restrict__ md,
                                                      const EnsTile* __restrict__ tiles, char* arena, int ntiles,
                                                      EnsStage sel,
                                                      const int* __restrict__ base, int koff,
                                                      unsigned* __restrict__ claim, unsigned* __restrict__ claim_next) {
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const int step = run_step(base, koff);
  if (!PERSIST) {
    (void)ens_tile<T, KIND, TY>(md, tiles, arena, sel, xcd_tile_dir(b, nb, static_cast<unsigned>(sel.rev)),
                                static_cast<int>(threadIdx.x), step);
    return;
  }
  if (b == 0 && threadIdx.x < kClaimGroups) claim_next[threadIdx.x * kClaimStride] = 0u;
  const EnsGroup gr = ens_group(b, nb, static_cast<unsigned>(ntiles));
  // the claimed tile, through LDS; the claim for tile k + 1 is issued at the
  // start of tile k, its result stored at the end, so its latency is hidden
  __shared__ unsigned s_tile[2];
  if (threadIdx.x == 0) s_tile[0] = gr.lo + atomicAdd(claim + gr.g * kClaimStride, 1u);
  for (unsigned k = 0;; ++k) {
    __syncthreads();  // s_tile[k & 1] is stored; the previous tile's LDS reads are done
    // uniform: the tile's loads go through the scalar unit
    const unsigned t = __builtin_amdgcn_readfirstlane(s_tile[k & 1u]);
    if (t >= gr.hi) break;
    unsigned nxt = 0;
    if (threadIdx.x == 0) nxt = gr.lo + atomicAdd(claim + gr.g * kClaimStride, 1u);
    // keep the thread-index arithmetic inside the loop: hoisting it out would
    // hold ~30 more VGPRs for the whole kernel and halve the occupancy
    int tid = static_cast<int>(threadIdx.x);
    asm volatile("" : "+v"(tid));
    if (!ens_tile<T, KIND, TY>(md, tiles, arena, sel, t, tid, step)) break;
    if (threadIdx.x == 0) s_tile[(k + 1u) & 1u] = nxt;
  }
}

// Rebase of the geopotential deviation: the exact sum Pref + dP becomes the
// new FP64 reference field, its cast the FP32 one, and dP restarts at zero.
// The represented state is unchanged (in FP64); keeping dP small keeps the
// rounding of each stage's increment relative to the increment, so the FP32
// state stops losing mass to round-off accumulated into a growing deviation.
template <typename T>
__global__ void __launch_bounds__(256) split_rebase(double* __restrict__ pref64, T* __restrict__ pref,
                                                    T* __restrict__ dP, unsigned long long np) {
  for (unsigned long long d = blockIdx.x * 256ull + threadIdx.x; d < np; d += gridDim.x * 256ull) {
    const double p = pref64[d] + static_cast<double>(dP[d]);
    pref64[d] = p;
    pref[d] = static_cast<T>(p);
    dP[d] = T(0);
  }
}

// The same for the members of an ensemble (blockIdx.y = position in the
// member table) whose own step count reaches a multiple of kRebaseSteps with
// step koff of the graph: done[] holds each member's steps since its upload,
// modulo kRebaseSteps, as of the start of the graph (ens_advance updates it
// at the graph's end).
template <typename T>
__global__ void __launch_bounds__(256) ens_rebase(const EnsMember<T>* __restrict__ md, const int* base, int koff,
                                                  const int* done) {
  const EnsMember<T>* M = md + blockIdx.y;
  const int c = run_step(base, koff);
  if (c >= M->target || (done[M->member] + koff + 1) % kRebaseSteps != 0) return;
  T* dP = M->buf[0][0];
  for (unsigned long long d = blockIdx.x * 256ull + threadIdx.x; d < M->np; d += gridDim.x * 256ull) {
    const double p = M->pref64[d] + static_cast<double>(dP[d]);
    M->pref64[d] = p;
    M->pref_rw[d] = static_cast<T>(p);
    dP[d] = T(0);
  }
}

// End of a graph of `steps` steps: every member's step count advances by the
// steps it took in the graph, then the run's step index (one workgroup; every
// thread reads the index before thread 0 advances it).
template <typename T>
__global__ void __launch_bounds__(256) ens_advance(const EnsMember<T>* __restrict__ md, int nm, int* base, int* done,
                                                   int steps) {
  const int c = *base;
  for (int k = threadIdx.x; k < nm; k += 256) {
    const int taken = min(max(md[k].target - c, 0), steps);
    done[md[k].member] = (done[md[k].member] + taken) % kRebaseSteps;
  }
  __syncthreads();
  if (threadIdx.x == 0) *base = c + steps;
}

// Start of a run: step index 0.
__global__ void ens_run_start(int* base) {
  if (threadIdx.x == 0) *base = 0;
}

// Compensated FP64 mass/energy of the split state.
struct KSum {
  double s, c;
  __device__ void add(double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
    s = t;
  }
};

constexpr int kDT = 256;

template <typename T>
__global__ void __launch_bounds__(kDT) split_diag(Prog<const T> q, const double* pref64, const double* sqrtG,
                                                  const double* phis, int n, int h, double a, double* partials) {
  __shared__ double sh[4][kDT];
  const int P = n + 2 * h;
  const long ni = static_cast<long>(n) * n * 6;
  KSum m{0., 0.}, e{0., 0.};
  const double a2 = a * a;
  for (long c = blockIdx.x * static_cast<long>(kDT) + threadIdx.x; c < ni; c += static_cast<long>(gridDim.x) * kDT) {
    const int i = static_cast<int>(c % n), j = static_cast<int>((c / n) % n), p = static_cast<int>(c / (static_cast<long>(n) * n));
    const long g = gidx(i, j, p, h, P);
    const double phiG = a2 * (pref64[g] + static_cast<double>(q.phiG[g]));
    const double u = a * static_cast<double>(q.u[g]), v = a * static_cast<double>(q.v[g]);
    const double sG = sqrtG[g];
    const double phi = phiG / sG;
    const double KE = 0.5 * phi * ((u * u + v * v) / sG);
    const double pt = phi + phis[g];
    m.add(phiG);
    e.add(sG * (KE + 0.5 * (pt * pt)));
  }
  sh[0][threadIdx.x] = m.s; sh[1][threadIdx.x] = m.c; sh[2][threadIdx.x] = e.s; sh[3][threadIdx.x] = e.c;
  __syncthreads();
  if (threadIdx.x == 0) {
    KSum M{0., 0.}, Ee{0., 0.};
    double cm = 0., ce = 0.;
    for (int t = 0; t < kDT; ++t) {
      M.add(sh[0][t]); cm += sh[1][t];
      Ee.add(sh[2][t]); ce += sh[3][t];
    }
    partials[4 * blockIdx.x + 0] = M.s;
    partials[4 * blockIdx.x + 1] = M.c + cm;
    partials[4 * blockIdx.x + 2] = Ee.s;
    partials[4 * blockIdx.x + 3] = Ee.c + ce;
  }
}

// nb <= 4 * kDT partial blocks; all threads stage them in LDS (one round of
// loads instead of nb dependent ones), thread 0 sums them in block order.
__global__ void __launch_bounds__(kDT) split_diag_final(const double* partials, int nb, double* result) {
  __shared__ double sp[4 * 4 * kDT];
  for (int k = threadIdx.x; k < 4 * nb; k += kDT) sp[k] = partials[k];
  __syncthreads();
  if (threadIdx.x != 0) return;
  KSum M{0., 0.}, E{0., 0.};
  double cm = 0., ce = 0.;
  for (int b = 0; b < nb; ++b) {
    M.add(sp[4 * b + 0]); cm += sp[4 * b + 1];
    E.add(sp[4 * b + 2]); ce += sp[4 * b + 3];
  }
  result[0] = M.s + (M.c + cm);
  result[1] = E.s + (E.c + ce);
}

// Multiplier of the stage kernel's tile decomposition: for d >= 1 and
// m = ceil(2^31 / d), hi32((2 t) m) = floor(t / d) whenever t d < 2^31 (the
// error of m d over 2^31 is below d); the host keeps tile counts below 2^20
// and d below 2^11.
unsigned tile_magic(unsigned d) { return static_cast<unsigned>((0x80000000ull + d - 1) / d); }

// A stage launch; with an event pair, the events are the dispatch's own start
// and end (hipExtLaunchKernelGGL), so they time the kernel alone.
template <typename K, typename... Args>
void launch_kernel(K kernel, dim3 grid, dim3 block, hipStream_t s, const hipEvent_t* ev, Args... args) {
  if (ev) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev[0], ev[1], 0, args...);
  else hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

template <typename T, int TY, int OP>
void launch_split_op(hipStream_t s, int kind, const SplitArgs<T>& a, const SplitStatics<T>& st,
                     const HaloScatter& hs, const hipEvent_t* ev) {
  const dim3 grid(static_cast<unsigned>(a.nbx * a.nby * 6));
  const dim3 block(TX * TY);
  switch (kind) {
    case kTendOnly: launch_kernel(stage_split<T, kTendOnly, TY, OP>, grid, block, s, ev, a, st, hs); break;
    case kRK3s1: launch_kernel(stage_split<T, kRK3s1, TY, OP>, grid, block, s, ev, a, st, hs); break;
    case kRK3s2: launch_kernel(stage_split<T, kRK3s2, TY, OP>, grid, block, s, ev, a, st, hs); break;
    case kRK3s3: launch_kernel(stage_split<T, kRK3s3, TY, OP>, grid, block, s, ev, a, st, hs); break;
    case kRK4a: launch_kernel(stage_split<T, kRK4a, TY, OP>, grid, block, s, ev, a, st, hs); break;
    default: launch_kernel(stage_split<T, kRK4f, TY, OP>, grid, block, s, ev, a, st, hs); break;
  }
}

template <typename T, int TY>
void launch_split_ty(hipStream_t s, int kind, int op, const SplitArgs<T>& a, const SplitStatics<T>& st,
                     const HaloScatter& hs, const hipEvent_t* ev) {
  switch (op) {
    case CCS_OP_CENTRED2: launch_split_op<T, TY, 1>(s, kind, a, st, hs, ev); break;
    case CCS_OP_CENTRED4: launch_split_op<T, TY, 2>(s, kind, a, st, hs, ev); break;
    case CCS_OP_CENTRED6: launch_split_op<T, TY, 3>(s, kind, a, st, hs, ev); break;
    default: launch_split_op<T, TY, 0>(s, kind, a, st, hs, ev); break;
  }
}

// Tile heights instantiated (8, 16, 32); the launch fills nbx, nby for the
// chosen one.  op: CCS_OP_WENO or a centred CCS_OP_CENTRED{2,4,6}.
template <typename T>
void launch_split(hipStream_t s, int kind, int op, SplitArgs<T> a, const SplitStatics<T>& st,
                  const HaloScatter& hs, int tile_y, const hipEvent_t* ev = nullptr) {
  a.nbx = (a.n + TX - 1) / TX;
  a.nby = (a.n + tile_y - 1) / tile_y;
  a.mbx = tile_magic(static_cast<unsigned>(a.nbx));
  a.mby = tile_magic(static_cast<unsigned>(a.nby));
  if (tile_y == 16) launch_split_ty<T, 16>(s, kind, op, a, st, hs, ev);
  else if (tile_y == 32) launch_split_ty<T, 32>(s, kind, op, a, st, hs, ev);
  else launch_split_ty<T, 8>(s, kind, op, a, st, hs, ev);
}

}  // namespace

// ------------------------------------------------------------------ host
struct SplitState {
  ccs_config cfg{};
  HaloScatter hs{};
  bool fp64 = false;
  int n = 0, h = 0;
  int tile_y = 16;  // stage-kernel tile height (8, 16 or 32)
  unsigned launches = 0;  // stage launches enqueued (their parity sets the tile order)
  std::size_t np = 0, ni = 0, esz = 4;
  void* dmem = nullptr;
  // typed device arrays (float or double)
  void *pref = nullptr, *sg = nullptr, *phis = nullptr, *fs = nullptr;
  void* buf[3][3]{};
  void* tend[9]{};
  double *pref64_d = nullptr, *sqrtG_d = nullptr, *phis_d = nullptr, *partials = nullptr, *result = nullptr;
  // host FP64 tables for output
  std::vector<double> sqrtG, IG11, IG12, IG21, IG22, M11, M12, M21, M22, pref64;
  bool mesh_set = false;
};

namespace {

constexpr int kSplitDiagBlocks = 256;

template <typename T>
SplitStatics<T> statics_of(const SplitState* s) {
  return SplitStatics<T>{static_cast<const T*>(s->pref), static_cast<const T*>(s->sg),
                         static_cast<const T*>(s->phis), static_cast<const T*>(s->fs)};
}

template <typename T>
Prog<T> prog_of(const SplitState* s, int b) {
  return Prog<T>{static_cast<T*>(s->buf[b][0]), static_cast<T*>(s->buf[b][1]), static_cast<T*>(s->buf[b][2])};
}

template <typename T>
Prog<const T> cprog_of(const SplitState* s, int b) {
  return Prog<const T>{static_cast<const T*>(s->buf[b][0]), static_cast<const T*>(s->buf[b][1]),
                       static_cast<const T*>(s->buf[b][2])};
}

template <typename T>
SplitArgs<T> base_args(const SplitState* s) {
  SplitArgs<T> a{};
  const double r = s->cfg.radius;
  a.n = s->n;
  a.h = s->h;
  a.rdx = static_cast<T>(1.0 / s->cfg.dx);
  a.dts = static_cast<T>(s->cfg.dt / r);
  a.epsF = static_cast<T>(1e-2 / (r * r)) * T(1.0 / 16);
  a.epsE = static_cast<T>(1e-2) * T(1.0 / 16);
  a.epsP = static_cast<T>(1e-2 / (r * r * r * r)) * T(1.0 / 64);
  return a;
}

template <typename T>
void enqueue_step_t(SplitState* s, hipStream_t st, int scheme, hipEvent_t* ev) {
  SplitArgs<T> a = base_args<T>(s);
  const SplitStatics<T> S = statics_of<T>(s);
  int stage = 0;
  static const char* const kStageName[4] = {"ccs stage 1", "ccs stage 2", "ccs stage 3", "ccs stage 4"};
  auto launch = [&](int kind) {
    a.rev = s->launches++ & 1u;  // alternate the tile order (xcd_tile_dir)
    roctxRangePush(kStageName[stage]);
    launch_split<T>(st, kind, s->cfg.spatial_op, a, S, s->hs, s->tile_y, ev ? ev + 2 * stage : nullptr);
    roctxRangePop();
    ++stage;
  };
  if (scheme == CCS_RK3_TVD) {
    a.in = cprog_of<T>(s, 0); a.base = cprog_of<T>(s, 0); a.out = prog_of<T>(s, 1);
    launch(kRK3s1);
    a.in = cprog_of<T>(s, 1); a.out = prog_of<T>(s, 2);
    launch(kRK3s2);
    a.in = cprog_of<T>(s, 2); a.out = prog_of<T>(s, 0);
    launch(kRK3s3);
  } else {
    const double r = s->cfg.radius;
    a.base = cprog_of<T>(s, 0);
    a.in = cprog_of<T>(s, 0); a.out = prog_of<T>(s, 1); a.incs = static_cast<T>(0.5 * s->cfg.dt / r);
    for (int k = 0; k < 3; ++k) a.tend[k] = static_cast<T*>(s->tend[k]);
    launch(kRK4a);
    a.in = cprog_of<T>(s, 1); a.out = prog_of<T>(s, 2);
    for (int k = 0; k < 3; ++k) a.tend[k] = static_cast<T*>(s->tend[3 + k]);
    launch(kRK4a);
    a.in = cprog_of<T>(s, 2); a.out = prog_of<T>(s, 1); a.incs = static_cast<T>(s->cfg.dt / r);
    for (int k = 0; k < 3; ++k) a.tend[k] = static_cast<T*>(s->tend[6 + k]);
    launch(kRK4a);
    a.in = cprog_of<T>(s, 1); a.out = prog_of<T>(s, 0);
    for (int k = 0; k < 9; ++k) a.tacc[k] = static_cast<const T*>(s->tend[k]);
    launch(kRK4f);
  }
}

template <typename T>
void upload_cast(hipStream_t st, void* dst, const std::vector<double>& src, std::vector<char>& staging) {
  staging.resize(src.size() * sizeof(T));
  T* t = reinterpret_cast<T*>(staging.data());
  for (std::size_t k = 0; k < src.size(); ++k) t[k] = static_cast<T>(src[k]);
  (void)hipMemcpyAsync(dst, staging.data(), staging.size(), hipMemcpyHostToDevice, st);
  (void)hipStreamSynchronize(st);
}

template <typename T>
void download_to_double(hipStream_t st, const void* src, std::size_t cnt, std::vector<double>& out) {
  std::vector<T> tmp(cnt);
  (void)hipMemcpyAsync(tmp.data(), src, cnt * sizeof(T), hipMemcpyDeviceToHost, st);
  (void)hipStreamSynchronize(st);
  out.resize(cnt);
  for (std::size_t k = 0; k < cnt; ++k) out[k] = static_cast<double>(tmp[k]);
}

}  // namespace

int split_create(const ccs_config& cfg, const HaloScatter& hs, hipStream_t st, SplitState** out, std::string* err) {
  SplitState* s = new SplitState();
  s->cfg = cfg;
  s->hs = hs;
  s->fp64 = (cfg.precision == CCS_FP64_FAST);
  s->esz = s->fp64 ? 8 : 4;
  s->n = cfg.n;
  s->h = cfg.halo;
  s->np = padded_cells(cfg.n, cfg.halo);
  s->ni = static_cast<std::size_t>(cfg.n) * cfg.n * 6;
  // the stage kernel addresses each array with a 32-bit byte offset
  if (s->np * s->esz >= (std::size_t{1} << 32)) {
    *err = "split_create: a padded field exceeds 4 GiB; n is too large for the split path";
    delete s;
    return CCS_ERR_INVALID;
  }
  if (const char* ty = std::getenv("CCS_SPLIT_TILE_Y")) {
    const int v = std::atoi(ty);
    if (v != 8 && v != 16 && v != 32) {
      *err = "CCS_SPLIT_TILE_Y must be 8, 16 or 32";
      delete s;
      return CCS_ERR_INVALID;
    }
    s->tile_y = v;
  }
  // the stage kernel's tile decomposition (tile_magic) is exact below 2^20
  // tiles (n up to about 6700 with 8-row tiles)
  {
    const std::size_t nbx = (static_cast<std::size_t>(s->n) + 31) / 32, nby = (static_cast<std::size_t>(s->n) + 7) / 8;
    if (nbx * nby * 6 >= (std::size_t{1} << 20)) {
      *err = "split_create: too many tiles for the stage kernel; n is too large for the split path";
      delete s;
      return CCS_ERR_INVALID;
    }
  }
  // One allocation: the FP64 block (Pref, sqrtG, phis, diagnostics partials),
  // then 4 statics + 9 state arrays (padded) and 9 tendencies (interior).
  // Every array starts on a 256-byte boundary, so the size is the sum of the
  // rounded sizes, and the carving below walks the same list.
  auto rup = [](std::size_t b) { return (b + 255) / 256 * 256; };
  const std::size_t dbl = 3 * s->np + 4 * kSplitDiagBlocks + 2;
  const std::size_t db = rup(dbl * sizeof(double)), fb = rup(s->np * s->esz), tb = rup(s->ni * s->esz);
  const std::size_t bytes = db + (4 + 9) * fb + 9 * tb;
  if (hipMalloc(&s->dmem, bytes) != hipSuccess) {
    *err = "split_create: device allocation failed";
    delete s;
    return CCS_ERR_NOMEM;
  }
  // cleared on the handle's stream, and waited for: a memset on the null
  // stream is not ordered before the uploads on a non-blocking stream
  if (hipMemsetAsync(s->dmem, 0, bytes, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    *err = "split_create: clearing the device buffers failed";
    (void)hipFree(s->dmem);
    delete s;
    return CCS_ERR_DEVICE;
  }
  char* const base = static_cast<char*>(s->dmem);
  char* p = base;
  auto take = [&p](std::size_t rounded) { void* r = p; p += rounded; return r; };
  double* d = reinterpret_cast<double*>(take(db));
  s->pref64_d = d; s->sqrtG_d = d + s->np; s->phis_d = d + 2 * s->np;
  s->partials = d + 3 * s->np; s->result = s->partials + 4 * kSplitDiagBlocks;
  s->pref = take(fb); s->sg = take(fb); s->phis = take(fb); s->fs = take(fb);
  for (int b = 0; b < 3; ++b)
    for (int k = 0; k < 3; ++k) s->buf[b][k] = take(fb);
  for (int k = 0; k < 9; ++k) s->tend[k] = take(tb);
  if (static_cast<std::size_t>(p - base) != bytes) {  // the carving and the size disagree
    *err = "split_create: internal error in the device buffer layout";
    (void)hipFree(s->dmem);
    delete s;
    return CCS_ERR_INVALID;
  }
  *out = s;
  return CCS_OK;
}

void split_destroy(SplitState* s) {
  if (!s) return;
  if (s->dmem) (void)hipFree(s->dmem);
  delete s;
}

int split_set_mesh(SplitState* s, hipStream_t st, double* const T[15], std::string* err) {
  const int n = s->n, h = s->h;
  const std::size_t np = s->np;
  const double r = s->cfg.radius;
  // The split path uses G = sqrtG I: reject grids that are not conformal.
  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - h; j <= n + h; ++j)
      for (int i = 1 - h; i <= n + h; ++i) {
        if (is_corner(n, i, j)) continue;
        const std::size_t d = pidx(n, h, i, j, p);
        const double g = T[0][d];
        const double off = std::fabs(T[2][d]) + std::fabs(T[3][d]);
        if (std::fabs(T[1][d] * g - 1.0) > 1e-9 || std::fabs(T[4][d] * g - 1.0) > 1e-9 || off * g > 1e-9) {
          *err = "split-precision path needs a conformal grid (G = sqrtG I); use CCS_FP64_PARITY";
          return CCS_ERR_INVALID;
        }
      }
  std::vector<double> sg(np), fs(np, 0.0);
  for (std::size_t d = 0; d < np; ++d) {
    sg[d] = T[0][d] / (r * r);
    fs[d] = r * sg[d] * T[6][d];
  }
  std::vector<double> phis(T[5], T[5] + np);
  // the WENO stage kernel stores its pointwise terms scaled by 1/8 (exact):
  // its surface-geopotential static carries the same factor
  std::vector<double> phis_k(phis);
  if (s->cfg.spatial_op == CCS_OP_WENO)
    for (double& x : phis_k) x *= 0.125;
  std::vector<char> stage;
  if (s->fp64) {
    upload_cast<double>(st, s->sg, sg, stage);
    upload_cast<double>(st, s->fs, fs, stage);
    upload_cast<double>(st, s->phis, phis_k, stage);
  } else {
    upload_cast<float>(st, s->sg, sg, stage);
    upload_cast<float>(st, s->fs, fs, stage);
    upload_cast<float>(st, s->phis, phis_k, stage);
  }
  std::vector<double> sq(T[0], T[0] + np);
  (void)hipMemcpyAsync(s->sqrtG_d, sq.data(), np * sizeof(double), hipMemcpyHostToDevice, st);
  (void)hipMemcpyAsync(s->phis_d, phis.data(), np * sizeof(double), hipMemcpyHostToDevice, st);
  if (hipStreamSynchronize(st) != hipSuccess) {
    *err = "split_set_mesh: upload failed";
    return CCS_ERR_DEVICE;
  }
  s->sqrtG = std::move(sq);
  s->IG11.assign(T[1], T[1] + np); s->IG12.assign(T[2], T[2] + np);
  s->IG21.assign(T[3], T[3] + np); s->IG22.assign(T[4], T[4] + np);
  s->M11.assign(T[7], T[7] + np); s->M12.assign(T[8], T[8] + np);
  s->M21.assign(T[9], T[9] + np); s->M22.assign(T[10], T[10] + np);
  s->mesh_set = true;
  return CCS_OK;
}

int split_set_state(SplitState* s, hipStream_t st, const double* phiG, const double* u, const double* v,
                    const double*, const double*, const double*, std::string* err) {
  const int n = s->n, h = s->h;
  const std::size_t np = s->np;
  const double r = s->cfg.radius;
  s->pref64.assign(np, 0.0);
  std::vector<double> U(np), V(np), dP(np, 0.0);
  for (std::size_t d = 0; d < np; ++d) {
    s->pref64[d] = phiG[d] / (r * r);
    U[d] = u[d] / r;
    V[d] = v[d] / r;
  }
  // Halo cells hold copies of interior cells (ghost_mod.f90:18-82).  After
  // the first stage the kernel writes a halo cell's deviation dP from its
  // source cell, so the halo's reference value must be the source's: halo
  // Pref = source Pref, as fill_ghost makes it, and the uploaded halo value
  // is kept for the first stage (which reads the state as uploaded) in the
  // halo's deviation.
  for (int p = 1; p <= 6; ++p)
    for (int side = 0; side < 4; ++side) {
      const HaloStrip& hsr = halo_strip(p, side);
      for (int k = 1; k <= h; ++k)
        for (int t = 1; t <= n; ++t) {
          int di, dj, si, sj;
          halo_dst(n, side, k, t, &di, &dj);
          halo_src(n, hsr, k, t, &si, &sj);
          const std::size_t dd = pidx(n, h, di, dj, p), ds = pidx(n, h, si, sj, hsr.q);
          const double src = phiG[ds] / (r * r);
          dP[dd] = s->pref64[dd] - src;
          s->pref64[dd] = src;
        }
    }
  std::vector<char> stage;
  if (s->fp64) {
    upload_cast<double>(st, s->pref, s->pref64, stage);
    upload_cast<double>(st, s->buf[0][0], dP, stage);
    upload_cast<double>(st, s->buf[0][1], U, stage);
    upload_cast<double>(st, s->buf[0][2], V, stage);
  } else {
    upload_cast<float>(st, s->pref, s->pref64, stage);
    upload_cast<float>(st, s->buf[0][0], dP, stage);
    upload_cast<float>(st, s->buf[0][1], U, stage);
    upload_cast<float>(st, s->buf[0][2], V, stage);
  }
  (void)hipMemcpyAsync(s->pref64_d, s->pref64.data(), np * sizeof(double), hipMemcpyHostToDevice, st);
  if (hipStreamSynchronize(st) != hipSuccess) {
    *err = "split_set_state: upload failed";
    return CCS_ERR_DEVICE;
  }
  return CCS_OK;
}

int split_enqueue_step(SplitState* s, hipStream_t st, int scheme, bool, hipEvent_t* ev) {
  if (s->fp64) enqueue_step_t<double>(s, st, scheme, ev);
  else enqueue_step_t<float>(s, st, scheme, ev);
  return CCS_OK;
}

void split_enqueue_rebase(SplitState* s, hipStream_t st) {
  const dim3 grid(static_cast<unsigned>(std::min<std::size_t>((s->np + 255) / 256, 2048))), block(256);
  if (s->fp64)
    hipLaunchKernelGGL(split_rebase<double>, grid, block, 0, st, s->pref64_d, static_cast<double*>(s->pref),
                       static_cast<double*>(s->buf[0][0]), static_cast<unsigned long long>(s->np));
  else
    hipLaunchKernelGGL(split_rebase<float>, grid, block, 0, st, s->pref64_d, static_cast<float*>(s->pref),
                       static_cast<float*>(s->buf[0][0]), static_cast<unsigned long long>(s->np));
}

int split_get_state(SplitState* s, hipStream_t st, bool, double* phiG, double* u, double* v, double* phi,
                    double* uc, double* vc, double* zonal, double* merid, std::string* err) {
  const int n = s->n, h = s->h;
  const std::size_t np = s->np;
  const double r = s->cfg.radius;
  std::vector<double> dP, U, V;
  // the device reference field (a rebase replaces it)
  if (hipMemcpyAsync(s->pref64.data(), s->pref64_d, np * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) {
    *err = "split_get_state: download failed";
    return CCS_ERR_DEVICE;
  }
  if (s->fp64) {
    download_to_double<double>(st, s->buf[0][0], np, dP);
    download_to_double<double>(st, s->buf[0][1], np, U);
    download_to_double<double>(st, s->buf[0][2], np, V);
  } else {
    download_to_double<float>(st, s->buf[0][0], np, dP);
    download_to_double<float>(st, s->buf[0][1], np, U);
    download_to_double<float>(st, s->buf[0][2], np, V);
  }
  if (hipGetLastError() != hipSuccess) {
    *err = "split_get_state: download failed";
    return CCS_ERR_DEVICE;
  }
  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - h; j <= n + h; ++j)
      for (int i = 1 - h; i <= n + h; ++i) {
        const std::size_t d = pidx(n, h, i, j, p);
        const bool corner = is_corner(n, i, j);
        const bool interior = i >= 1 && i <= n && j >= 1 && j <= n;
        const double pg = corner ? 0.0 : r * r * (s->pref64[d] + dP[d]);
        const double uu = corner ? 0.0 : r * U[d], vv = corner ? 0.0 : r * V[d];
        if (phiG) phiG[d] = pg;
        if (u) u[d] = uu;
        if (v) v[d] = vv;
        if (phi) phi[d] = corner ? 0.0 : pg / s->sqrtG[d];
        if (uc) uc[d] = s->IG11[d] * uu + s->IG12[d] * vv;
        if (vc) vc[d] = s->IG21[d] * uu + s->IG22[d] * vv;
        if (zonal) zonal[d] = interior ? s->M11[d] * uu + s->M12[d] * vv : 0.0;
        if (merid) merid[d] = interior ? s->M21[d] * uu + s->M22[d] * vv : 0.0;
      }
  return CCS_OK;
}

int split_tendency(SplitState* s, hipStream_t st, bool, double* tp, double* tu, double* tv, std::string* err) {
  const double r = s->cfg.radius;
  std::vector<double> t0, t1, t2;
  if (s->fp64) {
    SplitArgs<double> a = base_args<double>(s);
    a.in = cprog_of<double>(s, 0); a.base = a.in;
    for (int k = 0; k < 3; ++k) a.tend[k] = static_cast<double*>(s->tend[k]);
    launch_split<double>(st, kTendOnly, s->cfg.spatial_op, a, statics_of<double>(s), s->hs, s->tile_y);
    download_to_double<double>(st, s->tend[0], s->ni, t0);
    download_to_double<double>(st, s->tend[1], s->ni, t1);
    download_to_double<double>(st, s->tend[2], s->ni, t2);
  } else {
    SplitArgs<float> a = base_args<float>(s);
    a.in = cprog_of<float>(s, 0); a.base = a.in;
    for (int k = 0; k < 3; ++k) a.tend[k] = static_cast<float*>(s->tend[k]);
    launch_split<float>(st, kTendOnly, s->cfg.spatial_op, a, statics_of<float>(s), s->hs, s->tile_y);
    download_to_double<float>(st, s->tend[0], s->ni, t0);
    download_to_double<float>(st, s->tend[1], s->ni, t1);
    download_to_double<float>(st, s->tend[2], s->ni, t2);
  }
  if (hipGetLastError() != hipSuccess) {
    *err = "split_tendency: kernel or copy failed";
    return CCS_ERR_DEVICE;
  }
  // back to the reference variables: d(phiG)/dt = a dP/dtau, du/dt = dU/dtau
  for (std::size_t k = 0; k < s->ni; ++k) {
    if (tp) tp[k] = r * t0[k];
    if (tu) tu[k] = t1[k];
    if (tv) tv[k] = t2[k];
  }
  return CCS_OK;
}

int split_diagnostics(SplitState* s, hipStream_t st, bool, double* result, std::string* err) {
  if (s->fp64) {
    hipLaunchKernelGGL(split_diag<double>, dim3(kSplitDiagBlocks), dim3(kDT), 0, st, cprog_of<double>(s, 0),
                       s->pref64_d, s->sqrtG_d, s->phis_d, s->n, s->h, s->cfg.radius, s->partials);
  } else {
    hipLaunchKernelGGL(split_diag<float>, dim3(kSplitDiagBlocks), dim3(kDT), 0, st, cprog_of<float>(s, 0),
                       s->pref64_d, s->sqrtG_d, s->phis_d, s->n, s->h, s->cfg.radius, s->partials);
  }
  static_assert(kSplitDiagBlocks <= 4 * kDT, "partials fit the final block's LDS");
  hipLaunchKernelGGL(split_diag_final, dim3(1), dim3(kDT), 0, st, s->partials, kSplitDiagBlocks, s->result);
  if (hipMemcpyAsync(result, s->result, 2 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    *err = "split_diagnostics: kernel or copy failed";
    return CCS_ERR_DEVICE;
  }
  return CCS_OK;
}

// ------------------------------------------------------------ ensemble plan
struct EnsPlan {
  bool fp64 = false;
  int tile_y = 16, persistent = 0, grid = 0, ntiles = 0, nm = 0;
  int rebase_gx = 8;        // rebase grid: workgroups per member
  void* dmem = nullptr;     // member table, tile records, run step index, probe index, step counters
  void* d_desc = nullptr;
  EnsTile* d_tiles = nullptr;  // per tile: member slot, tile position, the member's target and arrays
  int* d_run = nullptr;     // step index within the current run at the start of the graph being replayed
  int* d_probe = nullptr;   // -1: the step index of the autotuner's launches (every member steps)
  // persistent grid: the claim counters of kClaimSlots launches, used in
  // turn: slot s by the s-th stage of a step (its kernel clears slot s + 1
  // modulo the stages of the step), kProbeSlot and kProbeSlot + 1 by the
  // autotuner's launches in turn
  unsigned* d_claim = nullptr;
  mutable int probe_turn = 0;
  int* d_done = nullptr;    // per member: steps since its upload, modulo kRebaseSteps
  std::vector<char> md_nat;           // member table in member order (EnsMember<T>)
  std::vector<int> tiles_per_member, nbx, nby;
  std::vector<EnsTile> mtile;         // per member: the fields of its tile records that do not vary
  std::vector<int> target;            // targets of the table on the device (empty: none uploaded)
};

namespace {

// The tile records of a member in table slot `slot` with `target` steps:
// x fastest, then y, then panel (the order of its cells in the padded
// layout).
void append_tiles(std::vector<EnsTile>& tiles, const EnsTile& m, int slot, int target, int nbx, int nby) {
  EnsTile r = m;
  r.slot = static_cast<unsigned>(slot);
  r.target = target;
  for (int p = 0; p < 6; ++p)
    for (int by = 0; by < nby; ++by)
      for (int bx = 0; bx < nbx; ++bx) {
        r.pk = tile_pack(p, bx, by);
        tiles.push_back(r);
      }
}

// The member's arrays at a fixed stride (EnsTile::fbase, here still the
// address: the offset from the arena, the member table's allocation, is
// taken once that is allocated), as split_create carves them; false if they
// are not.
bool member_arrays(const SplitState* s, EnsTile* r) {
  const char* b = static_cast<const char*>(s->pref);
  const std::ptrdiff_t d = static_cast<const char*>(s->sg) - b;
  if (d <= 0 || d >= (std::ptrdiff_t{1} << 32)) return false;
  const void* arrs[13] = {s->pref, s->sg, s->phis, s->fs, s->buf[0][0], s->buf[0][1], s->buf[0][2],
                          s->buf[1][0], s->buf[1][1], s->buf[1][2], s->buf[2][0], s->buf[2][1], s->buf[2][2]};
  for (int k = 0; k < 13; ++k)
    if (static_cast<const char*>(arrs[k]) != b + k * d) return false;
  r->fbase = static_cast<long long>(reinterpret_cast<std::uintptr_t>(b));
  r->fstride = static_cast<unsigned>(d);
  r->n = s->n;
  r->h = s->h;
  return true;
}

constexpr int kClaimSlots = 6;
constexpr int kProbeSlot = 4;

// claim counters of slot `slot` and of the slot its launch clears
struct Claims {
  unsigned *cur, *next;
};
Claims claims(const EnsPlan* e, int slot, int next) {
  return Claims{e->d_claim + slot * kClaimGroups * kClaimStride, e->d_claim + next * kClaimGroups * kClaimStride};
}

template <typename T, int TY, bool PERSIST>
void ens_launch_k(const EnsPlan* e, hipStream_t s, int kind, const EnsStage& sel, const int* run, int koff,
                  Claims c) {
  const auto* md = static_cast<const EnsMember<T>*>(e->d_desc);
  const dim3 grid(static_cast<unsigned>(e->grid)), block(TX * TY);
  const int nt = e->ntiles;
  switch (kind) {
    case kTendOnly: hipLaunchKernelGGL((stage_ens<T, kTendOnly, TY, PERSIST>), grid, block, 0, s, md, e->d_tiles, static_cast<char*>(e->d_desc), nt, sel, run, koff, c.cur, c.next); break;
    case kRK3s1: hipLaunchKernelGGL((stage_ens<T, kRK3s1, TY, PERSIST>), grid, block, 0, s, md, e->d_tiles, static_cast<char*>(e->d_desc), nt, sel, run, koff, c.cur, c.next); break;
    case kRK3s2: hipLaunchKernelGGL((stage_ens<T, kRK3s2, TY, PERSIST>), grid, block, 0, s, md, e->d_tiles, static_cast<char*>(e->d_desc), nt, sel, run, koff, c.cur, c.next); break;
    case kRK3s3: hipLaunchKernelGGL((stage_ens<T, kRK3s3, TY, PERSIST>), grid, block, 0, s, md, e->d_tiles, static_cast<char*>(e->d_desc), nt, sel, run, koff, c.cur, c.next); break;
    case kRK4a: hipLaunchKernelGGL((stage_ens<T, kRK4a, TY, PERSIST>), grid, block, 0, s, md, e->d_tiles, static_cast<char*>(e->d_desc), nt, sel, run, koff, c.cur, c.next); break;
    default: hipLaunchKernelGGL((stage_ens<T, kRK4f, TY, PERSIST>), grid, block, 0, s, md, e->d_tiles, static_cast<char*>(e->d_desc), nt, sel, run, koff, c.cur, c.next); break;
  }
}

template <typename T, int TY>
void ens_launch_ty(const EnsPlan* e, hipStream_t s, int kind, const EnsStage& sel, const int* run, int koff,
                   Claims c) {
  if (e->persistent) ens_launch_k<T, TY, true>(e, s, kind, sel, run, koff, c);
  else ens_launch_k<T, TY, false>(e, s, kind, sel, run, koff, c);
}

// stage `stage` of `nstages` of a step (claim slot `stage`, clearing the
// next stage's), or an autotuner launch (stage < 0: the probe slots in turn)
template <typename T>
void ens_launch(const EnsPlan* e, hipStream_t s, int kind, const EnsStage& sel, const int* run, int koff, int stage,
                int nstages) {
  Claims c;
  if (stage >= 0) {
    c = claims(e, stage, (stage + 1) % nstages);
  } else {
    const int k = e->probe_turn;
    e->probe_turn ^= 1;
    c = claims(e, kProbeSlot + k, kProbeSlot + (k ^ 1));
  }
  if (e->tile_y == 8) ens_launch_ty<T, 8>(e, s, kind, sel, run, koff, c);
  else if (e->tile_y == 32) ens_launch_ty<T, 32>(e, s, kind, sel, run, koff, c);
  else ens_launch_ty<T, 16>(e, s, kind, sel, run, koff, c);
}

template <typename T, int TY>
int ens_occupancy() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, stage_ens<T, kRK3s2, TY, true>, TX * TY, 0) != hipSuccess)
    return 0;
  return blocks;
}

template <typename T>
int ens_build(EnsPlan* e, SplitState* const* mem, int nm, int device, const long long* done0, std::string* err) {
  const int TY = e->tile_y;
  std::vector<EnsMember<T>> md(nm);
  std::vector<EnsTile> tiles;
  e->tiles_per_member.resize(nm);
  e->nbx.resize(nm);
  e->nby.resize(nm);
  e->mtile.resize(nm);
  unsigned long long np_max = 0;
  for (int m = 0; m < nm; ++m) {
    const SplitState* s = mem[m];
    EnsMember<T>& d = md[m];
    d.a = base_args<T>(s);
    d.a.nbx = (s->n + TX - 1) / TX;
    d.a.nby = (s->n + TY - 1) / TY;
    d.st = statics_of<T>(s);
    for (int b = 0; b < 3; ++b)
      for (int k = 0; k < 3; ++k) d.buf[b][k] = static_cast<T*>(s->buf[b][k]);
    for (int k = 0; k < 9; ++k) d.tend[k] = static_cast<T*>(s->tend[k]);
    d.pref64 = s->pref64_d;
    d.pref_rw = static_cast<T*>(s->pref);
    d.np = s->np;
    d.inc_half = static_cast<T>(0.5 * s->cfg.dt / s->cfg.radius);
    d.inc_full = static_cast<T>(s->cfg.dt / s->cfg.radius);
    d.target = 0;
    d.member = m;
    for (int q = 0; q < 6; ++q)
      for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 11; ++k) d.hs[q][r][k] = s->hs.r[q][r][k];
    e->tiles_per_member[m] = 6 * d.a.nbx * d.a.nby;
    e->nbx[m] = d.a.nbx;
    e->nby[m] = d.a.nby;
    np_max = std::max(np_max, d.np);
    if (!member_arrays(s, &e->mtile[m])) {
      *err = "ensemble: a member's device arrays are not laid out at a fixed stride";
      return CCS_ERR_INVALID;
    }
    append_tiles(tiles, e->mtile[m], m, 0, d.a.nbx, d.a.nby);
  }
  e->ntiles = static_cast<int>(tiles.size());
  // the rebase: about four cells per thread of the largest member
  e->rebase_gx = static_cast<int>(std::min<unsigned long long>(2048, std::max<unsigned long long>(8, (np_max + 1023) / 1024)));
  auto rup = [](std::size_t b) { return (b + 255) / 256 * 256; };
  const std::size_t bd = rup(md.size() * sizeof(EnsMember<T>)), bt = rup(tiles.size() * sizeof(EnsTile));
  const std::size_t bc = kClaimSlots * kClaimGroups * kClaimStride * sizeof(unsigned);
  const std::size_t bytes = bd + bt + 2 * rup(sizeof(int)) + rup(nm * sizeof(int)) + bc;
  if (hipMalloc(&e->dmem, bytes) != hipSuccess) {
    *err = "ensemble: device allocation failed";
    return CCS_ERR_NOMEM;
  }
  char* p = static_cast<char*>(e->dmem);
  e->d_desc = p;
  // the records hold each member's arrays as an offset from the arena (the
  // member table's address, passed to the stage kernel as a plain pointer)
  const long long table = static_cast<long long>(reinterpret_cast<std::uintptr_t>(e->d_desc));
  for (EnsTile& r : e->mtile) r.fbase -= table;
  for (EnsTile& r : tiles) r.fbase -= table;
  e->d_tiles = reinterpret_cast<EnsTile*>(p + bd);
  e->d_run = reinterpret_cast<int*>(p + bd + bt);
  e->d_probe = reinterpret_cast<int*>(p + bd + bt + rup(sizeof(int)));
  e->d_done = reinterpret_cast<int*>(p + bd + bt + 2 * rup(sizeof(int)));
  e->d_claim = reinterpret_cast<unsigned*>(p + bd + bt + 2 * rup(sizeof(int)) + rup(nm * sizeof(int)));
  const std::vector<unsigned> claim0(bc / sizeof(unsigned), 0u);
  e->md_nat.assign(reinterpret_cast<const char*>(md.data()), reinterpret_cast<const char*>(md.data() + nm));
  std::vector<int> done(nm, 0);
  for (int m = 0; m < nm && done0; ++m) done[m] = static_cast<int>(done0[m] % kRebaseSteps);
  const int minus_one = -1, zero = 0;
  if (hipMemcpy(e->d_desc, md.data(), md.size() * sizeof(EnsMember<T>), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_tiles, tiles.data(), tiles.size() * sizeof(EnsTile), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_run, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_probe, &minus_one, sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_done, done.data(), nm * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_claim, claim0.data(), bc, hipMemcpyHostToDevice) != hipSuccess) {
    *err = "ensemble: table upload failed";
    return CCS_ERR_DEVICE;
  }
  if (e->persistent) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) {
      *err = "ensemble: cannot query the CU count";
      return CCS_ERR_DEVICE;
    }
    const int occ = (TY == 8) ? ens_occupancy<T, 8>() : (TY == 32) ? ens_occupancy<T, 32>() : ens_occupancy<T, 16>();
    e->grid = std::min(e->ntiles, cus * std::max(occ, 1));
  } else {
    e->grid = e->ntiles;
  }
  return CCS_OK;
}

// Member table and tile map of a run: members ordered by decreasing target
// (stable, so equal targets keep the member order), tiles following the
// members.  Synchronous: the stream is drained first, then both tables are
// replaced.
template <typename T>
int ens_set_run_t(EnsPlan* e, hipStream_t st, const int* target, std::string* err) {
  const int nm = e->nm;
  std::vector<int> order(nm);
  for (int m = 0; m < nm; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return target[x] > target[y]; });
  const auto* nat = reinterpret_cast<const EnsMember<T>*>(e->md_nat.data());
  std::vector<EnsMember<T>> md(nm);
  std::vector<EnsTile> tiles;
  tiles.reserve(static_cast<std::size_t>(e->ntiles));
  for (int k = 0; k < nm; ++k) {
    const int m = order[k];
    md[k] = nat[m];
    md[k].target = target[m];
    append_tiles(tiles, e->mtile[m], k, target[m], e->nbx[m], e->nby[m]);
  }
  if (hipStreamSynchronize(st) != hipSuccess ||
      hipMemcpy(e->d_desc, md.data(), nm * sizeof(EnsMember<T>), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_tiles, tiles.data(), tiles.size() * sizeof(EnsTile), hipMemcpyHostToDevice) != hipSuccess) {
    *err = "ensemble: run table upload failed";
    return CCS_ERR_DEVICE;
  }
  e->target.assign(target, target + nm);
  return CCS_OK;
}

// `steps` (1 or kRebaseSteps) steps: the stage launches of each, the rebase
// after step k where bit k of rebase_mask is set, then the advance of the
// step counters.
template <typename T>
void ens_steps_t(const EnsPlan* e, hipStream_t st, int scheme, int steps, unsigned rebase_mask) {
  const int* run = e->d_run;
  const auto* md = static_cast<const EnsMember<T>*>(e->d_desc);
  for (int k = 0; k < steps; ++k) {
    if (scheme == CCS_RK3_TVD) {
      // tile order alternating from launch to launch (xcd_tile_dir)
      const int r = (3 * k) & 1;
      ens_launch<T>(e, st, kRK3s1, EnsStage{0, 0, 1, 0, 0, r}, run, k, 0, 3);
      ens_launch<T>(e, st, kRK3s2, EnsStage{1, 0, 2, 0, 0, r ^ 1}, run, k, 1, 3);
      ens_launch<T>(e, st, kRK3s3, EnsStage{2, 0, 0, 0, 0, r}, run, k, 2, 3);
    } else {
      ens_launch<T>(e, st, kRK4a, EnsStage{0, 0, 1, 0, 0, 0}, run, k, 0, 4);
      ens_launch<T>(e, st, kRK4a, EnsStage{1, 0, 2, 3, 0, 1}, run, k, 1, 4);
      ens_launch<T>(e, st, kRK4a, EnsStage{2, 0, 1, 6, 1, 0}, run, k, 2, 4);
      ens_launch<T>(e, st, kRK4f, EnsStage{1, 0, 0, 0, 0, 1}, run, k, 3, 4);
    }
    if ((rebase_mask >> k) & 1u)
      hipLaunchKernelGGL(ens_rebase<T>, dim3(static_cast<unsigned>(e->rebase_gx), static_cast<unsigned>(e->nm)),
                         dim3(256), 0, st, md, run, k, static_cast<const int*>(e->d_done));
  }
  hipLaunchKernelGGL(ens_advance<T>, dim3(1), dim3(256), 0, st, md, e->nm, e->d_run, e->d_done, steps);
}

}  // namespace

int ens_plan_create(SplitState* const* mem, int nm, int tile_y, int persistent, int device,
                    const long long* steps_done, EnsPlan** out, std::string* err) {
  if (nm <= 0 || !mem || (tile_y != 8 && tile_y != 16 && tile_y != 32)) {
    *err = "ensemble: invalid plan (members, tile height 8/16/32)";
    return CCS_ERR_INVALID;
  }
  const bool fp64 = mem[0]->fp64;
  for (int m = 0; m < nm; ++m)
    if (!mem[m] || mem[m]->fp64 != fp64 || !mem[m]->mesh_set) {
      *err = "ensemble: every member needs the same precision and a mesh";
      return CCS_ERR_ORDER;
    }
  EnsPlan* e = new EnsPlan();
  e->fp64 = fp64;
  e->tile_y = tile_y;
  e->persistent = persistent ? 1 : 0;
  e->nm = nm;
  const int rc = fp64 ? ens_build<double>(e, mem, nm, device, steps_done, err)
                      : ens_build<float>(e, mem, nm, device, steps_done, err);
  if (rc != CCS_OK) {
    ens_plan_destroy(e);
    return rc;
  }
  *out = e;
  return CCS_OK;
}

void ens_plan_destroy(EnsPlan* e) {
  if (!e) return;
  if (e->dmem) (void)hipFree(e->dmem);
  delete e;
}

int ens_plan_set_run(EnsPlan* e, hipStream_t st, const int* target, std::string* err) {
  for (int m = 0; m < e->nm; ++m)
    if (target[m] < 0) {
      *err = "ensemble: negative step count";
      return CCS_ERR_INVALID;
    }
  if (!std::equal(target, target + e->nm, e->target.begin(), e->target.end())) {
    const int rc = e->fp64 ? ens_set_run_t<double>(e, st, target, err) : ens_set_run_t<float>(e, st, target, err);
    if (rc != CCS_OK) return rc;
  }
  hipLaunchKernelGGL(ens_run_start, dim3(1), dim3(64), 0, st, e->d_run);
  return CCS_OK;
}

int ens_plan_reset_member(EnsPlan* e, hipStream_t st, int member) {
  const int zero = 0;
  if (hipStreamSynchronize(st) != hipSuccess ||
      hipMemcpy(e->d_done + member, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return CCS_ERR_DEVICE;
  return CCS_OK;
}

void ens_enqueue_steps(const EnsPlan* e, hipStream_t st, int scheme, int steps, unsigned rebase_mask) {
  if (e->fp64) ens_steps_t<double>(e, st, scheme, steps, rebase_mask);
  else ens_steps_t<float>(e, st, scheme, steps, rebase_mask);
}

void ens_enqueue_probe(const EnsPlan* e, hipStream_t st) {
  // SSP-RK3 stage 2: reads buffers 1 and 0, writes buffer 2, which holds
  // only a stage-2 result between the stages of a step, so the state (buffer
  // 0) is untouched
  if (e->fp64) ens_launch<double>(e, st, kRK3s2, EnsStage{1, 0, 2, 0, 0}, e->d_probe, 0, -1, 0);
  else ens_launch<float>(e, st, kRK3s2, EnsStage{1, 0, 2, 0, 0}, e->d_probe, 0, -1, 0);
}

void ens_enqueue_tendency(const EnsPlan* e, hipStream_t st) {
  if (e->fp64) ens_launch<double>(e, st, kTendOnly, EnsStage{0, 0, 0, 0, 0}, e->d_probe, 0, -1, 0);
  else ens_launch<float>(e, st, kTendOnly, EnsStage{0, 0, 0, 0, 0}, e->d_probe, 0, -1, 0);
}

int ens_plan_grid(const EnsPlan* e) { return e->grid; }
int ens_plan_tile_y(const EnsPlan* e) { return e->tile_y; }
int ens_plan_persistent(const EnsPlan* e) { return e->persistent; }
int ens_plan_tiles(const EnsPlan* e) { return e->ntiles; }

}  // namespace ccs
