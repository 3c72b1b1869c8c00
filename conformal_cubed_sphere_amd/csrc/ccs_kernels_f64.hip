// ccs_kernels_f64.hip — FP64 parity kernels for gfx950.
//
// One launch performs one Runge-Kutta stage of the reference time step:
//   spatial_operator (spatial_operators_mod.F90:13-181)
//   + the stage update (time_scheme_mod.F90:69-115)
//   + correct_bdy_ghost (time_scheme_mod.F90:117-125): the interior
//     covariant -> sphere -> covariant round trip and the edge-halo exchange,
//   so each stage reads the input state once and writes the output state once.
//
// Arithmetic follows the reference expression by expression (this file is
// compiled with -ffp-contract=off; divisions and square roots are the IEEE
// correctly rounded ones), which keeps the branchy WENO-Z weight selection
// (spatial_operators_mod.F90:348-355) on the same side as the CPU reference.
//
// Tiling: a 256-thread workgroup owns a TX x TY tile of one panel (TY = 8,
// or 7 where a grid of such tiles fits one round of the GPU's workgroup
// slots: 487 faces then take two passes of the workgroup, not three, which
// shortens a launch that is one round of workgroups long).  Phase 1
// evaluates the pointwise terms (phi, contravariant winds, Bernoulli function,
// mass fluxes) on the tile plus a 3-cell cross-shaped border into LDS.
// Phase 2 evaluates every x- and y-face of the tile once (six WENO-Z
// reconstructions and the Rusanov flux per face).  Phase 3 forms the
// divergence, the 6th-order vorticity and the stage update per cell, writes
// the cell and scatters edge-strip cells into the halos of the neighbouring
// panels (the inverse of ghost_mod.f90:34-76), rotating the wind through the
// destination panel's basis as convert_wind_SP2P does.
#include <hip/hip_ext.h>

#include "ccs_kernels.h"

namespace ccs {
namespace {

constexpr int TX = 32;
constexpr int NT = 256;
constexpr int EX = TX + 6;

__device__ __forceinline__ long gidx(int i, int j, int p, int h, int P) {
  return static_cast<long>(i + h) + static_cast<long>(P) * (static_cast<long>(j + h) + static_cast<long>(P) * p);
}

// RN(a / b) for a divisor b whose correctly rounded reciprocal y = RN(1/b)
// is known: q = RN(a y), r = a - b q (exact through the FMA), RN(q + r y).
// This is Markstein's correction step; with y correctly rounded it returns
// the correctly rounded quotient (bit-identical to the IEEE division) for
// operands whose quotient is normal, in 3 instructions instead of the ~11 of
// the general FP64 division sequence.
__device__ __forceinline__ double div_known(double a, double b, double y) {
  const double q = a * y;
  const double r = fma(-q, b, a);
  return fma(r, y, q);
}

// WENO_limiter with dir > 0 on the window (q1..q5) centred on q3
// (spatial_operators_mod.F90:282-361).  dir < 0 is the same expression on the
// reversed window, weno5(q5,q4,q3,q2,q1).
__device__ __forceinline__ double weno5(double q1, double q2, double q3, double q4, double q5) {
  const double s1 = div_known(q1, 3., 1. / 3.) - 7. / 6. * q2 + 11. / 6. * q3;
  const double s2 = -div_known(q2, 6., 1. / 6.) + 5. / 6. * q3 + 1. / 3. * q4;
  const double s3 = div_known(q3, 3., 1. / 3.) + 5. / 6. * q4 - 1. / 6. * q5;
  const double a1 = q1 - 2. * q2 + q3;
  const double a2 = q2 - 2. * q3 + q4;
  const double a3 = q3 - 2. * q4 + q5;
  const double b1 = q1 - 4. * q2 + 3. * q3;
  const double b2 = q2 - q4;
  const double b3 = 3. * q3 - 4. * q4 + q5;
  const double be1 = div_known(a1 * a1 * 13., 12., 1. / 12.) + b1 * b1 * 0.25;
  const double be2 = div_known(a2 * a2 * 13., 12., 1. / 12.) + b2 * b2 * 0.25;
  const double be3 = div_known(a3 * a3 * 13., 12., 1. / 12.) + b3 * b3 * 0.25;
  const double t40 = fabs(be1 - be2);
  const double t41 = fabs(be2 - be3);
  const double t5 = fabs(be3 - be1);
  double mb = be1;
  if (be2 < mb) mb = be2;
  if (be3 < mb) mb = be3;
  double w1, w2, w3;
  if (t40 <= mb && t41 > mb) {
    w1 = 1. / 3.; w2 = 2. / 3.; w3 = 0.;
  } else if (t40 > mb && t41 < mb) {
    w1 = 0.; w2 = 2. / 3.; w3 = 1. / 3.;
  } else {
    const double al1 = 0.1 * (1. + t5 / (1.E-2 + be1));
    const double al2 = 0.6 * (1. + t5 / (1.E-2 + be2));
    const double al3 = 0.3 * (1. + t5 / (1.E-2 + be3));
    const double sa = al1 + al2 + al3;
    // the three quotients by sa from its correctly rounded reciprocal
    const double rsa = 1. / sa;
    w1 = div_known(al1, sa, rsa); w2 = div_known(al2, sa, rsa); w3 = div_known(al3, sa, rsa);
  }
  return s1 * w1 + s2 * w2 + s3 * w3;
}

__device__ __forceinline__ double d6(double m3, double m2, double m1, double p1, double p2, double p3,
                                     double dh, double r60dh) {
  return div_known(-m3 + 9. * m2 - 45. * m1 + 45. * p1 - 9. * p2 + p3, 60. * dh, r60dh);
}

// Spatial operator variants (enum ccs_spatial_op): 0 WENO-Z, 1/2/3 centred
// 2nd/4th/6th order, 4 WENO-Z with the extended-lattice metric and upwind
// face winds in the vorticity.
template <int KIND, bool EXPL, int OP, int TY>
__global__ void __launch_bounds__(NT) stage_f64(StageArgsF64 a, StaticsF64 st, HaloScatter hs) {
  constexpr bool kWeno = (OP == 0 || OP == 4);
  constexpr int EY = TY + 6;
  constexpr int NE = EX * EY;
  static_assert(TX * TY <= NT, "one thread per tile cell");
  // LDS, sized so that four workgroups fit a CU.  Fields read along both
  // directions cover the extended tile (row-major [lj][li], EX x EY); fields
  // read only along x cover the tile's rows (XB: [lj-3][li], EX x TY), those
  // read only along y the tile's columns (YB: [lj][li-3], TX x EY).  The face
  // values are written after a barrier into the storage of sFx/sFy/sE, which
  // phase 3 no longer reads.  OP 4 takes its face speeds from a table, so the
  // storage of G^-1_11 / G^-1_22 holds its upwind face winds instead.
  constexpr int NXB = EX * TY, NYB = TX * EY;
  constexpr int NFX = (TX + 1) * TY;
  constexpr int NFY = TX * (TY + 1);
  static_assert(2 * NFX + 2 * NFY <= NXB + NYB + NE, "face values fit the dead field storage");
  static_assert((NT & (NT - 1)) == 0 && NT % 64 == 0, "the last face pass rotates whole waves");
  __shared__ double s_xy[NXB + NYB + NE];  // sFx | sFy | sE, later the face values
  __shared__ double sPtG[NE], sPt[NE];
  __shared__ double sUC[NXB], sG11[NXB], sV[NXB];
  __shared__ double sVC[NYB], sG22[NYB], sU[NYB];
  double* const sFx = s_xy;
  double* const sFy = s_xy + NXB;
  double* const sE = s_xy + NXB + NYB;
  double* const sFxF = s_xy;
  double* const sExF = s_xy + NFX;
  double* const sFyF = s_xy + 2 * NFX;
  double* const sEyF = s_xy + 2 * NFX + NFY;
  auto xb = [](int lj, int li) { return (lj - 3) * EX + li; };
  auto yb = [](int lj, int li) { return lj * TX + li - 3; };

  const int n = a.n, h = a.h, P = n + 2 * h;
  // XCD-aware tile order: workgroup b runs on XCD b % 8; each XCD gets a
  // contiguous run of tiles (x fastest, then y, then panel) so the borders
  // shared by neighbouring tiles are fetched into one L2.
  const unsigned nbx = (static_cast<unsigned>(a.n) + TX - 1) / TX, nby = (static_cast<unsigned>(a.n) + TY - 1) / TY;
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned xcd = b & 7u;
  const unsigned tf = xcd * (nb >> 3) + min(xcd, nb & 7u) + (b >> 3);
  // rev: this XCD's run of tiles backwards (StageArgsF64::rev)
  const unsigned len = (nb >> 3) + (xcd < (nb & 7u) ? 1u : 0u);
  const unsigned t = a.rev ? tf - (b >> 3) + (len - 1u - (b >> 3)) : tf;
  const unsigned tq = t / nbx;
  const int p = static_cast<int>(tq / nby);
  const int i0 = static_cast<int>(t - tq * nbx) * TX, j0 = static_cast<int>(tq - static_cast<unsigned>(p) * nby) * TY;
  const int tid = threadIdx.x;

  // ---- phase 1: pointwise terms on the cross-shaped extended tile
  constexpr int NX = EX * TY;       // rows of the tile, all extended columns
  constexpr int NCROSS = NX + 6 * TX;
  for (int e = tid; e < NCROSS; e += NT) {
    int li, lj;
    if (e < NX) {
      li = e % EX; lj = 3 + e / EX;
    } else {
      const int e2 = e - NX;
      li = 3 + e2 % TX;
      const int r = e2 / TX;
      lj = (r < 3) ? r : TY + r;
    }
    const int i = i0 + li - 3, j = j0 + lj - 3;
    if (i < -h || i >= n + h || j < -h || j >= n + h) continue;
    if ((i < 0 || i >= n) && (j < 0 || j >= n)) continue;  // panel corner: never read
    const long g = gidx(i, j, p, h, P);
    const double phiG = a.in.phiG[g], u = a.in.u[g], v = a.in.v[g];
    const double sg = st.sqrtG[g];
    const double g11 = st.IG11[g], g22 = st.IG22[g];
    // (OP 4 reads G^-1 only for the contravariant winds; OP 0 also for the face speeds)
    double phi, uc, vc;
    if (EXPL) {
      phi = a.phi0[g]; uc = a.uc0[g]; vc = a.vc0[g];
    } else {
      phi = phiG / sg;                           // time_scheme_mod.F90:81
      uc = g11 * u + st.IG12[g] * v;             // projection_mod.F90:145-146
      vc = st.IG21[g] * u + g22 * v;
    }
    const double phit = phi + st.phis[g];        // spatial_operators_mod.F90:71-78
    const double phitG = phit * sg;
    const double K = 0.5 * (u * uc + v * vc);
    const int l = lj * EX + li;
    sE[l] = phit + K;
    sPtG[l] = phitG;
    sPt[l] = phit;
    if (lj >= 3 && lj < TY + 3) {  // tile rows: x-direction fields
      const int lx = xb(lj, li);
      sFx[lx] = phiG * uc;
      sUC[lx] = uc;
      if (OP == 0) sG11[lx] = g11;
      sV[lx] = v;
    }
    if (li >= 3 && li < TX + 3) {  // tile columns: y-direction fields
      const int ly = yb(lj, li);
      sFy[ly] = phiG * vc;
      sVC[ly] = vc;
      if (OP == 0) sG22[ly] = g22;
      sU[ly] = u;
    }
  }
  __syncthreads();

  // ---- phase 2: faces (spatial_operators_mod.F90:80-149), held in registers
  // until every thread has finished reading the fields.  x and y faces run
  // the same code on strided LDS views, so no wave executes both variants.
  // Faces [2 NT, NFX + NFY) fill only part of a third pass; which wave takes
  // them rotates with the workgroup so that the SIMDs share the extra work.
  constexpr int NFACE_IT = kWeno ? (NFX + NFY + NT - 1) / NT : 0;
  double fF[NFACE_IT + 1], fE[NFACE_IT + 1], fW[NFACE_IT + 1];
#pragma unroll
  for (int it = 0; it < NFACE_IT; ++it) {
    const int e = it * NT + ((it == NFACE_IT - 1) ? ((tid + 64 * static_cast<int>(b & 3u)) & (NT - 1)) : tid);
    fF[it] = 0.;
    fE[it] = 0.;
    fW[it] = 0.;
    if (e >= NFX + NFY) continue;
    // band offset/stride (fields read along one direction) and full-tile
    // offset/stride of the face's left (lower) cell
    int ob, sb, of, sf;
    bool ok;
    long fq;  // OP 4: index of the face in its speed table
    if (e < NFX) {
      const int fi = e % (TX + 1), ty = e / (TX + 1);
      const int ig = i0 + fi - 1, j = j0 + ty;
      ok = j < n && ig >= -1 && ig <= n - 1;
      ob = ty * EX + fi + 2; sb = 1;           // x band: row ty of the tile
      of = (ty + 3) * EX + fi + 2; sf = 1;
      fq = (ig + 1) + static_cast<long>(n + 1) * (j + static_cast<long>(n) * p);
    } else {
      const int e2 = e - NFX;
      const int tx = e2 % TX, fj = e2 / TX;
      const int i = i0 + tx, jg = j0 + fj - 1;
      ok = i < n && jg >= -1 && jg <= n - 1;
      ob = (fj + 2) * TX + tx; sb = TX;        // y band: column tx of the tile
      of = (fj + 2) * EX + tx + 3; sf = EX;
      fq = i + static_cast<long>(n) * ((jg + 1) + static_cast<long>(n + 1) * p);
    }
    if (!ok) continue;
    const double* F = (e < NFX ? sFx : sFy) + ob;
    const double* U = (e < NFX ? sUC : sVC) + ob;
    const double* G = (e < NFX ? sG11 : sG22) + ob;
    const double* E = sE + of;
    const double* Q = sPtG + of;
    const double* Pt = sPt + of;
    const double Fp = weno5(F[-2 * sb], F[-sb], F[0], F[sb], F[2 * sb]);
    const double Fn = weno5(F[3 * sb], F[2 * sb], F[sb], F[0], F[-sb]);
    const double Ep = weno5(E[-2 * sf], E[-sf], E[0], E[sf], E[2 * sf]);
    const double En = weno5(E[3 * sf], E[2 * sf], E[sf], E[0], E[-sf]);
    const double Pp = weno5(Q[-2 * sf], Q[-sf], Q[0], Q[sf], Q[2 * sf]);
    const double Pn = weno5(Q[3 * sf], Q[2 * sf], Q[sf], Q[0], Q[-sf]);
    const double uB = 0.5 * (U[sb] + U[0]);
    const double pB = 0.5 * (Pt[sf] + Pt[0]);
    // OP 4: matrixIG_ext at lattice point (2i+1, j) / (i, 2j+1)
    // (spatial_operators_mod_WENO.F90:111-112,134-135)
    const double gq = (OP == 4) ? (e < NFX ? a.igx : a.igy)[fq] * pB : 0.5 * (G[sb] + G[0]) * pB;
    const double e1 = uB - sqrt(gq), e2 = uB + sqrt(gq);
    const double lam = fabs(e1) > fabs(e2) ? fabs(e1) : fabs(e2);
    fF[it] = 0.5 * (Fp + Fn - lam * (Pn - Pp));
    fE[it] = 0.5 * (Ep + En - lam * (0. - 0.));
    if (OP == 4) {
      // upwind face value of the tangential covariant wind: v at x faces, u at
      // y faces (spatial_operators_mod_WENO.F90:88-89,98-99,118,141)
      const double* W = (e < NFX ? sV : sU) + ob;
      const double Wp = weno5(W[-2 * sb], W[-sb], W[0], W[sb], W[2 * sb]);
      const double Wn = weno5(W[3 * sb], W[2 * sb], W[sb], W[0], W[-sb]);
      fW[it] = 0.5 * (Wp + Wn - copysign(1., uB) * (Wn - Wp));
    }
  }
  if (kWeno) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NFACE_IT; ++it) {
      const int e = it * NT + ((it == NFACE_IT - 1) ? ((tid + 64 * static_cast<int>(b & 3u)) & (NT - 1)) : tid);
      if (e < NFX) {
        sFxF[e] = fF[it];  // e = ty * (TX + 1) + fi
        sExF[e] = fE[it];
        if (OP == 4) sG11[e] = fW[it];
      } else if (e < NFX + NFY) {
        sFyF[e - NFX] = fF[it];  // fj * TX + tx
        sEyF[e - NFX] = fE[it];
        if (OP == 4) sG22[e - NFX] = fW[it];
      }
    }
    __syncthreads();
  }

  // ---- phase 3: cells
  const int tx = tid % TX, ty = tid / TX;
  const int i = i0 + tx, j = j0 + ty;
  if (ty >= TY || i >= n || j >= n) return;
  const double dx = a.dx;
  const double rdx = a.rdx, r60 = a.r60dx;
  const int lxc = ty * EX + (tx + 3);   // x band, this cell
  const int lyc = (ty + 3) * TX + tx;   // y band, this cell
  const int lc = (ty + 3) * EX + (tx + 3);
  double dfluxdx, dEdx, dvdx, dfluxdy, dEdy, dudy;
  if (kWeno) {
    const int fx = ty * (TX + 1) + tx, fy = ty * TX + tx;
    dfluxdx = div_known(sFxF[fx + 1] - sFxF[fx], dx, rdx);
    dEdx = div_known(sExF[fx + 1] - sExF[fx], dx, rdx);
    dfluxdy = div_known(sFyF[fy + TX] - sFyF[fy], dx, rdx);
    dEdy = div_known(sEyF[fy + TX] - sEyF[fy], dx, rdx);
    if (OP == 4) {  // spatial_operators_mod_WENO.F90:160,163
      dvdx = div_known(sG11[fx + 1] - sG11[fx], dx, rdx);
      dudy = div_known(sG22[fy + TX] - sG22[fy], dx, rdx);
    } else {
      dvdx = d6(sV[lxc - 3], sV[lxc - 2], sV[lxc - 1], sV[lxc + 1], sV[lxc + 2], sV[lxc + 3], dx, r60);
      dudy = d6(sU[lyc - 3 * TX], sU[lyc - 2 * TX], sU[lyc - TX], sU[lyc + TX], sU[lyc + 2 * TX],
                sU[lyc + 3 * TX], dx, r60);
    }
  } else {
    // centred differences of the cell values (x band: stride 1, y band:
    // stride TX, full tile: stride EX)
    auto cd = [&](const double* f, int l, int s) {
      if (OP == 1)  // spatial_operators_mod_2nd.F90:54-59
        return div_known(f[l + s] - f[l - s], a.dx2, a.rdx2);
      if (OP == 2)  // spatial_operators_mod_4th.F90:58-63
        return div_known(f[l - 2 * s] - 8. * f[l - s] + 8. * f[l + s] - f[l + 2 * s], a.dx12, a.rdx12);
      // spatial_operators_mod_6th.F90:64-69
      return d6(f[l - 3 * s], f[l - 2 * s], f[l - s], f[l + s], f[l + 2 * s], f[l + 3 * s], dx, r60);
    };
    dfluxdx = cd(sFx, lxc, 1);
    dEdx = cd(sE, lc, 1);
    dvdx = cd(sV, lxc, 1);
    dfluxdy = cd(sFy, lyc, TX);
    dEdy = cd(sE, lc, EX);
    dudy = cd(sU, lyc, TX);
  }
  const long g = gidx(i, j, p, h, P);
  const double vort = dvdx - dudy + st.sqrtG[g] * st.f[g];
  const double tp = -(dfluxdx + dfluxdy);
  const double tu = -dEdx + vort * sVC[lyc];
  const double tv = -dEdy - vort * sUC[lxc];

  const long ii = static_cast<long>(i) + static_cast<long>(n) * (static_cast<long>(j) + static_cast<long>(n) * p);
  if (KIND == kTendOnly) {
    a.tend[0][ii] = tp;
    a.tend[1][ii] = tu;
    a.tend[2][ii] = tv;
    return;
  }
  double pn, un, vn;
  if (KIND == kRK3s1) {
    pn = a.base.phiG[g] + a.dt * tp;
    un = a.base.u[g] + a.dt * tu;
    vn = a.base.v[g] + a.dt * tv;
  } else if (KIND == kRK3s2) {
    pn = 0.75 * a.base.phiG[g] + 0.25 * a.in.phiG[g] + 0.25 * a.dt * tp;
    un = 0.75 * a.base.u[g] + 0.25 * a.in.u[g] + 0.25 * a.dt * tu;
    vn = 0.75 * a.base.v[g] + 0.25 * a.in.v[g] + 0.25 * a.dt * tv;
  } else if (KIND == kRK3s3) {
    pn = div_known(a.base.phiG[g], 3., 1. / 3.) + 2. / 3. * a.in.phiG[g] + 2. / 3. * a.dt * tp;
    un = div_known(a.base.u[g], 3., 1. / 3.) + 2. / 3. * a.in.u[g] + 2. / 3. * a.dt * tu;
    vn = div_known(a.base.v[g], 3., 1. / 3.) + 2. / 3. * a.in.v[g] + 2. / 3. * a.dt * tv;
  } else if (KIND == kRK4a) {
    a.tend[0][ii] = tp;
    a.tend[1][ii] = tu;
    a.tend[2][ii] = tv;
    pn = a.base.phiG[g] + a.inc * tp;
    un = a.base.u[g] + a.inc * tu;
    vn = a.base.v[g] + a.inc * tv;
  } else {  // kRK4f
    const double T0 = div_known(a.tacc[0][ii] + 2. * a.tacc[3][ii] + 2. * a.tacc[6][ii] + tp, 6., 1. / 6.);
    const double T1 = div_known(a.tacc[1][ii] + 2. * a.tacc[4][ii] + 2. * a.tacc[7][ii] + tu, 6., 1. / 6.);
    const double T2 = div_known(a.tacc[2][ii] + 2. * a.tacc[5][ii] + 2. * a.tacc[8][ii] + tv, 6., 1. / 6.);
    pn = a.base.phiG[g] + a.dt * T0;
    un = a.base.u[g] + a.dt * T1;
    vn = a.base.v[g] + a.dt * T2;
  }
  // correct_bdy_ghost: convert_wind_P2SP on the interior cell ...
  const double zw = st.M11[g] * un + st.M12[g] * vn;
  const double mw = st.M21[g] * un + st.M22[g] * vn;
  // ... convert_wind_SP2P back on the same cell
  a.out.phiG[g] = pn;
  a.out.u[g] = st.N11[g] * zw + st.N12[g] * mw;
  a.out.v[g] = st.N21[g] * zw + st.N22[g] * mw;
  // ... and the halo copies of (phiG, zonal, meridional) rotated into the
  // destination panel basis.
  const int i1 = i + 1, j1 = j + 1;
  const bool strip[4] = {i1 <= h, i1 >= n - h + 1, j1 <= h, j1 >= n - h + 1};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (!strip[s]) continue;
    const int* r = hs.r[p][s];
    const int pd = r[0] - 1;
    const int id = r[1] + r[2] * i1 + r[3] * j1;
    const int jd = r[4] + r[5] * i1 + r[6] * j1;
    const long gd = gidx(id - 1, jd - 1, pd, h, P);
    a.out.phiG[gd] = pn;
    a.out.u[gd] = st.N11[gd] * zw + st.N12[gd] * mw;
    a.out.v[gd] = st.N21[gd] * zw + st.N22[gd] * mw;
  }
}

__global__ void derived_f64(Prog<const double> q, StaticsF64 st, int n, int h, double* phi, double* uc,
                            double* vc, double* zonal, double* merid) {
  const int P = n + 2 * h;
  const long np = static_cast<long>(P) * P * 6;
  for (long g = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; g < np;
       g += static_cast<long>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(g % P) - h;
    const int j = static_cast<int>((g / P) % P) - h;
    const bool corner = (i < 0 || i >= n) && (j < 0 || j >= n);
    const bool interior = i >= 0 && i < n && j >= 0 && j < n;
    if (corner) {
      if (phi) phi[g] = 0.;
      if (uc) uc[g] = 0.;
      if (vc) vc[g] = 0.;
      if (zonal) zonal[g] = 0.;
      if (merid) merid[g] = 0.;
      continue;
    }
    const double u = q.u[g], v = q.v[g];
    if (phi) phi[g] = q.phiG[g] / st.sqrtG[g];
    if (uc) uc[g] = st.IG11[g] * u + st.IG12[g] * v;
    if (vc) vc[g] = st.IG21[g] * u + st.IG22[g] * v;
    if (zonal) zonal[g] = interior ? st.M11[g] * u + st.M12[g] * v : 0.;
    if (merid) merid[g] = interior ? st.M21[g] * u + st.M22[g] * v : 0.;
  }
}

// Neumaier compensated accumulation.
struct KSum {
  double s, c;
  __device__ void add(double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
    s = t;
  }
  __device__ void merge(const KSum& o) {
    add(o.s);
    c += o.c;
  }
};

template <int NTH>
__device__ KSum block_reduce(KSum v, double* sh_s, double* sh_c) {
  sh_s[threadIdx.x] = v.s;
  sh_c[threadIdx.x] = v.c;
  __syncthreads();
  for (int off = NTH / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      KSum a{sh_s[threadIdx.x], sh_c[threadIdx.x]};
      a.merge(KSum{sh_s[threadIdx.x + off], sh_c[threadIdx.x + off]});
      sh_s[threadIdx.x] = a.s;
      sh_c[threadIdx.x] = a.c;
    }
    __syncthreads();
  }
  return KSum{sh_s[0], sh_c[0]};
}

constexpr int kDiagThreads = 256;

__global__ void __launch_bounds__(kDiagThreads)
    diag_partial_f64(Prog<const double> q, const double* phi0, const double* uc0, const double* vc0,
                     StaticsF64 st, int n, int h, double* partials) {
  __shared__ double sh_s[kDiagThreads], sh_c[kDiagThreads];
  const int P = n + 2 * h;
  const long ni = static_cast<long>(n) * n * 6;
  KSum m{0., 0.}, e{0., 0.};
  for (long c = blockIdx.x * static_cast<long>(kDiagThreads) + threadIdx.x; c < ni;
       c += static_cast<long>(gridDim.x) * kDiagThreads) {
    const int i = static_cast<int>(c % n), j = static_cast<int>((c / n) % n), p = static_cast<int>(c / (static_cast<long>(n) * n));
    const long g = gidx(i, j, p, h, P);
    const double phiG = q.phiG[g], u = q.u[g], v = q.v[g];
    double phi, uc, vc;
    if (phi0) {
      phi = phi0[g]; uc = uc0[g]; vc = vc0[g];
    } else {
      phi = phiG / st.sqrtG[g];
      uc = st.IG11[g] * u + st.IG12[g] * v;
      vc = st.IG21[g] * u + st.IG22[g] * v;
    }
    // diag_mod.F90:20-26 and :40-52
    const double KE = 0.5 * phi * (u * uc + v * vc);
    const double pt = phi + st.phis[g];
    const double PE = 0.5 * (pt * pt);
    m.add(phiG);
    e.add(st.sqrtG[g] * (KE + PE));
  }
  KSum rm = block_reduce<kDiagThreads>(m, sh_s, sh_c);
  __syncthreads();
  KSum re = block_reduce<kDiagThreads>(e, sh_s, sh_c);
  if (threadIdx.x == 0) {
    partials[4 * blockIdx.x + 0] = rm.s;
    partials[4 * blockIdx.x + 1] = rm.c;
    partials[4 * blockIdx.x + 2] = re.s;
    partials[4 * blockIdx.x + 3] = re.c;
  }
}

__global__ void __launch_bounds__(kDiagThreads) diag_final_f64(const double* partials, int nb, double* result) {
  __shared__ double sh_s[kDiagThreads], sh_c[kDiagThreads];
  KSum m{0., 0.}, e{0., 0.};
  for (int b = threadIdx.x; b < nb; b += kDiagThreads) {
    m.merge(KSum{partials[4 * b + 0], partials[4 * b + 1]});
    e.merge(KSum{partials[4 * b + 2], partials[4 * b + 3]});
  }
  KSum rm = block_reduce<kDiagThreads>(m, sh_s, sh_c);
  __syncthreads();
  KSum re = block_reduce<kDiagThreads>(e, sh_s, sh_c);
  if (threadIdx.x == 0) {
    result[0] = rm.s + rm.c;
    result[1] = re.s + re.c;
  }
}

}  // namespace

int stage_f64_tile_y(int n, int cus) {
  const long wg7 = static_cast<long>((n + TX - 1) / TX) * ((n + 6) / 7) * 6;
  return wg7 <= 4L * cus ? 7 : 8;  // four workgroups fit a CU
}

// With an event pair, the events are the dispatch's own start and end
// (hipExtLaunchKernelGGL), so they time the kernel alone.
template <typename K>
void launch_f64(K kernel, dim3 grid, dim3 block, hipStream_t s, const hipEvent_t* ev, const StageArgsF64& a,
                const StaticsF64& st, const HaloScatter& hs) {
  if (ev) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev[0], ev[1], 0, a, st, hs);
  else hipLaunchKernelGGL(kernel, grid, block, 0, s, a, st, hs);
}

void launch_stage_f64(hipStream_t s, int kind, bool expl, const StageArgsF64& a_in, const StaticsF64& st,
                      const HaloScatter& hs, int tile_y, const hipEvent_t* ev) {
  StageArgsF64 a = a_in;
  a.rdx = 1.0 / a.dx;            // correctly rounded reciprocals for div_known
  a.r60dx = 1.0 / (60.0 * a.dx);
  a.dx2 = 2.0 * a.dx;
  a.rdx2 = 1.0 / a.dx2;
  a.dx12 = 12.0 * a.dx;
  a.rdx12 = 1.0 / a.dx12;
  const int ty = tile_y == 7 ? 7 : 8;
  const dim3 grid(static_cast<unsigned>(((a.n + TX - 1) / TX) * ((a.n + ty - 1) / ty) * 6));
  const dim3 block(NT);
#define CCS_LAUNCH_TY(K, E, O)                                                           \
  do {                                                                                  \
    if (ty == 7) launch_f64(stage_f64<K, E, O, 7>, grid, block, s, ev, a, st, hs); \
    else launch_f64(stage_f64<K, E, O, 8>, grid, block, s, ev, a, st, hs);         \
  } while (0)
#define CCS_LAUNCH_OP(K, O)                       \
  do {                                           \
    if (expl) CCS_LAUNCH_TY(K, true, O);         \
    else CCS_LAUNCH_TY(K, false, O);             \
  } while (0)
#define CCS_LAUNCH(K)                      \
  do {                                     \
    switch (a.op) {                        \
      case 1: CCS_LAUNCH_OP(K, 1); break;  \
      case 2: CCS_LAUNCH_OP(K, 2); break;  \
      case 3: CCS_LAUNCH_OP(K, 3); break;  \
      case 4: CCS_LAUNCH_OP(K, 4); break;  \
      default: CCS_LAUNCH_OP(K, 0); break; \
    }                                      \
  } while (0)
  switch (kind) {
    case kTendOnly: CCS_LAUNCH(kTendOnly); break;
    case kRK3s1: CCS_LAUNCH(kRK3s1); break;
    case kRK3s2: CCS_LAUNCH(kRK3s2); break;
    case kRK3s3: CCS_LAUNCH(kRK3s3); break;
    case kRK4a: CCS_LAUNCH(kRK4a); break;
    default: CCS_LAUNCH(kRK4f); break;
  }
#undef CCS_LAUNCH
#undef CCS_LAUNCH_OP
#undef CCS_LAUNCH_TY
}

void launch_derived_f64(hipStream_t s, const Prog<const double>& q, const StaticsF64& st, int n, int h,
                        double* phi, double* uc, double* vc, double* zonal, double* merid) {
  hipLaunchKernelGGL(derived_f64, dim3(1024), dim3(256), 0, s, q, st, n, h, phi, uc, vc, zonal, merid);
}

void launch_diag_f64(hipStream_t s, const Prog<const double>& q, const double* phi, const double* uc,
                     const double* vc, const StaticsF64& st, int n, int h, double* partials,
                     double* result) {
  hipLaunchKernelGGL(diag_partial_f64, dim3(kDiagBlocks), dim3(kDiagThreads), 0, s, q, phi, uc, vc, st, n,
                     h, partials);
  hipLaunchKernelGGL(diag_final_f64, dim3(1), dim3(kDiagThreads), 0, s, partials, kDiagBlocks, result);
}

}  // namespace ccs
