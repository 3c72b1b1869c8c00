// ccs_internal.h — layout helpers and the panel-edge halo table shared by the
// host-side sources of libccs (mesh producer, initial conditions, dycore).
#ifndef CCS_INTERNAL_H
#define CCS_INTERNAL_H

#include <cstddef>
#include <cstdint>

namespace ccs {

// Padded index of reference cell (i, j, p), 1-based i, j in 1-h..n+h and
// panel p in 1..6, Fortran order (i fastest).
inline std::size_t pidx(int n, int h, int i, int j, int p) {
  const std::size_t P = static_cast<std::size_t>(n + 2 * h);
  return static_cast<std::size_t>(i + h - 1) +
         P * (static_cast<std::size_t>(j + h - 1) + P * static_cast<std::size_t>(p - 1));
}

// Interior index of reference cell (i, j, p), 1-based i, j in 1..n.
inline std::size_t iidx(int n, int i, int j, int p) {
  const std::size_t N = static_cast<std::size_t>(n);
  return static_cast<std::size_t>(i - 1) +
         N * (static_cast<std::size_t>(j - 1) + N * static_cast<std::size_t>(p - 1));
}

inline std::size_t padded_cells(int n, int h) {
  const std::size_t P = static_cast<std::size_t>(n + 2 * h);
  return P * P * 6;
}

inline bool is_corner(int n, int i, int j) {
  return (i < 1 || i > n) && (j < 1 || j > n);
}

enum Side { kWest = 0, kEast = 1, kSouth = 2, kNorth = 3 };

// One edge strip of the halo: destination panel p (1..6), side s, layer
// k = 1..h (1 nearest the edge), position t = 1..n along the edge.
// Source cell on panel q:  i_s = i0 + in*n + ik*k + it*t,  j_s likewise.
// Vector components and Jacobian columns of the source are re-expressed in
// the destination panel basis by a signed permutation:
//   dst component 1 = s1 * src component c1,  dst component 2 = s2 * src component c2.
struct HaloStrip {
  int q;
  int i0, in, ik, it;
  int j0, jn, jk, jt;
  int c1, s1, c2, s2;  // c in {0,1} (0-based component), s in {-1,+1}
};

// The edge-halo permutation of CubedSphereFillHalo / CubedSphereFillJab
// (reference dynamic_core/src/ghost_mod.f90:34-76 for positions and
// :100-153 for the signed column swaps).  Indexed [p-1][side].
const HaloStrip& halo_strip(int p, int side);

// Destination (1-based) of halo cell (side, k, t) on an n-cell panel.
inline void halo_dst(int n, int side, int k, int t, int* i, int* j) {
  switch (side) {
    case kWest:  *i = 1 - k; *j = t; break;
    case kEast:  *i = n + k; *j = t; break;
    case kSouth: *i = t; *j = 1 - k; break;
    default:     *i = t; *j = n + k; break;
  }
}

inline void halo_src(int n, const HaloStrip& r, int k, int t, int* i, int* j) {
  *i = r.i0 + r.in * n + r.ik * k + r.it * t;
  *j = r.j0 + r.jn * n + r.jk * k + r.jt * t;
}

// Inverse of the halo table, used by the fused stage kernel to scatter each
// updated edge-strip cell into the halos that copy it.  For source panel q
// (1..6) and the strip of width h along side sigma of q:
//   destination panel p, and 1-based destination indices
//     i_d = ei0 + eii*i_s + eij*j_s,  j_d = ej0 + eji*i_s + ejj*j_s,
//   with the vector re-expressed by (c1,s1,c2,s2) as in HaloStrip.
struct ScatterRule {
  int p;
  int ei0, eii, eij;
  int ej0, eji, ejj;
  int c1, s1, c2, s2;
};

// Build the 24 scatter rules for panel size n (indexed [q-1][sigma]).
// Returns false if the inversion of the table fails its self-check.
bool build_scatter_rules(int n, int h, ScatterRule rules[6][4]);

}  // namespace ccs

#endif  // CCS_INTERNAL_H
