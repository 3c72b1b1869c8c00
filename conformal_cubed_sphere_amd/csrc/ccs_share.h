// ccs_share.h — the persistent ensemble grid's partition of the global tile
// list over its workgroups.  Plain C++ apart from the host/device qualifier,
// so a CPU test can compile it with g++ and check the coverage.
#ifndef CCS_SHARE_H
#define CCS_SHARE_H

#ifdef __HIP__
#define CCS_HD __host__ __device__
#else
#define CCS_HD
#endif

namespace ccs {

// Spacing of the claim counters in 32-bit words: one counter per 128-byte
// line, so the groups' atomics do not share a cache line.
constexpr unsigned kClaimStride = 32;
// Counters per launch: one per group (at most 8 groups).
constexpr unsigned kClaimGroups = 8;

struct EnsGroup {
  unsigned g;       // group of the workgroup: its claim counter
  unsigned lo, hi;  // the group's contiguous tiles [lo, hi) of the global list
};

// Group of workgroup b of nb: the workgroups form min(nb, 8) groups (group
// g = b mod 8 runs on XCD g when nb >= 8, workgroups being dispatched round
// robin over the 8 XCDs), and group g owns the contiguous tiles [lo, hi).
// Its workgroups claim them one at a time from the group's counter (tile
// lo + the counter's old value, until that reaches hi), so every tile is
// claimed exactly once and a slow workgroup takes fewer tiles.  The ranges
// cover the list exactly, for every nb >= 1.
CCS_HD inline EnsGroup ens_group(unsigned b, unsigned nb, unsigned ntiles) {
  const unsigned ng = nb < kClaimGroups ? nb : kClaimGroups;
  const unsigned g = b % ng;
  EnsGroup s;
  s.g = g;
  s.lo = static_cast<unsigned>((static_cast<unsigned long long>(ntiles) * g) / ng);
  s.hi = static_cast<unsigned>((static_cast<unsigned long long>(ntiles) * (g + 1)) / ng);
  return s;
}

}  // namespace ccs

#endif
