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

struct EnsShare {
  unsigned first, end, stride;  // tiles first, first + stride, ... below end
};

// Share of workgroup b of nb: the workgroups form min(nb, 8) groups (group
// g = b mod 8 runs on XCD g when nb >= 8, workgroups being dispatched round
// robin over the 8 XCDs), group g owns the contiguous tiles [lo, hi) of the
// global list and its workgroups take them in turn (first tile lo + b / groups,
// stride the group's workgroup count).  Every tile belongs to exactly one
// group and is visited by exactly one of its workgroups, for every nb >= 1.
CCS_HD inline EnsShare ens_share(unsigned b, unsigned nb, unsigned ntiles) {
  const unsigned ng = nb < 8u ? nb : 8u;
  const unsigned g = b % ng;
  EnsShare s;
  s.first = static_cast<unsigned>((static_cast<unsigned long long>(ntiles) * g) / ng) + b / ng;
  s.end = static_cast<unsigned>((static_cast<unsigned long long>(ntiles) * (g + 1)) / ng);
  s.stride = nb / ng + (g < nb % ng ? 1u : 0u);
  return s;
}

}  // namespace ccs

#endif
