// Coverage check of the persistent ensemble grid's tile partition
// (conformal_cubed_sphere_amd/csrc/ccs_share.h), compiled with g++ by
// tests/test_ensemble_partition.py.  For every grid size nb and tile count:
// the groups' ranges must cover the tile list exactly once, every group must
// have a workgroup, and a simulated run in which the workgroups claim tiles
// from their group's counter in a scrambled order must process every tile
// exactly once.
#include <cstdio>
#include <vector>

#include "ccs_share.h"

int main() {
  int bad = 0;
  for (unsigned nb = 1; nb <= 80 && bad < 10; ++nb)
    for (unsigned ntiles = 1; ntiles <= 300 && bad < 10; ++ntiles) {
      std::vector<int> owner(ntiles, 0), members(ccs::kClaimGroups, 0);
      for (unsigned b = 0; b < nb; ++b) {
        const ccs::EnsGroup s = ccs::ens_group(b, nb, ntiles);
        if (s.g >= ccs::kClaimGroups || s.lo > s.hi || s.hi > ntiles) {
          std::printf("nb %u ntiles %u: workgroup %u has group %u range [%u, %u)\n", nb, ntiles, b, s.g, s.lo, s.hi);
          ++bad;
          break;
        }
        if (members[s.g]++ == 0)
          for (unsigned t = s.lo; t < s.hi; ++t) ++owner[t];
      }
      for (unsigned t = 0; t < ntiles; ++t)
        if (owner[t] != 1) {
          std::printf("nb %u ntiles %u: tile %u in %d group ranges\n", nb, ntiles, t, owner[t]);
          ++bad;
          break;
        }
      // simulated claims: workgroups take turns in a scrambled order; a
      // workgroup stops once its claim reaches its group's end
      std::vector<unsigned> counter(ccs::kClaimGroups, 0);
      std::vector<int> seen(ntiles, 0);
      std::vector<bool> done(nb, false);
      unsigned live = nb, turn = 0;
      while (live > 0) {
        // consecutive turns visit every residue modulo nb (7919 is prime)
        const unsigned b = (turn * 7919u + ntiles) % nb;
        ++turn;
        if (done[b]) continue;
        const ccs::EnsGroup s = ccs::ens_group(b, nb, ntiles);
        const unsigned t = s.lo + counter[s.g]++;
        if (t >= s.hi) {
          done[b] = true;
          --live;
          continue;
        }
        ++seen[t];
      }
      for (unsigned t = 0; t < ntiles; ++t)
        if (seen[t] != 1) {
          std::printf("nb %u ntiles %u: tile %u processed %d times\n", nb, ntiles, t, seen[t]);
          ++bad;
          break;
        }
    }
  std::printf(bad ? "FAIL\n" : "OK\n");
  return bad ? 1 : 0;
}
