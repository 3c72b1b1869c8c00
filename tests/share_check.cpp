// Coverage check of the persistent ensemble grid's tile partition
// (conformal_cubed_sphere_amd/csrc/ccs_share.h), compiled with g++ by
// tests/test_ensemble_partition.py: for every grid size nb and tile count,
// each tile must be visited by exactly one workgroup.
#include <cstdio>
#include <vector>

#include "ccs_share.h"

int main() {
  int bad = 0;
  for (unsigned nb = 1; nb <= 80 && bad < 10; ++nb)
    for (unsigned ntiles = 1; ntiles <= 300 && bad < 10; ++ntiles) {
      std::vector<int> seen(ntiles, 0);
      for (unsigned b = 0; b < nb; ++b) {
        const ccs::EnsShare s = ccs::ens_share(b, nb, ntiles);
        for (unsigned t = s.first; t < s.end; t += s.stride) {
          if (t >= ntiles) {
            std::printf("nb %u ntiles %u: workgroup %u visits tile %u out of range\n", nb, ntiles, b, t);
            ++bad;
            break;
          }
          ++seen[t];
        }
      }
      for (unsigned t = 0; t < ntiles; ++t)
        if (seen[t] != 1) {
          std::printf("nb %u ntiles %u: tile %u visited %d times\n", nb, ntiles, t, seen[t]);
          ++bad;
          break;
        }
    }
  std::printf(bad ? "FAIL\n" : "OK\n");
  return bad ? 1 : 0;
}
