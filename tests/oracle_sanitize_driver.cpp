// Driver for the sanitizer test of the CPU oracle (tests/test_host_sanitizers.py):
// built with -fsanitize=address,undefined from oracle/ccs_oracle.c and the
// library's host mesh code, it runs the oracle's metrics, initial cases,
// halo fill, spatial operator (all five operator sources), RK3 and RK4 steps
// and diagnostics at small sizes and both halo widths.
#include <cmath>
#include <cstdio>
#include <vector>

#include "ccs.h"
#include "ccs_oracle.h"

#define FAIL() (++fails, std::printf("check failed: line %d, n=%d, halo=%d\n", __LINE__, n, h))

int main() {
  int fails = 0;
  oracle_set_threads(2);
  for (int n : {4, 8, 16}) {
    for (int h : {3, 4}) {
      if (h > n) continue;
      const size_t ni = (size_t)n * n * 6, P = n + 2 * h, np = P * P * 6;
      std::vector<double> lond(ni), latd(ni), xyz(3 * ni), jab(6 * ni);
      if (ccs_mesh_geometry(n, lond.data(), latd.data(), xyz.data(), jab.data()) != 0) { FAIL(); continue; }
      std::vector<double> lon(np), lat(np), sinlat(np), coslat(np), sqrtG(np), f(np), area(np);
      std::vector<double> G(4 * np), IG(4 * np), A(4 * np), IA(4 * np);
      double dx = 0;
      if (oracle_metrics(n, h, 6.371e6, 7.292e-5, lond.data(), latd.data(), jab.data(), &dx, lon.data(), lat.data(),
                         sinlat.data(), coslat.data(), sqrtG.data(), f.data(), area.data(), G.data(), IG.data(),
                         A.data(), IA.data()) != 0) { FAIL(); continue; }
      const size_t ne = (size_t)(2 * n + 1) * (2 * n + 1) * 6;
      std::vector<double> ig_ext(4 * ne);
      if (ccs_mesh_ext_metric(n, 6.371e6, nullptr, ig_ext.data()) != 0) FAIL();
      for (int cn : {2, 5, 6, 7}) {
        std::vector<double> s[9];
        for (auto& v : s) v.assign(np, 0.0);
        std::vector<double> phis(np), phit(np);
        if (oracle_init_case(n, h, 6.371e6, 7.292e-5, 9.80616, cn, 3, 8, 1, 20.0, 10000.0, 100.0, 5e5, 1.5e6, 2000.0,
                             1.5e6, lon.data(), lat.data(), sinlat.data(), coslat.data(), sqrtG.data(), G.data(),
                             IA.data(), s[0].data(), s[1].data(), s[2].data(), s[3].data(), s[4].data(), s[5].data(),
                             s[6].data(), s[7].data(), phis.data()) != 0) { FAIL(); continue; }
        for (int op = 0; op <= 4; ++op) {
          oracle_mesh m{};
          m.n = n; m.h = h; m.dx = dx; m.dt = 30.0;
          m.sqrtG = sqrtG.data(); m.G = G.data(); m.IG = IG.data(); m.A = A.data(); m.IA = IA.data();
          m.f = f.data(); m.phis = phis.data(); m.op = op; m.IG_ext = ig_ext.data();
          std::vector<double> w[9];
          for (int k = 0; k < 8; ++k) w[k] = s[k];
          w[8] = phit;
          oracle_stat st{w[0].data(), w[1].data(), w[2].data(), w[3].data(), w[4].data(), w[5].data(),
                         w[6].data(), w[7].data(), w[8].data()};
          std::vector<double> t0(np), t1(np), t2(np);
          oracle_spatial_operator(&m, &st, t0.data(), t1.data(), t2.data());
          for (int scheme : {0, 1})
            if (oracle_steps(&m, &st, 2, scheme) != 0) FAIL();
          double a, b, c, d;
          oracle_diagnostics(&m, &st, &a, &b, &c, &d);
          if (!std::isfinite(c) || !std::isfinite(d)) FAIL();
        }
      }
      std::vector<double> fld(np, 1.0);
      oracle_fill_halo(n, h, fld.data());
    }
  }
  std::printf("oracle sanitizer driver: %d failures\n", fails);
  return fails != 0;
}
