// Host-code driver for the sanitizer test (tests/test_host_sanitizers.py):
// built with g++ -fsanitize=address,undefined from the library's host sources
// (no HIP), it runs the mesh map, the metric tables, the four initial cases,
// the extended-lattice metric, the halo scatter rules and the netCDF mesh and
// history files at several sizes and halo widths, and feeds truncated mesh
// files to the reader, which must reject them without reading past the end.
// argv[1]: a directory for the files it writes.
#include <cstdio>
#include <vector>
#include <string>
#include "ccs.h"
#include "ccs_internal.h"
using namespace ccs;
#define FAIL() (++fails, std::printf("check failed: line %d, n=%d, halo=%d\n", __LINE__, n, h))

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  int fails = 0;
  for (int n : {4, 6, 16, 34, 48}) {
    for (int h : {3, 4}) {
      if (h > n) continue;
      const size_t ni = (size_t)n * n * 6, P = n + 2 * h, np = P * P * 6;
      std::vector<double> lon(ni), lat(ni), xyz(3 * ni), jab(6 * ni);
      if (ccs_mesh_geometry(n, lon.data(), lat.data(), xyz.data(), jab.data()) != 0) { FAIL(); continue; }
      std::vector<double> a[9]; for (auto& v : a) v.assign(np, 0.0);
      std::vector<double> G(4 * np), IG(4 * np), A(4 * np), IA(4 * np);
      ccs_metrics m{};
      m.n = n; m.halo = h; m.radius = 6.371e6; m.omega = 7.292e-5;
      m.lon = a[0].data(); m.lat = a[1].data(); m.sinlon = a[2].data(); m.coslon = a[3].data();
      m.sinlat = a[4].data(); m.coslat = a[5].data(); m.sqrtG = a[6].data(); m.f = a[7].data(); m.areaCell = a[8].data();
      m.G = G.data(); m.IG = IG.data(); m.A = A.data(); m.IA = IA.data();
      if (ccs_mesh_metrics(lon.data(), lat.data(), jab.data(), &m) != 0) { FAIL(); continue; }
      for (int cn : {2, 5, 6, 7}) {
        ccs_case_params c; ccs_case_defaults(&c); c.case_num = cn; c.mountain = 1;
        std::vector<double> s[9]; for (auto& v : s) v.assign(np, 0.0);
        if (ccs_init_case(&m, &c, s[0].data(), s[1].data(), s[2].data(), s[3].data(), s[4].data(), s[5].data(),
                          s[6].data(), s[7].data(), s[8].data()) != 0) FAIL();
        if (cn == 7) {
          std::string hp = dir + "/hist_" + std::to_string(n) + ".nc";
          ccs_history_desc d{}; d.n = n; d.halo = h; d.case_num = cn; d.dx_deg = m.dx * 57.29577951308232; d.dt = 60.0;
          ccs_history* hh = nullptr;
          if (ccs_history_create(hp.c_str(), &d, m.lon, m.lat, m.areaCell, s[8].data(), &hh) != 0) { FAIL(); }
          else {
            for (int t = 1; t <= 3; ++t)
              if (ccs_history_write(hh, t, s[3].data(), s[8].data(), s[1].data(), s[2].data(), s[4].data(), s[5].data(),
                                    s[6].data(), s[7].data()) != 0) FAIL();
            if (ccs_history_close(hh) != 0) FAIL();
          }
        }
      }
      const size_t ne = (size_t)(2 * n + 1) * (2 * n + 1) * 6;
      std::vector<double> ig(4 * ne);
      if (ccs_mesh_ext_metric(n, m.radius, nullptr, ig.data()) != 0) FAIL();
      ScatterRule rules[6][4];
      if (!build_scatter_rules(n, h, rules)) FAIL();
      std::string mp = dir + "/mesh_" + std::to_string(n) + ".nc";
      if (ccs_mesh_file_write(mp.c_str(), n) != 0) { FAIL(); continue; }
      int nr = 0;
      std::vector<double> lon2(ni), lat2(ni), jab2(6 * ni), jabx(6 * ne);
      if (ccs_mesh_file_read(mp.c_str(), &nr, lon2.data(), lat2.data(), jab2.data()) != 0 || nr != n) FAIL();
      if (ccs_mesh_file_read_ext(mp.c_str(), &nr, jabx.data()) != 0 || nr != n) FAIL();
      if (ccs_mesh_ext_metric(n, m.radius, jabx.data(), ig.data()) != 0) FAIL();
      // truncated copies of the file must be rejected, never read past the end
      std::vector<char> buf;
      if (FILE* f = std::fopen(mp.c_str(), "rb")) {
        for (int ch; (ch = std::fgetc(f)) != EOF;) buf.push_back(static_cast<char>(ch));
        std::fclose(f);
      }
      if (buf.size() < 64) { FAIL(); continue; }
      for (size_t cut : {(size_t)8, (size_t)40, buf.size() / 3, buf.size() - 4}) {
        const std::string tp = mp + ".cut";
        FILE* g = std::fopen(tp.c_str(), "wb");
        if (!g) { FAIL(); continue; }
        std::fwrite(buf.data(), 1, cut, g);
        std::fclose(g);
        if (ccs_mesh_file_read(tp.c_str(), &nr, lon2.data(), lat2.data(), jab2.data()) == 0) FAIL();
      }
    }
  }
  std::printf("host sanitizer driver: %d failures\n", fails);
  return fails != 0;
}
