// ccs_netcdf.cpp — the reference's on-disk formats, written and read without
// a netCDF library.
//
// Files use the netCDF classic data model in the 64-bit-offset encoding
// (CDF-2): header (magic "CDF\x02", record count, dimension, attribute and
// variable lists) followed by the fixed-size variables and then the records
// of the variables that run along the unlimited dimension, all big-endian.
// Any netCDF reader opens them.
//
//  - Mesh file: the schema the mesh generator writes and the dycore reads
//    (mesh_generation/src/output_mod.f90:25-83, io_mod.F90:37-117): the
//    conformal map on the (2N+1)^2 "extended" points of each panel (cell
//    centres at even indices, edges and corners at odd ones), xi/eta in map
//    units, x/y/z on the unit sphere, lon/lat in degrees, jab per radian,
//    global attributes dx, dy (degrees) and ids..ife.  The reference writes
//    this file as netCDF-4/HDF5; the dycore's nf90_open reads CDF-2 as well.
//  - History file: ccsm_output.nc of io_mod.F90:309-541 (CDF-2, as there):
//    lon/lat/areaCell/phis on the padded grid and, per output, time and
//    phi, phit, u, v, uc, vc, zonal_wind, meridional_wind.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ccs.h"
#include "ccs_internal.h"

namespace ccs {
void map_points(const std::vector<double>& coord, std::vector<std::complex<double>>& w,
                std::vector<std::complex<double>>& dwdz);
void panel_point(int p, std::complex<double> w, std::complex<double> dwdz, double g[3], double jab[6]);
}  // namespace ccs

namespace {

constexpr double kFill = -9999999999999999.;  // constants_mod.F90:14
constexpr int NC_CHAR = 2, NC_INT = 4, NC_DOUBLE = 6;

// ---------------------------------------------------------------- encoding
struct Buf {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back(static_cast<unsigned char>(v >> s));
  }
  void u64(uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back(static_cast<unsigned char>(v >> s));
  }
  void pad() {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string& s) {
    u32(static_cast<uint32_t>(s.size()));
    b.insert(b.end(), s.begin(), s.end());
    pad();
  }
};

void put_be64(unsigned char* p, double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  for (int k = 0; k < 8; ++k) p[k] = static_cast<unsigned char>(u >> (56 - 8 * k));
}

struct Att {
  std::string name;
  int type;
  std::string text;
  std::vector<double> d;
  std::vector<int32_t> i;
};

Att att_text(const std::string& n, const std::string& t) { return Att{n, NC_CHAR, t, {}, {}}; }
Att att_double(const std::string& n, double v) { return Att{n, NC_DOUBLE, "", {v}, {}}; }
Att att_int(const std::string& n, int v) { return Att{n, NC_INT, "", {}, {v}}; }

void put_atts(Buf& h, const std::vector<Att>& atts) {
  if (atts.empty()) {
    h.u32(0);
    h.u32(0);
    return;
  }
  h.u32(0x0C);
  h.u32(static_cast<uint32_t>(atts.size()));
  for (const Att& a : atts) {
    h.name(a.name);
    h.u32(static_cast<uint32_t>(a.type));
    if (a.type == NC_CHAR) {
      h.u32(static_cast<uint32_t>(a.text.size()));
      h.b.insert(h.b.end(), a.text.begin(), a.text.end());
    } else if (a.type == NC_DOUBLE) {
      h.u32(static_cast<uint32_t>(a.d.size()));
      for (double v : a.d) {
        unsigned char t[8];
        put_be64(t, v);
        h.b.insert(h.b.end(), t, t + 8);
      }
    } else {
      h.u32(static_cast<uint32_t>(a.i.size()));
      for (int32_t v : a.i) h.u32(static_cast<uint32_t>(v));
    }
    h.pad();
  }
}

struct Var {
  std::string name;
  std::vector<int> dims;  // dimension ids, slowest first
  int type;
  std::vector<Att> atts;
  uint64_t vsize = 0, begin = 0;
  bool record = false;
};

// Header and layout of a CDF-2 file.  Dimension 0 may be the unlimited one.
struct Layout {
  std::vector<std::pair<std::string, uint32_t>> dims;  // length 0 = unlimited
  std::vector<Att> gatts;
  std::vector<Var> vars;
  uint64_t recsize = 0, data_begin = 0;
  uint32_t numrecs = 0;

  std::vector<unsigned char> header() {
    // two passes: the header size fixes the data offsets
    for (int pass = 0; pass < 2; ++pass) {
      Buf h;
      h.b = {'C', 'D', 'F', 2};
      h.u32(numrecs);
      h.u32(0x0A);
      h.u32(static_cast<uint32_t>(dims.size()));
      for (auto& d : dims) {
        h.name(d.first);
        h.u32(d.second);
      }
      put_atts(h, gatts);
      h.u32(0x0B);
      h.u32(static_cast<uint32_t>(vars.size()));
      for (Var& v : vars) {
        h.name(v.name);
        h.u32(static_cast<uint32_t>(v.dims.size()));
        for (int id : v.dims) h.u32(static_cast<uint32_t>(id));
        put_atts(h, v.atts);
        h.u32(static_cast<uint32_t>(v.type));
        h.u32(static_cast<uint32_t>(std::min<uint64_t>(v.vsize, 0xFFFFFFFFu)));
        h.u64(v.begin);
      }
      if (pass == 1) return h.b;
      // assign offsets
      uint64_t off = h.b.size();
      data_begin = off;
      recsize = 0;
      for (Var& v : vars) {
        uint64_t elems = 1;
        v.record = !v.dims.empty() && dims[v.dims[0]].second == 0;
        for (std::size_t k = v.record ? 1 : 0; k < v.dims.size(); ++k) elems *= dims[v.dims[k]].second;
        const uint64_t esz = (v.type == NC_DOUBLE) ? 8 : 4;
        v.vsize = (elems * esz + 3) / 4 * 4;
        if (!v.record) {
          v.begin = off;
          off += v.vsize;
        }
      }
      for (Var& v : vars)
        if (v.record) {
          v.begin = off + recsize;
          recsize += v.vsize;
        }
    }
    return {};
  }
};

bool write_at(FILE* f, uint64_t off, const void* p, std::size_t n) {
  if (fseeko(f, static_cast<off_t>(off), SEEK_SET) != 0) return false;
  return fwrite(p, 1, n, f) == n;
}

bool write_doubles(FILE* f, uint64_t off, const std::vector<double>& v) {
  std::vector<unsigned char> b(v.size() * 8);
  for (std::size_t k = 0; k < v.size(); ++k) put_be64(b.data() + 8 * k, v[k]);
  return write_at(f, off, b.data(), b.size());
}

// ---------------------------------------------------------------- decoding
struct Reader {
  std::vector<unsigned char> d;
  std::size_t p = 0;
  bool ok = true;
  int version = 0;
  uint32_t u32() {
    if (p + 4 > d.size()) { ok = false; return 0; }
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) v = (v << 8) | d[p++];
    return v;
  }
  uint64_t u64() {
    if (p + 8 > d.size()) { ok = false; return 0; }
    uint64_t v = 0;
    for (int k = 0; k < 8; ++k) v = (v << 8) | d[p++];
    return v;
  }
  std::string name() {
    const uint32_t n = u32();
    if (!ok || p + n > d.size()) { ok = false; return ""; }
    std::string s(reinterpret_cast<const char*>(&d[p]), n);
    p += (n + 3) / 4 * 4;
    return s;
  }
  double be64(std::size_t off) const {
    uint64_t u = 0;
    for (int k = 0; k < 8; ++k) u = (u << 8) | d[off + k];
    double v;
    std::memcpy(&v, &u, 8);
    return v;
  }
};

struct RVar {
  std::string name;
  std::vector<uint32_t> shape;
  int type = 0;
  uint64_t begin = 0;
};

struct RAtt {
  int type = 0;
  std::vector<double> vals;
};

bool read_atts(Reader& r, std::vector<std::pair<std::string, RAtt>>* out) {
  const uint32_t tag = r.u32(), n = r.u32();
  if (tag == 0 && n == 0) return r.ok;
  if (tag != 0x0C) return false;
  for (uint32_t k = 0; k < n && r.ok; ++k) {
    std::string nm = r.name();
    RAtt a;
    a.type = static_cast<int>(r.u32());
    const uint32_t cnt = r.u32();
    const int esz = (a.type == 1 || a.type == 2) ? 1 : (a.type == 3) ? 2 : (a.type == 6) ? 8 : 4;
    const std::size_t nb = static_cast<std::size_t>(cnt) * esz;
    if (r.p + nb > r.d.size()) return false;
    for (uint32_t e = 0; e < cnt && (a.type == 4 || a.type == 6); ++e) {
      if (a.type == 6) a.vals.push_back(r.be64(r.p + 8 * e));
      else {
        uint32_t v = 0;
        for (int b = 0; b < 4; ++b) v = (v << 8) | r.d[r.p + 4 * e + b];
        a.vals.push_back(static_cast<double>(static_cast<int32_t>(v)));
      }
    }
    r.p += (nb + 3) / 4 * 4;
    if (out) out->emplace_back(nm, a);
  }
  return r.ok;
}

}  // namespace

using namespace ccs;

// ------------------------------------------------------------ mesh file
extern "C" int ccs_mesh_file_write(const char* path, int n) {
  if (!path || n < 4 || (n % 2) != 0) return CCS_ERR_INVALID;
  const int m = 2 * n + 1;  // extended points per panel edge
  std::vector<double> coord(m);
  for (int k = 0; k < m; ++k) coord[k] = -1.0 + static_cast<double>(k) / n;
  std::vector<std::complex<double>> w, dwdz;
  map_points(coord, w, dwdz);
  const std::size_t np = static_cast<std::size_t>(m) * m;
  std::vector<double> xi(6 * np), eta(6 * np), x(6 * np), y(6 * np), z(6 * np), lon(6 * np), lat(6 * np),
      jab(36 * np);
  const double r2d = 45.0 / std::atan(1.0);
  for (int p = 0; p < 6; ++p)
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        const std::size_t c = static_cast<std::size_t>(j) * m + i, q = c + np * p;
        double g[3], jb[6];
        panel_point(p, w[c], dwdz[c], g, jb);
        xi[q] = coord[i];
        eta[q] = coord[j];
        x[q] = g[0];
        y[q] = g[1];
        z[q] = g[2];
        lon[q] = std::atan2(g[1], g[0]) * r2d;
        lat[q] = std::atan2(g[2], std::sqrt(g[0] * g[0] + g[1] * g[1])) * r2d;
        for (int e = 0; e < 6; ++e) jab[6 * q + e] = jb[e];
      }
  Layout L;
  L.dims = {{"three", 3}, {"two", 2}, {"lon", static_cast<uint32_t>(m)}, {"lat", static_cast<uint32_t>(m)},
            {"nPatch", 6}};
  const double dx = 90.0 / n;
  L.gatts = {att_double("dx", dx), att_double("dy", dx), att_int("ids", 1), att_int("ide", m), att_int("jds", 1),
             att_int("jde", m), att_int("ifs", 1), att_int("ife", 6)};
  const std::vector<int> d3 = {4, 3, 2};
  auto axis = att_text("_CoordinateAxisTypes", "lon lat nPatch");
  L.vars = {{"xi", d3, NC_DOUBLE, {axis}},
            {"eta", d3, NC_DOUBLE, {axis}},
            {"x", d3, NC_DOUBLE, {axis}},
            {"y", d3, NC_DOUBLE, {axis}},
            {"z", d3, NC_DOUBLE, {axis}},
            {"lon", d3, NC_DOUBLE,
             {att_text("units", "degree_east"), att_text("long_name", "longitude on sphere coordinate for Cells"), axis}},
            {"lat", d3, NC_DOUBLE,
             {att_text("units", "degree_north"), att_text("long_name", "latitude on sphere coordinate for Cells"), axis}},
            {"jab", {4, 3, 2, 1, 0}, NC_DOUBLE, {att_text("_CoordinateAxisTypes", "three two lon lat nPatch")}}};
  const std::vector<unsigned char> hdr = L.header();
  FILE* f = std::fopen(path, "wb");
  if (!f) return CCS_ERR_INVALID;
  bool ok = write_at(f, 0, hdr.data(), hdr.size());
  const std::vector<double>* data[8] = {&xi, &eta, &x, &y, &z, &lon, &lat, &jab};
  for (int k = 0; k < 8 && ok; ++k) ok = write_doubles(f, L.vars[k].begin, *data[k]);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? CCS_OK : CCS_ERR_INVALID;
}

namespace {
int read_mesh(const char* path, int* n_out, double* lon_deg, double* lat_deg, double* jab, double* jab_ext) {
  if (!path || !n_out) return CCS_ERR_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return CCS_ERR_INVALID;
  Reader r;
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  r.d.resize(sz > 0 ? static_cast<std::size_t>(sz) : 0);
  const bool got = !r.d.empty() && std::fread(r.d.data(), 1, r.d.size(), f) == r.d.size();
  std::fclose(f);
  // classic (CDF-1) and 64-bit-offset (CDF-2) only; netCDF-4 is HDF5
  if (!got || r.d.size() < 8 || r.d[0] != 'C' || r.d[1] != 'D' || r.d[2] != 'F' || (r.d[3] != 1 && r.d[3] != 2))
    return CCS_ERR_INVALID;
  r.version = r.d[3];
  r.p = 4;
  r.u32();  // numrecs
  std::vector<uint32_t> dimlen;
  {
    const uint32_t tag = r.u32(), nd = r.u32();
    if (tag == 0x0A)
      for (uint32_t k = 0; k < nd && r.ok; ++k) {
        r.name();
        dimlen.push_back(r.u32());
      }
  }
  std::vector<std::pair<std::string, RAtt>> gatts;
  if (!read_atts(r, &gatts)) return CCS_ERR_INVALID;
  std::vector<RVar> vars;
  {
    const uint32_t tag = r.u32(), nv = r.u32();
    if (tag != 0x0B) return CCS_ERR_INVALID;
    for (uint32_t k = 0; k < nv && r.ok; ++k) {
      RVar v;
      v.name = r.name();
      const uint32_t nd = r.u32();
      for (uint32_t e = 0; e < nd; ++e) {
        const uint32_t id = r.u32();
        v.shape.push_back(id < dimlen.size() ? dimlen[id] : 0);
      }
      if (!read_atts(r, nullptr)) return CCS_ERR_INVALID;
      v.type = static_cast<int>(r.u32());
      r.u32();  // vsize
      v.begin = (r.version == 2) ? r.u64() : r.u32();
      vars.push_back(v);
    }
  }
  if (!r.ok) return CCS_ERR_INVALID;
  auto gatt = [&](const char* nm, double* v) {
    for (auto& a : gatts)
      if (a.first == nm && !a.second.vals.empty()) {
        *v = a.second.vals[0];
        return true;
      }
    return false;
  };
  double ids = 0, ide = 0, jds = 0, jde = 0;
  if (!gatt("ids", &ids) || !gatt("ide", &ide) || !gatt("jds", &jds) || !gatt("jde", &jde)) return CCS_ERR_INVALID;
  // lattice extents from the file: keep the int conversions in range
  constexpr double kMaxExt = 1 << 20;
  if (!(ide - ids >= 8 && ide - ids <= kMaxExt && jde - jds >= 8 && jde - jds <= kMaxExt)) return CCS_ERR_INVALID;
  const int n = static_cast<int>(ide - ids) / 2;  // io_mod.F90:49-52
  if (n < 4 || static_cast<int>(jde - jds) / 2 != n) return CCS_ERR_INVALID;
  *n_out = n;
  const int m = static_cast<int>(ide - ids) + 1;
  const RVar* vl = nullptr;
  const RVar* vt = nullptr;
  const RVar* vj = nullptr;
  for (const RVar& v : vars) {
    if (v.name == "lon") vl = &v;
    if (v.name == "lat") vt = &v;
    if (v.name == "jab") vj = &v;
  }
  if (!vl || !vt || !vj || vl->type != NC_DOUBLE || vt->type != NC_DOUBLE || vj->type != NC_DOUBLE)
    return CCS_ERR_INVALID;
  const std::size_t np = static_cast<std::size_t>(m) * m;
  // data ranges inside the file (offsets come from the file: no wrap-around)
  auto fits = [&r](uint64_t begin, std::size_t bytes) {
    return begin <= r.d.size() && bytes <= r.d.size() - begin;
  };
  if (!fits(vl->begin, 6 * np * 8) || !fits(vt->begin, 6 * np * 8) || !fits(vj->begin, 36 * np * 8))
    return CCS_ERR_INVALID;
  // cell centres at extended indices 2i, 2j (1-based; io_mod.F90:122-135)
  for (int p = 0; p < 6; ++p)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        const std::size_t e = static_cast<std::size_t>(2 * i - 1) + static_cast<std::size_t>(m) * (2 * j - 1) + np * p;
        const std::size_t c = static_cast<std::size_t>(i - 1) + static_cast<std::size_t>(n) * ((j - 1) + n * p);
        if (lon_deg) lon_deg[c] = r.be64(vl->begin + 8 * e);
        if (lat_deg) lat_deg[c] = r.be64(vt->begin + 8 * e);
        if (jab)
          for (int k = 0; k < 6; ++k) jab[6 * c + k] = r.be64(vj->begin + 8 * (6 * e + k));
      }
  if (jab_ext)  // the whole extended lattice, (3,2,lon,lat,nPatch) as stored
    for (std::size_t k = 0; k < 36 * np; ++k) jab_ext[k] = r.be64(vj->begin + 8 * k);
  return CCS_OK;
}
}  // namespace

extern "C" int ccs_mesh_file_read(const char* path, int* n_out, double* lon_deg, double* lat_deg, double* jab) {
  return read_mesh(path, n_out, lon_deg, lat_deg, jab, nullptr);
}

extern "C" int ccs_mesh_file_read_ext(const char* path, int* n_out, double* jab_ext) {
  return read_mesh(path, n_out, nullptr, nullptr, nullptr, jab_ext);
}

// ------------------------------------------------------------ history file
struct ccs_history {
  FILE* f = nullptr;
  Layout L;
  int n = 0, h = 0;
  std::size_t np = 0;
};

extern "C" int ccs_history_create(const char* path, const ccs_history_desc* d, const double* lon,
                                  const double* lat, const double* areaCell, const double* phis,
                                  ccs_history** out) {
  if (!path || !d || !lon || !lat || !areaCell || !phis || !out) return CCS_ERR_INVALID;
  if (d->n < 4 || d->halo < 1) return CCS_ERR_INVALID;
  *out = nullptr;
  ccs_history* H = new ccs_history();
  H->n = d->n;
  H->h = d->halo;
  const int n = d->n, h = d->halo, P = n + 2 * h;
  H->np = padded_cells(n, h);
  Layout& L = H->L;
  L.dims = {{"time", 0}, {"lon", static_cast<uint32_t>(P)}, {"lat", static_cast<uint32_t>(P)}, {"nPatch", 6}};
  // io_mod.F90:349-369 (DOF = 3, constants_mod.F90:4; index ranges io_mod.F90:49-83)
  const int DOF = 3;
  L.gatts = {att_int("MCV_ORDER", DOF), att_double("dx", d->dx_deg), att_double("dy", d->dx_deg),
             att_double("dt", d->dt), att_int("xhalo", h), att_int("yhalo", h), att_int("case_num", d->case_num),
             att_int("ics", 1 - h), att_int("ice", n + h), att_int("jcs", 1 - h), att_int("jce", n + h),
             att_int("ips", 1 - h * (DOF - 1)), att_int("ipe", n * (DOF - 1) + 1 + h * (DOF - 1)),
             att_int("jps", 1 - h * (DOF - 1)), att_int("jpe", n * (DOF - 1) + 1 + h * (DOF - 1)),
             att_int("its", 1), att_int("ite", n), att_int("jts", 1), att_int("jte", n), att_int("ifs", 1),
             att_int("ife", 6)};
  const std::vector<int> d3 = {3, 2, 1}, d4 = {0, 3, 2, 1};
  auto axis3 = att_text("_CoordinateAxisTypes", "lon lat nPatch");
  auto axis4 = att_text("_CoordinateAxisTypes", "lon lat nPatch time");
  auto fill = att_double("_FillValue", kFill);
  auto rec = [&](const char* nm, const char* units, const char* ln) {
    return Var{nm, d4, NC_DOUBLE, {att_text("units", units), att_text("long_name", ln), axis4, fill}};
  };
  L.vars = {{"lon", d3, NC_DOUBLE,
             {att_text("units", "degree_east"), att_text("long_name", "longitude on sphere coordinate for Cells"),
              axis3}},
            {"lat", d3, NC_DOUBLE,
             {att_text("units", "degree_north"), att_text("long_name", "latitude on sphere coordinate for Cells"),
              axis3}},
            {"areaCell", d3, NC_DOUBLE,
             {att_text("units", "m^2"), att_text("long_name", "area of cells"),
              att_text("_CoordinateAxisTypes", "lon lat nPatch time")}},
            {"phis", d3, NC_DOUBLE,
             {att_text("units", "m^2/s^2"), att_text("long_name", "surface height on points"), axis3, fill}},
            {"time", {0}, NC_INT, {att_text("units", "seconds"), att_text("long_name", "time")}},
            rec("phi", "m^2/s^2", "geopotential height on points"),
            rec("phit", "m^2/s^2", "total geopotential height on points"),
            rec("u", "m/s", "covariant wind vector on x direcvtion"),
            rec("v", "m/s", "covariant wind vector on y direcvtion"),
            rec("uc", "m/s", "contravariant wind vector on x direcvtion"),
            rec("vc", "m/s", "contravariant wind vector on y direcvtion"),
            rec("zonal_wind", "m/s", "zonal wind"),
            rec("meridional_wind", "m/s", "meridional wind")};
  const std::vector<unsigned char> hdr = L.header();
  H->f = std::fopen(path, "w+b");
  if (!H->f) {
    delete H;
    return CCS_ERR_INVALID;
  }
  // lon/lat in degrees; corners (never filled, io_mod.F90:118-119) get FillValue
  const double r2d = 45.0 / std::atan(1.0);
  std::vector<double> vlon(H->np), vlat(H->np), va(areaCell, areaCell + H->np), vp(phis, phis + H->np);
  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - h; j <= n + h; ++j)
      for (int i = 1 - h; i <= n + h; ++i) {
        const std::size_t c = pidx(n, h, i, j, p);
        const bool corner = is_corner(n, i, j);
        vlon[c] = corner ? kFill : lon[c] * r2d;
        vlat[c] = corner ? kFill : lat[c] * r2d;
      }
  bool ok = write_at(H->f, 0, hdr.data(), hdr.size());
  ok = ok && write_doubles(H->f, L.vars[0].begin, vlon) && write_doubles(H->f, L.vars[1].begin, vlat) &&
       write_doubles(H->f, L.vars[2].begin, va) && write_doubles(H->f, L.vars[3].begin, vp);
  ok = ok && std::fflush(H->f) == 0;
  if (!ok) {
    ccs_history_close(H);
    return CCS_ERR_INVALID;
  }
  *out = H;
  return CCS_OK;
}

extern "C" int ccs_history_write(ccs_history* H, int time_slot, const double* phi, const double* phis,
                                 const double* u, const double* v, const double* uc, const double* vc,
                                 const double* zonal, const double* merid) {
  if (!H || !H->f || time_slot < 1 || !phi || !phis || !u || !v || !uc || !vc || !zonal || !merid)
    return CCS_ERR_INVALID;
  Layout& L = H->L;
  const uint64_t rec0 = static_cast<uint64_t>(time_slot - 1) * L.recsize;
  // time (int) = output index, as history_write_stat writes it (io_mod.F90:478)
  unsigned char t[4];
  for (int k = 0; k < 4; ++k) t[k] = static_cast<unsigned char>(static_cast<uint32_t>(time_slot) >> (24 - 8 * k));
  bool ok = write_at(H->f, L.vars[4].begin + rec0, t, 4);
  std::vector<double> phit(H->np);
  for (std::size_t k = 0; k < H->np; ++k) phit[k] = phi[k] + phis[k];
  const double* fields[8] = {phi, phit.data(), u, v, uc, vc, zonal, merid};
  for (int k = 0; k < 8 && ok; ++k)
    ok = write_doubles(H->f, L.vars[5 + k].begin + rec0, std::vector<double>(fields[k], fields[k] + H->np));
  if (ok && static_cast<uint32_t>(time_slot) > L.numrecs) {
    L.numrecs = static_cast<uint32_t>(time_slot);
    unsigned char c[4];
    for (int k = 0; k < 4; ++k) c[k] = static_cast<unsigned char>(L.numrecs >> (24 - 8 * k));
    ok = write_at(H->f, 4, c, 4);
  }
  ok = ok && std::fflush(H->f) == 0;
  return ok ? CCS_OK : CCS_ERR_INVALID;
}

extern "C" int ccs_history_close(ccs_history* H) {
  if (!H) return CCS_ERR_INVALID;
  int rc = CCS_OK;
  if (H->f && std::fclose(H->f) != 0) rc = CCS_ERR_INVALID;
  delete H;
  return rc;
}
