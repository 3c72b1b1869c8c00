// ccs_api.hip — C ABI of the time-step path (include/ccs.h): device state,
// buffer rotation, HIP-graph replay and the host/device copies.
//
// State buffers.  The reference keeps nIntegralSubSteps+2 copies of all nine
// stat_field arrays and copies stat(new) into stat(old) every step
// (stat_mod.F90:22-65, ccsm.F90:66).  Here a run keeps three prognostic
// buffers (phiG,u,v) A, B, C in HBM and rotates them so that every step ends
// in A without a copy:
//   RK3:  A -> B (stage 1),  B,A -> C (stage 2),  C,A -> A (stage 3, in place)
//   RK4:  A -> B, B -> C, C -> B, B -> A (tendencies of stages 1-3 kept)
// phi, uc, vc are derived on the fly inside each stage kernel from phiG, u, v
// exactly as update_stat/correct_bdy_ghost define them; only the very first
// stage after ccs_set_state reads the uploaded companions (initTestCase's).
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ccs.h"
#include "ccs_internal.h"
#include "ccs_kernels.h"
#include "ccs_split.h"

using namespace ccs;

struct ccs_handle {
  ccs_config cfg{};
  int n = 0, h = 0, P = 0;
  std::size_t np = 0, ni = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool mesh_set = false, state_set = false, explicit_valid = false, ext_set = false;
  HaloScatter hs{};
  // FP64 parity path
  double* dmem = nullptr;
  StaticsF64 st{};
  Prog<double> buf[3]{};
  double *phi0 = nullptr, *uc0 = nullptr, *vc0 = nullptr;
  double* tend[9]{};       // RK4 stage tendencies / ccs_tendency output (interior)
  double* derived[5]{};    // get_state scratch: phi, uc, vc, zonal, merid
  double* partials = nullptr;
  double* result = nullptr;
  double* face_speed = nullptr;  // CCS_OP_WENO_EXT: x-face then y-face G^-1 entries
  int f64_tile_y = 8;            // rows per tile of the FP64 stage kernel (stage_f64_tile_y)
  unsigned launches = 0;         // FP64 stage launches enqueued (their parity sets the tile order)
  // split-precision path
  SplitState* split = nullptr;
  // Step graphs.  Every step leaves the state in the same buffer, so a graph
  // of several steps is the one-step sequence repeated.  The split path
  // folds its geopotential deviation into the FP64 reference field after
  // every step whose count since ccs_set_state is a multiple of kRebaseSteps,
  // whatever the grouping of steps into calls: graph[0] is one step, graph[1]
  // one step and the rebase, graph[2 + phase] kGraphSteps steps starting at
  // a step count of phase modulo kRebaseSteps (with the rebase after step
  // kRebaseSteps - phase).  All are captured by the first ccs_step, so no
  // later call pays for a capture.
  static constexpr int kGraphs = 2 + kRebaseSteps;
  hipGraph_t graph[kGraphs]{};
  hipGraphExec_t graph_exec[kGraphs]{};
  bool graphs_ready = false;
  long long steps_done = 0;  // steps since the last ccs_set_state
};

namespace {

const char* kVersion = "ccs-mi355x 0.1.0 (gfx950)";

// Steps per replay of the long graph: one launch per kGraphSteps steps keeps
// the host submission rate far above what the GPU consumes.
constexpr int kGraphSteps = 16;
static_assert(kGraphSteps == kRebaseSteps, "a long graph holds exactly one rebase");

// A roctx range over a scope: the host-side span of an entry point or a
// stage enqueue, seen by rocprofv3 --marker-trace.
struct Range {
  explicit Range(const char* what) { roctxRangePush(what); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
  ~Range() { roctxRangePop(); }
};

// HIP events destroyed on every exit path.
struct EventSet {
  std::vector<hipEvent_t> ev;
  explicit EventSet(std::size_t n) : ev(n, nullptr) {}
  EventSet(const EventSet&) = delete;
  EventSet& operator=(const EventSet&) = delete;
  ~EventSet() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  hipError_t create() {
    for (hipEvent_t& e : ev) {
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  }
  hipEvent_t& operator[](std::size_t k) { return ev[k]; }
};

int fail(ccs_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(hd, call)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail((hd), CCS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

bool valid_precision(int p) { return p == CCS_FP64_PARITY || p == CCS_FP32_SPLIT || p == CCS_FP64_FAST; }

void op_args(const ccs_handle* hd, StageArgsF64& a) {
  a.op = hd->cfg.spatial_op;
  if (hd->face_speed) {
    a.igx = hd->face_speed;
    a.igy = hd->face_speed + static_cast<std::size_t>(hd->n + 1) * hd->n * 6;
  }
}

int enqueue_step_f64(ccs_handle* hd, bool expl, hipEvent_t* ev) {
  StageArgsF64 a{};
  a.n = hd->n;
  a.h = hd->h;
  op_args(hd, a);
  a.dx = hd->cfg.dx;
  a.dt = hd->cfg.dt;
  a.phi0 = hd->phi0;
  a.uc0 = hd->uc0;
  a.vc0 = hd->vc0;
  auto cst = [](const Prog<double>& b) { return Prog<const double>{b.phiG, b.u, b.v}; };
  Prog<double>& A = hd->buf[0];
  Prog<double>& B = hd->buf[1];
  Prog<double>& C = hd->buf[2];
  int stage = 0;
  static const char* const kStageName[4] = {"ccs stage 1", "ccs stage 2", "ccs stage 3", "ccs stage 4"};
  auto launch = [&](int kind, bool e) {
    const Range r(kStageName[stage]);
    a.rev = hd->launches++ & 1u;  // alternate the tile order (StageArgsF64::rev)
    launch_stage_f64(hd->stream, kind, e, a, hd->st, hd->hs, hd->f64_tile_y, ev ? ev + 2 * stage : nullptr);
    ++stage;
  };
  if (hd->cfg.scheme == CCS_RK3_TVD) {
    a.in = cst(A); a.base = cst(A); a.out = B;
    launch(kRK3s1, expl);
    a.in = cst(B); a.base = cst(A); a.out = C;
    launch(kRK3s2, false);
    a.in = cst(C); a.base = cst(A); a.out = A;
    launch(kRK3s3, false);
  } else {
    const double half = 0.5 * hd->cfg.dt;
    a.in = cst(A); a.base = cst(A); a.out = B; a.inc = half;
    a.tend[0] = hd->tend[0]; a.tend[1] = hd->tend[1]; a.tend[2] = hd->tend[2];
    launch(kRK4a, expl);
    a.in = cst(B); a.out = C; a.inc = half;
    a.tend[0] = hd->tend[3]; a.tend[1] = hd->tend[4]; a.tend[2] = hd->tend[5];
    launch(kRK4a, false);
    a.in = cst(C); a.out = B; a.inc = hd->cfg.dt;
    a.tend[0] = hd->tend[6]; a.tend[1] = hd->tend[7]; a.tend[2] = hd->tend[8];
    launch(kRK4a, false);
    a.in = cst(B); a.out = A;
    for (int k = 0; k < 9; ++k) a.tacc[k] = hd->tend[k];
    launch(kRK4f, false);
  }
  return CCS_OK;
}

int enqueue_step(ccs_handle* hd, bool expl, hipEvent_t* ev) {
  if (hd->cfg.precision == CCS_FP64_PARITY) return enqueue_step_f64(hd, expl, ev);
  return split_enqueue_step(hd->split, hd->stream, hd->cfg.scheme, expl, ev);
}

int stages_per_step(const ccs_handle* hd) { return hd->cfg.scheme == CCS_RK3_TVD ? 3 : 4; }

bool rebase_after(const ccs_handle* hd, long long step_index) {  // step_index: 1-based count after the step
  return hd->cfg.precision != CCS_FP64_PARITY && step_index % kRebaseSteps == 0;
}

// one step, then the rebase if the step count reaches a multiple of kGraphSteps
int enqueue_counted_step(ccs_handle* hd, bool expl, hipEvent_t* ev) {
  const int rc = enqueue_step(hd, expl, ev);
  if (rc != CCS_OK) return rc;
  ++hd->steps_done;
  if (rebase_after(hd, hd->steps_done)) split_enqueue_rebase(hd->split, hd->stream);
  return CCS_OK;
}

// Capture graph g of the handle (see ccs_handle::graph).
int capture_graph(ccs_handle* hd, int g) {
  if (hd->graph_exec[g]) return CCS_OK;
  const bool split = hd->cfg.precision != CCS_FP64_PARITY;
  const int nsteps = g < 2 ? 1 : kGraphSteps;
  // the step after which the rebase runs (1-based within the graph), 0: none
  const int rebase_at = !split ? 0 : g == 0 ? 0 : g == 1 ? 1 : kRebaseSteps - (g - 2);
  HIPCHK(hd, hipStreamBeginCapture(hd->stream, hipStreamCaptureModeThreadLocal));
  int rc = CCS_OK;
  for (int s = 1; s <= nsteps && rc == CCS_OK; ++s) {
    rc = enqueue_step(hd, false, nullptr);
    if (rc == CCS_OK && s == rebase_at) split_enqueue_rebase(hd->split, hd->stream);
  }
  hipGraph_t gr = nullptr;
  const hipError_t ec = hipStreamEndCapture(hd->stream, &gr);
  if (rc != CCS_OK) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  if (ec != hipSuccess) return fail(hd, CCS_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  hd->graph[g] = gr;
  HIPCHK(hd, hipGraphInstantiate(&hd->graph_exec[g], hd->graph[g], nullptr, nullptr, 0));
  return CCS_OK;
}

// Every graph a ccs_step may replay: 1-step, 1-step + rebase and the long
// graph of each phase (split path), 1-step and one long graph (FP64 parity).
int capture_graphs(ccs_handle* hd) {
  if (hd->graphs_ready) return CCS_OK;
  const bool split = hd->cfg.precision != CCS_FP64_PARITY;
  for (int g = 0; g < ccs_handle::kGraphs; ++g) {
    if (!split && (g == 1 || g > 2)) continue;
    const int rc = capture_graph(hd, g);
    if (rc != CCS_OK) return rc;
  }
  hd->graphs_ready = true;
  return CCS_OK;
}

}  // namespace

extern "C" const char* ccs_version(void) { return kVersion; }

extern "C" const char* ccs_last_error(const ccs_handle* h) {
  if (!h) return "null handle";
  return h->err.c_str();
}

extern "C" int ccs_create(const ccs_config* cfg, ccs_handle** out) {
  if (!cfg || !out) return CCS_ERR_INVALID;
  *out = nullptr;
  if (cfg->n < 4 || cfg->halo < 3 || cfg->halo > cfg->n) return CCS_ERR_INVALID;
  if (!valid_precision(cfg->precision)) return CCS_ERR_INVALID;
  if (cfg->scheme != CCS_RK3_TVD && cfg->scheme != CCS_RK4) return CCS_ERR_INVALID;
  if (cfg->eflux != CCS_EFLUX_NONE) return CCS_ERR_INVALID;
  if (!(cfg->dt > 0.0) || !(cfg->dx > 0.0) || !(cfg->radius > 0.0)) return CCS_ERR_INVALID;
  if (cfg->spatial_op < CCS_OP_WENO || cfg->spatial_op > CCS_OP_WENO_EXT) return CCS_ERR_INVALID;
  // the split-precision kernels implement the default and the centred
  // operators; the extended-lattice metric of WENO_EXT is parity-path only
  if (cfg->spatial_op == CCS_OP_WENO_EXT && cfg->precision != CCS_FP64_PARITY) return CCS_ERR_INVALID;
  ccs_handle* hd = new ccs_handle();
  hd->cfg = *cfg;
  hd->n = cfg->n;
  hd->h = cfg->halo;
  hd->P = cfg->n + 2 * cfg->halo;
  hd->np = padded_cells(hd->n, hd->h);
  hd->ni = static_cast<std::size_t>(hd->n) * hd->n * 6;
  ScatterRule rules[6][4];
  if (!build_scatter_rules(hd->n, hd->h, rules)) {
    delete hd;
    return CCS_ERR_INVALID;
  }
  for (int q = 0; q < 6; ++q)
    for (int s = 0; s < 4; ++s) {
      const ScatterRule& r = rules[q][s];
      const int v[11] = {r.p, r.ei0, r.eii, r.eij, r.ej0, r.eji, r.ejj, r.c1, r.s1, r.c2, r.s2};
      for (int k = 0; k < 11; ++k) hd->hs.r[q][s][k] = v[k];
    }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) {
    delete hd;
    return CCS_ERR_DEVICE;
  }
  if (hipSetDevice(cfg->device) != hipSuccess || hipStreamCreateWithFlags(&hd->stream, hipStreamNonBlocking) != hipSuccess) {
    delete hd;
    return CCS_ERR_DEVICE;
  }
  int rc = CCS_OK;
  if (cfg->precision == CCS_FP64_PARITY) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, cfg->device) != hipSuccess || cus <= 0) {
      ccs_destroy(hd);
      return CCS_ERR_DEVICE;
    }
    hd->f64_tile_y = stage_f64_tile_y(hd->n, cus);
    // statics 15 + state 9 + explicit 3 + derived 5 (padded), tendencies 9
    // (interior), diagnostics partials and result; the carving below walks
    // the same list and must end exactly at the end of the allocation
    const std::size_t np = hd->np, ni = hd->ni;
    const std::size_t total = (15 + 9 + 3 + 5) * np + 9 * ni + 4 * kDiagBlocks + 2;
    if (hipMalloc(&hd->dmem, total * sizeof(double)) != hipSuccess) {
      ccs_destroy(hd);
      return CCS_ERR_NOMEM;
    }
    // cleared on the handle's stream, and waited for: a memset on the null
    // stream is not ordered before the uploads on a non-blocking stream
    if (hipMemsetAsync(hd->dmem, 0, total * sizeof(double), hd->stream) != hipSuccess ||
        hipStreamSynchronize(hd->stream) != hipSuccess) {
      ccs_destroy(hd);
      return CCS_ERR_DEVICE;
    }
    double* p = hd->dmem;
    auto take = [&p](std::size_t cnt) { double* r = p; p += cnt; return r; };
    hd->st.sqrtG = take(np); hd->st.IG11 = take(np); hd->st.IG12 = take(np); hd->st.IG21 = take(np);
    hd->st.IG22 = take(np); hd->st.phis = take(np); hd->st.f = take(np);
    hd->st.M11 = take(np); hd->st.M12 = take(np); hd->st.M21 = take(np); hd->st.M22 = take(np);
    hd->st.N11 = take(np); hd->st.N12 = take(np); hd->st.N21 = take(np); hd->st.N22 = take(np);
    for (int b = 0; b < 3; ++b) {
      hd->buf[b].phiG = take(np); hd->buf[b].u = take(np); hd->buf[b].v = take(np);
    }
    hd->phi0 = take(np); hd->uc0 = take(np); hd->vc0 = take(np);
    for (int k = 0; k < 5; ++k) hd->derived[k] = take(np);
    for (int k = 0; k < 9; ++k) hd->tend[k] = take(ni);
    hd->partials = take(4 * kDiagBlocks);
    hd->result = take(2);
    if (static_cast<std::size_t>(p - hd->dmem) != total) {  // the carving and the size disagree
      ccs_destroy(hd);
      return CCS_ERR_INVALID;
    }
  } else {
    rc = split_create(*cfg, hd->hs, hd->stream, &hd->split, &hd->err);
    if (rc != CCS_OK) {
      ccs_destroy(hd);
      return rc;
    }
  }
  *out = hd;
  return CCS_OK;
}

extern "C" void ccs_destroy(ccs_handle* hd) {
  if (!hd) return;
  // teardown: errors here have nowhere to go
  for (int g = 0; g < ccs_handle::kGraphs; ++g) {
    if (hd->graph_exec[g]) (void)hipGraphExecDestroy(hd->graph_exec[g]);
    if (hd->graph[g]) (void)hipGraphDestroy(hd->graph[g]);
  }
  if (hd->stream) (void)hipStreamSynchronize(hd->stream);
  if (hd->dmem) (void)hipFree(hd->dmem);
  if (hd->face_speed) (void)hipFree(hd->face_speed);
  if (hd->split) split_destroy(hd->split);
  if (hd->stream) (void)hipStreamDestroy(hd->stream);
  delete hd;
}

extern "C" int ccs_set_mesh(ccs_handle* hd, const double* sqrtG, const double* G, const double* IG,
                            const double* A, const double* IA, const double* f, const double* phis) {
  if (!hd) return CCS_ERR_INVALID;
  if (!sqrtG || !G || !IG || !A || !IA || !f || !phis) return fail(hd, CCS_ERR_INVALID, "ccs_set_mesh: null array");
  HIPCHK(hd, hipSetDevice(hd->cfg.device));
  const int n = hd->n, h = hd->h;
  const std::size_t np = hd->np;
  // Host-side SoA tables; corners stay zero.
  std::vector<double> t(15 * np, 0.0);
  double* T[15];
  for (int k = 0; k < 15; ++k) T[k] = t.data() + k * np;
  for (int p = 1; p <= 6; ++p)
    for (int j = 1 - h; j <= n + h; ++j)
      for (int i = 1 - h; i <= n + h; ++i) {
        if (is_corner(n, i, j)) continue;
        const std::size_t d = pidx(n, h, i, j, p);
        const double* g = G + 4 * d;
        const double* ig = IG + 4 * d;
        const double* a = A + 4 * d;
        const double* ia = IA + 4 * d;
        if (!(sqrtG[d] > 0.0) || !std::isfinite(sqrtG[d]))
          return fail(hd, CCS_ERR_INVALID, "ccs_set_mesh: sqrtG must be positive and finite off the corners");
        T[0][d] = sqrtG[d];
        T[1][d] = ig[0]; T[2][d] = ig[2]; T[3][d] = ig[1]; T[4][d] = ig[3];
        T[5][d] = phis[d];
        T[6][d] = (i >= 1 && i <= n && j >= 1 && j <= n) ? f[d] : 0.0;
        // M = matmul(A, IG) (projection_mod.F90:90), N = matmul(G, IA) (:71)
        T[7][d] = a[0] * ig[0] + a[2] * ig[1];   // M11
        T[8][d] = a[0] * ig[2] + a[2] * ig[3];   // M12
        T[9][d] = a[1] * ig[0] + a[3] * ig[1];   // M21
        T[10][d] = a[1] * ig[2] + a[3] * ig[3];  // M22
        T[11][d] = g[0] * ia[0] + g[2] * ia[1];  // N11
        T[12][d] = g[0] * ia[2] + g[2] * ia[3];  // N12
        T[13][d] = g[1] * ia[0] + g[3] * ia[1];  // N21
        T[14][d] = g[1] * ia[2] + g[3] * ia[3];  // N22
      }
  if (hd->cfg.precision == CCS_FP64_PARITY) {
    double* dst = const_cast<double*>(hd->st.sqrtG);  // statics are one contiguous block
    HIPCHK(hd, hipMemcpyAsync(dst, t.data(), 15 * np * sizeof(double), hipMemcpyHostToDevice, hd->stream));
    HIPCHK(hd, hipStreamSynchronize(hd->stream));
  } else {
    const int rc = split_set_mesh(hd->split, hd->stream, T, &hd->err);
    if (rc != CCS_OK) return rc;
  }
  hd->mesh_set = true;
  return CCS_OK;
}

extern "C" int ccs_set_mesh_ext(ccs_handle* hd, const double* ig_ext) {
  if (!hd) return CCS_ERR_INVALID;
  if (!ig_ext) return fail(hd, CCS_ERR_INVALID, "ccs_set_mesh_ext: null array");
  if (hd->cfg.spatial_op != CCS_OP_WENO_EXT)
    return fail(hd, CCS_ERR_INVALID, "ccs_set_mesh_ext: only CCS_OP_WENO_EXT reads matrixIG_ext");
  HIPCHK(hd, hipSetDevice(hd->cfg.device));
  const int n = hd->n;
  const std::size_t m = 2 * static_cast<std::size_t>(n) + 1, nf = static_cast<std::size_t>(n + 1) * n * 6;
  // matrixIG_ext(r,c,a,b,p), 1-based lattice indices
  auto ext = [&](int r, int c, int ia, int ib, int p) {
    return ig_ext[(r - 1) + 2 * (c - 1) + 4 * ((ia - 1) + m * ((ib - 1) + m * (p - 1)))];
  };
  std::vector<double> t(2 * nf);
  for (int p = 1; p <= 6; ++p)
    for (int j = 1; j <= n; ++j)
      for (int i = 0; i <= n; ++i)  // x-face (i+1/2, j): lattice (2i+1, j)
        t[i + static_cast<std::size_t>(n + 1) * ((j - 1) + static_cast<std::size_t>(n) * (p - 1))] =
            ext(1, 1, 2 * i + 1, j, p);
  for (int p = 1; p <= 6; ++p)
    for (int j = 0; j <= n; ++j)
      for (int i = 1; i <= n; ++i)  // y-face (i, j+1/2): lattice (i, 2j+1)
        t[nf + (i - 1) + static_cast<std::size_t>(n) * (j + static_cast<std::size_t>(n + 1) * (p - 1))] =
            ext(2, 2, i, 2 * j + 1, p);
  for (double x : t)
    if (!std::isfinite(x) || x < 0.0)
      return fail(hd, CCS_ERR_INVALID, "ccs_set_mesh_ext: face entries of matrixIG_ext must be finite and >= 0");
  if (!hd->face_speed && hipMalloc(&hd->face_speed, 2 * nf * sizeof(double)) != hipSuccess)
    return fail(hd, CCS_ERR_NOMEM, "ccs_set_mesh_ext: device allocation failed");
  HIPCHK(hd, hipMemcpyAsync(hd->face_speed, t.data(), 2 * nf * sizeof(double), hipMemcpyHostToDevice, hd->stream));
  HIPCHK(hd, hipStreamSynchronize(hd->stream));
  hd->ext_set = true;
  return CCS_OK;
}

extern "C" int ccs_set_state(ccs_handle* hd, const double* phiG, const double* u, const double* v,
                             const double* phi, const double* uc, const double* vc) {
  if (!hd) return CCS_ERR_INVALID;
  const Range range("ccs_set_state");
  if (!phiG || !u || !v) return fail(hd, CCS_ERR_INVALID, "ccs_set_state: phiG, u, v are required");
  const bool expl = phi && uc && vc;
  if (!expl && (phi || uc || vc)) return fail(hd, CCS_ERR_INVALID, "ccs_set_state: give all of phi, uc, vc or none");
  if (!hd->mesh_set) return fail(hd, CCS_ERR_ORDER, "ccs_set_state before ccs_set_mesh");
  HIPCHK(hd, hipSetDevice(hd->cfg.device));
  const std::size_t np = hd->np;
  for (std::size_t d = 0; d < np; ++d)
    if (!std::isfinite(phiG[d]) || !std::isfinite(u[d]) || !std::isfinite(v[d]))
      return fail(hd, CCS_ERR_NONFINITE, "ccs_set_state: non-finite value in the state");
  if (hd->cfg.precision == CCS_FP64_PARITY) {
    const std::size_t bytes = np * sizeof(double);
    HIPCHK(hd, hipMemcpyAsync(hd->buf[0].phiG, phiG, bytes, hipMemcpyHostToDevice, hd->stream));
    HIPCHK(hd, hipMemcpyAsync(hd->buf[0].u, u, bytes, hipMemcpyHostToDevice, hd->stream));
    HIPCHK(hd, hipMemcpyAsync(hd->buf[0].v, v, bytes, hipMemcpyHostToDevice, hd->stream));
    if (expl) {
      HIPCHK(hd, hipMemcpyAsync(hd->phi0, phi, bytes, hipMemcpyHostToDevice, hd->stream));
      HIPCHK(hd, hipMemcpyAsync(hd->uc0, uc, bytes, hipMemcpyHostToDevice, hd->stream));
      HIPCHK(hd, hipMemcpyAsync(hd->vc0, vc, bytes, hipMemcpyHostToDevice, hd->stream));
    }
    HIPCHK(hd, hipStreamSynchronize(hd->stream));
  } else {
    const int rc = split_set_state(hd->split, hd->stream, phiG, u, v, phi, uc, vc, &hd->err);
    if (rc != CCS_OK) return rc;
  }
  hd->explicit_valid = expl;
  hd->state_set = true;
  hd->steps_done = 0;
  return CCS_OK;
}

extern "C" int ccs_step(ccs_handle* hd, int nsteps) {
  if (!hd) return CCS_ERR_INVALID;
  const Range range("ccs_step");
  if (nsteps < 0) return fail(hd, CCS_ERR_INVALID, "ccs_step: nsteps < 0");
  if (!hd->mesh_set || !hd->state_set) return fail(hd, CCS_ERR_ORDER, "ccs_step before ccs_set_mesh/ccs_set_state");
  if (hd->cfg.spatial_op == CCS_OP_WENO_EXT && !hd->ext_set)
    return fail(hd, CCS_ERR_ORDER, "ccs_step: CCS_OP_WENO_EXT needs ccs_set_mesh_ext first");
  HIPCHK(hd, hipSetDevice(hd->cfg.device));
  int k = 0;
  if (nsteps > 0 && hd->explicit_valid) {
    const int rc = enqueue_counted_step(hd, true, nullptr);
    if (rc != CCS_OK) return rc;
    HIPCHK(hd, hipGetLastError());
    hd->explicit_valid = false;
    k = 1;
  }
  if (hd->cfg.use_graph && k < nsteps) {
    const int rc = capture_graphs(hd);
    if (rc != CCS_OK) return rc;
    const bool split = hd->cfg.precision != CCS_FP64_PARITY;
    while (k < nsteps) {
      int g, adv;
      if (nsteps - k >= kGraphSteps) {
        g = 2 + (split ? static_cast<int>(hd->steps_done % kRebaseSteps) : 0);
        adv = kGraphSteps;
      } else {
        g = rebase_after(hd, hd->steps_done + 1) ? 1 : 0;
        adv = 1;
      }
      HIPCHK(hd, hipGraphLaunch(hd->graph_exec[g], hd->stream));
      hd->steps_done += adv;
      k += adv;
    }
  }
  for (; k < nsteps; ++k) {
    const int rc = enqueue_counted_step(hd, false, nullptr);
    if (rc != CCS_OK) return rc;
  }
  HIPCHK(hd, hipGetLastError());
  return CCS_OK;
}

extern "C" int ccs_synchronize(ccs_handle* hd) {
  if (!hd) return CCS_ERR_INVALID;
  HIPCHK(hd, hipStreamSynchronize(hd->stream));
  return CCS_OK;
}

extern "C" int ccs_get_state(ccs_handle* hd, double* phiG, double* u, double* v, double* phi, double* uc,
                             double* vc, double* zonal, double* merid) {
  if (!hd) return CCS_ERR_INVALID;
  const Range range("ccs_get_state");
  if (!hd->state_set) return fail(hd, CCS_ERR_ORDER, "ccs_get_state before ccs_set_state");
  HIPCHK(hd, hipSetDevice(hd->cfg.device));
  if (hd->cfg.precision != CCS_FP64_PARITY)
    return split_get_state(hd->split, hd->stream, hd->explicit_valid, phiG, u, v, phi, uc, vc, zonal, merid, &hd->err);
  const std::size_t bytes = hd->np * sizeof(double);
  const Prog<const double> q{hd->buf[0].phiG, hd->buf[0].u, hd->buf[0].v};
  launch_derived_f64(hd->stream, q, hd->st, hd->n, hd->h, hd->derived[0], hd->derived[1], hd->derived[2],
                     hd->derived[3], hd->derived[4]);
  HIPCHK(hd, hipGetLastError());
  if (phiG) HIPCHK(hd, hipMemcpyAsync(phiG, hd->buf[0].phiG, bytes, hipMemcpyDeviceToHost, hd->stream));
  if (u) HIPCHK(hd, hipMemcpyAsync(u, hd->buf[0].u, bytes, hipMemcpyDeviceToHost, hd->stream));
  if (v) HIPCHK(hd, hipMemcpyAsync(v, hd->buf[0].v, bytes, hipMemcpyDeviceToHost, hd->stream));
  const double* src_phi = hd->explicit_valid ? hd->phi0 : hd->derived[0];
  const double* src_uc = hd->explicit_valid ? hd->uc0 : hd->derived[1];
  const double* src_vc = hd->explicit_valid ? hd->vc0 : hd->derived[2];
  if (phi) HIPCHK(hd, hipMemcpyAsync(phi, src_phi, bytes, hipMemcpyDeviceToHost, hd->stream));
  if (uc) HIPCHK(hd, hipMemcpyAsync(uc, src_uc, bytes, hipMemcpyDeviceToHost, hd->stream));
  if (vc) HIPCHK(hd, hipMemcpyAsync(vc, src_vc, bytes, hipMemcpyDeviceToHost, hd->stream));
  if (zonal) HIPCHK(hd, hipMemcpyAsync(zonal, hd->derived[3], bytes, hipMemcpyDeviceToHost, hd->stream));
  if (merid) HIPCHK(hd, hipMemcpyAsync(merid, hd->derived[4], bytes, hipMemcpyDeviceToHost, hd->stream));
  HIPCHK(hd, hipStreamSynchronize(hd->stream));
  return CCS_OK;
}

extern "C" int ccs_tendency(ccs_handle* hd, double* tp, double* tu, double* tv) {
  if (!hd) return CCS_ERR_INVALID;
  if (!hd->mesh_set || !hd->state_set) return fail(hd, CCS_ERR_ORDER, "ccs_tendency before set_mesh/set_state");
  if (hd->cfg.spatial_op == CCS_OP_WENO_EXT && !hd->ext_set)
    return fail(hd, CCS_ERR_ORDER, "ccs_tendency: CCS_OP_WENO_EXT needs ccs_set_mesh_ext first");
  HIPCHK(hd, hipSetDevice(hd->cfg.device));
  if (hd->cfg.precision != CCS_FP64_PARITY)
    return split_tendency(hd->split, hd->stream, hd->explicit_valid, tp, tu, tv, &hd->err);
  StageArgsF64 a{};
  a.n = hd->n; a.h = hd->h; a.dx = hd->cfg.dx; a.dt = hd->cfg.dt;
  op_args(hd, a);
  a.phi0 = hd->phi0; a.uc0 = hd->uc0; a.vc0 = hd->vc0;
  a.in = Prog<const double>{hd->buf[0].phiG, hd->buf[0].u, hd->buf[0].v};
  a.base = a.in;
  a.tend[0] = hd->tend[0]; a.tend[1] = hd->tend[1]; a.tend[2] = hd->tend[2];
  launch_stage_f64(hd->stream, kTendOnly, hd->explicit_valid, a, hd->st, hd->hs, hd->f64_tile_y);
  HIPCHK(hd, hipGetLastError());
  const std::size_t bytes = hd->ni * sizeof(double);
  if (tp) HIPCHK(hd, hipMemcpyAsync(tp, hd->tend[0], bytes, hipMemcpyDeviceToHost, hd->stream));
  if (tu) HIPCHK(hd, hipMemcpyAsync(tu, hd->tend[1], bytes, hipMemcpyDeviceToHost, hd->stream));
  if (tv) HIPCHK(hd, hipMemcpyAsync(tv, hd->tend[2], bytes, hipMemcpyDeviceToHost, hd->stream));
  HIPCHK(hd, hipStreamSynchronize(hd->stream));
  return CCS_OK;
}

extern "C" int ccs_diagnostics(ccs_handle* hd, double* mass, double* energy) {
  if (!hd) return CCS_ERR_INVALID;
  const Range range("ccs_diagnostics");
  if (!hd->mesh_set || !hd->state_set) return fail(hd, CCS_ERR_ORDER, "ccs_diagnostics before set_mesh/set_state");
  HIPCHK(hd, hipSetDevice(hd->cfg.device));
  double r[2] = {0.0, 0.0};
  if (hd->cfg.precision != CCS_FP64_PARITY) {
    const int rc = split_diagnostics(hd->split, hd->stream, hd->explicit_valid, r, &hd->err);
    if (rc != CCS_OK) return rc;
  } else {
    const Prog<const double> q{hd->buf[0].phiG, hd->buf[0].u, hd->buf[0].v};
    launch_diag_f64(hd->stream, q, hd->explicit_valid ? hd->phi0 : nullptr, hd->explicit_valid ? hd->uc0 : nullptr,
                    hd->explicit_valid ? hd->vc0 : nullptr, hd->st, hd->n, hd->h, hd->partials, hd->result);
    HIPCHK(hd, hipGetLastError());
    HIPCHK(hd, hipMemcpyAsync(r, hd->result, 2 * sizeof(double), hipMemcpyDeviceToHost, hd->stream));
    HIPCHK(hd, hipStreamSynchronize(hd->stream));
  }
  if (!std::isfinite(r[0]) || !std::isfinite(r[1])) {
    if (mass) *mass = r[0];
    if (energy) *energy = r[1];
    return fail(hd, CCS_ERR_NONFINITE, "ccs_diagnostics: non-finite mass or energy (state blew up)");
  }
  if (mass) *mass = r[0];
  if (energy) *energy = r[1];
  return CCS_OK;
}

extern "C" int ccs_time_steps(ccs_handle* hd, int nsteps, double* total_ms, double* kernel_ms) {
  if (!hd) return CCS_ERR_INVALID;
  if (nsteps <= 0) return fail(hd, CCS_ERR_INVALID, "ccs_time_steps: nsteps must be positive");
  if (!hd->mesh_set || !hd->state_set) return fail(hd, CCS_ERR_ORDER, "ccs_time_steps before set_mesh/set_state");
  if (hd->cfg.spatial_op == CCS_OP_WENO_EXT && !hd->ext_set)
    return fail(hd, CCS_ERR_ORDER, "ccs_time_steps: CCS_OP_WENO_EXT needs ccs_set_mesh_ext first");
  HIPCHK(hd, hipSetDevice(hd->cfg.device));
  if (hd->explicit_valid) {  // the first step reads the uploaded companions: keep it out of the timing
    const int rc = ccs_step(hd, 1);
    if (rc != CCS_OK) return rc;
  }
  EventSet ends(2);
  HIPCHK(hd, ends.create());
  HIPCHK(hd, hipEventRecord(ends[0], hd->stream));
  int rc = ccs_step(hd, nsteps);
  HIPCHK(hd, hipEventRecord(ends[1], hd->stream));
  HIPCHK(hd, hipEventSynchronize(ends[1]));
  float ms = 0.f;
  HIPCHK(hd, hipEventElapsedTime(&ms, ends[0], ends[1]));
  if (total_ms) *total_ms = ms;
  if (rc == CCS_OK && kernel_ms) {
    // Second pass: every stage launch (no graph) carries its own start and end
    // events (the dispatch's timestamps), mean duration.
    const int ns = stages_per_step(hd);
    EventSet ev(2 * ns);
    HIPCHK(hd, ev.create());
    double acc = 0.0;
    for (int k = 0; k < nsteps && rc == CCS_OK; ++k) {
      rc = enqueue_counted_step(hd, false, ev.ev.data());
      HIPCHK(hd, hipEventSynchronize(ev[2 * ns - 1]));
      for (int s = 0; s < ns; ++s) {
        float t = 0.f;
        HIPCHK(hd, hipEventElapsedTime(&t, ev[2 * s], ev[2 * s + 1]));
        acc += t;
      }
    }
    *kernel_ms = acc / (static_cast<double>(nsteps) * ns);
  }
  return rc;
}

// ------------------------------------------------------------------ ensemble
// Members are ordinary split-path handles (own buffers, statics, dt); the
// ensemble owns the batched launch plan, a stream and the step graphs.
struct ccs_ensemble {
  ccs_ensemble_config cfg{};
  std::vector<ccs_handle*> m;
  hipStream_t stream = nullptr;
  std::string err;
  EnsPlan* plan = nullptr;
  int tile_y = 0, persistent = 0, tuned_from_cache = 0;
  double stage_ms = 0.0;
  // one step of a run, and kGraphSteps steps per rebase mask (the steps of
  // the graph after which some member may rebase; see ens_enqueue_steps):
  // every replay makes the same launches (the run's step index lives on the
  // device)
  hipGraph_t graph1[2]{};            // one step: without, with the rebase
  hipGraphExec_t graph1_exec[2]{};
  std::map<unsigned, std::pair<hipGraph_t, hipGraphExec_t>> graphs16;
  std::vector<long long> done;  // per member: steps since its last ccs_ensemble_set_state
};

namespace {

int efail(ccs_ensemble* e, int code, const std::string& msg) {
  if (e) e->err = msg;
  return code;
}

#define ENSCHK(e, call)                                                                       \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return efail((e), CCS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

void ens_drop_graphs(ccs_ensemble* e) {
  for (int r = 0; r < 2; ++r) {
    if (e->graph1_exec[r]) (void)hipGraphExecDestroy(e->graph1_exec[r]);
    if (e->graph1[r]) (void)hipGraphDestroy(e->graph1[r]);
    e->graph1_exec[r] = nullptr;
    e->graph1[r] = nullptr;
  }
  for (auto& kv : e->graphs16) {
    if (kv.second.second) (void)hipGraphExecDestroy(kv.second.second);
    if (kv.second.first) (void)hipGraphDestroy(kv.second.first);
  }
  e->graphs16.clear();
}

// Capture `steps` steps with the given rebase mask into a graph.
int ens_capture(ccs_ensemble* e, int steps, unsigned mask, hipGraph_t* g, hipGraphExec_t* x) {
  ENSCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  ens_enqueue_steps(e->plan, e->stream, e->cfg.scheme, steps, mask);
  hipGraph_t gr = nullptr;
  ENSCHK(e, hipStreamEndCapture(e->stream, &gr));
  *g = gr;
  ENSCHK(e, hipGraphInstantiate(x, gr, nullptr, nullptr, 0));
  return CCS_OK;
}

// Steps of the kGraphSteps steps from run step c0 after which some member
// reaches a multiple of kRebaseSteps of its own steps: member m, at its
// step count done[m] before the run, takes step c0 + k as its
// (done[m] + c0 + k + 1)-th if c0 + k < target[m].
unsigned ens_rebase_mask(const ccs_ensemble* e, const int* target, int c0) {
  unsigned mask = 0;
  for (std::size_t m = 0; m < e->m.size(); ++m) {
    if (c0 >= target[m]) continue;
    const int k = static_cast<int>((kRebaseSteps - 1 - (e->done[m] + c0) % kRebaseSteps) % kRebaseSteps);
    if (c0 + k < target[m]) mask |= 1u << k;
  }
  return mask;
}

int ens_member(ccs_ensemble* e, int member) {
  if (!e) return CCS_ERR_INVALID;
  if (member < 0 || member >= static_cast<int>(e->m.size()))
    return efail(e, CCS_ERR_INVALID, "ensemble: member index out of range");
  return CCS_OK;
}

// pass a member's error text up to the ensemble
int ens_forward(ccs_ensemble* e, int member, int rc) {
  if (rc != CCS_OK) e->err = "member " + std::to_string(member) + ": " + e->m[member]->err;
  return rc;
}

std::vector<SplitState*> ens_states(const ccs_ensemble* e) {
  std::vector<SplitState*> s;
  for (ccs_handle* h : e->m) s.push_back(h->split);
  return s;
}

// ---- autotune cache: the choice per (device, precision, scheme, halo,
// member sizes, constraints of the config), kept for the life of the process
// and, when CCS_AUTOTUNE_CACHE names a file, in that file (one line per
// choice: key tile_y persistent stage_ms).
struct TuneChoice {
  int tile_y = 16, persistent = 0;
  double stage_ms = 0.0;
};
std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune;
bool g_tune_loaded = false;

std::string tune_key(const ccs_ensemble* e) {
  hipDeviceProp_t prop{};
  std::string arch = "dev";
  if (hipGetDeviceProperties(&prop, e->cfg.device) == hipSuccess)
    arch = std::string(prop.gcnArchName) + "x" + std::to_string(prop.multiProcessorCount);
  for (char& c : arch)
    if (c == ' ' || c == ':') c = '_';
  std::map<int, int> sizes;
  for (const ccs_handle* h : e->m) ++sizes[h->n];
  std::string k = arch + ":p" + std::to_string(e->cfg.precision) + ":s" + std::to_string(e->cfg.scheme) + ":h" +
                  std::to_string(e->cfg.halo) + ":t" + std::to_string(e->cfg.tile_y) + ":g" +
                  std::to_string(e->cfg.persistent) + ":";
  for (const auto& kv : sizes) k += std::to_string(kv.first) + "x" + std::to_string(kv.second) + ",";
  return k;
}

void tune_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* path = std::getenv("CCS_AUTOTUNE_CACHE");
  if (!path || !*path) return;
  std::FILE* f = std::fopen(path, "r");
  if (!f) return;
  char key[4096];
  TuneChoice c;
  while (std::fscanf(f, "%4095s %d %d %lf", key, &c.tile_y, &c.persistent, &c.stage_ms) == 4)
    if ((c.tile_y == 8 || c.tile_y == 16 || c.tile_y == 32) && (c.persistent == 0 || c.persistent == 1))
      g_tune[key] = c;
  std::fclose(f);
}

bool tune_lookup(const std::string& key, TuneChoice* c) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load_locked();
  const auto it = g_tune.find(key);
  if (it == g_tune.end()) return false;
  *c = it->second;
  return true;
}

void tune_store(const std::string& key, const TuneChoice& c) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load_locked();
  g_tune[key] = c;
  const char* path = std::getenv("CCS_AUTOTUNE_CACHE");
  // keys the loader's 4095-character field cannot hold stay in memory only
  if (!path || !*path || key.size() > 4000) return;
  if (std::FILE* f = std::fopen(path, "a")) {
    std::fprintf(f, "%s %d %d %.9g\n", key.c_str(), c.tile_y, c.persistent, c.stage_ms);
    std::fclose(f);
  }
}

// Advance member m by target[m] steps: one batched launch per RK stage for
// all members still stepping, replayed from the one-step and kGraphSteps-step
// graphs.
int ens_run(ccs_ensemble* e, const int* target) {
  const Range range("ccs_ensemble run");
  for (ccs_handle* h : e->m)
    if (!h->mesh_set || !h->state_set)
      return efail(e, CCS_ERR_ORDER, "ensemble run before every member's mesh and state are set");
  int steps = 0;
  for (std::size_t k = 0; k < e->m.size(); ++k) {
    if (target[k] < 0) return efail(e, CCS_ERR_INVALID, "ensemble: negative step count");
    steps = std::max(steps, target[k]);
  }
  if (!e->plan) {
    const int rc = ccs_ensemble_autotune(e, nullptr, nullptr, nullptr);
    if (rc != CCS_OK) return rc;
  }
  ENSCHK(e, hipSetDevice(e->cfg.device));
  if (steps == 0) return CCS_OK;
  // members' own streams (uploads) are idle: every member call synchronises
  int rc = ens_plan_set_run(e->plan, e->stream, target, &e->err);
  if (rc != CCS_OK) return rc;
  const int scheme = e->cfg.scheme;
  int k = 0;
  for (; steps - k >= kGraphSteps; k += kGraphSteps) {
    const unsigned mask = ens_rebase_mask(e, target, k);
    if (!e->cfg.use_graph) {
      ens_enqueue_steps(e->plan, e->stream, scheme, kGraphSteps, mask);
      continue;
    }
    auto it = e->graphs16.find(mask);
    if (it == e->graphs16.end()) {
      hipGraph_t g = nullptr;
      hipGraphExec_t x = nullptr;
      rc = ens_capture(e, kGraphSteps, mask, &g, &x);
      if (rc != CCS_OK) return rc;
      it = e->graphs16.emplace(mask, std::make_pair(g, x)).first;
    }
    ENSCHK(e, hipGraphLaunch(it->second.second, e->stream));
  }
  for (; k < steps; ++k) {
    // with the rebase only where some member's step count calls for it
    const unsigned r = ens_rebase_mask(e, target, k) & 1u;
    if (!e->cfg.use_graph) {
      ens_enqueue_steps(e->plan, e->stream, scheme, 1, r);
      continue;
    }
    if (!e->graph1_exec[r]) {
      rc = ens_capture(e, 1, r, &e->graph1[r], &e->graph1_exec[r]);
      if (rc != CCS_OK) return rc;
    }
    ENSCHK(e, hipGraphLaunch(e->graph1_exec[r], e->stream));
  }
  ENSCHK(e, hipGetLastError());
  for (std::size_t m = 0; m < e->m.size(); ++m) e->done[m] += target[m];
  return CCS_OK;
}

}  // namespace

extern "C" const char* ccs_ensemble_last_error(const ccs_ensemble* e) {
  if (!e) return "null ensemble";
  return e->err.c_str();
}

extern "C" int ccs_ensemble_create(const ccs_ensemble_config* cfg, const int* n, const double* dt, const double* dx,
                                   ccs_ensemble** out) {
  if (!cfg || !n || !dt || !dx || !out) return CCS_ERR_INVALID;
  *out = nullptr;
  if (cfg->n_members < 1 || cfg->halo < 3) return CCS_ERR_INVALID;
  if (cfg->precision != CCS_FP32_SPLIT && cfg->precision != CCS_FP64_FAST) return CCS_ERR_INVALID;
  if (cfg->scheme != CCS_RK3_TVD && cfg->scheme != CCS_RK4) return CCS_ERR_INVALID;
  if (cfg->tile_y != 0 && cfg->tile_y != 8 && cfg->tile_y != 16 && cfg->tile_y != 32) return CCS_ERR_INVALID;
  if (cfg->persistent < -1 || cfg->persistent > 1) return CCS_ERR_INVALID;
  for (int k = 0; k < cfg->n_members; ++k)
    if (n[k] < 4 || n[k] < cfg->halo || !(dt[k] > 0.0) || !(dx[k] > 0.0)) return CCS_ERR_INVALID;
  ccs_ensemble* e = new ccs_ensemble();
  e->cfg = *cfg;
  e->done.assign(cfg->n_members, 0);
  for (int k = 0; k < cfg->n_members; ++k) {
    ccs_config c{};
    c.n = n[k];
    c.halo = cfg->halo;
    c.precision = cfg->precision;
    c.scheme = cfg->scheme;
    c.eflux = CCS_EFLUX_NONE;
    c.device = cfg->device;
    c.use_graph = 0;
    c.dt = dt[k];
    c.dx = dx[k];
    c.radius = cfg->radius;
    c.gravity = cfg->gravity;
    c.omega = cfg->omega;
    ccs_handle* h = nullptr;
    const int rc = ccs_create(&c, &h);
    if (rc != CCS_OK) {
      ccs_ensemble_destroy(e);
      return rc;
    }
    e->m.push_back(h);
  }
  if (hipSetDevice(cfg->device) != hipSuccess ||
      hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    ccs_ensemble_destroy(e);
    return CCS_ERR_DEVICE;
  }
  *out = e;
  return CCS_OK;
}

extern "C" void ccs_ensemble_destroy(ccs_ensemble* e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  ens_drop_graphs(e);
  if (e->plan) ens_plan_destroy(e->plan);
  for (ccs_handle* h : e->m) ccs_destroy(h);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

extern "C" int ccs_ensemble_set_mesh(ccs_ensemble* e, int member, const double* sqrtG, const double* G,
                                     const double* IG, const double* A, const double* IA, const double* f,
                                     const double* phis) {
  int rc = ens_member(e, member);
  if (rc != CCS_OK) return rc;
  ENSCHK(e, hipStreamSynchronize(e->stream));
  return ens_forward(e, member, ccs_set_mesh(e->m[member], sqrtG, G, IG, A, IA, f, phis));
}

extern "C" int ccs_ensemble_set_state(ccs_ensemble* e, int member, const double* phiG, const double* u,
                                      const double* v) {
  int rc = ens_member(e, member);
  if (rc != CCS_OK) return rc;
  ENSCHK(e, hipStreamSynchronize(e->stream));
  rc = ens_forward(e, member, ccs_set_state(e->m[member], phiG, u, v, nullptr, nullptr, nullptr));
  if (rc != CCS_OK) return rc;
  // this member's rebase points count from this upload; the others keep theirs
  e->done[member] = 0;
  if (e->plan && ens_plan_reset_member(e->plan, e->stream, member) != CCS_OK)
    return efail(e, CCS_ERR_DEVICE, "ccs_ensemble_set_state: step counter reset failed");
  return CCS_OK;
}

extern "C" int ccs_ensemble_get_state(ccs_ensemble* e, int member, double* phiG, double* u, double* v, double* phi,
                                      double* uc, double* vc, double* zonal, double* merid) {
  int rc = ens_member(e, member);
  if (rc != CCS_OK) return rc;
  ENSCHK(e, hipStreamSynchronize(e->stream));
  return ens_forward(e, member, ccs_get_state(e->m[member], phiG, u, v, phi, uc, vc, zonal, merid));
}

extern "C" int ccs_ensemble_diagnostics(ccs_ensemble* e, int member, double* mass, double* energy) {
  int rc = ens_member(e, member);
  if (rc != CCS_OK) return rc;
  ENSCHK(e, hipStreamSynchronize(e->stream));
  return ens_forward(e, member, ccs_diagnostics(e->m[member], mass, energy));
}

extern "C" int ccs_ensemble_steps_done(const ccs_ensemble* e, int member, long long* steps) {
  if (!e || !steps || member < 0 || member >= static_cast<int>(e->m.size())) return CCS_ERR_INVALID;
  *steps = e->done[member];
  return CCS_OK;
}

extern "C" int ccs_ensemble_autotune(ccs_ensemble* e, int* tile_y, int* persistent, double* stage_ms) {
  if (!e) return CCS_ERR_INVALID;
  for (ccs_handle* h : e->m)
    if (!h->mesh_set) return efail(e, CCS_ERR_ORDER, "ccs_ensemble_autotune before every member's mesh is set");
  ENSCHK(e, hipSetDevice(e->cfg.device));
  ENSCHK(e, hipStreamSynchronize(e->stream));
  std::vector<SplitState*> st = ens_states(e);
  const int nm = static_cast<int>(st.size());
  const std::string key = tune_key(e);
  TuneChoice cached;
  EnsPlan* best = nullptr;
  double best_ms = 0.0;
  int rc = CCS_OK;
  if (tune_lookup(key, &cached)) {
    rc = ens_plan_create(st.data(), nm, cached.tile_y, cached.persistent, e->cfg.device, e->done.data(), &best, &e->err);
    if (rc != CCS_OK) return rc;
    best_ms = cached.stage_ms;
    e->tuned_from_cache = 1;
  } else {
    std::vector<int> tys, pers;
    if (e->cfg.tile_y) tys.push_back(e->cfg.tile_y); else tys = {16, 8, 32};
    if (e->cfg.persistent >= 0) pers.push_back(e->cfg.persistent); else pers = {0, 1};
    EventSet ev(2);
    ENSCHK(e, ev.create());
    constexpr int kReps = 10;
    for (int ty : tys) {
      for (int ps : pers) {
        EnsPlan* p = nullptr;
        rc = ens_plan_create(st.data(), nm, ty, ps, e->cfg.device, e->done.data(), &p, &e->err);
        if (rc != CCS_OK) break;
        // a stage-2 launch has the cost of a stage and changes no state
        ens_enqueue_probe(p, e->stream);
        (void)hipEventRecord(ev[0], e->stream);
        for (int r = 0; r < kReps; ++r) ens_enqueue_probe(p, e->stream);
        (void)hipEventRecord(ev[1], e->stream);
        float ms = 0.f;
        if (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) {
          ens_plan_destroy(p);
          rc = efail(e, CCS_ERR_DEVICE, "ccs_ensemble_autotune: timing failed");
          break;
        }
        const double mean = ms / kReps;
        // a plan must beat the best so far by 1 %: ties keep the earlier one
        // (16-row tiles, one workgroup per tile come first)
        if (!best || mean < 0.99 * best_ms) {
          if (best) ens_plan_destroy(best);
          best = p;
          best_ms = mean;
        } else {
          ens_plan_destroy(p);
        }
      }
      if (rc != CCS_OK) break;
    }
    if (rc != CCS_OK) {
      if (best) ens_plan_destroy(best);
      return rc;
    }
    tune_store(key, TuneChoice{ens_plan_tile_y(best), ens_plan_persistent(best), best_ms});
    e->tuned_from_cache = 0;
  }
  ens_drop_graphs(e);
  if (e->plan) ens_plan_destroy(e->plan);
  e->plan = best;
  e->stage_ms = best_ms;
  e->tile_y = ens_plan_tile_y(best);
  e->persistent = ens_plan_persistent(best);
  if (tile_y) *tile_y = e->tile_y;
  if (persistent) *persistent = e->persistent;
  if (stage_ms) *stage_ms = e->stage_ms;
  return CCS_OK;
}

extern "C" int ccs_ensemble_autotune_source(const ccs_ensemble* e, int* from_cache) {
  if (!e || !from_cache) return CCS_ERR_INVALID;
  *from_cache = e->tuned_from_cache;
  return CCS_OK;
}

extern "C" int ccs_ensemble_step(ccs_ensemble* e, int nsteps) {
  if (!e) return CCS_ERR_INVALID;
  if (nsteps < 0) return efail(e, CCS_ERR_INVALID, "ccs_ensemble_step: nsteps < 0");
  const std::vector<int> target(e->m.size(), nsteps);
  return ens_run(e, target.data());
}

extern "C" int ccs_ensemble_step_members(ccs_ensemble* e, const int* nsteps) {
  if (!e) return CCS_ERR_INVALID;
  if (!nsteps) return efail(e, CCS_ERR_INVALID, "ccs_ensemble_step_members: null step counts");
  return ens_run(e, nsteps);
}

extern "C" int ccs_ensemble_run_until(ccs_ensemble* e, double run_time) {
  if (!e) return CCS_ERR_INVALID;
  if (!(run_time >= 0.0)) return efail(e, CCS_ERR_INVALID, "ccs_ensemble_run_until: run_time must be >= 0");
  std::vector<int> target(e->m.size());
  for (std::size_t k = 0; k < e->m.size(); ++k) {
    // integer conversion of run_time / dt, as parameters_mod.F90:149 derives nsteps
    const double q = run_time / e->m[k]->cfg.dt;
    if (!(q < 2147483647.0)) return efail(e, CCS_ERR_INVALID, "ccs_ensemble_run_until: too many steps");
    const long long total = static_cast<long long>(q);
    target[k] = static_cast<int>(std::max(0LL, total - e->done[k]));
  }
  return ens_run(e, target.data());
}

extern "C" int ccs_ensemble_time_steps(ccs_ensemble* e, int nsteps, double* total_ms) {
  if (!e) return CCS_ERR_INVALID;
  if (nsteps <= 0) return efail(e, CCS_ERR_INVALID, "ccs_ensemble_time_steps: nsteps must be positive");
  if (!e->plan) {
    const int rc = ccs_ensemble_autotune(e, nullptr, nullptr, nullptr);
    if (rc != CCS_OK) return rc;
  }
  // a run of nsteps for every member is set up (and its tables uploaded)
  // before the timed region
  const std::vector<int> target(e->m.size(), nsteps);
  int rc = ens_plan_set_run(e->plan, e->stream, target.data(), &e->err);
  if (rc != CCS_OK) return rc;
  EventSet ev(2);
  ENSCHK(e, ev.create());
  ENSCHK(e, hipEventRecord(ev[0], e->stream));
  rc = ccs_ensemble_step(e, nsteps);
  ENSCHK(e, hipEventRecord(ev[1], e->stream));
  ENSCHK(e, hipEventSynchronize(ev[1]));
  float ms = 0.f;
  ENSCHK(e, hipEventElapsedTime(&ms, ev[0], ev[1]));
  if (total_ms) *total_ms = ms;
  return rc;
}

extern "C" int ccs_ensemble_synchronize(ccs_ensemble* e) {
  if (!e) return CCS_ERR_INVALID;
  ENSCHK(e, hipStreamSynchronize(e->stream));
  return CCS_OK;
}
