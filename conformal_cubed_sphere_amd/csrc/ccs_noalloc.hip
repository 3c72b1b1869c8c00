// ccs_noalloc.hip — total mass and energy of a state held in caller-owned
// device memory, without any allocation: for callers that cannot allocate
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).
//
// The kernels are those of ccs_diagnostics on the FP64 parity path
// (diag_mod.F90:10-53 as compensated FP64 sums, launch_diag_f64), so the two
// entry points agree bit for bit on the same arrays.  Their only device
// memory beyond the caller's inputs and output is the per-block partial sums,
// kept here in a file-scope device array, `scratch`: each GPU the library
// runs on holds its own copy until the process ends.  Every call on one GPU
// shares that GPU's copy, so calls that overlap in time on one GPU race; the
// partial-sum kernel overwrites every slot it uses, so nothing needs clearing
// between calls.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "ccs.h"
#include "ccs_internal.h"
#include "ccs_kernels.h"

namespace {

// kDiagBlocks partial sums of (mass, its compensation, energy, its
// compensation): 512 x 4 doubles = 16 KiB, whatever the grid size.
constexpr int kScratchDoubles = 4 * ccs::kDiagBlocks;
__device__ __align__(256) double scratch[kScratchDoubles];

// Device address of each GPU's copy of scratch (the array's name is a host
// stand-in, not a device address), looked up once per device.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

void say(char* err, std::size_t len, const char* msg) {
  if (err && len) std::snprintf(err, len, "%s", msg);
}

}  // namespace

extern "C" int ccs_diagnostics_noalloc(int n, int halo, const double* phiG, const double* u, const double* v,
                                       const double* sqrtG, const double* ig11, const double* ig12,
                                       const double* ig21, const double* ig22, const double* phis,
                                       double* result, void* stream, char* err, size_t err_len) {
  if (n < 4 || halo < 3 || halo > n) {
    say(err, err_len, "ccs_diagnostics_noalloc: need n >= 4 and 3 <= halo <= n");
    return CCS_ERR_INVALID;
  }
  // the largest grid: padded arrays of at most 2^31 - 1 cells (about n = 18900)
  if (ccs::padded_cells(n, halo) > 0x7fffffffull) {
    say(err, err_len, "ccs_diagnostics_noalloc: grid beyond the maximum of 2^31 - 1 padded cells");
    return CCS_ERR_INVALID;
  }
  if (!phiG || !u || !v || !sqrtG || !ig11 || !ig12 || !ig21 || !ig22 || !phis || !result) {
    say(err, err_len, "ccs_diagnostics_noalloc: null device array");
    return CCS_ERR_INVALID;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    say(err, err_len, "ccs_diagnostics_noalloc: no current device");
    return CCS_ERR_DEVICE;
  }
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        say(err, err_len, "ccs_diagnostics_noalloc: hipGetSymbolAddress of the scratch array failed");
        return CCS_ERR_DEVICE;
      }
      g_scratch[dev] = static_cast<double*>(p);
    }
    part = g_scratch[dev];
  }
  ccs::StaticsF64 st{};
  st.sqrtG = sqrtG;
  st.IG11 = ig11;
  st.IG12 = ig12;
  st.IG21 = ig21;
  st.IG22 = ig22;
  st.phis = phis;
  const ccs::Prog<const double> q{phiG, u, v};
  ccs::launch_diag_f64(static_cast<hipStream_t>(stream), q, nullptr, nullptr, nullptr, st, n, halo, part, result);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    say(err, err_len, hipGetErrorString(e));
    return CCS_ERR_DEVICE;
  }
  return CCS_OK;
}
