// ccs_split.hip — split-precision path (placeholder until implemented).
#include "ccs_split.h"

namespace ccs {

struct SplitState {};

int split_create(const ccs_config&, const HaloScatter&, SplitState**, std::string* err) {
  *err = "split-precision path not available in this build";
  return CCS_ERR_INVALID;
}
void split_destroy(SplitState* s) { delete s; }
int split_set_mesh(SplitState*, hipStream_t, double* const*, std::string*) { return CCS_ERR_INVALID; }
int split_set_state(SplitState*, hipStream_t, const double*, const double*, const double*, const double*,
                    const double*, const double*, std::string*) { return CCS_ERR_INVALID; }
int split_enqueue_step(SplitState*, hipStream_t, int, bool, hipEvent_t*) { return CCS_ERR_INVALID; }
int split_get_state(SplitState*, hipStream_t, bool, double*, double*, double*, double*, double*, double*,
                    double*, double*, std::string*) { return CCS_ERR_INVALID; }
int split_tendency(SplitState*, hipStream_t, bool, double*, double*, double*, std::string*) { return CCS_ERR_INVALID; }
int split_diagnostics(SplitState*, hipStream_t, bool, double*, std::string*) { return CCS_ERR_INVALID; }

}  // namespace ccs
