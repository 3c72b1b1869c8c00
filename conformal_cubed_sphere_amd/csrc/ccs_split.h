// ccs_split.h — split-precision (FP32 state, FP64-derived statics) path of
// the time step, and its FP64 instantiation for throughput comparison.
#ifndef CCS_SPLIT_H
#define CCS_SPLIT_H

#include <hip/hip_runtime.h>

#include <string>

#include "ccs.h"
#include "ccs_kernels.h"

namespace ccs {

struct SplitState;

int split_create(const ccs_config& cfg, const HaloScatter& hs, hipStream_t st, SplitState** out, std::string* err);
void split_destroy(SplitState* s);
// T: the 15 host-side SoA tables built by ccs_set_mesh (sqrtG, IG11, IG12,
// IG21, IG22, phis, f, M11, M12, M21, M22, N11, N12, N21, N22), padded.
int split_set_mesh(SplitState* s, hipStream_t st, double* const T[15], std::string* err);
int split_set_state(SplitState* s, hipStream_t st, const double* phiG, const double* u, const double* v,
                    const double* phi, const double* uc, const double* vc, std::string* err);
int split_enqueue_step(SplitState* s, hipStream_t st, int scheme, bool first, hipEvent_t* ev);
// Fold the geopotential deviation into the FP64 reference field (state at a
// step boundary, buffer 0); the represented state is unchanged.
void split_enqueue_rebase(SplitState* s, hipStream_t st);
int split_get_state(SplitState* s, hipStream_t st, bool first, double* phiG, double* u, double* v,
                    double* phi, double* uc, double* vc, double* zonal, double* merid, std::string* err);
int split_tendency(SplitState* s, hipStream_t st, bool first, double* tp, double* tu, double* tv,
                   std::string* err);
int split_diagnostics(SplitState* s, hipStream_t st, bool first, double* result, std::string* err);

// Steps between two folds of the geopotential deviation into the FP64
// reference field, counted from the upload of the state (ccs_set_state).
constexpr int kRebaseSteps = 16;

// Batched ensemble of split-path members: one launch per RK stage covers the
// tiles of every member (tile height 8, 16 or 32; persistent: a grid of
// CUs x resident workgroups that loops over the tiles, else one workgroup
// per tile).  Members keep their own buffers, statics and time step.
//
// A run advances member m by target[m] steps.  The launches of one step are
// the same for every step of a run (so one captured graph serves them all):
// a device counter holds the step index within the run, every stage kernel
// skips the tiles of members that have taken their steps, the rebase kernel
// folds the deviation of members whose own step count since the upload of
// their state reaches a multiple of kRebaseSteps, and a one-workgroup kernel
// advances the per-member step counts and the run's step index.
struct EnsPlan;
// steps_done[m]: the member's steps since the upload of its state (NULL: 0).
int ens_plan_create(SplitState* const* mem, int nm, int tile_y, int persistent, int device,
                    const long long* steps_done, EnsPlan** out, std::string* err);
void ens_plan_destroy(EnsPlan* e);
// Set up a run: member m takes target[m] >= 0 steps.  Members are ordered by
// decreasing target so that the tiles still stepping are a prefix of the tile
// list; a change of targets drains st and uploads the new tables.  Then the
// run's step index is reset on st (no graph may be capturing on it).
int ens_plan_set_run(EnsPlan* e, hipStream_t st, const int* target, std::string* err);
// The member's state was uploaded again: its step count restarts at zero
// (drains st).
int ens_plan_reset_member(EnsPlan* e, hipStream_t st, int member);
void ens_enqueue_step(const EnsPlan* e, hipStream_t st, int scheme);
void ens_enqueue_tendency(const EnsPlan* e, hipStream_t st);
int ens_plan_grid(const EnsPlan* e);
int ens_plan_tile_y(const EnsPlan* e);
int ens_plan_persistent(const EnsPlan* e);
int ens_plan_tiles(const EnsPlan* e);

}  // namespace ccs

#endif
