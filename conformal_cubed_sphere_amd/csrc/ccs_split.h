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
// A run advances member m by target[m] steps, from graphs of 1 and of
// kRebaseSteps steps whose launches are the same for every replay: a device
// counter holds the run's step index at the start of the graph (advanced by
// a one-workgroup kernel at the graph's end, with the per-member step
// counts), every stage kernel skips the tiles of members that have taken
// their steps, and the rebase kernel folds the deviation of members whose own
// step count since the upload of their state reaches a multiple of
// kRebaseSteps.  In a graph of kRebaseSteps steps the rebase kernel follows
// only the steps named by a mask, which the host derives from the members'
// step counts (one step of the graph when every member counts from the same
// phase).
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
// Enqueue `steps` (1 or kRebaseSteps) steps; rebase_mask: for a graph of
// kRebaseSteps steps, bit k set where some member may reach its rebase count
// with step k (a superset is correct); a single step always carries it.
void ens_enqueue_steps(const EnsPlan* e, hipStream_t st, int scheme, int steps, unsigned rebase_mask);
void ens_enqueue_tendency(const EnsPlan* e, hipStream_t st);
// One stage launch that leaves every member's state as it is (the stage-2
// kernel, whose output buffer is dead between steps): the autotuner's probe.
void ens_enqueue_probe(const EnsPlan* e, hipStream_t st);
int ens_plan_grid(const EnsPlan* e);
int ens_plan_tile_y(const EnsPlan* e);
int ens_plan_persistent(const EnsPlan* e);
int ens_plan_tiles(const EnsPlan* e);

}  // namespace ccs

#endif
