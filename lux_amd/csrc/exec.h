// Persistent executor of rectified requests (exec.hip): every stage of a request with byte path volumes (census
// codes, the eight path chains, the border replays, the winner-take-all) becomes work items of a device work list
// that resident wavefronts pull by ticket, an item starting once the stages it reads are complete. A whole batch of
// pairs runs in one launch, or in one launch per pass group.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "engine.h"

namespace lux {

// One pair of a batch: its parameters, its prepared workspace (engine.h: prepare) and its device images and outputs.
struct ExecPair {
    const lux_sgm_rect_params* p = nullptr;
    const Prepared* w = nullptr;
    const void* left = nullptr;
    const void* right = nullptr;
    uint16_t* disp = nullptr;  // required
    float* sub = nullptr;      // may be null
    uint8_t* valid = nullptr;  // may be null
};

// Whether the executor runs this request: a census request with byte path volumes (path_bytes) whose schedule
// writes pass +1's half 0 (every height above one tile row does).
bool exec_supported(const lux_sgm_rect_params* p, const Prepared& w);
// The same from the parameters alone (checked ones: check_rect).
bool exec_params_supported(const lux_sgm_rect_params* p);

// A batch's work lists under one schedule, uploaded to the device (opaque outside exec.hip).
struct ExecBatch;
// groups: 1 (every stage of every pair in one launch) or 7 (one launch per pass group: census, pass +1's row paths,
// its sweeps and border chains, its border replays, pass -1's sweeps and chains, its row paths, its replays and the
// finish). order 0: stage by stage, pairs in batch order; 1: the pairs of a stage with the longest chains first.
// Allocates and uploads synchronously (not capturable).
int exec_create(const ExecPair* pairs, int n, int groups, int order, int device, ExecBatch** out, std::string* err);
void exec_destroy(ExecBatch* b);
// Enqueues one run on s: the counters cleared by a kernel, then the launches. Capturable into a graph.
// waves: resident wavefronts of each launch (0: as many as the device holds at once).
hipError_t exec_enqueue(ExecBatch* b, int waves, hipStream_t s);
// After s has drained: LUX_SGM_EDEVICE if a wait gave up (bounded spins), else 0.
int exec_check(ExecBatch* b, hipStream_t s, std::string* err);
// Launches one run enqueues.
int exec_launches(const ExecBatch* b);

}  // namespace lux
