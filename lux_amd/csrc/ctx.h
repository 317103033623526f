// The context behind lux_sgm_ctx* and the error helpers of the C ABI, shared by the translation units that
// implement entry points (api.hip, sweep.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lux_sgm.h"
#include "engine.h"

namespace lux {

constexpr int kMaxBranches = 8;

struct ExecBatch;
struct GraphEntry {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t gen = 0;
    ExecBatch* xb = nullptr;  // the executor's work lists when the graph runs the executor
};

}  // namespace lux

struct lux_sgm_ctx {
    int device = 0;
    std::string err;
    std::mutex mu;
    hipStream_t stream = nullptr;
    lux::Budget budget;
    lux::Workspace ws;                                   // single requests and layers
    std::vector<std::unique_ptr<lux::Workspace>> bws;     // one per pair of a batch
    hipStream_t side[lux::kMaxBranches] = {};
    hipEvent_t fork = nullptr, join[lux::kMaxBranches] = {};
    hipEvent_t chain_fork = nullptr;                     // the border chains' fork (they run on side[3])
    std::map<std::string, lux::GraphEntry> graphs;        // batch descriptor -> instantiated graph
    std::map<std::string, lux_sgm_schedule> tuned;        // batch shapes -> schedule
    uint64_t gen = 0;                                     // bumped when a batch workspace is reallocated
    int opt_bands = 0;
    // stage timing (lux_sgm_set_profiling): events recorded between the launches of one request
    bool profiling = false;
    hipEvent_t ev[16] = {};
    int nev = 0;
};

namespace lux {

// The context's side streams and events, made on first use (default priority, non-blocking).
inline int ensure_side_streams(lux_sgm_ctx* c)
{
    if (c->fork) return 0;
    hipError_t e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->chain_fork, hipEventDisableTiming);
    for (int k = 0; e == hipSuccess && k < kMaxBranches; k++) {
        e = hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join[k], hipEventDisableTiming);
    }
    return e == hipSuccess ? 0 : hip_status(e, "side streams", &c->err);
}

// The context's side stream and events as the border chains take them (side 3; the batch form uses the side streams
// as its branches; calls on a context are serialised by its mutex). Without side streams the chains stay in order.
inline Conc conc_of(lux_sgm_ctx* c)
{
    Conc k;
    if (ensure_side_streams(c)) return k;
    k.chain = c->side[3];
    k.chain_fork = c->chain_fork;
    k.chain_join = c->join[3];
    return k;
}

inline int fail(lux_sgm_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

inline int hip_fail(lux_sgm_ctx* c, hipError_t e, const char* where) { return hip_status(e, where, c ? &c->err : nullptr); }

}  // namespace lux
