// Request engine shared by the entry points of liblux_sgm.so: pooled device workspaces under an HBM budget,
// the schedule-plan cache, and the enqueue of one rectified request or one layer onto a stream.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lux_sgm.h"
#include "kernels.h"
#include "plan.h"

namespace lux {

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
};

// Device bytes held by the workspaces of one context or server against its HBM budget (0 = unbounded):
// the device-side form of the reference's memory admission (DenseStereoSgm.cs:142-156, :330,339-344).
struct Budget {
    uint64_t limit = 0;
    uint64_t used = 0;
    uint64_t peak = 0;
    std::mutex mu;
    // Account for growing an allocation from `old_n` to `new_n` bytes; false if it would exceed the limit.
    bool grow(uint64_t old_n, uint64_t new_n);
    void release(uint64_t n);
};

// The device buffers of one request in flight. Buffers are grow-only and reused by the next request with the
// same or a smaller shape (the ScratchPool pattern, Lux.Engine/Pipeline/ScratchPool.cs:6-15).
struct Workspace {
    int device = 0;
    Budget* budget = nullptr;
    std::map<std::string, DevBuf> bufs;
    std::set<std::string> filled;    // buffers whose contents outlive a call (cleared when reallocated or released)
    void* sync_zeroed = nullptr;     // band flag array known to hold no live epoch
    std::map<std::string, std::vector<float>> uploaded;  // host copies of float buffers uploaded once (when filled)
    int epoch = 0;
    int force_literal = 0;           // LUX_SGM_OPT_LITERAL: layers of <= 1024 lanes take the line-buffer-literal path
    int lit_resident[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};  // resident literal wavefronts [npl 2/4/0][adaptive]

    // Arena mode (arena != nullptr): get() hands out 256-byte aligned pieces of [arena, arena + arena_cap) instead of
    // allocating, and release() frees nothing. The no-allocation entry point gives a fresh workspace the library's
    // static scratch this way, so a request's buffers need no hipMalloc.
    uint8_t* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;

    Workspace() = default;
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    ~Workspace() { release(); }
    int get(const char* name, size_t bytes, void** out, std::string* err);
    // A small pinned host buffer (grow-only) for read-backs the host waits for.
    int pinned(size_t bytes, void** out, std::string* err);
    void* host_pinned = nullptr;
    size_t host_pinned_n = 0;
    uint64_t held() const;
    void release();
};

// One request prepared on a workspace: geometry and the kernel-side buffer table.
struct Prepared {
    Geometry g;
    SgmBuffers b;
    uint8_t* raw = nullptr;
    uint16_t* agg = nullptr;
};

// The caller's side stream and events for the border chains (SgmBuffers::chain), reused per call.
struct Conc {
    hipStream_t chain = nullptr;
    hipEvent_t chain_fork = nullptr, chain_join = nullptr;
};
// Gives w the three path-cost volumes when they fit the budget (ranged layers then run each pass as one launch);
// otherwise the paths run one after another. Returns whether it did.
bool try_path_volumes(Workspace& ws, Prepared* w);
// Runs w's border chains on c's chain stream, beside the sweeps (no-op without one).
void attach_chain(Prepared* w, const Conc* c);

struct RectOut {
    uint16_t* disp = nullptr;
    float* sub = nullptr;
    uint8_t* valid = nullptr;
    uint16_t* agg = nullptr;  // [H*W*cap]
};

struct EngineOpts {
    int bands = 0;
    // The launches may be captured into a graph and replayed: band flags are cleared by a kernel before each
    // band launch instead of relying on a fresh epoch per request.
    bool replayable = false;
    // Other requests share the device at the same time (batch branches, server slots): the chains keep four-wavefront
    // workgroups instead of one-wavefront ones.
    bool shared = false;
};

// The named device buffers of one request: prepare() allocates exactly these, and the server admits a request
// by the growth they need in a slot's workspace. host_img_bytes > 0 adds the staging of host images and outputs.
struct BufferNeed {
    std::string name;
    size_t bytes;
};
// pv8: the request runs with byte path volumes (path_bytes).
std::vector<BufferNeed> request_buffers(const Geometry& g, const Plan& plan, bool census, size_t host_img_bytes = 0,
                                        bool pv8 = false);
// A rectified request whose eight path costs all fit a byte (census cost, no padding planes, census maximum + P2 <=
// 255, at most 256 planes): it runs with one byte volume per path instead of a u16 aggregate read and written by
// every path, 22 bytes a cell instead of 36, with the same results.
bool path_bytes(const lux_sgm_rect_params* p);
// Bytes by which ws must grow to hold `need` (0: it already does).
uint64_t growth(const Workspace& ws, const std::vector<BufferNeed>& need);

int check_rect(const lux_sgm_rect_params* p, std::string* err);
int check_layer(const lux_sgm_layer_params* p, std::string* err);
// Device bytes one rectified request holds in a workspace (inputs and outputs excluded; 0 if its schedule is
// unsupported).
uint64_t rect_workspace_bytes(const lux_sgm_rect_params* p);

// Plans are pure functions of (W, H, T, xsw): one process-wide cache.
int plan_for(int W, int H, int T, int xsw, std::shared_ptr<Plan>* out, std::string* err);

// The buffers of a request in ws (Workspace::get: allocated, or placed in an arena) and w's buffer table over them,
// with no device work. visits (may be null): the plan's visit lists [plus | minus | cellvis plus | cellvis minus]
// are already on the device there, else they get a workspace buffer of their own (not filled here). named (may be
// null) receives every buffer by name (the staging buffers img_l, img_r, out_* when host_img_bytes > 0).
int prepare_layout(Workspace& ws, const Geometry& g, const Plan& plan, bool census, bool pv8, size_t host_img_bytes,
                   const int* visits, Prepared* w, std::map<std::string, void*>* named, std::string* err);
// The plan's visit lists into vis (plan_ints ints) on s.
int upload_plan(const Plan& plan, int* vis, hipStream_t s, std::string* err);
size_t plan_ints(const Plan& plan);
// Allocate the workspace for geometry g and upload the plan's visit lists (copies from host memory on s).
// Nothing here may be captured into a graph; enqueue_* below may.
int prepare(Workspace& ws, const Geometry& g, int T, int xsw, bool census, hipStream_t s, Prepared* w, std::string* err,
            bool pv8 = false);

// The line-buffer-literal path (launch_literal). prepare_literal gives w the raw and aggregate volumes of g (the
// workspace buffers prepare() uses, so a layer's cost can be written into them first). run_literal runs one layer
// whose largest cap is maxcap and synchronises s: LUX_SGM_EUNSUPPORTED above 1024 lanes or when the device cannot
// hold the wavefronts the schedule's waits need, LUX_SGM_EDEVICE when a wait gave up.
int prepare_literal(Workspace& ws, const Geometry& g, Prepared* w, std::string* err);
int run_literal(Workspace& ws, LiteralLayer l, int maxcap, uint16_t* agg_pass1, uint16_t* wta, hipStream_t s,
                std::string* err);

// Enqueue one rectified request on s: cost volume, the eight paths, WTA, subpixel, L/R. Kernel launches and
// device-to-device copies only. ev (if not null) receives 2 + kSgmMarks events for the stage times.
int enqueue_rect(Workspace& ws, Prepared& w, const lux_sgm_rect_params* p, const void* dl, const void* dr, const RectOut& o,
                 const EngineOpts& opt, hipStream_t s, hipEvent_t* ev, std::string* err);

// Coarse-to-fine requests (lux_sgm_pyramid_params): level geometry, checks, buffers and the enqueue.
struct PyrLevel {
    int W = 0, H = 0, stride = 0, D = 0, radius = 0;  // radius 0: the coarsest level, all D planes
};
std::vector<PyrLevel> pyramid_levels(const lux_sgm_pyramid_params* p);
int check_pyramid(const lux_sgm_pyramid_params* p, std::string* err);
int pyramid_buffers(const lux_sgm_pyramid_params* p, size_t host_img_bytes, std::vector<BufferNeed>* out, std::string* err);
int enqueue_pyramid(Workspace& ws, const lux_sgm_pyramid_params* p, const void* dl, const void* dr, uint16_t* disp,
                    float* sub, hipStream_t s, std::string* err, const Conc* conc = nullptr);

// DenseUpsampleLayer (upsample.hip): prev depth (and confidence, may be null) pw x ph and RGBA8 guide W x H (device)
// -> depth (and confidence when both confidence pointers are set) W x H.
int enqueue_upsample(Workspace& ws, const float* d_prev, const float* d_prev_conf, int pw, int ph, const uint8_t* d_guide,
                     int W, int H, float sigma, float* d_depth, float* d_conf, hipStream_t s, std::string* err);

// After s has drained: a band tile that gave up waiting for its neighbours leaves an error word set.
int check_band_error(Workspace& ws, hipStream_t s, std::string* err);

int hip_status(hipError_t e, const char* where, std::string* err);

// Looks up the device address of the no-allocation entry point's scratch array on `device` (current) once per device
// (noalloc.hip). lux_sgm_create calls it; the entry point calls it too, for a device it has not been called on yet.
int noalloc_setup(int device, std::string* err);

}  // namespace lux
