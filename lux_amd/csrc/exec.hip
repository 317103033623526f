// Persistent executor of rectified requests (declared in exec.h).
//
// A batch of pairs becomes one work list per launch: tickets map to (pair, stage, item) through a short segment
// table, and resident one-wavefront workgroups take tickets from a counter until the list is exhausted. The items
// are the same per-chain device functions the plain launch sequence runs (one row path, one sweep chain, one border
// visit, one census row block ...), so every result is the plain request's bit for bit. An item whose stage reads
// another stage's results waits on that stage's completion counter (same pair) with a bounded spin: tickets are
// handed out in an order where every prerequisite's items come first, so the smallest unfinished ticket can always
// run and every wait ends. Across launches the kernel boundary orders results. Within a launch they cross XCDs,
// whose L2 caches write back lazily and are not coherent with each other, through flush epochs (XSync): the item
// that completes a stage raises an epoch, each XCD writes its L2 back once per epoch (the next of its wavefronts
// to reach a scheduling point does it) and acknowledges, and a consumer starts once every XCD has acknowledged the
// epoch of the stage it reads (or has no wavefront left). That is one L2 write-back per stage and XCD instead of
// one per item. No consumer needs an invalidate: within a run no line is read before the item that writes it has
// finished (each pass has its own border-replay buffers), so no cache holds a stale copy.
#include "exec.h"
#include "kernels_impl.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>

namespace lux {

namespace {

// The stages of one request, numbered so that every prerequisite has a smaller number.
enum XStage : int {
    X_CENSUS = 0,  // census codes: blocks of kXCensusRows rows of one image (left blocks, then right blocks)
    X_FWD,         // pass +1 row path (->), census cost fused, raw volume and path volume 0 written
    X_CHAIN_P,     // pass +1 border chain (path 2 of column W-1)
    X_SW_P1,       // pass +1 sweeps: down-right, down, down-left
    X_SW_P2,
    X_SW_P3,
    X_RPATH_P,     // pass +1 border visits, paths 0, 1 and 3
    X_RSEED_P,     // pass +1 seed-table entries of the border cell
    X_RADD_P,      // pass +1 border costs summed per row
    X_CHAIN_M,     // pass -1 border chain (path 2 of column 0)
    X_SW_M1,       // pass -1 sweeps: up-left, up, up-right
    X_SW_M2,
    X_SW_M3,
    X_BWD,         // pass -1 row path (<-): the eight-path sum, winner-take-all, subpixel, right-image disparity
    X_RPATH_M,     // pass -1 border visits
    X_RADD_M,      // pass -1 border costs summed per row
    X_FINISH,      // column 0's winner-take-all and every pixel's validity
    X_NSTAGES
};
constexpr int kXCensusRows = 8;
constexpr uint32_t xbit(int s) { return 1u << s; }

// The stages whose items must all be complete before an item of stage s starts (same request): what the plain
// launch sequence orders by its stream. A stage completes only after its own prerequisites did, so direct ones
// suffice.
__host__ __device__ constexpr uint32_t exec_prereq(int s)
{
    switch (s) {
        case X_FWD: return xbit(X_CENSUS);
        case X_CHAIN_P: case X_SW_P1: case X_SW_P2: case X_SW_P3: return xbit(X_FWD);  // the raw volume
        // the saved neighbour-column states (row path, down-right sweep)
        case X_RPATH_P: case X_RSEED_P: return xbit(X_SW_P1) | xbit(X_SW_P2) | xbit(X_SW_P3);
        case X_RADD_P: return xbit(X_RPATH_P) | xbit(X_CHAIN_P);
        // pass +1's seed table: interior entries from its sweeps, the border cell's from its chain and replay; the
        // pass -1 chain also reuses the pass +1 chain's cost buffer, which pass +1's border sums read
        case X_SW_M1: case X_SW_M2: case X_SW_M3: return xbit(X_RSEED_P) | xbit(X_CHAIN_P);
        case X_CHAIN_M: return xbit(X_RSEED_P) | xbit(X_RADD_P);
        case X_BWD: return xbit(X_SW_M1) | xbit(X_SW_M2) | xbit(X_SW_M3) | xbit(X_RADD_P);
        case X_RPATH_M: return xbit(X_BWD);
        case X_RADD_M: return xbit(X_RPATH_M) | xbit(X_CHAIN_M);
        case X_FINISH: return xbit(X_RADD_M);
        default: return 0u;
    }
}

// The launches of the two groupings, as stage sets.
constexpr uint32_t kAll = (1u << X_NSTAGES) - 1;
constexpr uint32_t kG0 = xbit(X_CENSUS);
constexpr uint32_t kG1 = xbit(X_FWD);
constexpr uint32_t kG2 = xbit(X_CHAIN_P) | xbit(X_SW_P1) | xbit(X_SW_P2) | xbit(X_SW_P3);
constexpr uint32_t kG3 = xbit(X_RPATH_P) | xbit(X_RSEED_P) | xbit(X_RADD_P);
constexpr uint32_t kG4 = xbit(X_CHAIN_M) | xbit(X_SW_M1) | xbit(X_SW_M2) | xbit(X_SW_M3);
constexpr uint32_t kG5 = xbit(X_BWD);
constexpr uint32_t kG6 = xbit(X_RPATH_M) | xbit(X_RADD_M) | xbit(X_FINISH);
constexpr uint32_t kGroups7[7] = {kG0, kG1, kG2, kG3, kG4, kG5, kG6};

// One request of a work list.
struct ExecJob {
    int variant;      // bit 0: two packed pairs a lane (NP = 2), bit 1: padded lanes (cap < 128 * NP)
    int img16, radius, stride;
    int W, H, tol, pad_;
    const void* left;
    const void* right;
    uint64_t* cl;     // census codes [H][W]
    uint64_t* cr;
    uint8_t* valid;   // may be null
    uint32_t n[X_NSTAGES];  // items of each stage
    RowArgs rf, rb;   // the row paths of pass +1 and pass -1 (rb, with pass -1's border sums, also the finish)
    SweepArgs sp[3];  // pass +1 sweeps (paths 1, 2, 3)
    SweepArgs sm[3];  // pass -1 sweeps
    ReplayArgs cp, cm;
    const int* cellvis_p;
    const int* cellvis_m;
};

// Tickets [t0, next segment's t0) of a launch run items first, first + 1, ... of one stage of one job.
struct ExecSeg {
    uint32_t t0;
    int job, stage, first;
};

// Flush epochs (see the file comment): req counts the epochs raised, ack[x] is the last epoch XCD x has written its
// L2 back for, live[x] its running wavefronts (an XCD with none holds nothing unwritten: each leaves after a
// write-back).
constexpr int kXcds = 8;
struct XSync {
    uint32_t req;
    uint32_t ack[kXcds];
    int32_t live[kXcds];
};

__device__ __forceinline__ int xcc_id()
{
    uint32_t v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(v));
    return (int)(v & (kXcds - 1));
}

__device__ __forceinline__ uint32_t ald(const uint32_t* p)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane(
        (int)__hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// This XCD's share of the protocol: write the L2 back if an epoch was raised that it has not acknowledged.
__device__ __forceinline__ void xsync_poll(XSync* y, int x, int lane)
{
    const uint32_t r = ald(&y->req);
    if (ald(&y->ack[x]) >= r) return;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // buffer_wbl2: every dirty line of this XCD's L2 to memory
    if (lane == 0) __hip_atomic_fetch_max(&y->ack[x], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Whether epoch e has reached memory from every XCD.
__device__ __forceinline__ bool xsync_flushed(const XSync* y, uint32_t e)
{
    for (int x = 0; x < kXcds; x++)
        if (ald(&y->ack[x]) < e && (int)ald((const uint32_t*)&y->live[x]) > 0) return false;
    return true;
}

struct ExecLaunch {
    const ExecJob* jobs;
    const ExecSeg* segs;
    uint32_t* ticket;  // this launch's ticket counter (zero before the launch)
    uint32_t* done;    // [job][X_NSTAGES] completed items (zero before the first launch of a run)
    uint32_t* epoch;   // [job][X_NSTAGES] the flush epoch raised when the stage completed (0: not yet)
    XSync* sync;
    int* err;          // a wait gave up
    uint32_t total;    // tickets of this launch
    int nseg;
    uint32_t wait;     // stages of this launch (prerequisites outside it are complete at the launch boundary)
    uint32_t release;  // stages of this launch that a later stage of the same launch reads
};

// A wave-uniform copy of a struct in device memory (every word through readfirstlane, so it lives in SGPRs).
template <class T>
__device__ __forceinline__ T uload(const T* p)
{
    static_assert(sizeof(T) % 4 == 0, "word-sized structs");
    constexpr int N = (int)(sizeof(T) / 4);
    uint32_t w[N];
    const uint32_t* s = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
    for (int k = 0; k < N; k++) w[k] = (uint32_t)__builtin_amdgcn_readfirstlane((int)s[k]);
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

constexpr uint32_t kSpinLimit = 1u << 23;  // polls of one wait, about 2 s with the sleeps below

// Waits until every stage in `pre` of job `job` is complete and flushed from every XCD; false if a wait gave up
// (the error word is set then). Waiting wavefronts keep their own XCD's share of the protocol going.
__device__ __forceinline__ bool exec_wait(const ExecLaunch& L, const ExecJob* J, int job, uint32_t pre, int x, int lane)
{
    for (int s = 0; s < X_NSTAGES; s++) {
        if (!(pre & xbit(s))) continue;
        const uint32_t need = (uint32_t)__builtin_amdgcn_readfirstlane((int)J->n[s]);
        const size_t k0 = (size_t)job * X_NSTAGES + s;
        for (uint32_t k = 0;; k++) {
            xsync_poll(L.sync, x, lane);
            if (ald(L.done + k0) >= need) {
                const uint32_t e = ald(L.epoch + k0);
                if (e && xsync_flushed(L.sync, e)) break;
            }
            if (ald((const uint32_t*)L.err)) return false;
            if (k >= kSpinLimit) {
                if (lane == 0) __hip_atomic_store(L.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            __builtin_amdgcn_s_sleep(4);
        }
    }
    return true;
}

// An item is done: its stores are in this XCD's L2 (waitcnt), then the count; the item that completes a stage
// another stage of this launch reads raises a flush epoch and records it with the stage.
__device__ __forceinline__ void item_done(const ExecLaunch& L, const ExecJob* J, int job, int stage, int lane)
{
    __builtin_amdgcn_s_waitcnt(0);
    if (lane == 0) {
        const size_t k0 = (size_t)job * X_NSTAGES + stage;
        const uint32_t old = __hip_atomic_fetch_add(L.done + k0, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((L.release & xbit(stage)) && old + 1 == J->n[stage]) {
            const uint32_t e = __hip_atomic_fetch_add(&L.sync->req, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
            __hip_atomic_store(L.epoch + k0, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Rows [y0, y0 + kXCensusRows) of one image's census codes, 64 pixels of a row at a time.
template <typename T, int R>
__device__ __forceinline__ void census_rows(const T* img, int W, int H, int stride, uint64_t* out, int y0, int lane)
{
    const int y1 = min(y0 + kXCensusRows, H);
    for (int y = y0; y < y1; y++)
        for (int x = lane; x < W; x += 64) out[(size_t)y * W + x] = census_at<T, R>(img, W, H, stride, x, y);
}

__device__ __forceinline__ void census_item(const ExecJob* J, int i, int lane)
{
    const int W = __builtin_amdgcn_readfirstlane(J->W), H = __builtin_amdgcn_readfirstlane(J->H);
    const int stride = __builtin_amdgcn_readfirstlane(J->stride);
    const int img16 = __builtin_amdgcn_readfirstlane(J->img16), R = __builtin_amdgcn_readfirstlane(J->radius);
    const int nb = (H + kXCensusRows - 1) / kXCensusRows;
    const bool right = i >= nb;
    const void* img = uload(right ? &J->right : &J->left);
    uint64_t* out = uload(right ? &J->cr : &J->cl);
    const int y0 = (right ? i - nb : i) * kXCensusRows;
#define LUX_XCEN(T, RR) census_rows<T, RR>((const T*)img, W, H, stride, out, y0, lane)
    if (img16) {
        if (R == 1) LUX_XCEN(uint16_t, 1); else if (R == 2) LUX_XCEN(uint16_t, 2); else LUX_XCEN(uint16_t, 3);
    } else {
        if (R == 1) LUX_XCEN(uint8_t, 1); else if (R == 2) LUX_XCEN(uint8_t, 2); else LUX_XCEN(uint8_t, 3);
    }
#undef LUX_XCEN
}

// Item i of a path or replay stage: the device functions of the plain sequence (kernels_impl.h), with the
// diagonal chains longest first.
#define LUX_XITEM(NAME) template <int NP, bool MASKED> __device__ __forceinline__ void NAME(const ExecJob* J, int i, int lane)
LUX_XITEM(x_fwd) { row_fwd_wave<NP, MASKED, true, false, false, true>(uload(&J->rf), i, lane); }
LUX_XITEM(x_chain_p)
{
    const ReplayArgs a = uload(&J->cp);
    replay_chain_wave<NP, MASKED, false, false>(a, a.visits, lane);
}
LUX_XITEM(x_sw_p1)
{
    const SweepArgs a = uload(&J->sp[0]);
    sweep_wave<NP, 1, 1, false, MASKED, false, false, true>(a, diag_order<1>(i, a.W, a.H), lane);
}
LUX_XITEM(x_sw_p2) { sweep_wave<NP, 1, 2, false, MASKED, false, false, true>(uload(&J->sp[1]), i, lane); }
LUX_XITEM(x_sw_p3)
{
    const SweepArgs a = uload(&J->sp[2]);
    sweep_wave<NP, 1, 3, false, MASKED, false, false, true>(a, diag_order<-1>(i, a.W, a.H), lane);
}
LUX_XITEM(x_rpath_p) { replay_paths_wave<NP, MASKED, false>(uload(&J->cp), i, lane); }
LUX_XITEM(x_rseed_p)
{
    const ReplayArgs a = uload(&J->cp);
    replay_seeds_wave<NP, MASKED, false, false>(a, a.visits, lane);
}
LUX_XITEM(x_radd_p) { replay_add_wave<NP, false>(uload(&J->cp), uload(&J->cellvis_p), i, lane); }
LUX_XITEM(x_chain_m)
{
    const ReplayArgs a = uload(&J->cm);
    replay_chain_wave<NP, MASKED, false, false>(a, a.visits, lane);
}
LUX_XITEM(x_sw_m1)
{
    const SweepArgs a = uload(&J->sm[0]);
    sweep_wave<NP, -1, 1, false, MASKED, false, false, true>(a, diag_order<-1>(i, a.W, a.H), lane);
}
LUX_XITEM(x_sw_m2) { sweep_wave<NP, -1, 2, false, MASKED, false, false, true>(uload(&J->sm[1]), i, lane); }
LUX_XITEM(x_sw_m3)
{
    const SweepArgs a = uload(&J->sm[2]);
    sweep_wave<NP, -1, 3, false, MASKED, false, false, true>(a, diag_order<1>(i, a.W, a.H), lane);
}
LUX_XITEM(x_bwd) { row_bwd_wave<NP, MASKED, false, false, true, true, 0>(uload(&J->rb), i, lane); }
LUX_XITEM(x_rpath_m) { replay_paths_wave<NP, MASKED, false>(uload(&J->cm), i, lane); }
LUX_XITEM(x_radd_m) { replay_add_wave<NP, false>(uload(&J->cm), uload(&J->cellvis_m), i, lane); }
LUX_XITEM(x_finish)
{
    wta_finish_wave<NP, false>(uload(&J->rb), __builtin_amdgcn_readfirstlane(J->tol), uload(&J->valid), i, lane);
}
#undef LUX_XITEM

// SM: the stages this kernel holds (the others' code is left out).
template <uint32_t SM, int NP, bool MASKED>
__device__ __forceinline__ void chain_item(const ExecJob* J, int stage, int i, int lane)
{
#define LUX_HAS(S) ((SM & xbit(S)) && stage == (S))
    if (LUX_HAS(X_FWD)) x_fwd<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_CHAIN_P)) x_chain_p<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_SW_P1)) x_sw_p1<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_SW_P2)) x_sw_p2<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_SW_P3)) x_sw_p3<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_RPATH_P)) x_rpath_p<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_RSEED_P)) x_rseed_p<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_RADD_P)) x_radd_p<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_CHAIN_M)) x_chain_m<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_SW_M1)) x_sw_m1<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_SW_M2)) x_sw_m2<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_SW_M3)) x_sw_m3<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_BWD)) x_bwd<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_RPATH_M)) x_rpath_m<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_RADD_M)) x_radd_m<NP, MASKED>(J, i, lane);
    else if (LUX_HAS(X_FINISH)) x_finish<NP, MASKED>(J, i, lane);
#undef LUX_HAS
}

template <uint32_t SM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SM == kAll ? 4 : 1, 8))) void sgm_exec(ExecLaunch L)
{
    const int lane = threadIdx.x;
    const int x = xcc_id();
    if (lane == 0) __hip_atomic_fetch_add(&L.sync->live[x], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        if (L.release) xsync_poll(L.sync, x, lane);
        uint32_t t = 0;
        if (lane == 0) t = __hip_atomic_fetch_add(L.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
        if (t >= L.total) break;
        int lo = 0, hi = L.nseg - 1;  // the last segment starting at or before t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((uint32_t)__builtin_amdgcn_readfirstlane((int)L.segs[mid].t0) <= t) lo = mid;
            else hi = mid - 1;
        }
        const ExecSeg sg = uload(&L.segs[lo]);
        const ExecJob* J = L.jobs + sg.job;
        const int item = sg.first + (int)(t - sg.t0);
        const uint32_t pre = exec_prereq(sg.stage) & L.wait;
        if (pre && !exec_wait(L, J, sg.job, pre, x, lane)) continue;  // gave up: the error word tells the host
        if ((SM & xbit(X_CENSUS)) && sg.stage == X_CENSUS) {
            census_item(J, item, lane);
        } else {
            switch (__builtin_amdgcn_readfirstlane(J->variant)) {
                case 0: chain_item<SM, 1, false>(J, sg.stage, item, lane); break;
                case 1: chain_item<SM, 2, false>(J, sg.stage, item, lane); break;
                case 2: chain_item<SM, 1, true>(J, sg.stage, item, lane); break;
                default: chain_item<SM, 2, true>(J, sg.stage, item, lane); break;
            }
        }
        item_done(L, J, sg.job, sg.stage, lane);
    }
    // leave with nothing of this wavefront's unwritten in the L2 (an XCD with no wavefront left counts as flushed)
    if (L.release) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) __hip_atomic_fetch_add(&L.sync->live[x], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

const void* exec_kernel(uint32_t sm)
{
    switch (sm) {
        case kAll: return (const void*)sgm_exec<kAll>;
        case kG0: return (const void*)sgm_exec<kG0>;
        case kG1: return (const void*)sgm_exec<kG1>;
        case kG2: return (const void*)sgm_exec<kG2>;
        case kG3: return (const void*)sgm_exec<kG3>;
        case kG4: return (const void*)sgm_exec<kG4>;
        case kG5: return (const void*)sgm_exec<kG5>;
        case kG6: return (const void*)sgm_exec<kG6>;
        default: return nullptr;
    }
}

hipError_t launch_exec(uint32_t sm, unsigned grid, const ExecLaunch& L, hipStream_t s)
{
    switch (sm) {
        case kAll: sgm_exec<kAll><<<grid, 64, 0, s>>>(L); break;
        case kG0: sgm_exec<kG0><<<grid, 64, 0, s>>>(L); break;
        case kG1: sgm_exec<kG1><<<grid, 64, 0, s>>>(L); break;
        case kG2: sgm_exec<kG2><<<grid, 64, 0, s>>>(L); break;
        case kG3: sgm_exec<kG3><<<grid, 64, 0, s>>>(L); break;
        case kG4: sgm_exec<kG4><<<grid, 64, 0, s>>>(L); break;
        case kG5: sgm_exec<kG5><<<grid, 64, 0, s>>>(L); break;
        case kG6: sgm_exec<kG6><<<grid, 64, 0, s>>>(L); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Resident one-wavefront workgroups of a kernel on the device (per device and kernel, cached).
int resident_waves(int device, uint32_t sm)
{
    static std::mutex mu;
    static std::map<std::pair<int, uint32_t>, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(device, sm);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, exec_kernel(sm), 64, 0) != hipSuccess) per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 0;
    const int n = std::max(1, per_cu) * std::max(1, cus);
    cache.emplace(key, n);
    return n;
}

}  // namespace

struct ExecBatch {
    int device = 0;
    std::vector<ExecJob> jobs;
    std::vector<uint32_t> stages;            // each launch's stage set
    std::vector<std::vector<ExecSeg>> segs;  // each launch's segments
    std::vector<uint32_t> totals;
    std::vector<int> resident;               // each launch's kernel: one-wavefront workgroups the device holds
    void* dev = nullptr;                     // jobs | segments | done counters | tickets | error word
    size_t seg_off = 0, done_off = 0, epoch_off = 0, sync_off = 0, ticket_off = 0, err_off = 0, bytes = 0;
};

bool exec_supported(const lux_sgm_rect_params* p, const Prepared& w)
{
    return p->cost == LUX_SGM_COST_CENSUS && w.b.pv[0] != nullptr && w.b.cl != nullptr && w.b.seed_row >= 0 &&
           w.g.dpl <= 4 && !w.b.vstart && !w.b.ap.on;
}

bool exec_params_supported(const lux_sgm_rect_params* p)
{
    if (!path_bytes(p)) return false;
    std::shared_ptr<Plan> plan;
    return plan_for(p->width, p->height, p->tile, p->xs_clamp_w, &plan, nullptr) == 0 && plan->seed_row >= 0;
}

namespace {

ExecJob make_job(const ExecPair& q)
{
    const lux_sgm_rect_params* p = q.p;
    const Prepared& w = *q.w;
    const Geometry& g = w.g;
    const SgmBuffers& b = w.b;
    ExecJob j{};
    const int np = g.dpl / 2;
    j.variant = (np == 2 ? 1 : 0) | (g.cap != 64 * g.dpl ? 2 : 0);
    j.img16 = p->bits > 8;
    j.radius = p->census_radius;
    j.stride = p->stride;
    j.W = g.W;
    j.H = g.H;
    j.tol = p->lr_tolerance;
    j.left = q.left;
    j.right = q.right;
    j.cl = const_cast<uint64_t*>(b.cl);
    j.cr = const_cast<uint64_t*>(b.cr);
    j.valid = q.valid;
    SgmOutputs o;
    o.disp = q.disp ? q.disp : b.disp_tmp;
    o.sub = q.sub;
    o.valid = q.valid;
    o.tol = p->lr_tolerance;
    o.write_agg = 0;
    SweepArgs a;
    RowArgs ra;
    ReplayArgs r;
    sgm_base_args(g, b, o, p->p1, p->p2, true, a, ra, r);
    const size_t colsz = (size_t)2 * g.H * g.cs, colminsz = (size_t)2 * g.H;
    // pass +1 (the base arguments' column slots and border replay)
    j.rf = ra;
    for (int k = 0; k < 3; k++) {
        j.sp[k] = a;
        j.sp[k].pvs = b.pvs;
        j.sp[k].pv = b.pv[1 + k];
    }
    j.cp = r;
    // pass -1
    a.cols = b.cols + colsz;
    a.colmins = b.colmins + colminsz;
    ra.cols = b.cols + colsz;
    ra.colmins = b.colmins + colminsz;
    ra.bagg_minus = b.bagg[1];  // read by the finish only
    j.rb = ra;
    for (int k = 0; k < 3; k++) {
        j.sm[k] = a;
        j.sm[k].pvs = b.pvs;
        j.sm[k].pv = b.pv[4 + k];
    }
    r.dir = -1;
    r.col = 0;
    r.visits = b.visits_minus;
    r.nvis = b.n_minus;
    r.colagg = b.bagg[1];
    r.cols = b.cols + colsz;
    r.colmins = b.colmins + colminsz;
    // pass -1's border replay buffers: the second half (prepare() gives two, one per pass)
    const size_t maxvis = (size_t)std::max(b.n_plus, b.n_minus) + 1;
    r.partial = b.partial + maxvis * g.cs;
    r.chainp = b.chainp + maxvis * g.cs;
    j.cm = r;
    j.cellvis_p = b.cellvis_plus;
    j.cellvis_m = b.cellvis_minus;
    const uint32_t H = (uint32_t)g.H, nv = (uint32_t)g.W, nd = (uint32_t)(g.W + g.H - 1);
    j.n[X_CENSUS] = 2 * ((H + kXCensusRows - 1) / kXCensusRows);
    j.n[X_FWD] = H;
    j.n[X_CHAIN_P] = 1;
    j.n[X_SW_P1] = nd;
    j.n[X_SW_P2] = nv;
    j.n[X_SW_P3] = nd;
    j.n[X_RPATH_P] = (uint32_t)b.n_plus;
    j.n[X_RSEED_P] = 1;
    j.n[X_RADD_P] = H;
    j.n[X_CHAIN_M] = 1;
    j.n[X_SW_M1] = nd;
    j.n[X_SW_M2] = nv;
    j.n[X_SW_M3] = nd;
    j.n[X_BWD] = H;
    j.n[X_RPATH_M] = (uint32_t)b.n_minus;
    j.n[X_RADD_M] = H;
    j.n[X_FINISH] = H;
    return j;
}

// Relative cost of one step of a chain of job j (lanes carry one or two packed pairs).
double step_cost(const ExecJob& j) { return (j.variant & 1) ? 1.6 : 1.0; }

}  // namespace

int exec_create(const ExecPair* pairs, int n, int groups, int order, int device, ExecBatch** out, std::string* err)
{
    if (groups != 1 && groups != 7) {
        if (err) *err = "executor groups must be 1 or 7";
        return LUX_SGM_EINVAL;
    }
    hipError_t e0 = hipSetDevice(device);
    if (e0 != hipSuccess) return hip_status(e0, "executor device", err);
    auto* bt = new ExecBatch();
    bt->device = device;
    for (int i = 0; i < n; i++) {
        if (!exec_supported(pairs[i].p, *pairs[i].w)) {
            delete bt;
            if (err) *err = "pair " + std::to_string(i) + " is not an executor request";
            return LUX_SGM_EUNSUPPORTED;
        }
        bt->jobs.push_back(make_job(pairs[i]));
    }
    if (groups == 1) bt->stages.push_back(kAll);
    else bt->stages.assign(kGroups7, kGroups7 + 7);
    // jobs of a stage: batch order, or the longest chains first
    std::vector<int> jorder(n);
    std::iota(jorder.begin(), jorder.end(), 0);
    if (order == 1)
        std::stable_sort(jorder.begin(), jorder.end(), [&](int x, int y) {
            const ExecJob &a = bt->jobs[x], &b = bt->jobs[y];
            return std::max(a.W, a.H) * step_cost(a) > std::max(b.W, b.H) * step_cost(b);
        });
    for (uint32_t sm : bt->stages) {
        std::vector<ExecSeg> segs;
        uint32_t t = 0;
        for (int s = 0; s < X_NSTAGES; s++) {
            if (!(sm & xbit(s))) continue;
            for (int k : jorder) {
                const uint32_t cnt = bt->jobs[k].n[s];
                if (!cnt) continue;
                segs.push_back(ExecSeg{t, k, s, 0});
                t += cnt;
            }
        }
        bt->segs.push_back(segs);
        bt->totals.push_back(t);
        bt->resident.push_back(resident_waves(device, sm));
    }
    // one allocation: jobs, segments, done counters, one ticket per launch, the error word
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    take(sizeof(ExecJob) * (size_t)n);
    size_t nsegs = 0;
    for (auto& v : bt->segs) nsegs += v.size();
    bt->seg_off = take(sizeof(ExecSeg) * std::max<size_t>(nsegs, 1));
    bt->done_off = take(sizeof(uint32_t) * (size_t)n * X_NSTAGES);
    bt->epoch_off = take(sizeof(uint32_t) * (size_t)n * X_NSTAGES);
    bt->sync_off = take(sizeof(XSync) * bt->stages.size());
    bt->ticket_off = take(sizeof(uint32_t) * bt->stages.size());
    bt->err_off = take(sizeof(int));
    bt->bytes = off;
    hipError_t e = hipMalloc(&bt->dev, bt->bytes);
    std::vector<uint8_t> host(bt->bytes, 0);
    std::memcpy(host.data(), bt->jobs.data(), sizeof(ExecJob) * (size_t)n);
    size_t so = bt->seg_off;
    for (auto& v : bt->segs) {
        std::memcpy(host.data() + so, v.data(), sizeof(ExecSeg) * v.size());
        so += sizeof(ExecSeg) * v.size();
    }
    if (e == hipSuccess) e = hipMemcpy(bt->dev, host.data(), bt->bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (bt->dev) (void)hipFree(bt->dev);
        delete bt;
        return hip_status(e, "executor upload", err);
    }
    *out = bt;
    return 0;
}

void exec_destroy(ExecBatch* b)
{
    if (!b) return;
    if (b->dev) (void)hipFree(b->dev);
    delete b;
}

int exec_launches(const ExecBatch* b) { return b ? (int)b->stages.size() : 0; }

hipError_t exec_enqueue(ExecBatch* b, int waves, hipStream_t s)
{
    uint8_t* base = (uint8_t*)b->dev;
    const size_t nclear = (b->bytes - b->done_off) / 4;  // done counters, tickets and the error word
    hipError_t e = launch_fill_u32((uint32_t*)(base + b->done_off), nclear, 0u, s);
    size_t so = b->seg_off;
    for (size_t k = 0; e == hipSuccess && k < b->stages.size(); k++) {
        ExecLaunch L;
        L.jobs = (const ExecJob*)base;
        L.segs = (const ExecSeg*)(base + so);
        L.ticket = (uint32_t*)(base + b->ticket_off) + k;
        L.done = (uint32_t*)(base + b->done_off);
        L.epoch = (uint32_t*)(base + b->epoch_off);
        L.sync = (XSync*)(base + b->sync_off) + k;
        L.err = (int*)(base + b->err_off);
        L.total = b->totals[k];
        L.nseg = (int)b->segs[k].size();
        L.wait = b->stages[k];
        L.release = 0;
        for (int st = 0; st < X_NSTAGES; st++)  // a stage another stage of this launch reads
            for (int u = 0; u < X_NSTAGES; u++)
                if ((b->stages[k] & xbit(st)) && (b->stages[k] & xbit(u)) && (exec_prereq(u) & xbit(st))) L.release |= xbit(st);
        so += sizeof(ExecSeg) * b->segs[k].size();
        if (!L.total) continue;
        const int res = waves > 0 ? waves : b->resident[k];
        const unsigned grid = (unsigned)std::min<uint32_t>(L.total, (uint32_t)res);
        e = launch_exec(b->stages[k], grid, L, s);
    }
    return e;
}

int exec_check(ExecBatch* b, hipStream_t s, std::string* err)
{
    int v = 0;
    hipError_t e = hipMemcpyAsync(&v, (uint8_t*)b->dev + b->err_off, sizeof v, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e, "executor error word", err);
    if (v) {
        if (err) *err = "executor: a stage wait gave up";
        return LUX_SGM_EDEVICE;
    }
    return 0;
}

}  // namespace lux
