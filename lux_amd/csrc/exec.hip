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
#include "exec_impl.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>

namespace lux {

namespace {

// Tickets [t0, next segment's t0) of a launch run items first, first + 1, ... of one stage of one job.
// Every stage of a batch's pairs (their images and outputs are device memory: no staging).
constexpr uint32_t kBatch = kAll & ~(xbit(X_IN) | xbit(X_OUT));

struct ExecSeg {
    uint32_t t0;
    int job, stage, first;
};

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

template <uint32_t SM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SM == kBatch ? 4 : 1, 8))) void sgm_exec(ExecLaunch L)
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
        run_item<SM>(J, sg.stage, item, lane);
        item_done(L, J, sg.job, sg.stage, lane);
    }
    // leave with nothing of this wavefront's unwritten in the L2 (an XCD with no wavefront left counts as flushed)
    if (L.release) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) __hip_atomic_fetch_add(&L.sync->live[x], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

const void* exec_kernel(uint32_t sm)
{
    switch (sm) {
        case kBatch: return (const void*)sgm_exec<kBatch>;
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
        case kBatch: sgm_exec<kBatch><<<grid, 64, 0, s>>>(L); break;
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
    if (groups == 1) bt->stages.push_back(kBatch);
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
