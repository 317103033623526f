// The device-resident request loop (declared in dloop.h): the served form with its scheduler on the device.
//
// One persistent kernel holds the device. Its first wavefront is the scheduler, the first of every other XCD's
// wavefronts keeps that XCD's cache write-backs current (the flush epochs below), the others are workers.
//  - The host writes each request into a ring in coherent pinned memory: the job descriptor of the executor
//    (exec_impl.h) with every pointer into the request's workspace given as an offset, a bit per 64-bit word that
//    says which words those are, and the workspace size. Host images are staged in pinned memory beside it.
//  - The scheduler admits ring entries in arrival order while a slot is free and the request's workspace fits in
//    the device pool (one allocation of the HBM budget, first fit): it copies the descriptor into the slot,
//    relocates it to the block it chose, publishes the slot, and marks the entry admitted. When the last stage of
//    a slot's request has completed and is flushed it marks the entry done and frees the block.
//  - Workers pull items from the live slots, smallest request first: the next item of a slot runs once the stages
//    it reads are complete and flushed from every XCD (the flush epochs of exec_impl.h), claimed by one atomic add
//    on the stage's (generation, items claimed) word, so every item runs exactly once and none waits.
//    Requests therefore overlap on the device at stage granularity (continuous batching), and each item is the
//    same device function the plain request runs, with the same results.
//  - A block of the pool a finished request held may be handed to a new one: a worker invalidates its caches
//    before its first item of each admission, so no stale line of the old request is read.
//  - With nothing to do for a while, or when the host asks it to stop, the kernel leaves; the host launches it
//    again with the next request (a flag in pinned memory, checked on both sides, means no request is left behind).
// Requests whose path costs do not fit a byte, and coarse-to-fine requests, are the host dispatcher's (server.hip).
#include "dloop.h"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "engine.h"
#include "exec_impl.h"
#include "pinned.h"

namespace lux {

namespace {

int fail(int code, const std::string& msg, std::string* err)
{
    if (err) *err = msg;
    return code;
}

constexpr int kDSlots = 16;
constexpr int kRing = 1024;
constexpr int kJobWords = (int)(sizeof(ExecJob) / 8);
constexpr int kRelocMasks = (kJobWords + 31) / 32;
static_assert(sizeof(ExecJob) % 8 == 0, "the descriptor is relocated 64-bit word by word");
constexpr uint32_t kDLoopStages = kAll;
constexpr uint64_t kPoolAlign = 4096;
// Workers that always look for work; the others look only while a live request has stages at least half as wide as
// their index (claims beyond a stage's items only contend on its counter, and every idle look reads the slot tables).
constexpr uint32_t kAwake = 1024;

// Items one claim takes: the short items (staging chunks, border visits and sums, finish rows) in eights, the chains
// (row paths, sweeps, border chains) and census row blocks one at a time.
__host__ __device__ constexpr uint32_t stage_batch(int s)
{
    return (s == X_IN || s == X_RPATH_P || s == X_RADD_P || s == X_RPATH_M || s == X_RADD_M || s == X_FINISH ||
            s == X_OUT)
               ? 8u
               : 1u;
}

enum : uint32_t { R_FREE = 0, R_SUBMITTED = 1, R_ADMITTED = 2, R_DONE = 3 };

// One ring entry (coherent pinned host memory). The host writes everything but the last four fields, then state.
struct DReq {
    ExecJob job;                   // workspace pointers as offsets from the request's pool block
    uint32_t reloc[kRelocMasks];   // bit w: 64-bit word w of job is such an offset
    uint64_t ws_bytes;
    uint32_t prio;                 // smaller first (cells)
    uint32_t state;                // R_*
    uint32_t slot;                 // the device writes these
    uint32_t pad_;
    uint64_t t_admit, t_done;      // device wall clock
    uint64_t t_stage[X_NSTAGES];   // when each stage completed (device wall clock; 0: no items)
};

struct DSlot {                     // device memory; written by the scheduler, published by `live`
    uint64_t cnt[X_NSTAGES];       // per stage: generation << 32 | items claimed (workers add one to claim one)
    uint64_t ready;                // generation << 32 | stages whose inputs are complete and flushed
    uint32_t gen;                  // admissions of this slot
    uint32_t live, total, admit_epoch, ring, prio;
    uint32_t width;                // claims the widest stage takes (items over the stage's batch)
    uint32_t start[X_NSTAGES + 1]; // first item of each stage
    uint32_t done[X_NSTAGES];
    uint32_t epoch[X_NSTAGES];     // flush epoch raised when the stage completed
    uint64_t base, bytes;          // pool block
    ExecJob job;                   // relocated
};

struct DState {                    // device memory, kept across launches of the loop
    XSync sync;
    uint32_t quit;                 // the scheduler tells the workers to leave
    uint32_t flusher[kXcds];       // an XCD's flusher is taken (cleared with quit before each launch)
    uint32_t inv;                  // admissions so far (the invalidate epochs)
    uint32_t admitted;             // ring entries admitted so far
    uint32_t nlive;
    uint32_t progress;             // items completed (the scheduler's watchdog)
    uint32_t work;                 // bumped when work may have become ready (admission, stage completed or flushed)
    uint32_t fault;                // the watchdog saw no item complete for kStallTicks with requests running
    uint32_t order[kDSlots];       // live slots, smallest request first
    uint32_t span;                 // the widest live request's width: workers at or past kAwake and 2 * span park
    uint64_t used, peak;           // pool bytes
    DSlot slot[kDSlots];
};

struct DHost {                     // coherent pinned host memory
    uint32_t stop;                 // host: leave once nothing runs
    uint32_t alive;                // 1 while a loop kernel owns the ring (the kernel clears it when it leaves)
};

struct DArgs {
    DState* st;
    DHost* host;
    DReq* ring;
    uint8_t* pool;
    uint64_t pool_bytes;
    int nslots;
    uint64_t idle_ticks;           // the loop leaves after this long with nothing to do (device wall clock)
    uint64_t stall_ticks;          // ... and gives up after this long with requests running and no item completing
};

__device__ __forceinline__ uint32_t sys_ld(const uint32_t* p)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane(
        (int)__hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
}
__device__ __forceinline__ void sys_st(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint64_t ald64(const uint64_t* p)
{
    const uint64_t v = __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
}
__device__ __forceinline__ void ast(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Rewrites the priority order of the live slots (scheduler): repeatedly the next (prio, slot) above the last one.
__device__ __forceinline__ void dl_order(const DArgs& A, int lane)
{
    DState* S = A.st;
    uint64_t last = 0;  // (prio << 8 | slot) + 1 of the previous entry
    int n = 0;
    uint32_t span = 0;
    for (;; n++) {
        uint64_t best = ~0ull;
        for (int k = 0; k < A.nslots; k++) {
            if (!ald(&S->slot[k].live)) continue;
            const uint64_t key = ((uint64_t)ald(&S->slot[k].prio) << 8 | (uint64_t)k) + 1;
            if (key > last && key < best) best = key;
        }
        if (best == ~0ull) break;
        if (lane == 0) ast(&S->order[n], (uint32_t)((best - 1) & 0xff));
        span = max(span, ald(&S->slot[(best - 1) & 0xff].width));
        last = best;
    }
    __builtin_amdgcn_s_waitcnt(0);
    if (lane == 0) {
        ast(&S->span, span);
        ast(&S->nlive, (uint32_t)n);
    }
}

// First fit of `bytes` in the pool among the live slots' blocks; ~0 if none.
__device__ __forceinline__ uint64_t dl_fit(const DArgs& A, uint64_t bytes)
{
    DState* S = A.st;
    uint64_t best = ~0ull;
    for (int c = -1; c < A.nslots; c++) {
        uint64_t at = 0;
        if (c >= 0) {
            if (!ald(&S->slot[c].live)) continue;
            at = ald64(&S->slot[c].base) + ald64(&S->slot[c].bytes);
            at = (at + kPoolAlign - 1) & ~(kPoolAlign - 1);
        }
        if (at + bytes > A.pool_bytes || at >= best) continue;
        bool ok = true;
        for (int k = 0; k < A.nslots && ok; k++) {
            if (!ald(&S->slot[k].live)) continue;
            const uint64_t b0 = ald64(&S->slot[k].base), b1 = b0 + ald64(&S->slot[k].bytes);
            ok = at + bytes <= b0 || at >= b1;
        }
        if (ok) best = at;
    }
    return best;
}

__device__ __forceinline__ void dl_scheduler(const DArgs& A, int x, int lane)
{
    DState* S = A.st;
    uint32_t seq = ald(&S->admitted);
    uint64_t idle_from = wall_clock64(), moved_at = idle_from;
    uint32_t moved = ald(&S->progress);
    for (;;) {
        xsync_poll(&S->sync, x, lane);
        bool busy = false, changed = false;
        // completions: the last stage with items complete and flushed from every XCD
        for (int k = 0; k < A.nslots; k++) {
            DSlot* D = &S->slot[k];
            if (!ald(&D->live)) continue;
            busy = true;
            const int last = ald(&D->start[X_OUT + 1]) > ald(&D->start[X_OUT]) ? X_OUT : X_FINISH;
            if (ald(&D->done[last]) < ald(&D->start[last + 1]) - ald(&D->start[last])) continue;
            const uint32_t e = ald(&D->epoch[last]);
            if (!e || !xsync_flushed(&S->sync, e)) continue;
            DReq* R = A.ring + ald(&D->ring);
            if (lane == 0) {
                __hip_atomic_store(&R->t_done, wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
                sys_st(&R->state, R_DONE);
                __hip_atomic_fetch_add(&S->used, (uint64_t)0 - ald64(&D->bytes), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                ast(&D->live, 0u);
            }
            changed = true;
        }
        if (changed) dl_order(A, lane);
        // admission, in arrival order (the head waits for a slot and for pool room); none once stopping
        while (!sys_ld(&A.host->stop)) {
            int k = -1;
            for (int i = 0; i < A.nslots && k < 0; i++)
                if (!ald(&S->slot[i].live)) k = i;
            if (k < 0) break;
            DReq* R = A.ring + (seq % kRing);
            if (sys_ld(&R->state) != R_SUBMITTED) break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // the host's words of this entry
            const uint64_t bytes = __hip_atomic_load(&R->ws_bytes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            const uint64_t base = dl_fit(A, bytes);
            if (base == ~0ull) break;
            DSlot* D = &S->slot[k];
            // the descriptor, relocated to the block (64-bit words, one a lane)
            const uint64_t add = (uint64_t)(A.pool + base);
            const uint64_t* src = reinterpret_cast<const uint64_t*>(&R->job);
            uint64_t* dst = reinterpret_cast<uint64_t*>(&D->job);
            for (int w = lane; w < kJobWords; w += 64) {
                uint64_t v = src[w];
                if ((R->reloc[w >> 5] >> (w & 31)) & 1u) v += add;
                dst[w] = v;
            }
            uint32_t total = 0, width = 0;
            const uint32_t gen = ald(&D->gen) + 1;
            if (lane == 0) {
                for (int s = 0; s < X_NSTAGES; s++) {
                    D->start[s] = total;
                    total += R->job.n[s];
                    width = max(width, (R->job.n[s] + stage_batch(s) - 1) / stage_batch(s));
                    D->done[s] = 0;
                    D->epoch[s] = 0;
                    D->cnt[s] = (uint64_t)gen << 32;
                }
                D->gen = gen;
                D->ready = (uint64_t)gen << 32;
                D->start[X_NSTAGES] = total;
                D->prio = R->prio;
                D->width = width;
                D->ring = seq % kRing;
                D->base = base;
                D->bytes = bytes;
                D->admit_epoch = ++S->inv;
            }
            total = (uint32_t)__builtin_amdgcn_readfirstlane((int)total);
            // publish: descriptor and counters to memory first, then the slot (each store complete before the next)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (lane == 0) {
                ast(&D->total, total);
                __builtin_amdgcn_s_waitcnt(0);
                ast(&D->live, 1u);
                const uint64_t u = __hip_atomic_fetch_add(&S->used, bytes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + bytes;
                __hip_atomic_fetch_max(&S->peak, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&R->t_admit, wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                sys_st(&R->slot, (uint32_t)k);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
                sys_st(&R->state, R_ADMITTED);
            }
            seq++;
            if (lane == 0) ast(&S->admitted, seq);
            dl_order(A, lane);
            if (lane == 0) __hip_atomic_fetch_add(&S->work, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            busy = true;
        }
        const uint64_t now = wall_clock64();
        if (busy) {
            idle_from = now;
            const uint32_t pr = ald(&S->progress);
            if (pr != moved) {
                moved = pr;
                moved_at = now;
            } else if (now - moved_at > A.stall_ticks) {  // nothing completes: give up (the host reports it)
                if (lane == 0) ast(&S->fault, 1u);
                break;
            }
        } else {
            if (sys_ld(&A.host->stop)) break;
            if (now - idle_from > A.idle_ticks) {
                // leave unless a request came in meanwhile: clear `alive`, then look once more (the host writes its
                // entry, then looks at `alive`, so one of the two sees the other)
                if (lane == 0) sys_st(&A.host->alive, 0u);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "");
                if (sys_ld(&A.ring[seq % kRing].state) != R_SUBMITTED) break;
                if (lane == 0) sys_st(&A.host->alive, 1u);
                idle_from = now;
            }
        }
        __builtin_amdgcn_s_sleep(127);
    }
    if (lane == 0) ast(&S->quit, 1u);
}

// Whether stage s of slot D may start: every stage it reads complete and flushed (a stage with no items stands for
// its own prerequisites). st: lane l holds start[l].
__device__ __forceinline__ bool dl_ready(const DState* S, const DSlot* D, int s, uint32_t st)
{
    uint32_t pre = exec_prereq(s), seen = 0;
    while (pre) {
        const int p = __builtin_ctz(pre);
        pre &= pre - 1;
        if (seen & xbit(p)) continue;
        seen |= xbit(p);
        const uint32_t first = (uint32_t)__builtin_amdgcn_readlane((int)st, p);
        const uint32_t n = (uint32_t)__builtin_amdgcn_readlane((int)st, p + 1) - first;
        if (!n) {
            pre |= exec_prereq(p) & ~seen;
            continue;
        }
        if (ald(&D->done[p]) < n) return false;
        const uint32_t e = ald(&D->epoch[p]);
        if (!e || !xsync_flushed(&S->sync, e)) return false;
    }
    return true;
}

// A claim [t, t + batch) on stage s of slot D that landed on an admission other than the one the claimer looked at:
// the items are that request's, so they run once its stage s is ready (or not at all past the stage's end).
// st and n receive the request's stage starts and sizes (lane l: stage l).
__device__ __forceinline__ bool dl_adopt(DState* S, DSlot* D, int s, uint32_t t, uint32_t& st, uint32_t& n, int lane)
{
    st = lane <= X_NSTAGES ? __hip_atomic_load(&D->start[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0xffffffffu;
    n = (uint32_t)__shfl_down((int)st, 1) - st;
    if (t >= (uint32_t)__builtin_amdgcn_readlane((int)n, s)) return false;
    while (!dl_ready(S, D, s, st)) {
        if (ald(&S->quit)) return false;
        __builtin_amdgcn_s_sleep(127);
    }
    return true;
}

__device__ __forceinline__ uint64_t add64(uint64_t* p, uint64_t v, int lane)
{
    uint64_t r = 0;
    if (lane == 0) r = __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(r >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)r);
}

// Finds and claims the next ready items of the live slots, smallest request first; false if none is ready. A stage
// is taken when every stage it reads is complete and flushed; its items are claimed by atomic adds on the stage's
// counter, whose high word is the slot's generation (a claim that lands on a newer admission of the slot belongs to
// that request: dl_adopt).
__device__ __forceinline__ bool dl_claim(const DArgs& A, int lane, int* slot, int* stage, uint32_t* item, uint32_t* count,
                                         uint32_t* size, uint32_t* gen_out)
{
    DState* S = A.st;
    const uint32_t nl = min(ald(&S->nlive), (uint32_t)kDSlots);
    for (uint32_t q = 0; q < nl; q++) {
        const int k = (int)(ald(&S->order[q]) & (kDSlots - 1));
        DSlot* D = &S->slot[k];
        if (!ald(&D->live)) continue;
        const uint32_t gen = ald(&D->gen);
        // lane l: stage l's first item, its items and the items claimed so far
        uint32_t st = lane <= X_NSTAGES ? __hip_atomic_load(&D->start[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                        : 0xffffffffu;
        uint32_t n = (uint32_t)__shfl_down((int)st, 1) - st;
        uint64_t c = lane < X_NSTAGES ? __hip_atomic_load(&D->cnt[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                      : ~0ull;
        uint64_t left = __builtin_amdgcn_ballot_w64(lane < X_NSTAGES && (uint32_t)(c >> 32) == gen && (uint32_t)c < n);
        const uint64_t rw = ald64(&D->ready);
        const uint32_t rdy = (uint32_t)(rw >> 32) == gen ? (uint32_t)rw : 0u;
        while (left) {
            const int s = __builtin_ctzll(left);
            left &= left - 1;
            if (!(rdy & xbit(s))) {
                if (!dl_ready(S, D, s, st)) continue;
                // remember it for this admission only: a worker that looked at an earlier request of the slot must
                // not mark the new one's stage
                if (lane == 0) {
                    uint64_t cur = rw;
                    while ((uint32_t)(cur >> 32) == gen &&
                           !__hip_atomic_compare_exchange_strong(&D->ready, &cur, cur | xbit(s), __ATOMIC_RELAXED,
                                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    }
                }
            }
            const uint32_t batch = stage_batch(s);
            const uint64_t v = add64(&D->cnt[s], batch, lane);
            const uint32_t g = (uint32_t)(v >> 32), t = (uint32_t)v;
            if (g != gen) {
                if (!dl_adopt(S, D, s, t, st, n, lane)) return false;
            } else if (t >= (uint32_t)__builtin_amdgcn_readlane((int)n, s)) {
                continue;  // the stage ran out meanwhile
            }
            const uint32_t ns = (uint32_t)__builtin_amdgcn_readlane((int)n, s);
            *slot = k;
            *stage = s;
            *item = t;
            *count = min(batch, ns - t);
            *size = ns;
            *gen_out = g;
            return true;
        }
    }
    return false;
}

// This XCD's share of the flush protocol; an acknowledgement may make a stage ready, so it bumps the work counter.
__device__ __forceinline__ bool dl_flush(DState* S, int x, int lane)
{
    if (!xsync_poll(&S->sync, x, lane)) return false;
    if (lane == 0) __hip_atomic_fetch_add(&S->work, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// One wavefront per XCD does nothing but its XCD's write-backs, so a completed stage is flushed within microseconds
// even while every other wavefront of the XCD is inside a long chain.
__device__ __forceinline__ void dl_flusher(const DArgs& A, int x, int lane)
{
    DState* S = A.st;
    while (!ald(&S->quit)) {
        dl_flush(S, x, lane);
        __builtin_amdgcn_s_sleep(32);
    }
}

__device__ __forceinline__ void dl_worker(const DArgs& A, int x, int lane, uint32_t w)
{
    DState* S = A.st;
    uint32_t my_inv = 0, backoff = 1;
    int fk = -1, fs = 0;     // the stage of the last claim: tried first (one atomic add while it lasts)
    uint32_t fgen = 0, fn = 0;
    for (;;) {
        dl_flush(S, x, lane);
        if (ald(&S->quit)) break;
        int k = 0, s = 0;
        uint32_t item = 0, cnt = 0, n = 0, gen = 0;
        bool got = false;
        if (fk >= 0) {
            DSlot* D = &S->slot[fk];
            const uint32_t batch = stage_batch(fs);
            const uint64_t v = add64(&D->cnt[fs], batch, lane);
            const uint32_t g = (uint32_t)(v >> 32), t = (uint32_t)v;
            uint32_t st = 0, nn = 0;
            if (g == fgen && t < fn) {
                got = true;
                n = fn;
            } else if (g != fgen && dl_adopt(S, D, fs, t, st, nn, lane)) {
                got = true;
                n = (uint32_t)__builtin_amdgcn_readlane((int)nn, fs);
            }
            if (got) {
                k = fk;
                s = fs;
                item = t;
                cnt = min(batch, n - t);
                gen = g;
            } else {
                fk = -1;
            }
        }
        if (!got && w >= kAwake && w >= 2 * ald(&S->span)) {
            // parked: a look about every 80 us, with no read of the work counter or the slot tables
            for (int i = 0; i < 16; i++) __builtin_amdgcn_s_sleep(127);
            continue;
        }
        const uint32_t seen = ald(&S->work);
        if (!got) got = dl_claim(A, lane, &k, &s, &item, &cnt, &n, &gen);
        if (!got) {
            // nothing ready: sleep until the work counter moves (checked every 5 us, and a look anyway after a while
            // that doubles up to about 170 us), so that idle wavefronts read one word instead of the slot tables
            for (uint32_t w = 0; w < backoff; w++) {
                __builtin_amdgcn_s_sleep(127);
                if (ald(&S->work) != seen || ald(&S->quit)) break;
            }
            backoff = min(backoff * 2, 32u);
            continue;
        }
        backoff = 1;
        k = __builtin_amdgcn_readfirstlane(k);
        s = __builtin_amdgcn_readfirstlane(s);
        fk = k;
        fs = s;
        fgen = gen;
        fn = n;
        DSlot* D = &S->slot[k];
        const uint32_t ae = ald(&D->admit_epoch);
        if (ae > my_inv) {  // no stale line of an earlier request in this block
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            my_inv = ae;
        }
        for (uint32_t j = 0; j < cnt; j++) run_item<kDLoopStages>(&D->job, s, (int)(item + j), lane);
        __builtin_amdgcn_s_waitcnt(0);
        if (lane == 0) {
            __hip_atomic_fetch_add(&S->progress, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t o = __hip_atomic_fetch_add(&D->done[s], cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (o + cnt == n) {
                const uint32_t e = __hip_atomic_fetch_add(&S->sync.req, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
                ast(&D->epoch[s], e);
                __hip_atomic_store(&A.ring[ald(&D->ring)].t_stage[s], wall_clock64(), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_fetch_add(&S->work, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void sgm_dloop(DArgs A)
{
    const int lane = threadIdx.x, x = xcc_id();
    if (lane == 0) __hip_atomic_fetch_add(&A.st->sync.live[x], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t won = 0;
    if (blockIdx.x != 0 && lane == 0) {
        uint32_t zero = 0;
        won = __hip_atomic_compare_exchange_strong(&A.st->flusher[x], &zero, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
    }
    if (blockIdx.x == 0) dl_scheduler(A, x, lane);
    else if (__builtin_amdgcn_readfirstlane((int)won)) dl_flusher(A, x, lane);
    else dl_worker(A, x, lane, blockIdx.x);
    // leave with nothing unwritten in the L2 (an XCD with no wavefront left counts as flushed)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) __hip_atomic_fetch_add(&A.st->sync.live[x], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

using Clock = std::chrono::steady_clock;

struct DTicket {
    uint32_t idx = 0;
    int mem = 0;
    uint16_t* disp = nullptr;
    float* sub = nullptr;
    uint8_t* valid = nullptr;
    void* hin = nullptr;
    size_t hin_n = 0;
    void* hout = nullptr;
    size_t hout_n = 0;
    size_t px = 0;
    Clock::time_point t_submit;
};

}  // namespace

struct DLoop {
    int device = 0;
    uint64_t budget = 0;
    int nslots = 0;
    hipStream_t stream = nullptr, copy = nullptr;
    uint8_t* pool = nullptr;
    DState* st = nullptr;
    DHost* host = nullptr;
    DReq* ring = nullptr;
    int grid = 0;
    double clock_khz = 100000.0;
    std::mutex mu;
    uint64_t seq = 0;
    std::map<uint64_t, DTicket> tickets;
    std::map<std::tuple<int, int, int, int>, int*> plans;  // the visit lists of each schedule, on the device
    PinnedPool pinned;
    uint64_t submitted = 0, completed = 0, failed = 0;
};

namespace {

hipError_t dl_launch(DLoop* d)
{
    // quit is cleared after the previous loop kernel left (same stream); the kernel itself raises it when it leaves
    hipError_t e = hipMemsetAsync(&d->st->quit, 0, sizeof(uint32_t) * (1 + kXcds), d->stream);  // quit, flushers
    __atomic_store_n(&d->host->alive, 1u, __ATOMIC_SEQ_CST);
    DArgs A;
    A.st = d->st;
    A.host = d->host;
    A.ring = d->ring;
    A.pool = d->pool;
    A.pool_bytes = d->budget;
    A.nslots = d->nslots;
    A.idle_ticks = (uint64_t)(d->clock_khz * 20.0);      // 20 ms
    A.stall_ticks = (uint64_t)(d->clock_khz * 10000.0);  // 10 s
    if (e == hipSuccess) {
        sgm_dloop<<<d->grid, 64, 0, d->stream>>>(A);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace

int dloop_create(int device, uint64_t budget, int slots, DLoop** out, std::string* err)
{
    if (!budget) return fail(LUX_SGM_EINVAL, "the device-resident loop needs an HBM budget (its pool)", err);
    if (slots < 1 || slots > kDSlots) return fail(LUX_SGM_EINVAL, "slots must be 1..16", err);
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_status(e, "hipSetDevice", err);
    auto* d = new DLoop();
    d->device = device;
    d->budget = budget;
    d->nslots = slots;
    d->pinned.flags = hipHostMallocCoherent;
    int per_cu = 0, cus = 0, khz = 0;
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&d->copy, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&d->pool, budget);
    if (e == hipSuccess) e = hipMalloc(&d->st, sizeof(DState));
    if (e == hipSuccess) e = hipMemset(d->st, 0, sizeof(DState));
    if (e == hipSuccess) e = hipHostMalloc(&d->host, sizeof(DHost), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostMalloc(&d->ring, sizeof(DReq) * kRing, hipHostMallocCoherent);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)sgm_dloop, 64, 0);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess && hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        d->clock_khz = khz;
    if (e != hipSuccess) {
        dloop_destroy(d);
        return hip_status(e, "device loop set-up", err);
    }
    std::memset(d->host, 0, sizeof(DHost));
    std::memset((void*)d->ring, 0, sizeof(DReq) * kRing);
    d->grid = std::max(2, std::max(1, per_cu) * std::max(1, cus));
    *out = d;
    return 0;
}

void dloop_destroy(DLoop* d)
{
    if (!d) return;
    (void)hipSetDevice(d->device);
    if (d->host) __atomic_store_n(&d->host->stop, 1u, __ATOMIC_SEQ_CST);
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    for (auto& kv : d->tickets) {
        d->pinned.put(kv.second.hin, kv.second.hin_n);
        d->pinned.put(kv.second.hout, kv.second.hout_n);
    }
    d->pinned.clear();
    for (auto& kv : d->plans) (void)hipFree(kv.second);
    if (d->ring) (void)hipHostFree(d->ring);
    if (d->host) (void)hipHostFree(d->host);
    if (d->st) (void)hipFree(d->st);
    if (d->pool) (void)hipFree(d->pool);
    if (d->copy) (void)hipStreamDestroy(d->copy);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

namespace {

size_t img_bytes(const lux_sgm_rect_params& p) { return (size_t)p.height * p.stride * (p.bits > 8 ? 2 : 1); }

// The descriptor of request p with its workspace placed at `base` (no allocation: an arena at a made-up address).
int dl_job(const lux_sgm_rect_params* p, const Plan& plan, const int* visits, int mem, const void* left,
           const void* right, uint16_t* disp, float* sub, uint8_t* valid, const DTicket& tk, uint64_t base, ExecJob* out,
           uint64_t* bytes, std::string* err)
{
    Workspace ws;
    ws.arena = (uint8_t*)base;
    ws.arena_cap = 1ull << 44;
    const Geometry g = geometry(p->width, p->height, p->disparities, p->align);
    const size_t imgb = img_bytes(*p), px = (size_t)p->width * p->height;
    Prepared w;
    std::map<std::string, void*> named;
    int rc = prepare_layout(ws, g, plan, true, true, mem == LUX_SGM_MEM_HOST ? imgb : 0, visits, &w, &named, err);
    if (rc) return rc;
    ExecPair q;
    q.p = p;
    q.w = &w;
    if (mem == LUX_SGM_MEM_HOST) {
        q.left = named["img_l"];
        q.right = named["img_r"];
        q.disp = disp ? (uint16_t*)named["out_disp"] : nullptr;
        q.sub = sub ? (float*)named["out_sub"] : nullptr;
        q.valid = valid ? (uint8_t*)named["out_valid"] : nullptr;
    } else {
        q.left = left;
        q.right = right;
        q.disp = disp;
        q.sub = sub;
        q.valid = valid;
    }
    ExecJob j = make_job(q);
    if (mem == LUX_SGM_MEM_HOST) {
        j.hin = (const uint8_t*)tk.hin;
        j.hout = (uint8_t*)tk.hout;
        j.din[0] = (uint8_t*)named["img_l"];
        j.din[1] = (uint8_t*)named["img_r"];
        j.dout[0] = (const uint8_t*)q.disp;
        j.dout[1] = (const uint8_t*)q.sub;
        j.dout[2] = (const uint8_t*)q.valid;
        j.imgb = imgb;
        j.px = px;
        auto chunks = [](uint64_t b) { return (uint32_t)((b + kXChunk - 1) / kXChunk); };
        j.n[X_IN] = 2 * chunks(imgb);
        j.n[X_OUT] = chunks(2 * px) + chunks(4 * px) + chunks(px);
    }
    *out = j;
    *bytes = (ws.arena_used + kPoolAlign - 1) & ~(kPoolAlign - 1);
    return 0;
}

}  // namespace

int dloop_submit(DLoop* d, const lux_sgm_rect_params* p, int mem, const void* left, const void* right, uint16_t* disp,
                 float* sub, uint8_t* valid, uint64_t* ticket, std::string* err)
{
    int rc = check_rect(p, err);
    if (rc) return rc;
    if (mem != LUX_SGM_MEM_HOST && mem != LUX_SGM_MEM_DEVICE) return fail(LUX_SGM_EINVAL, "unknown memory kind", err);
    if (!left || !right) return fail(LUX_SGM_EINVAL, "null image", err);
    if (!exec_params_supported(p))
        return fail(LUX_SGM_EUNSUPPORTED, "the device-resident loop takes census requests whose path costs fit a byte",
                    err);
    std::shared_ptr<Plan> plan;
    if ((rc = plan_for(p->width, p->height, p->tile, p->xs_clamp_w, &plan, err))) return rc;
    std::unique_lock<std::mutex> lk(d->mu);
    (void)hipSetDevice(d->device);
    // the next ring entry frees when its earlier request is collected (wait): block for it a while
    const auto t_full = Clock::now();
    for (;;) {
        const uint32_t idx = (uint32_t)(d->seq % kRing);
        if (__atomic_load_n(&d->ring[idx].state, __ATOMIC_ACQUIRE) == R_FREE) break;
        if (std::chrono::duration<double>(Clock::now() - t_full).count() > 60.0)
            return fail(LUX_SGM_ENOMEM, "request ring full for 60 s: collect earlier requests first", err);
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        lk.lock();
    }
    const uint32_t idx = (uint32_t)(d->seq % kRing);
    // the schedule's visit lists, on the device once per shape
    const auto key = std::make_tuple(p->width, p->height, p->tile, p->xs_clamp_w);
    auto it = d->plans.find(key);
    if (it == d->plans.end()) {
        int* vis = nullptr;
        hipError_t e = hipMalloc(&vis, plan_ints(*plan) * sizeof(int) + 16);
        if (e != hipSuccess) return hip_status(e, "plan upload", err);
        rc = upload_plan(*plan, vis, d->copy, err);
        if (!rc && (e = hipStreamSynchronize(d->copy)) != hipSuccess) rc = hip_status(e, "plan upload", err);
        if (rc) {
            (void)hipFree(vis);
            return rc;
        }
        it = d->plans.emplace(key, vis).first;
    }
    DTicket tk;
    tk.idx = idx;
    tk.mem = mem;
    tk.disp = disp;
    tk.sub = sub;
    tk.valid = valid;
    tk.px = (size_t)p->width * p->height;
    const size_t imgb = img_bytes(*p);
    if (mem == LUX_SGM_MEM_HOST) {
        tk.hin = d->pinned.get(2 * imgb, &tk.hin_n);
        tk.hout = d->pinned.get(7 * tk.px, &tk.hout_n);
        if (!tk.hin || !tk.hout) {
            d->pinned.put(tk.hin, tk.hin_n);
            d->pinned.put(tk.hout, tk.hout_n);
            return fail(LUX_SGM_ENOMEM, "pinned staging", err);
        }
        std::memcpy(tk.hin, left, imgb);
        std::memcpy((char*)tk.hin + imgb, right, imgb);
    }
    // the descriptor at two made-up bases: the words that differ by the bases' difference point into the workspace
    constexpr uint64_t kB0 = 1ull << 44, kB1 = 2ull << 44;
    ExecJob j0, j1;
    uint64_t n0 = 0, n1 = 0;
    if ((rc = dl_job(p, *plan, it->second, mem, left, right, disp, sub, valid, tk, kB0, &j0, &n0, err)) ||
        (rc = dl_job(p, *plan, it->second, mem, left, right, disp, sub, valid, tk, kB1, &j1, &n1, err))) {
        d->pinned.put(tk.hin, tk.hin_n);
        d->pinned.put(tk.hout, tk.hout_n);
        return rc;
    }
    if (n0 > d->budget) {
        d->pinned.put(tk.hin, tk.hin_n);
        d->pinned.put(tk.hout, tk.hout_n);
        return fail(LUX_SGM_ENOMEM, "request larger than the server's HBM budget", err);
    }
    DReq& R = d->ring[idx];
    uint64_t w0[kJobWords], w1[kJobWords];
    std::memcpy(w0, &j0, sizeof j0);
    std::memcpy(w1, &j1, sizeof j1);
    uint32_t mask[kRelocMasks] = {};
    for (int k = 0; k < kJobWords; k++) {
        if (w1[k] == w0[k]) continue;
        if (w1[k] - w0[k] != kB1 - kB0) {
            d->pinned.put(tk.hin, tk.hin_n);
            d->pinned.put(tk.hout, tk.hout_n);
            return fail(LUX_SGM_EINVAL, "descriptor word that is neither an address nor a workspace offset", err);
        }
        mask[k >> 5] |= 1u << (k & 31);
        w0[k] -= kB0;
    }
    std::memcpy((void*)&R.job, w0, sizeof w0);
    std::memcpy((void*)R.reloc, mask, sizeof mask);
    R.ws_bytes = n0;
    const uint64_t cells = (uint64_t)p->width * p->height * p->disparities;
    R.prio = (uint32_t)std::min<uint64_t>(cells, 0xffffffffu);
    R.slot = 0;
    R.t_admit = R.t_done = 0;
    std::memset((void*)R.t_stage, 0, sizeof R.t_stage);
    tk.t_submit = Clock::now();
    __atomic_store_n(&R.state, (uint32_t)R_SUBMITTED, __ATOMIC_SEQ_CST);
    // then look at `alive` (the loop clears it, then looks at the ring): launch a loop if none owns the ring
    if (__atomic_load_n(&d->host->alive, __ATOMIC_SEQ_CST) == 0) {
        hipError_t e = dl_launch(d);
        if (e != hipSuccess) return hip_status(e, "device loop launch", err);
    }
    d->tickets[d->seq + 1] = tk;
    *ticket = d->seq + 1;
    d->seq++;
    d->submitted++;
    return 0;
}

int dloop_poll(DLoop* d, uint64_t ticket, int32_t* state, std::string* err)
{
    std::lock_guard<std::mutex> lk(d->mu);
    auto it = d->tickets.find(ticket);
    if (it == d->tickets.end()) return fail(LUX_SGM_EINVAL, "unknown ticket", err);
    const uint32_t s = __atomic_load_n(&d->ring[it->second.idx].state, __ATOMIC_ACQUIRE);
    *state = s == R_DONE ? LUX_SGM_REQ_DONE : s == R_ADMITTED ? LUX_SGM_REQ_RUNNING : LUX_SGM_REQ_PENDING;
    return 0;
}

int dloop_wait(DLoop* d, uint64_t ticket, int32_t timeout_ms, lux_sgm_request_stats* stats, std::string* err)
{
    DTicket tk;
    {
        std::lock_guard<std::mutex> lk(d->mu);
        auto it = d->tickets.find(ticket);
        if (it == d->tickets.end()) return fail(LUX_SGM_EINVAL, "unknown ticket", err);
        tk = it->second;
    }
    DReq& R = d->ring[tk.idx];
    const auto t0 = Clock::now();
    for (uint32_t k = 0; __atomic_load_n(&R.state, __ATOMIC_ACQUIRE) != R_DONE; k++) {
        hipError_t e = hipStreamQuery(d->stream);
        if (e != hipSuccess && e != hipErrorNotReady) return hip_status(e, "device loop", err);
        if (e == hipSuccess) {
            uint32_t fault = 0;
            if (hipMemcpy(&fault, &d->st->fault, sizeof fault, hipMemcpyDeviceToHost) == hipSuccess && fault)
                return fail(LUX_SGM_EDEVICE, "device loop: no work item completed for 10 s", err);
        }
        if (e == hipSuccess && __atomic_load_n(&d->host->alive, __ATOMIC_SEQ_CST) == 0) {
            // no loop running and the request not done: start one (a second one only runs after the first leaves)
            std::lock_guard<std::mutex> lk(d->mu);
            if (__atomic_load_n(&R.state, __ATOMIC_ACQUIRE) != R_DONE && (e = dl_launch(d)) != hipSuccess)
                return hip_status(e, "device loop launch", err);
        }
        if (timeout_ms >= 0 &&
            std::chrono::duration<double, std::milli>(Clock::now() - t0).count() > (double)timeout_ms)
            return fail(LUX_SGM_ECANCELED, "timed out waiting for the request", err);
        if (k > 64) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    const auto t_done = Clock::now();
    if (tk.mem == LUX_SGM_MEM_HOST) {
        const char* so = (const char*)tk.hout;
        if (tk.disp) std::memcpy(tk.disp, so, tk.px * 2);
        if (tk.sub) std::memcpy(tk.sub, so + tk.px * 2, tk.px * 4);
        if (tk.valid) std::memcpy(tk.valid, so + tk.px * 6, tk.px);
    }
    if (std::getenv("LUX_SGM_DLOOP_TRACE")) {  // when each stage completed, ms after admission
        std::string line = "dloop request " + std::to_string(ticket) + ":";
        for (int k = 0; k < X_NSTAGES; k++)
            if (R.t_stage[k])
                line += " " + std::to_string(k) + "@" + std::to_string((double)(R.t_stage[k] - R.t_admit) / d->clock_khz);
        std::fprintf(stderr, "%s\n", line.c_str());
    }
    if (stats) {
        const double total = std::chrono::duration<double, std::milli>(Clock::now() - tk.t_submit).count();
        const double run = (double)(R.t_done - R.t_admit) / d->clock_khz;
        stats->run_ms = run;
        stats->total_ms = total;
        stats->queued_ms = std::max(0.0, std::chrono::duration<double, std::milli>(t_done - tk.t_submit).count() - run);
        stats->slot = (int32_t)R.slot;
        stats->status = 0;
    }
    std::lock_guard<std::mutex> lk(d->mu);
    d->pinned.put(tk.hin, tk.hin_n);
    d->pinned.put(tk.hout, tk.hout_n);
    __atomic_store_n(&R.state, (uint32_t)R_FREE, __ATOMIC_RELEASE);
    d->tickets.erase(ticket);
    d->completed++;
    return 0;
}

int dloop_info(DLoop* d, lux_sgm_server_info* info)
{
    std::lock_guard<std::mutex> lk(d->mu);
    std::memset(info, 0, sizeof *info);
    info->submitted = d->submitted;
    info->completed = d->completed;
    info->failed = d->failed;
    for (auto& kv : d->tickets) {
        const uint32_t s = __atomic_load_n(&d->ring[kv.second.idx].state, __ATOMIC_ACQUIRE);
        info->pending += s == R_SUBMITTED;
        info->running += s == R_ADMITTED;
    }
    info->slots = d->nslots;
    info->budget = d->budget;
    info->bytes_held = d->budget;  // the pool
    uint64_t peak = 0;
    if (hipMemcpy(&peak, &d->st->peak, sizeof peak, hipMemcpyDeviceToHost) == hipSuccess) info->bytes_peak = peak;
    return 0;
}

}  // namespace lux
