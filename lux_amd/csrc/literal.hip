// Line-buffer-literal aggregation of one DenseLayer (DenseStereoSgm.cs:158-166 buffers, :206-224 schedule,
// Pixel :234-317): the reference's two line-buffer halves, path segments, guard words and margin slots
// held word for word in HBM, and every visit of the sheared T x T tile schedule run against them. It serves
// the layers the per-direction sweeps do not model:
//   - padding that spills: a range with start - MinLo + cap > MaxHi writes its last planes past its path
//     segment, into the guard word and the low words of the next segment (the next slot's lead word and
//     path-0 segment for path 3), and its stale cleanup later refills the same words (:254-264, :280-313);
//     successors whose ranges touch those words read them (and Rd's zero past the end of the buffer, :324);
//   - odd tiles T, whose halves do not alternate row by row across band boundaries (:219);
//   - ranges of 513 to 1024 planes.
//
// Schedule. The reference runs visits in the order of the key (band j, tile i, tile row r), row r of tile i
// covering x in [iT - r, iT + T - r) (mirrored in pass -1), clamped to the image, and tile gwN-1 visiting the
// last column (the first in pass -1) a second time when (gwN-1)T - r >= W. Row y = jT + r (in pass order)
// therefore meets its cells in column order, tile i = (x~ + r) / T for column position x~, then possibly the
// second visit. A visit reads and writes line-buffer slots within two columns of its own, in its own half
// and the other (rows y-2 .. y+2 use them), so it only has to follow, in the reference's order, the visits
// of those rows within that reach. One wavefront runs one row's visits in order; before each it waits until
// each of rows y-2, y-1, y+1, y+2 has completed every visit with a smaller key within two columns (a
// prefix of that row's visits, so one progress counter per row says it). Rows of an earlier band come
// wholly before, those of a later band wholly after.
//
// The second visit of the border column is the exception: it reads the buffers as every earlier visit of its
// band left them, and the last writer near the border can be any row of the band (the rows below it whose
// border cells fall in the previous tile), so it waits on every row of its band.
//
// Progress. Waits only point to smaller keys, and a row never waits on a row of a later band, so a wait for a
// later row reaches at most T rows ahead. Rows are taken by ticket, so with T + 4 wavefronts resident (on the home
// XCD when the pass runs on one; the host checks one XCD's occupancy) every wait ends. A bounded spin sets the error
// word instead of hanging, and later waits are skipped, so the grid always drains.
//
// Memory. All line-buffer, minimum, range, guidance and progress words go through buffer loads and stores with a
// cache policy chosen per launch (LOC), and a visit's stores are drained (s_waitcnt) before its progress store.
//   - One XCD (LOC), when it holds every row of the layer at once: the first wavefront to start names its XCD the
//     pass's home, and wavefronts that start on another leave at once. The words then go through that XCD's L2, which
//     all its CUs share: loads and stores with sc0 and nt (the loads miss in the CU's L1; with sc0 alone a load may
//     hit a line the L1 kept from an earlier read). A hand-over between two wavefronts (a store, then a poll that
//     sees it) measured 0.21 us this way.
//   - All XCDs otherwise (more rows in flight than one XCD holds): sc1, coherent across XCDs, 0.53 us a hand-over.
//   LUX_SGM_LIT_LOCAL=0 or 1 forces either. Other kernels see the words through the kernel boundaries' write-back and
//   invalidation, as any kernel's output. The aggregate of a cell is read and written by its own row's wavefront
//   only, with plain accesses (the second visit of the border column takes the first visit's sums from registers).
// One wavefront per row, planes across lanes (NPL per lane; NPL 0: 64-plane chunks in turn, caps up to 1024).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "kernels.h"
#include "p2_dev.h"

namespace lux {
namespace {

constexpr int kLitReach = 2;   // columns either side whose visits may share line-buffer words with a visit
constexpr int kLitGrid = 2048; // persistent single-wavefront workgroups of the home XCD per pass (rows go by ticket)
constexpr int kLitXcds = 8;    // workgroups go to the XCDs in turn: the grid holds kLitGrid for each
constexpr int kLitXcdCus = 32; // CUs of one XCD

struct LitArgs {
    int W, H, T, gwN, rev_plus, align;
    int minlo, M, stride;            // M = MaxHi rounded to the alignment; stride = 4M + 5 (:160)
    long halfL, Llen;
    int halfM, halfR;
    uint16_t* L;
    uint16_t* Min;
    uint16_t* RS;
    uint16_t* RC;
    uint32_t* Pix;                   // guidance of the slot's cell, RGBA8 (0 for the initial fill)
    const uint16_t* start;           // null: every cell [0, ucap)
    const uint16_t* cap;
    int ucap;
    const uint8_t* raw;              // [H][W][cs]
    uint16_t* agg;                   // [H][W][cs]
    int cs;
    AdaptP2 ap;
    int p1, p2;
    uint32_t* progress;              // [H] visits completed, per row in pass order
    uint32_t* ticket;
    uint32_t* err;
    uint32_t* home;                  // the XCD the pass runs on (~0 until the first wavefront names its own)
};

__device__ __forceinline__ int lit_xcc()
{
    uint32_t v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(v));
    return (int)(v & 7u);
}
// A wave-uniform value into a scalar register (uniform words then take scalar arithmetic and branches).
__device__ __forceinline__ int lit_u(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Whether tile row r ends with the second visit of the last column (first column in pass -1): the last
// tile's row starts past the image (:213-216); LUX_SGM_XSW's clamp to W removes it in pass +1.
__device__ __forceinline__ int lit_rev(const LitArgs& a, int dir, int r)
{
    return (dir < 0 || a.rev_plus) && (a.gwN - 1) * a.T - r >= a.W;
}

// Visits of row2 (pass order) whose key precedes (band j, tile i, tile row r).
__device__ __forceinline__ int lit_before(const LitArgs& a, int dir, int j, int i, int r, int row2)
{
    const int jj = row2 / a.T, rr = row2 - jj * a.T;
    const int rev = lit_rev(a, dir, rr);
    if (jj < j) return a.W + rev;
    if (jj > j) return 0;
    if (rr < r && i == a.gwN - 1) return a.W + rev;
    const long lim = (long)(rr < r ? i + 1 : i) * a.T - rr;
    return (int)std::min(std::max(lim, 0L), (long)a.W);
}

// A visit's reads of the shared words: loads with the launch's cache policy through buffer resources. To the
// compiler they are plain loads, so the reads of a visit issue together after its acquire instead of one round trip
// each (and a poll loop keeps its load inside the loop only behind a compiler barrier: lit_poll_barrier). A read past
// a resource's end returns 0: for the line buffer that is Rd's rule past the end of the buffer (:318-324); every index
// read is >= 0 (the lowest is a lead word).
constexpr int kLitLoc = 3;  // cache policy on one XCD (LOC): sc0 | nt
constexpr int kLitDev = 16; // across XCDs: sc1
__device__ __forceinline__ void lit_poll_barrier() { asm volatile("" ::: "memory"); }
struct LitRes {
    __amdgpu_buffer_rsrc_t L, Min, RS, RC, Pix, P;
};
__device__ __forceinline__ __amdgpu_buffer_rsrc_t lit_rsrc(const void* base, long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ LitRes lit_res(const LitArgs& a)
{
    LitRes r;
    r.L = lit_rsrc(a.L, a.Llen * 2);
    r.Min = lit_rsrc(a.Min, 2L * a.halfM * 2);
    r.RS = lit_rsrc(a.RS, 2L * a.halfR * 2);
    r.RC = lit_rsrc(a.RC, 2L * a.halfR * 2);
    r.Pix = lit_rsrc(a.Pix, 2L * a.halfR * 4);
    r.P = lit_rsrc(a.progress, (long)a.H * 4);
    return r;
}
template <bool LOC>
__device__ __forceinline__ int bld16(__amdgpu_buffer_rsrc_t r, int i)
{
    return (int)__builtin_amdgcn_raw_buffer_load_b16(r, (int)(i * 2), 0, LOC ? kLitLoc : kLitDev);
}
template <bool LOC>
__device__ __forceinline__ uint32_t bld32(__amdgpu_buffer_rsrc_t r, int i)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)(i * 4), 0, LOC ? kLitLoc : kLitDev);
}

// Stores of the shared words: buffer stores with the launch's cache policy, issued with no branch; a store that
// should not happen gets an offset past every resource, which the range check drops.
constexpr uint32_t kLitDrop = 0x80000000u;
template <bool LOC>
__device__ __forceinline__ void bst16(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, uint32_t v)
{
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, r, (int)byte_off, 0, LOC ? kLitLoc : kLitDev);
}
template <bool LOC>
__device__ __forceinline__ void bst32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, uint32_t v)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)byte_off, 0, LOC ? kLitLoc : kLitDev);
}

// Minimum over the 64 lanes of two u16 halves at once (DPP row shifts and broadcasts), wave-uniform.
typedef unsigned short lit_u16x2 __attribute__((ext_vector_type(2)));
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t lit_dpp_pmin(uint32_t v)
{
    const uint32_t o = (uint32_t)__builtin_amdgcn_update_dpp((int)0xffffffff, (int)v, CTRL, ROWMASK, 0xf, false);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(lit_u16x2, v),
                                                                  __builtin_bit_cast(lit_u16x2, o)));
}
__device__ __forceinline__ uint32_t wave_pmin(uint32_t v)
{
    v = lit_dpp_pmin<0x111, 0xf>(v);  // row_shr:1
    v = lit_dpp_pmin<0x112, 0xf>(v);  // row_shr:2
    v = lit_dpp_pmin<0x114, 0xf>(v);  // row_shr:4
    v = lit_dpp_pmin<0x118, 0xf>(v);  // row_shr:8: lane 15 of each row holds the row's minimum
    v = lit_dpp_pmin<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = lit_dpp_pmin<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3: lane 63 holds the minimum
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// The cell's own words, which no other row writes: loaded one visit ahead (LitCell), so a visit's line-buffer
// reads are its only round trip after its waits.
struct LitCell {
    int start, cap;
    uint32_t g;
    int prev_spill;  // the cell two rows before in pass +1 (the slot's previous occupant) writes past its path-3 segment
};
// Whether a range's path-3 write (and so the stale cleanup of it) runs past its segment into the next slot.
__device__ __forceinline__ bool lit_spills(const LitArgs& a, int start, int cap) { return start - a.minlo + cap > a.M + 1; }
// cell2: the cell two rows before in pass order at the same column, or -1.
template <bool ADP>
__device__ __forceinline__ LitCell lit_cell(const LitArgs& a, size_t cell, long cell2)
{
    LitCell c;
    c.start = a.start ? (int)a.start[cell] : 0;
    c.cap = a.start ? (int)a.cap[cell] : a.ucap;
    c.g = ADP ? a.ap.guid[cell] : 0u;
    c.prev_spill = 0;
    if (cell2 >= 0) c.prev_spill = a.start ? lit_spills(a, (int)a.start[cell2], (int)a.cap[cell2]) : lit_spills(a, 0, a.ucap);
    return c;
}

// One Pixel call (:234-317) at (x, y).
// re: the second visit of the border column, whose aggregate the first visit (the one before) left in `last`.
template <int DIR, int NPL, bool ADP, bool LOC>
__device__ __forceinline__ void lit_visit(const LitArgs& a, int x, int y, int cur, int oth, bool re, const LitCell& cc,
                                          int (&last)[NPL])
{
    const int lane = threadIdx.x;
    const int slot = x + 1;
    const size_t cell = (size_t)y * a.W + x;
    const int start = cc.start, cap = cc.cap;
    const int q = start - a.minlo;
    const uint32_t g = cc.g;
    const LitRes R = lit_res(a);
    const int sp = bld16<LOC>(R.RS, cur * a.halfR + slot), cp = bld16<LOC>(R.RC, cur * a.halfR + slot);
    const size_t cbase = cell * (size_t)a.cs;

    // every read first: the visit never reads a word it writes (path 0 reads another slot of its half, paths
    // 1-3 the other half, and path 3's spill into the next slot comes after path 0's read in pass -1)
    // All loads are issued before any of them is used (one round trip for the visit), then the P2 terms.
    int minPrev[4], p2p[4];
    uint32_t gq[4];
    int lm[4][NPL], l0[4][NPL], lp[4][NPL];
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const int ns = slot + (p == 2 ? 0 : p == 3 ? DIR : -DIR);
        const int half = p == 0 ? cur : oth;
        minPrev[p] = bld16<LOC>(R.Min, half * a.halfM + ns * 4 + p);
        gq[p] = ADP ? bld32<LOC>(R.Pix, half * a.halfR + ns) : 0u;
        const int nb = half * (int)a.halfL + ns * a.stride + 1 + p * (a.M + 1) + q;  // below 2^30 (host check)
        // no branch around the loads: lanes past the cap read their wave's first word and drop it
#pragma unroll
        for (int h = 0; h < NPL; h++) {
            const int i0 = nb + (lane + 64 * h < cap ? lane + 64 * h : 0);
            lm[p][h] = bld16<LOC>(R.L, i0 - 1);
            l0[p][h] = bld16<LOC>(R.L, i0);
            lp[p][h] = bld16<LOC>(R.L, i0 + 1);
        }
    }
    // the cell's own raw costs and aggregate (no other row touches them), in the same round trip
    int rb[NPL], ag[NPL];
#pragma unroll
    for (int h = 0; h < NPL; h++) {
        const int k = lane + 64 * h, kk = k < cap ? k : 0;
        const int r0 = (int)a.raw[cbase + kk], a0 = (int)a.agg[cbase + kk];
        rb[h] = k < cap ? r0 : 0;
        ag[h] = k < cap ? (re ? last[h] : a0) : 0;
    }
#pragma unroll
    for (int p = 0; p < 4; p++) {
        minPrev[p] = lit_u(minPrev[p]);  // wave-uniform words: scalar arithmetic from here on
        int P2 = a.p2;
        if constexpr (ADP) P2 = (int)p2_step(a.ap, (uint32_t)a.p2, g, (uint32_t)lit_u((int)gq[p]));
        p2p[p] = std::min(P2 + minPrev[p], 65535);
#pragma unroll
        for (int h = 0; h < NPL; h++) {
            const bool in = lane + 64 * h < cap;
            lm[p][h] = in ? lm[p][h] : 0;
            l0[p][h] = in ? l0[p][h] : 0;
            lp[p][h] = in ? lp[p][h] : 0;
        }
    }

    // the recurrence (:288-298 / :302-312)
    int lc[4][NPL];
    uint32_t mins01 = 0xffffffffu, mins23 = 0xffffffffu;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        uint32_t run = 0xffffu;
#pragma unroll
        for (int h = 0; h < NPL; h++) {
            const int m = std::min(std::min(std::min(lm[p][h] + a.p1, 65535), l0[p][h]),
                                   std::min(std::min(lp[p][h] + a.p1, 65535), p2p[p]));
            const int v = std::max(std::min(rb[h] + m, 65535) - minPrev[p], 0);
            lc[p][h] = v;
            if (lane + 64 * h < cap) run = std::min(run, (uint32_t)v);
        }
        if (p == 0) mins01 = (mins01 & 0xffff0000u) | run;
        if (p == 1) mins01 = (mins01 & 0xffffu) | (run << 16);
        if (p == 2) mins23 = (mins23 & 0xffff0000u) | run;
        if (p == 3) mins23 = (mins23 & 0xffffu) | (run << 16);
    }
    mins01 = wave_pmin(mins01);
    mins23 = wave_pmin(mins23);

    // Pixel's stores in its order: cleanup(0), write(0), cleanup(1), write(1), ... Each is an interval of
    // words; one that runs past its segment continues into the next (the guard word, then the next segment,
    // or for path 3 the next slot's lead word and path-0 segment). A word is stored once, by the last
    // interval that covers it, so no two stores of a visit race.
    // Word indexes fit 31 bits (the host checks the buffer's size); a word past the end of the buffer is not
    // stored (the range check drops it).
    const int B = cur * (int)a.halfL + slot * a.stride;
    const int q0 = lit_u(sp) - a.minlo, Llen = (int)a.Llen, cpu = lit_u(cp);
    int lo[8], hi[8];
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const int bc = B + 1 + p * (a.M + 1);
        lo[2 * p] = bc + std::max(q0, 0);
        hi[2 * p] = bc + std::min(q0 + cpu, Llen - bc);  // k1 = min(cp, L.Length - baseCur - q0) (:262)
        lo[2 * p + 1] = bc + q;
        hi[2 * p + 1] = bc + q + cap;
    }
    // When neither the cleanup nor the write of any path runs past its segment (the usual visit), the intervals of
    // different paths are disjoint and only a path's own write overrides its cleanup: the same stores, with 2 lane
    // tests a word instead of up to 15.
    const bool own = q + cap <= a.M + 1 && q0 + cpu <= a.M + 1;
    if (own) {
#pragma unroll
        for (int o = 0; o < 8; o++) {
#pragma unroll
            for (int h = 0; h < NPL; h++) {
                const int w = lo[o] + lane + 64 * h;
                bool st = w < hi[o];
                if ((o & 1) == 0) st = st & !((w >= lo[o + 1]) & (w < hi[o + 1]));
                bst16<LOC>(R.L, st ? (uint32_t)w * 2u : kLitDrop, (o & 1) ? (uint32_t)lc[o >> 1][h] : 2000u);
            }
        }
    } else {
#pragma unroll
        for (int o = 0; o < 8; o++) {
#pragma unroll
            for (int h = 0; h < NPL; h++) {
                const int w = lo[o] + lane + 64 * h;
                bool st = w < hi[o];
#pragma unroll
                for (int o2 = o + 1; o2 < 8; o2++) st = st & !((w >= lo[o2]) & (w < hi[o2]));  // no branch
                bst16<LOC>(R.L, st ? (uint32_t)w * 2u : kLitDrop, (o & 1) ? (uint32_t)lc[o >> 1][h] : 2000u);
            }
        }
    }
#pragma unroll
    for (int h = 0; h < NPL; h++) {
        const int k = lane + 64 * h;
        last[h] = std::min(ag[h] + lc[0][h] + lc[1][h] + lc[2][h] + lc[3][h], 65535);
        if (k < cap) a.agg[cbase + k] = (uint16_t)last[h];
    }
    if (lane < 4) {
        const uint32_t pair = lane < 2 ? mins01 : mins23;
        const int rm = (lane & 1) ? (int)(pair >> 16) : (int)(pair & 0xffffu);
        bst16<LOC>(R.Min, (uint32_t)(cur * a.halfM + slot * 4 + lane) * 2u, (uint32_t)(cap == 0 ? 0xffff : rm));  // :275, :314
    }
    if (lane == 0) {
        bst16<LOC>(R.RS, (uint32_t)(cur * a.halfR + slot) * 2u, (uint32_t)start);  // :316
        bst16<LOC>(R.RC, (uint32_t)(cur * a.halfR + slot) * 2u, (uint32_t)cap);
        if constexpr (ADP) bst32<LOC>(R.Pix, (uint32_t)(cur * a.halfR + slot) * 4u, g);  // :239
    }
}

// The same visit for caps up to 1024 (NPL 0): 64-plane chunks one after another, the path costs kept in LDS for
// the store phase.
constexpr int kLitMaxCap = 1024;

template <int DIR, bool ADP, bool LOC>
__device__ __forceinline__ void lit_visit_lds(const LitArgs& a, int x, int y, int cur, int oth, bool re, const LitCell& cc,
                                              uint16_t (*lcs)[kLitMaxCap], uint16_t* aggl)
{
    const int lane = threadIdx.x;
    const int slot = x + 1;
    const size_t cell = (size_t)y * a.W + x;
    const int start = cc.start, cap = cc.cap;
    const int q = start - a.minlo;
    const uint32_t g = cc.g;
    const LitRes R = lit_res(a);
    const int sp = bld16<LOC>(R.RS, cur * a.halfR + slot), cp = bld16<LOC>(R.RC, cur * a.halfR + slot);
    const size_t cbase = cell * (size_t)a.cs;
    int minPrev[4], p2p[4];
    int nb[4];
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const int ns = slot + (p == 2 ? 0 : p == 3 ? DIR : -DIR);
        const int half = p == 0 ? cur : oth;
        minPrev[p] = bld16<LOC>(R.Min, half * a.halfM + ns * 4 + p);
        int P2 = a.p2;
        if constexpr (ADP) P2 = (int)p2_step(a.ap, (uint32_t)a.p2, g, bld32<LOC>(R.Pix, half * a.halfR + ns));
        p2p[p] = std::min(P2 + minPrev[p], 65535);
        nb[p] = half * (int)a.halfL + ns * a.stride + 1 + p * (a.M + 1) + q;
    }
    uint32_t run[4] = {0xffffu, 0xffffu, 0xffffu, 0xffffu};
    for (int k = lane; k < ((cap + 63) & ~63); k += 64) {
        const bool act = k < cap;
        int lm[4], l0[4], lp[4];
#pragma unroll
        for (int p = 0; p < 4; p++) {  // loads without a branch, as in lit_visit
            const int i0 = nb[p] + (act ? k : 0);
            const int vm = bld16<LOC>(R.L, i0 - 1), v0 = bld16<LOC>(R.L, i0), vp = bld16<LOC>(R.L, i0 + 1);
            lm[p] = act ? vm : 0;
            l0[p] = act ? v0 : 0;
            lp[p] = act ? vp : 0;
        }
        const int rb = act ? (int)a.raw[cbase + k] : 0;
        int sum = act ? (re ? (int)aggl[k] : (int)a.agg[cbase + k]) : 0;
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int m = std::min(std::min(std::min(lm[p] + a.p1, 65535), l0[p]),
                                   std::min(std::min(lp[p] + a.p1, 65535), p2p[p]));
            const int v = std::max(std::min(rb + m, 65535) - minPrev[p], 0);
            if (act) {
                lcs[p][k] = (uint16_t)v;
                run[p] = std::min(run[p], (uint32_t)v);
                sum += v;
            }
        }
        if (act) {
            const uint16_t v = (uint16_t)std::min(sum, 65535);
            a.agg[cbase + k] = v;
            aggl[k] = v;
        }
    }
    const uint32_t mins01 = wave_pmin(run[0] | (run[1] << 16)), mins23 = wave_pmin(run[2] | (run[3] << 16));
    __syncthreads();
    const int B = cur * (int)a.halfL + slot * a.stride;  // word indexes as in lit_visit
    const int q0 = sp - a.minlo, Llen = (int)a.Llen;
    int lo[8], hi[8];
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const int bc = B + 1 + p * (a.M + 1);
        lo[2 * p] = bc + std::max(q0, 0);
        hi[2 * p] = bc + std::min(q0 + cp, Llen - bc);
        lo[2 * p + 1] = bc + q;
        hi[2 * p + 1] = bc + q + cap;
    }
#pragma unroll
    for (int o = 0; o < 8; o++) {
        for (int k = lane; k < hi[o] - lo[o]; k += 64) {
            const int w = lo[o] + k;
            bool st = true;
#pragma unroll
            for (int o2 = o + 1; o2 < 8; o2++) st = st & !((w >= lo[o2]) & (w < hi[o2]));  // no branch
            bst16<LOC>(R.L, st ? (uint32_t)w * 2u : kLitDrop, (o & 1) ? (uint32_t)lcs[o >> 1][k] : 2000u);
        }
    }
    if (lane < 4) {
        const uint32_t pair = lane < 2 ? mins01 : mins23;
        const int rm = (lane & 1) ? (int)(pair >> 16) : (int)(pair & 0xffffu);
        bst16<LOC>(R.Min, (uint32_t)(cur * a.halfM + slot * 4 + lane) * 2u, (uint32_t)(cap == 0 ? 0xffff : rm));
    }
    if (lane == 0) {
        bst16<LOC>(R.RS, (uint32_t)(cur * a.halfR + slot) * 2u, (uint32_t)start);
        bst16<LOC>(R.RC, (uint32_t)(cur * a.halfR + slot) * 2u, (uint32_t)cap);
        if constexpr (ADP) bst32<LOC>(R.Pix, (uint32_t)(cur * a.halfR + slot) * 4u, g);
    }
}

// The waits run on the whole wavefront with wave-uniform values (scalar registers and branches): every lane reads
// the same progress word, and the bookkeeping stays off the vector unit.
// A spin past the bound (or another wavefront's) sets the error word, after which no wait is made again.
__device__ __forceinline__ bool lit_give_up(const LitArgs& a, int spins)
{
    if ((spins & 63) != 0) return false;
    const bool bad = lit_u((int)__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0 ||
                     spins > (1 << 22);
    if (bad && threadIdx.x == 0) atomicOr(a.err, 1u);
    return bad;
}

// Waits until row2 has completed `need` visits.
template <bool LOC>
__device__ __forceinline__ void lit_wait(const LitArgs& a, int row2, int need, int& seen, bool& dead)
{
    if (dead || seen >= need) return;
    const __amdgpu_buffer_rsrc_t P = lit_res(a).P;
    int spins = 0;
    for (;;) {
        lit_poll_barrier();
        seen = lit_u((int)bld32<LOC>(P, row2));
        if (seen >= need) return;
        if (lit_give_up(a, ++spins)) {
            dead = true;
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Waits until each of four rows has completed need[d] visits (need 0: nothing to wait for). The progress words of
// all four are read together, so one round trip serves them all, and the error word only every 64th round, so a
// progress store is seen one round trip after it lands rather than two.
template <bool LOC>
__device__ __forceinline__ void lit_wait4(const LitArgs& a, const int (&row2)[4], const int (&need)[4], int (&seen)[4],
                                          bool& dead)
{
    if (dead) return;
    if (seen[0] >= need[0] && seen[1] >= need[1] && seen[2] >= need[2] && seen[3] >= need[3]) return;
    const __amdgpu_buffer_rsrc_t P = lit_res(a).P;
    int spins = 0;
    for (;;) {
        lit_poll_barrier();
        uint32_t v[4];
#pragma unroll
        for (int d = 0; d < 4; d++) v[d] = bld32<LOC>(P, row2[d]);
        bool ok = true;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            seen[d] = lit_u((int)v[d]);
            ok = ok && seen[d] >= need[d];
        }
        if (ok) return;
        if (lit_give_up(a, ++spins)) {
            dead = true;
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// reach: columns past c (in pass order) whose visits of row2 the visit waits for (kLitReach, or less when the
// visit's ranges show that fewer can share a word with it: lit_reach).
__device__ __forceinline__ int lit_need(const LitArgs& a, int dir, int j, int i, int r, int c, int row2, int reach = kLitReach)
{
    int mp = std::min(c + reach, a.W - 1);
    if (lit_rev(a, dir, row2 % a.T) && c + reach >= a.W - 1) mp = a.W;
    return std::min(lit_before(a, dir, j, i, r, row2), mp + 1);
}
// A neighbour row2 of row (band j, tile row r), packed in one word for all of row's visits (so a visit makes no
// division and holds one scalar register a neighbour): bits 0-15 row2's tile row rr, bit 16 its second-visit flag,
// bit 17 rr < r, bits 18-19 its band against j (0 earlier, 1 later, 2 the same), bit 20 row2 inside the image.
__device__ __forceinline__ int lit_nb(const LitArgs& a, int dir, int j, int r, int row2)
{
    const int in = row2 >= 0 && row2 < a.H;
    const int rw = in ? row2 : 0, jj = rw / a.T, rr = rw - jj * a.T;
    return rr | (lit_rev(a, dir, rr) << 16) | ((rr < r) << 17) | ((jj < j ? 0 : jj > j ? 1 : 2) << 18) | (in << 20);
}
// The same as lit_need for a neighbour packed by lit_nb (0 for one outside the image).
__device__ __forceinline__ int lit_need_nb(const LitArgs& a, int i, int c, int nb, int reach)
{
    if (!(nb >> 20)) return 0;
    const int rr = nb & 0xffff, rev = (nb >> 16) & 1, lt = (nb >> 17) & 1, band = (nb >> 18) & 3;
    int before;
    if (band == 0) before = a.W + rev;
    else if (band == 1) before = 0;
    else if (lt && i == a.gwN - 1) before = a.W + rev;
    else before = std::min(std::max((lt ? i + 1 : i) * a.T - rr, 0), a.W);
    int mp = std::min(c + reach, a.W - 1);
    if (rev && c + reach >= a.W - 1) mp = a.W;
    return std::min(before, mp + 1);
}

template <int DIR, int NPL, bool ADP, bool LOC>
__global__ __launch_bounds__(64) void lit_pass(LitArgs a)
{
    const int lane = threadIdx.x;
    const int W = a.W, T = a.T;
    __shared__ uint16_t lcs[NPL == 0 ? 4 : 1][NPL == 0 ? kLitMaxCap : 1];
    __shared__ uint16_t aggl[NPL == 0 ? kLitMaxCap : 1];
    int last[NPL == 0 ? 1 : NPL];
    bool dead = false;  // lane 0: the error word is set, waits are skipped so the grid drains
    // the pass's home XCD: named by the first wavefront to get here; the others of other XCDs leave
    int stay = 1;
    if (LOC && lane == 0) {
        const uint32_t me = (uint32_t)lit_xcc();
        uint32_t seen_home = 0xffffffffu;
        __hip_atomic_compare_exchange_strong(a.home, &seen_home, me, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        stay = seen_home == 0xffffffffu || seen_home == me;
    }
    if (!__shfl(stay, 0)) return;
    for (;;) {
        int row = 0;
        if (lane == 0) row = (int)atomicAdd(a.ticket, 1u);
        row = lit_u(__shfl(row, 0));
        if (row >= a.H) break;
        const int j = row / T, r = row - j * T;
        const int y = DIR > 0 ? row : a.H - 1 - row;
        const int cur = (r & 1) == 0 ? 1 : 0, oth = 1 - cur;  // :219
        const int nvis = W + lit_rev(a, DIR, r);
        int seen[4] = {0, 0, 0, 0};  // progress last read of rows row-2, row-1, row+1, row+2
        // those rows (this row for one outside the image, which is waited for by nobody), packed (lit_nb)
        int rows2[4], nb2[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const int row2 = row + (d < 2 ? d - 2 : d - 1);
            rows2[d] = row2 >= 0 && row2 < a.H ? row2 : row;
            nb2[d] = lit_nb(a, DIR, j, r, row2);
        }
        // pass +1 with even T: the slot's previous occupant is the cell two rows up (halves alternate row by row)
        const bool prev2 = DIR > 0 && (T & 1) == 0 && y >= 2;
        const auto cell_at = [&](int cc0) { return (size_t)y * W + (DIR > 0 ? cc0 : W - 1 - cc0); };
        LitCell next = lit_cell<ADP>(a, cell_at(0), prev2 ? (long)cell_at(0) - 2L * W : -1L);
        int ti = 0, tm = r;  // (pos + r) / T and (pos + r) % T, kept step by step
        for (int pos = 0; pos < nvis; pos++) {
            const int c = pos < W ? pos : W - 1;
            const int i = pos < W ? ti : a.gwN - 1;
            if (++tm == T) {
                tm = 0;
                ti++;
            }
            LitCell cc = next;
            cc.start = lit_u(cc.start);
            cc.cap = lit_u(cc.cap);
            cc.g = (uint32_t)lit_u((int)cc.g);
            cc.prev_spill = lit_u(cc.prev_spill);
            if (pos + 1 < nvis) {  // in flight during this visit's waits
                const int c1 = pos + 1 < W ? pos + 1 : W - 1;
                next = lit_cell<ADP>(a, cell_at(c1), prev2 ? (long)cell_at(c1) - 2L * W : -1L);
            }
            if (pos < W) {
                // The reach this visit needs (with even T, whose halves alternate row by row; odd T keeps
                // kLitReach). Pass -1: one column, since the words a visit shares beyond its neighbours' slots (the
                // next slot in memory) lie behind it in pass order. Pass +1: towards rows y+-1, one column, two when
                // its path 3 reads past its segment into the next slot (the slot two columns on, written by that
                // column's visit); towards rows y+-2 (the same half), two when its path-3 write or its cleanup of
                // the previous occupant's range spills into the next slot, which the visit two columns on reads
                // through path 0.
                int reach1 = kLitReach, reach2 = kLitReach;
                if ((T & 1) == 0) {
                    if (DIR < 0) {
                        reach1 = reach2 = 1;
                    } else {
                        const int q = cc.start - a.minlo;
                        reach1 = 1 + (q + cc.cap >= a.M + 1 ? 1 : 0);
                        reach2 = 1 + (lit_spills(a, cc.start, cc.cap) || cc.prev_spill ? 1 : 0);
                    }
                }
                int need[4];
#pragma unroll
                for (int d = 0; d < 4; d++) need[d] = lit_need_nb(a, i, c, nb2[d], (d == 1 || d == 2) ? reach1 : reach2);
                lit_wait4<LOC>(a, rows2, need, seen, dead);
            } else {
                // the second visit of the border column reads the buffers as the band's earlier visits left them,
                // whichever rows made those: it waits on every row of its band (and rows row-2, row-1 before it)
                const int b1 = std::min(a.H, j * T + T);
                for (int row2 = std::max(0, std::min(row - 2, j * T)); row2 < b1; row2++) {
                    int s = 0;
                    if (row2 != row) lit_wait<LOC>(a, row2, lit_need(a, DIR, j, i, r, c, row2), s, dead);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            if constexpr (NPL == 0) lit_visit_lds<DIR, ADP, LOC>(a, DIR > 0 ? c : W - 1 - c, y, cur, oth, pos == W, cc, lcs, aggl);
            else lit_visit<DIR, NPL, ADP, LOC>(a, DIR > 0 ? c : W - 1 - c, y, cur, oth, pos == W, cc, last);
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (lane == 0) bst32<LOC>(lit_res(a).P, (uint32_t)row * 4u, (uint32_t)(pos + 1));
        }
    }
}

// First-minimum WTA (:225-231): the plane index start + arg.
__global__ void lit_wta(const uint16_t* agg, int cs, const uint16_t* start, const uint16_t* count, int ucount, size_t cells,
                        uint16_t* wta)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const int n = start ? (int)count[i] : ucount;
    const uint16_t* p = agg + i * (size_t)cs;
    int best = 0xffff, arg = 0;
    for (int k = 0; k < n; k++) {
        const int v = p[k];
        if (v < best) {
            best = v;
            arg = k;
        }
    }
    wta[i] = (uint16_t)((start ? (int)start[i] : 0) + arg);
}

__global__ void lit_fill16(uint16_t* p, size_t n, uint16_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void lit_fill32(uint32_t* p, size_t n, uint32_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

struct LitLayout {
    size_t L, Min, RS, RC, Pix, progress, ctl, total;  // byte offsets; ctl = ticket, err, home
};
LitLayout lit_layout(int W, int H, int M)
{
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t slots = (size_t)W + 2, stride = (size_t)4 * M + 5;
    LitLayout l;
    size_t o = 0;
    l.L = o;
    o += up(2 * slots * stride * 2 + 4);
    l.Min = o;
    o += up(2 * slots * 4 * 2);
    l.RS = o;
    o += up(2 * slots * 2);
    l.RC = o;
    o += up(2 * slots * 2);
    l.Pix = o;
    o += up(2 * slots * 4);
    l.progress = o;
    o += up((size_t)H * 4);
    l.ctl = o;
    o += 256;
    l.total = o;
    return l;
}

template <int DIR, int NPL, bool ADP>
hipError_t lit_launch(const LitArgs& a, bool loc, int grid, hipStream_t s)
{
    if (loc) lit_pass<DIR, NPL, ADP, true><<<grid, 64, 0, s>>>(a);
    else lit_pass<DIR, NPL, ADP, false><<<grid, 64, 0, s>>>(a);
    return hipGetLastError();
}

template <int NPL, bool ADP>
hipError_t lit_pass_dispatch(int dir, const LitArgs& a, bool loc, int grid, hipStream_t s)
{
    return dir > 0 ? lit_launch<1, NPL, ADP>(a, loc, grid, s) : lit_launch<-1, NPL, ADP>(a, loc, grid, s);
}

hipError_t lit_run_pass(int dir, int npl, bool adp, const LitArgs& a, bool loc, int grid, hipStream_t s)
{
    if (npl == 2) return adp ? lit_pass_dispatch<2, true>(dir, a, loc, grid, s) : lit_pass_dispatch<2, false>(dir, a, loc, grid, s);
    if (npl == 4) return adp ? lit_pass_dispatch<4, true>(dir, a, loc, grid, s) : lit_pass_dispatch<4, false>(dir, a, loc, grid, s);
    return adp ? lit_pass_dispatch<0, true>(dir, a, loc, grid, s) : lit_pass_dispatch<0, false>(dir, a, loc, grid, s);
}

}  // namespace

size_t literal_scratch_bytes(int W, int H, int maxhi) { return lit_layout(W, H, maxhi).total; }

int literal_depth(int W, int T) { return T + 4; }

int literal_resident(int npl, bool adp)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    const void* k = npl == 2   ? (adp ? (const void*)lit_pass<1, 2, true, true> : (const void*)lit_pass<1, 2, false, true>)
                    : npl == 4 ? (adp ? (const void*)lit_pass<1, 4, true, true> : (const void*)lit_pass<1, 4, false, true>)
                               : (adp ? (const void*)lit_pass<1, 0, true, true> : (const void*)lit_pass<1, 0, false, true>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, 64, 0) != hipSuccess) return 0;
    return per * std::min(cus, kLitXcdCus);  // a pass runs on one XCD
}

hipError_t launch_literal(const LiteralLayer& l, void* scratch, uint16_t* agg_pass1, uint16_t* wta, hipStream_t s)
{
    const LitLayout lay = lit_layout(l.W, l.H, l.maxhi);
    char* base = (char*)scratch;
    LitArgs a{};
    a.W = l.W;
    a.H = l.H;
    a.T = l.T;
    a.gwN = (l.W + 2 * l.T - 1) / l.T;
    a.rev_plus = !l.xsw;
    a.align = l.align;
    a.minlo = l.minlo;
    a.M = l.maxhi;
    a.stride = 4 * l.maxhi + 5;
    a.halfL = (long)(l.W + 2) * a.stride;
    a.Llen = 2 * a.halfL;
    if (a.Llen * 2 > 0x7fffffffL) return hipErrorInvalidValue;  // the visits' buffer loads take 32-bit byte offsets
    a.halfM = (l.W + 2) * 4;
    a.halfR = l.W + 2;
    a.L = (uint16_t*)(base + lay.L);
    a.Min = (uint16_t*)(base + lay.Min);
    a.RS = (uint16_t*)(base + lay.RS);
    a.RC = (uint16_t*)(base + lay.RC);
    a.Pix = (uint32_t*)(base + lay.Pix);
    a.progress = (uint32_t*)(base + lay.progress);
    a.ticket = (uint32_t*)(base + lay.ctl);
    a.err = a.ticket + 1;
    a.home = a.ticket + 2;
    a.start = l.start;
    a.cap = l.cap;
    a.ucap = l.ucap;
    a.raw = l.raw;
    a.agg = l.agg;
    a.cs = l.cs;
    a.ap = l.ap;
    a.p1 = l.p1;
    a.p2 = l.p2;
    const size_t cells = (size_t)l.W * l.H;
    // AllocLineBuffers (:158-165): L and Min 2000, range starts 0, range caps Align, guidance 0; Agg 0 (:56)
    const unsigned fb = 1024;
    lit_fill32<<<fb, 256, 0, s>>>((uint32_t*)a.L, (size_t)(a.Llen + 1) / 2, 0x07d007d0u);
    lit_fill32<<<fb, 256, 0, s>>>((uint32_t*)a.Min, (size_t)a.halfM, 0x07d007d0u);
    lit_fill16<<<fb, 256, 0, s>>>(a.RS, (size_t)2 * a.halfR, 0);
    lit_fill16<<<fb, 256, 0, s>>>(a.RC, (size_t)2 * a.halfR, (uint16_t)l.align);
    lit_fill32<<<fb, 256, 0, s>>>(a.Pix, (size_t)2 * a.halfR, 0u);
    lit_fill16<<<fb, 256, 0, s>>>(l.agg, cells * l.cs, 0);
    lit_fill32<<<1, 64, 0, s>>>(a.ticket, 2, 0u);
    hipError_t e = hipGetLastError();
    // one XCD when it holds every row at once (rows never wait for a wavefront to come free), else all of them
    const char* ev = getenv("LUX_SGM_LIT_LOCAL");
    const bool loc = ev && *ev ? atoi(ev) != 0 : l.H <= literal_resident(l.npl, l.ap.on != 0);
    const int grid = std::min(l.H, kLitGrid) * (loc ? kLitXcds : 1);
    for (int pass = 0; pass < 2 && e == hipSuccess; pass++) {
        lit_fill32<<<fb, 256, 0, s>>>(a.progress, (size_t)l.H, 0u);
        lit_fill32<<<1, 64, 0, s>>>(a.ticket, 1, 0u);  // the error word (ticket + 1) stays set across passes
        lit_fill32<<<1, 64, 0, s>>>(a.home, 1, 0xffffffffu);
        e = lit_run_pass(pass == 0 ? 1 : -1, l.npl, l.ap.on != 0, a, loc, grid, s);
        if (e == hipSuccess && pass == 0 && agg_pass1)
            e = hipMemcpyAsync(agg_pass1, l.agg, cells * l.cs * 2, hipMemcpyDeviceToDevice, s);
    }
    if (e == hipSuccess && wta) {
        lit_wta<<<(unsigned)((cells + 255) / 256), 256, 0, s>>>(l.agg, l.cs, l.start, l.count, l.ucount, cells, wta);
        e = hipGetLastError();
    }
    return e;
}

hipError_t literal_error(void* scratch, int W, int H, int maxhi, hipStream_t s, uint32_t* err)
{
    const LitLayout lay = lit_layout(W, H, maxhi);
    hipError_t e = hipMemcpyAsync(err, (char*)scratch + lay.ctl + 4, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

}  // namespace lux
