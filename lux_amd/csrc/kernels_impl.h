// gfx950 kernels of the dense-stereo SGM hot path (DenseStereoSgm.cs:181-232 and its cost front end): the
// device code and the launch sequence of one request, included by the translation units that instantiate it
// (kernels.hip: constant P2; kernels_adp.hip: image-adaptive P2), so the two compile in parallel.
//
// Layout: one cell = the cs planes of one pixel, contiguous (cell-major [y][x][cs]); a 64-lane wavefront
// owns one cell per step, lane l holding planes [l*DPL, (l+1)*DPL) as DPL/2 packed u16 pairs. Every path
// step is then one coalesced cs-byte raw load and a cs*2-byte aggregate read-modify-write, neighbour planes
// d-1 / d+1 come from v_alignbit inside a lane and DPP wave_shr/wave_shl across lanes, and the per-step
// minimum over planes is a DPP row_shr / row_bcast reduction. All arithmetic is the reference's exact u16
// saturating form (DenseStereoSgm.cs:288-298) on v_pk_add_u16/v_pk_sub_u16 with clamp and v_pk_min_u16.
#pragma once
#include "kernels.h"
#include "p2_dev.h"

#include <algorithm>
#include <type_traits>

namespace lux {

namespace {

constexpr uint32_t kInit2 = 0x07d007d0u;  // 2000 in both halves: the line buffers' fill value (:161-162)
constexpr int kBlock = 256;                // four wavefronts per workgroup

typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t as1(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t padd(uint32_t a, uint32_t b) { return as1(__builtin_elementwise_add_sat(as2(a), as2(b))); }
__device__ __forceinline__ uint32_t psub(uint32_t a, uint32_t b) { return as1(__builtin_elementwise_sub_sat(as2(a), as2(b))); }
__device__ __forceinline__ uint32_t pmin(uint32_t a, uint32_t b) { return as1(__builtin_elementwise_min(as2(a), as2(b))); }

// DPP cross-lane moves. `old` is what a lane without a source lane receives.
__device__ __forceinline__ uint32_t shr1(uint32_t v, uint32_t old)  // lane i <- lane i-1 (wave_shr:1)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t shl1(uint32_t v, uint32_t old)  // lane i <- lane i+1 (wave_shl:1)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x130, 0xf, 0xf, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp_min(uint32_t v)
{
    uint32_t o = (uint32_t)__builtin_amdgcn_update_dpp((int)0xffffffff, (int)v, CTRL, ROWMASK, 0xf, false);
    return v < o ? v : o;
}
// Minimum of a u32 over the 64 lanes, returned wave-uniform.
__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
    v = dpp_min<0x111, 0xf>(v);  // row_shr:1
    v = dpp_min<0x112, 0xf>(v);  // row_shr:2
    v = dpp_min<0x114, 0xf>(v);  // row_shr:4
    v = dpp_min<0x118, 0xf>(v);  // row_shr:8  -> lane 15 of each row holds the row minimum
    v = dpp_min<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_min<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3 -> lane 63 holds the minimum
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// ------------------------------------------------------------------ lane-local cell data
template <int NP>
struct Cell {
    uint32_t v[NP];
};

// Planes [lane*2*NP, +2*NP) of a cell: raw bytes widened to packed u16 pairs.
template <int NP>
__device__ __forceinline__ void load_raw(const uint8_t* p, bool active, Cell<NP>& r)
{
    if constexpr (NP == 1) {
        uint32_t b = active ? (uint32_t)(*(const uint16_t*)p) : 0u;
        r.v[0] = __builtin_amdgcn_perm(0u, b, 0x0c010c00u);
    } else {
        uint32_t w[NP / 2];
#pragma unroll
        for (int k = 0; k < NP / 2; k++) w[k] = active ? ((const uint32_t*)p)[k] : 0u;
#pragma unroll
        for (int k = 0; k < NP / 2; k++) {
            r.v[2 * k] = __builtin_amdgcn_perm(0u, w[k], 0x0c010c00u);
            r.v[2 * k + 1] = __builtin_amdgcn_perm(0u, w[k], 0x0c030c02u);
        }
    }
}
// The same planes as packed bytes, loaded without widening: a prefetch that stays in flight until use.
template <int NP>
struct RawW {
    uint32_t w[NP == 1 ? 1 : NP / 2];
};
template <int NP>
__device__ __forceinline__ void load_raw_w(const uint8_t* p, bool active, RawW<NP>& r)
{
    if constexpr (NP == 1) {
        r.w[0] = active ? (uint32_t)(*(const uint16_t*)p) : 0u;
    } else {
#pragma unroll
        for (int k = 0; k < NP / 2; k++) r.w[k] = active ? ((const uint32_t*)p)[k] : 0u;
    }
}
// Unconditional forms for prefetch rings: the caller passes an address that is always in bounds (idle lanes
// read lane 0's planes), so the load needs no select that would wait for it.
template <int NP>
__device__ __forceinline__ void ld_raw_w(const uint8_t* p, RawW<NP>& r)
{
    if constexpr (NP == 1) {
        r.w[0] = (uint32_t)(*(const uint16_t*)p);
    } else {
#pragma unroll
        for (int k = 0; k < NP / 2; k++) r.w[k] = ((const uint32_t*)p)[k];
    }
}
template <int NP>
__device__ __forceinline__ void ld_u16(const uint16_t* p, Cell<NP>& r)
{
#pragma unroll
    for (int k = 0; k < NP; k++) r.v[k] = ((const uint32_t*)p)[k];
}
// Agent-scope relaxed atomic accesses (global_load / global_store with sc1): the band-boundary states go
// through these so they are coherent across XCDs without a full L2 writeback / invalidate per tile.
template <int NP>
__device__ __forceinline__ void store_u16_agent(uint16_t* p, bool active, const Cell<NP>& r)
{
    if (active) {
#pragma unroll
        for (int k = 0; k < NP; k++)
            __hip_atomic_store((uint32_t*)p + k, r.v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
template <int NP>
__device__ __forceinline__ void load_u16_agent(const uint16_t* p, bool active, Cell<NP>& r)
{
#pragma unroll
    for (int k = 0; k < NP; k++)
        r.v[k] = active ? __hip_atomic_load((const uint32_t*)p + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
}
template <int NP>
__device__ __forceinline__ void widen(const RawW<NP>& b, Cell<NP>& r)
{
    if constexpr (NP == 1) {
        r.v[0] = __builtin_amdgcn_perm(0u, b.w[0], 0x0c010c00u);
    } else {
#pragma unroll
        for (int k = 0; k < NP / 2; k++) {
            r.v[2 * k] = __builtin_amdgcn_perm(0u, b.w[k], 0x0c010c00u);
            r.v[2 * k + 1] = __builtin_amdgcn_perm(0u, b.w[k], 0x0c030c02u);
        }
    }
}
template <int NP>
__device__ __forceinline__ void load_u16(const uint16_t* p, bool active, Cell<NP>& r)
{
#pragma unroll
    for (int k = 0; k < NP; k++) r.v[k] = active ? ((const uint32_t*)p)[k] : 0u;
}
template <int NP>
__device__ __forceinline__ void store_u16(uint16_t* p, bool active, const Cell<NP>& r)
{
    if (active) {
#pragma unroll
        for (int k = 0; k < NP; k++) ((uint32_t*)p)[k] = r.v[k];
    }
}

// A path cost vector as bytes (byte path volumes): the low byte of each u16 plane, planes in order. The caller
// guarantees the active planes' costs are below 256 (cost + P2 <= 255 bounds every lc, see PathBytes below).
template <int NP>
__device__ __forceinline__ void store_bytes(uint8_t* p, bool active, const Cell<NP>& c)
{
    if (!active) return;
    if constexpr (NP == 1) {
        *(uint16_t*)p = (uint16_t)__builtin_amdgcn_perm(0u, c.v[0], 0x0c0c0200u);
    } else {
#pragma unroll
        for (int k = 0; k < NP / 2; k++) ((uint32_t*)p)[k] = __builtin_amdgcn_perm(c.v[2 * k + 1], c.v[2 * k], 0x06040200u);
    }
}
template <int NP>
__device__ __forceinline__ void store_zero_bytes(uint8_t* p, bool active)
{
    if (!active) return;
    if constexpr (NP == 1) {
        *(uint16_t*)p = 0;
    } else {
#pragma unroll
        for (int k = 0; k < NP / 2; k++) ((uint32_t*)p)[k] = 0u;
    }
}
// acc += the bytes of w widened to packed u16 pairs (the inverse layout of store_bytes).
template <int NP>
__device__ __forceinline__ void add_bytes(const RawW<NP>& w, Cell<NP>& acc)
{
    Cell<NP> t;
    widen<NP>(w, t);
#pragma unroll
    for (int k = 0; k < NP; k++) acc.v[k] += t.v[k];  // no carry between the halves: sums stay far below 65536
}

// Branch-free stores for the step loops. A store inside a condition (a lane mask, a border column, a flush every
// 64 cells) splits the loop body, and the compiler's wait counting then has to assume the store may be missing on
// one side: it waits for prefetched loads several steps early (s_waitcnt vmcnt far below the loads in flight). These
// stores always issue, as raw buffer stores: a store that should not happen gets an offset past the buffer's size,
// and the hardware's range check drops it.
constexpr int kRsrcWord3 = 0x00020000;      // raw buffer, 32-bit data format (the gfx9 resource word 3)
constexpr uint32_t kDrop = 0x80000000u;     // an offset past every buffer these stores address
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kRsrcWord3);
}
__device__ __forceinline__ void bstore_u16(uint32_t v, __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, r, (int)off, 0, 0);
}
__device__ __forceinline__ void bstore_u32(uint32_t v, __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)off, 0, 0);
}
__device__ __forceinline__ void bstore_f32(float v, __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)off, 0, 0);
}
// A cell's NP packed pairs (u16 planes) at byte offset off (or dropped).
template <int NP>
__device__ __forceinline__ void bstore_cell(const Cell<NP>& c, __amdgpu_buffer_rsrc_t r, uint32_t off, bool on)
{
#pragma unroll
    for (int k = 0; k < NP; k++) bstore_u32(c.v[k], r, on ? off + 4u * k : kDrop);
}
// A cell's planes as bytes (the layout of store_bytes) at byte offset off (or dropped).
template <int NP>
__device__ __forceinline__ void bstore_bytes(const Cell<NP>& c, __amdgpu_buffer_rsrc_t r, uint32_t off, bool on)
{
    if constexpr (NP == 1) {
        bstore_u16(__builtin_amdgcn_perm(0u, c.v[0], 0x0c0c0200u), r, on ? off : kDrop);
    } else {
#pragma unroll
        for (int k = 0; k < NP / 2; k++)
            bstore_u32(__builtin_amdgcn_perm(c.v[2 * k + 1], c.v[2 * k], 0x06040200u), r, on ? off + 4u * k : kDrop);
    }
}

// One path step of DenseLayer.Pixel (:276-314) for the planes this lane owns.
//   m  = min(min(sat(Lm+P1), L0), min(sat(Lp+P1), P2p)),  P2p = min(P2 + minPrev, 65535)
//   lc = satsub(sat(raw + m), minPrev)
// Lm/Lp of the range ends read the guard words (2000). min(sat(a+P1), sat(b+P1)) = sat(min(a,b)+P1) as
// saturating addition of a constant is monotone. Planes beyond cap (mask 0) hold 2000 like the guards and
// are left out of the minimum. Returns the new minPrev (wave-uniform).
// edge_l / edge_r (both halves equal) are the neighbour planes just below lane 0's first plane and just above
// lane 63's last: the guard value 2000 unless a realigned range (below) puts a stored plane there.
template <int NP, bool MASKED>
__device__ __forceinline__ uint32_t path_step(const Cell<NP>& prev, uint32_t minprev, const Cell<NP>& raw,
                                              uint32_t p1x2, uint32_t p2, const Cell<NP>& mask, Cell<NP>& lc,
                                              uint32_t edge_l = kInit2, uint32_t edge_r = kInit2)
{
    uint32_t p2p = p2 + minprev;
    p2p = p2p > 65535u ? 65535u : p2p;
    const uint32_t p2x2 = p2p | (p2p << 16), minx2 = minprev | (minprev << 16);
    const uint32_t left = shr1(prev.v[NP - 1], edge_l);   // plane l*DPL-1 in the high half
    const uint32_t right = shl1(prev.v[0], edge_r);       // plane (l+1)*DPL in the low half
    uint32_t run = 0xffffffffu;
#pragma unroll
    for (int k = 0; k < NP; k++) {
        const uint32_t lm = __builtin_amdgcn_alignbit(prev.v[k], k == 0 ? left : prev.v[k - 1], 16);
        const uint32_t lp = __builtin_amdgcn_alignbit(k == NP - 1 ? right : prev.v[k + 1], prev.v[k], 16);
        uint32_t m = pmin(padd(pmin(lm, lp), p1x2), prev.v[k]);
        m = pmin(m, p2x2);
        uint32_t t = psub(padd(raw.v[k], m), minx2);
        if constexpr (MASKED) {
            t = (t & mask.v[k]) | (kInit2 & ~mask.v[k]);
            run = pmin(run, t | ~mask.v[k]);
        } else {
            run = pmin(run, t);
        }
        lc.v[k] = t;
    }
    const uint32_t lo = run & 0xffffu, hi = run >> 16;
    return wave_min(lo < hi ? lo : hi);
}

template <int NP>
__device__ __forceinline__ void make_mask(int lane, int cap, Cell<NP>& mask, bool& active_any)
{
    const int e0 = lane * 2 * NP;
#pragma unroll
    for (int k = 0; k < NP; k++) {
        uint32_t m = 0;
        if (e0 + 2 * k < cap) m |= 0xffffu;
        if (e0 + 2 * k + 1 < cap) m |= 0xffff0000u;
        mask.v[k] = m;
    }
    active_any = e0 < cap;
}

template <int NP>
__device__ __forceinline__ void set_init(Cell<NP>& c)
{
#pragma unroll
    for (int k = 0; k < NP; k++) c.v[k] = kInit2;
}

// ------------------------------------------------------------------ per-pixel plane ranges
// With per-pixel ranges (DenseLayer Start/Count/Cap, DenseStereoSgm.cs:15-26, :43-58) a cell stores planes
// [start, start+cap) of its path vector, lane l holding relative planes 2l, 2l+1 (one packed pair per lane,
// cap <= 128). The line-buffer slot a successor reads (:265, :288-298) then holds those planes and 2000
// everywhere else: the stale cleanup (:254-264) refills the previous occupant's range before each write, and
// the guard words sit between the path segments. That holds as long as no pixel's range reaches the last word
// of its segment (start - MinLo + cap <= MaxHi), which the host checks. So a successor starting at s sees
// the predecessor (start sp, cap cp) through this realignment: relative plane j of the successor is plane
// j + delta (delta = s - sp) of the predecessor, 2000 outside [0, cp); edge_l / edge_r are planes -1 and
// 128 in the successor's frame.
struct Range {
    int s = 0, c = 0;  // start, cap of the cell whose state a vector holds (c = 0: the initial fill)
};
__device__ __forceinline__ uint32_t rel_plane(uint32_t v, int r, int cp)
{
    if (r < 0 || r >= cp) return 2000u;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)v, r >> 1);
    return (r & 1) ? (w >> 16) : (w & 0xffffu);
}
__device__ __forceinline__ void realign(Cell<1>& st, const Range& from, int s, int lane, uint32_t& edge_l,
                                        uint32_t& edge_r)
{
    const int delta = s - from.s, cp = from.c;
    const uint32_t v = st.v[0];
    const int r0 = 2 * lane + delta;
    uint32_t out;
    if ((delta & 1) == 0) {
        const int sl = lane + (delta >> 1);
        out = (uint32_t)__builtin_amdgcn_ds_bpermute((sl & 63) << 2, (int)v);
    } else {
        const int sa = lane + ((delta - 1) >> 1);
        const uint32_t a = (uint32_t)__builtin_amdgcn_ds_bpermute((sa & 63) << 2, (int)v);
        const uint32_t b = (uint32_t)__builtin_amdgcn_ds_bpermute(((sa + 1) & 63) << 2, (int)v);
        out = __builtin_amdgcn_alignbit(b, a, 16);
    }
    const bool ok0 = r0 >= 0 && r0 < cp, ok1 = r0 + 1 >= 0 && r0 + 1 < cp;
    st.v[0] = (ok0 ? (out & 0xffffu) : 2000u) | (ok1 ? (out & 0xffff0000u) : (2000u << 16));
    const uint32_t el = rel_plane(v, delta - 1, cp), er = rel_plane(v, 128 + delta, cp);
    edge_l = el | (el << 16);
    edge_r = er | (er << 16);
}
// NP pairs a lane (P = 2 NP planes; lane l holds planes P*l + 2k, P*l + 2k + 1 in v[k]): plane r of a state,
// read wave-uniformly, 2000 outside [0, cp).
template <int NP>
__device__ __forceinline__ uint32_t rel_plane_np(const Cell<NP>& c, int r, int cp)
{
    if (r < 0 || r >= cp) return 2000u;
    const int ln = r / (2 * NP), e = r - ln * 2 * NP;
    uint32_t w = c.v[0];
#pragma unroll
    for (int k = 1; k < NP; k++) w = (e >> 1) == k ? c.v[k] : w;
    w = (uint32_t)__builtin_amdgcn_readlane((int)w, ln);
    return (e & 1) ? (w >> 16) : (w & 0xffffu);
}
// The realignment with NP pairs a lane: relative plane j of the new frame is plane j + delta of the old one.
// With delta = P*dl + de (0 <= de < P), lane l's pairs come from lanes l+dl and l+dl+1 (2 NP ds_bpermute), the
// pair offset de/2 selects among them and an odd de shifts by one plane (v_alignbit).
template <int NP>
__device__ __forceinline__ void realign_np(Cell<NP>& st, const Range& from, int s, int lane, uint32_t& edge_l,
                                           uint32_t& edge_r)
{
    constexpr int P = 2 * NP;
    const int delta = s - from.s, cp = from.c;
    const Cell<NP> v = st;
    const int dl = delta >= 0 ? delta / P : -((-delta + P - 1) / P);
    const int de = delta - dl * P, h = de >> 1;
    const int sa = ((lane + dl) & 63) << 2, sb = ((lane + dl + 1) & 63) << 2;
    uint32_t S[2 * NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
        S[k] = (uint32_t)__builtin_amdgcn_ds_bpermute(sa, (int)v.v[k]);
        S[NP + k] = (uint32_t)__builtin_amdgcn_ds_bpermute(sb, (int)v.v[k]);
    }
#pragma unroll
    for (int k = 0; k < NP; k++) {
        uint32_t p0 = S[0], p1 = S[1];
#pragma unroll
        for (int j = 1; j < 2 * NP; j++) p0 = k + h == j ? S[j] : p0;
#pragma unroll
        for (int j = 2; j < 2 * NP; j++) p1 = k + h + 1 == j ? S[j] : p1;
        const uint32_t out = (de & 1) ? __builtin_amdgcn_alignbit(p1, p0, 16) : p0;
        const int r0 = P * lane + 2 * k + delta;
        const bool ok0 = r0 >= 0 && r0 < cp, ok1 = r0 + 1 >= 0 && r0 + 1 < cp;
        st.v[k] = (ok0 ? (out & 0xffffu) : 2000u) | (ok1 ? (out & 0xffff0000u) : (2000u << 16));
    }
    const uint32_t el = rel_plane_np<NP>(v, delta - 1, cp), er = rel_plane_np<NP>(v, 64 * P + delta, cp);
    edge_l = el | (el << 16);
    edge_r = er | (er << 16);
}
// Move a chain's state into the frame of the next cell (start s); a no-op when the start does not change
// (the planes past the predecessor's cap already hold 2000).
template <int NP, bool VAR>
__device__ __forceinline__ void reframe(Cell<NP>& st, Range& from, int s, int c, int lane, uint32_t& edge_l,
                                        uint32_t& edge_r)
{
    edge_l = kInit2;
    edge_r = kInit2;
    if constexpr (VAR) {
        if constexpr (NP == 1) {
            if (s != from.s) realign(*(Cell<1>*)&st, from, s, lane, edge_l, edge_r);
        } else {
            if (s != from.s) realign_np<NP>(st, from, s, lane, edge_l, edge_r);
        }
        from.s = s;
        from.c = c;
    }
}
// A state loaded from a saved table (seeds, border columns) holds only the lanes of the stored stride: give
// every plane past the stored cap the 2000 the line buffer has there.
template <int NP>
__device__ __forceinline__ void clip_to_cap(Cell<NP>& st, int cap, int lane)
{
    Cell<NP> m;
    bool any;
    make_mask<NP>(lane, cap, m, any);
#pragma unroll
    for (int k = 0; k < NP; k++) st.v[k] = (st.v[k] & m.v[k]) | (kInit2 & ~m.v[k]);
}

// Plane q of a cell held in lanes (2*NP planes per lane), read wave-uniformly.
template <int NP>
__device__ __forceinline__ uint32_t plane_of(const Cell<NP>& c, int q)
{
    const int ln = q / (2 * NP), e = q - ln * 2 * NP;
    uint32_t w = c.v[0];
#pragma unroll
    for (int k = 1; k < NP; k++) w = (e >> 1) == k ? c.v[k] : w;
    w = (uint32_t)__builtin_amdgcn_readlane((int)w, ln);
    return (e & 1) ? (w >> 16) : (w & 0xffffu);
}

// Chains of one path direction: the cells whose predecessor (x-DX, y-DY) lies outside the image start one.
template <int DX, int DY>
__device__ __forceinline__ int chain_count(int W, int H) { return DY == 0 ? H : (DX == 0 ? W : W + H - 1); }

template <int DX, int DY>
__device__ __forceinline__ void chain_start(int c, int W, int H, int& x0, int& y0, int& n)
{
    if constexpr (DY == 0) {
        y0 = c; x0 = DX > 0 ? 0 : W - 1; n = W;
    } else if constexpr (DX == 0) {
        x0 = c; y0 = DY > 0 ? 0 : H - 1; n = H;
    } else {
        if (c < W) {
            x0 = c; y0 = DY > 0 ? 0 : H - 1;
        } else {
            const int k = c - W + 1;
            y0 = DY > 0 ? k : H - 1 - k; x0 = DX > 0 ? 0 : W - 1;
        }
        const int nx = DX > 0 ? W - x0 : x0 + 1, ny = DY > 0 ? H - y0 : y0 + 1;
        n = nx < ny ? nx : ny;
    }
}

// The diagonal chains (DY != 0, DX != 0) longest first: the full-length ones, then the two of each shorter length (one
// starting on the first row, one on the first column). A pass launch holds more wavefronts than are resident at once,
// so the ones that start last are then the short ones. A permutation of [0, W+H-1); indexes past it are kept.
template <int DX>
__device__ __forceinline__ int diag_order(int w, int W, int H)
{
    if (w >= W + H - 1) return w;
    int c;
    if (W >= H) {
        const int F = W - H + 1;  // first-row starts 0 .. W-H: length H
        if (w < F) {
            c = w;
        } else {
            const int t = w - F, j = t / 2 + 1;
            c = (t & 1) ? W - 1 + j : W - H + j;
        }
    } else {
        const int F = H - W + 1;  // the first-row start 0 and the first-column starts 1 .. H-W: length W
        if (w < F) {
            c = w == 0 ? 0 : W - 1 + w;
        } else {
            const int t = w - F, j = t / 2 + 1;
            c = (t & 1) ? W - 1 + (H - W) + j : j;
        }
    }
    return (DX < 0 && c < W) ? W - 1 - c : c;  // leftward chains from the first row grow the other way
}

struct SweepArgs {
    const uint8_t* raw;
    uint16_t* agg;
    int W, H, cs, cap;
    uint32_t p1x2, p2;
    uint16_t* seeds;     // [3][W][cs]
    uint16_t* seedmins;  // [3][W]
    int seed_row;
    uint16_t* cols;      // [2][H][cs] (this pass)
    uint16_t* colmins;   // [2][H]
    const uint16_t* vstart;  // per-pixel ranges: start and cap of every cell [H][W] (null: uniform)
    const uint16_t* vcap;
    uint16_t* seedstart;     // [3][W] start / cap of the cell each seed-table entry belongs to
    uint16_t* seedcap;
    AdaptP2 ap;
    uint32_t* seedguid;      // [3][W] (adaptive P2)
    uint8_t* pv = nullptr;   // byte path volume of this path (PV sweeps): cell c's planes at pv + c * pvs
    int pvs = 0;
};

// Per-cell words of a chain (the cell's range, its guidance), 64 chain positions in one VGPR (lane k: position
// i0 + k, one vector load), read out per step with v_readlane, the next 64 in flight: one load per 64 cells and
// word instead of one per cell (the row and sweep loops issue about a third of the vector loads they did).
// Block k (positions 64k .. 64k+63) sits in register b[k & 1]; when block k is done, that register is reloaded
// with block k+2 in place (a load into a register that is then copied elsewhere would wait for the load at once).
struct ChainWords {
    uint32_t b0 = 0, b1 = 0;
    template <typename F>
    __device__ __forceinline__ void init(const F& load, int lane)
    {
        b0 = load(lane);
        b1 = load(64 + lane);
    }
    __device__ __forceinline__ uint32_t at(int pos) const
    {
        // read both and pick the scalar (a select between the registers becomes a dynamically indexed array)
        const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)b0, pos & 63);
        const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)b1, pos & 63);
        return (pos & 64) ? r1 : r0;
    }
    // after the last use of position pos: at the end of its block, reload that block's register with the block
    // after next
    template <typename F>
    __device__ __forceinline__ void advance(int pos, const F& load, int lane)
    {
        if ((pos & 63) == 63) {
            if (pos & 64) b1 = load(pos + 65 + lane);
            else b0 = load(pos + 65 + lane);
        }
    }
};

// The adaptive P2 of a chain's steps, 64 positions a VGPR. Step i's P2 depends only on the guidance words of
// positions i and i-1 (of the chain's start state for i = 0), so when a block's words are in (loaded a block
// earlier, by ChainWords) the 64 lanes compute the block's P2 at once (lane j: position 64k+j, its predecessor's
// word from lane j-1 by DPP, lane 0 taking the previous block's last word) and each step reads its own with
// v_readlane: the same arithmetic as p2_step per step, a 64th of the instructions.
struct ChainP2 {
    ChainWords g;      // the guidance words (g.at(pos) is position pos's word)
    uint32_t p2 = 0;   // P2 of the current block's positions
    template <typename F>
    __device__ __forceinline__ void init(const F& load, int lane, const AdaptP2& q, uint32_t p2c, uint32_t gprev0)
    {
        g.init(load, lane);
        p2 = p2_step(q, p2c, g.b0, shr1(g.b0, gprev0));
    }
    __device__ __forceinline__ uint32_t at(int pos) const
    {
        return (uint32_t)__builtin_amdgcn_readlane((int)p2, pos & 63);
    }
    // after the last use of position pos: at the end of its block, the next block's P2 from its words, then the
    // freed register takes the block after next
    template <typename F>
    __device__ __forceinline__ void advance(int pos, const F& load, int lane, const AdaptP2& q, uint32_t p2c)
    {
        if ((pos & 63) == 63) {
            if (pos & 64) {
                p2 = p2_step(q, p2c, g.b0, shr1(g.b0, (uint32_t)__builtin_amdgcn_readlane((int)g.b1, 63)));
                g.b1 = load(pos + 65 + lane);
            } else {
                p2 = p2_step(q, p2c, g.b1, shr1(g.b1, (uint32_t)__builtin_amdgcn_readlane((int)g.b0, 63)));
                g.b0 = load(pos + 65 + lane);
            }
        }
    }
};

// One path direction of one pass. Path P of pass DIR steps by (DX, DY) = DIR * {(1,0),(1,1),(0,1),(-1,1)}[P]
// (the neighbour slots of :265). Each wavefront runs whole chains. The border column bx (W-1 in pass +1,
// 0 in pass -1) is left to the replay kernel; FIRST writes the aggregate instead of adding to it. PV writes the
// path's costs as bytes into its own volume a.pv (0 in the border column) and leaves the aggregate alone.
template <int NP, int DIR, int P, bool FIRST, bool MASKED, bool VAR, bool ADP, bool PV = false>
__device__ __forceinline__ void sweep_wave(const SweepArgs& a, int chain, int lane)
{
    constexpr int DX = DIR * (P == 0 ? 1 : P == 1 ? 1 : P == 2 ? 0 : -1);
    constexpr int DY = DIR * (P == 0 ? 0 : 1);
    const int W = a.W, H = a.H, cs = a.cs;
    if (chain >= chain_count<DX, DY>(W, H)) return;
    int x, y, n;
    chain_start<DX, DY>(chain, W, H, x, y, n);
    const int bx = DIR > 0 ? W - 1 : 0;
    if (P == 2 && x == bx) {  // the border column's vertical chain is the replay's
        if (FIRST || PV) {    // a path-cost volume of its own still needs its zeros there
            const int lo = lane * 2 * NP;
            Cell<NP> z;
#pragma unroll
            for (int k = 0; k < NP; k++) z.v[k] = 0;
            for (int i = 0; i < n; i++) {
                const size_t c = (size_t)(y + i * DY) * W + (x + i * DX);
                if constexpr (PV) store_zero_bytes<NP>(a.pv + c * a.pvs + lo, lo < cs);
                else store_u16<NP>(a.agg + c * cs + lo, lo < cs, z);
            }
        }
        return;
    }
    Cell<NP> mask;
    bool active;
    make_mask<NP>(lane, a.cap, mask, active);
    const int lo = lane * 2 * NP;

    Cell<NP> state;
    uint32_t minprev = 2000, gprev = 0;  // gprev: guidance of the start state's cell (0: the initial fill)
    Range from;
    set_init(state);
    if (DIR < 0 && P != 0 && y == H - 1) {
        // pass -1's first row reads half 0 as pass +1 left it (:28-29, :219)
        const int sc = x - DX;
        if (sc >= 0 && sc < W) {
            load_u16<NP>(a.seeds + ((size_t)(P - 1) * W + sc) * cs + lo, lo < cs, state);
            if (VAR) {
                from.s = a.seedstart[(P - 1) * W + sc];
                from.c = a.seedcap[(P - 1) * W + sc];
                clip_to_cap<NP>(state, from.c, lane);
            } else if (MASKED) {
#pragma unroll
                for (int k = 0; k < NP; k++) state.v[k] = (state.v[k] & mask.v[k]) | (kInit2 & ~mask.v[k]);
            }
            minprev = a.seedmins[(P - 1) * W + sc];
            if (ADP) gprev = a.seedguid[(P - 1) * W + sc];
        }
    }
    // software pipeline: PF cells in flight; raw stays packed until its step. Prefetch addresses are clamped
    // to the chain's last cell and idle lanes read lane 0's planes, so every load is unconditional.
    constexpr int PF = 8;
    const ptrdiff_t step = (ptrdiff_t)DY * W + DX;
    const size_t cell0 = (size_t)y * W + x;
    const int lol = (VAR ? lo < cs : active) ? lo : 0;
    size_t cell = cell0;
    RawW<NP> rb[PF];
    Cell<NP> ab[PF];
    // VAR / ADP: the chain's range (start | cap << 16) and guidance words, 64 positions a VGPR (ChainWords)
    const auto rs_at = [&](int pos) {
        const size_t c = cell0 + (ptrdiff_t)min(pos, n - 1) * step;
        return (uint32_t)a.vstart[c] | ((uint32_t)a.vcap[c] << 16);
    };
    const auto gq_at = [&](int pos) { return a.ap.guid[cell0 + (ptrdiff_t)min(pos, n - 1) * step]; };
    ChainWords rs;
    ChainP2 gq;
    if constexpr (VAR) rs.init(rs_at, lane);
    if constexpr (ADP) gq.init(gq_at, lane, a.ap, a.p2, gprev);
#pragma unroll
    for (int j = 0; j < PF; j++) {
        const size_t cj = cell0 + (ptrdiff_t)min(j, n - 1) * step;
        ld_raw_w<NP>(a.raw + cj * cs + lol, rb[j]);
        if (!FIRST && !PV) ld_u16<NP>(a.agg + cj * cs + lol, ab[j]);
    }
    Cell<NP> seedv, colv;
    uint32_t seedm = 0, colm = 0, sgd = 0;
    int xs = 0, yc = 0, scs = 0, scc = 0;
    bool has_seed = false, has_col = false;
    auto body = [&](int i, int j) {
        Cell<NP> raw;
        widen<NP>(rb[j], raw);
        const Cell<NP> agg = ab[j];
        uint32_t el = kInit2, er = kInit2;
        int cs_ = 0, cc_ = 0;
        if constexpr (VAR) {
            const uint32_t r = rs.at(i);
            cs_ = (int)(r & 0xffffu);
            cc_ = (int)(r >> 16);
            reframe<NP, VAR>(state, from, cs_, cc_, lane, el, er);
            make_mask<NP>(lane, cc_, mask, active);
        }
        const uint32_t gc = ADP ? gq.g.at(i) : 0u;
        const uint32_t p2 = ADP ? gq.at(i) : a.p2;
        const size_t cp = cell0 + (ptrdiff_t)min(i + PF, n - 1) * step;
        ld_raw_w<NP>(a.raw + cp * cs + lol, rb[j]);
        if (!FIRST && !PV) ld_u16<NP>(a.agg + cp * cs + lol, ab[j]);
        if constexpr (VAR) rs.advance(i, rs_at, lane);
        if constexpr (ADP) gq.advance(i, gq_at, lane, a.ap, a.p2);
        Cell<NP> lc;
        minprev = path_step<NP, MASKED>(state, minprev, raw, a.p1x2, p2, mask, lc, el, er);
        const bool bord = x == bx;
        if constexpr (PV) {
            // every step stores (the border column its zeros): no memory access of the step loop depends on a
            // condition, which would cost the prefetches their early issue (see bstore_*)
            Cell<NP> out;
#pragma unroll
            for (int k = 0; k < NP; k++) out.v[k] = bord ? 0u : lc.v[k];
            store_bytes<NP>(a.pv + cell * a.pvs + lo, MASKED ? active : true, out);
        } else if (!bord) {
            Cell<NP> out;
#pragma unroll
            for (int k = 0; k < NP; k++) out.v[k] = FIRST ? lc.v[k] : padd(agg.v[k], lc.v[k]);
            store_u16<NP>(a.agg + cell * cs + lo, active, out);
        } else if (FIRST) {
            Cell<NP> z;
#pragma unroll
            for (int k = 0; k < NP; k++) z.v[k] = 0;
            store_u16<NP>(a.agg + cell * cs + lo, active, z);
        }
        // the seed-table entry (pass +1 in the seed row) and the border replay's column state (path 0 / 1 next to
        // the border column): a chain meets each at most once, so they are kept and stored after the loop
        if (DIR > 0 && P != 0 && y == a.seed_row && !bord) {
            seedv = lc;
            seedm = minprev;
            xs = x;
            scs = cs_;
            scc = cc_;
            sgd = gc;
            has_seed = true;
        }
        if (P <= 1 && x == bx - DIR) {
            colv = lc;
            colm = minprev;
            yc = y;
            has_col = true;
        }
        state = lc;
        cell += step;
        x += DX;
        y += DY;
    };
    int i0 = 0;
    for (; i0 + PF <= n; i0 += PF) {
#pragma unroll
        for (int j = 0; j < PF; j++) body(i0 + j, j);
    }
#pragma unroll
    for (int j = 0; j < PF; j++)
        if (i0 + j < n) body(i0 + j, j);
    if (has_seed) {
        store_u16<NP>(a.seeds + ((size_t)(P - 1) * W + xs) * cs + lo, lo < cs, seedv);
        if (lane == 0) {
            a.seedmins[(P - 1) * W + xs] = (uint16_t)seedm;
            if (VAR) {
                a.seedstart[(P - 1) * W + xs] = (uint16_t)scs;
                a.seedcap[(P - 1) * W + xs] = (uint16_t)scc;
            }
            if (ADP) a.seedguid[(P - 1) * W + xs] = sgd;
        }
    }
    if (has_col) {  // the border replay reads this column's path 0 / 1 states
        store_u16<NP>(a.cols + ((size_t)P * H + yc) * cs + lo, lo < cs, colv);
        if (lane == 0) a.colmins[P * H + yc] = (uint16_t)colm;
    }
}
template <int NP, int DIR, int P, bool FIRST, bool MASKED, bool VAR, bool ADP>
__global__ __launch_bounds__(kBlock) void sgm_sweep(SweepArgs a)
{
    sweep_wave<NP, DIR, P, FIRST, MASKED, VAR, ADP>(
        a, __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)), threadIdx.x & 63);
}
// The three sweeps of one pass of a byte-volume request in one launch: wavefront w runs chain w / 3 of path 1 + w % 3,
// so every SIMD holds chains of all three directions. Alone, the vertical sweep has only W chains (about four
// wavefronts a SIMD at 4160 columns), too few to hide its step's latency; together the pass streams at the rate
// of the diagonals. Results are those of the three launches (the chains are independent). The rightward diagonals run
// longest first (their chains by index run long, short, then long again, so the last long ones would start late); the
// leftward ones keep the index order, which measured faster for them.
struct SweepArgs3 {
    SweepArgs s[3];
};
// One of the three alone (requests that share the device with others: their launches interleave with the other
// requests' kernels instead).
template <int NP, int DIR, int P, bool MASKED>
__global__ __launch_bounds__(kBlock) void sgm_sweep_pv(SweepArgs a)
{
    const int c = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    constexpr int DX = DIR * (P == 1 ? 1 : P == 3 ? -1 : 0);
    sweep_wave<NP, DIR, P, false, MASKED, false, false, true>(a, DX > 0 ? diag_order<1>(c, a.W, a.H) : c,
                                                                threadIdx.x & 63);
}
template <int NP, int DIR, bool MASKED>
__global__ __launch_bounds__(kBlock) void sgm_sweeps_pv(SweepArgs3 a)
{
    const int w = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int p = w % 3, c = w / 3, lane = threadIdx.x & 63;
    if (p == 0) sweep_wave<NP, DIR, 1, false, MASKED, false, false, true>(a.s[0], DIR > 0 ? diag_order<1>(c, a.s[0].W, a.s[0].H) : c, lane);
    else if (p == 1) sweep_wave<NP, DIR, 2, false, MASKED, false, false, true>(a.s[1], c, lane);
    else sweep_wave<NP, DIR, 3, false, MASKED, false, false, true>(a.s[2], DIR < 0 ? diag_order<1>(c, a.s[2].W, a.s[2].H) : c, lane);
}

// ------------------------------------------------------------------ horizontal paths (rows)
struct RowArgs {
    const uint8_t* raw;     // raw volume (read; written when the cost is fused)
    uint8_t* raw_out;       // fused census cost: where the raw bytes go
    const uint64_t* cl;     // fused census cost: census codes of the left / right images
    const uint64_t* cr;
    uint16_t* agg;
    int W, H, cs, cap, D;
    uint32_t p1x2, p2;
    uint16_t* cols;         // [2][H][cs] of this pass (slot 0: path 0)
    uint16_t* colmins;      // [2][H]
    uint16_t* disp;         // WTA outputs
    float* sub;
    uint16_t* dr;           // [H][W] right-image disparity
    uint32_t* dr0;          // [H] dR(0) without cell (0, 0)
    int write_agg;          // store the final aggregate too
    const uint16_t* vstart; // per-pixel ranges [H][W]: start, cap and count (null: uniform)
    const uint16_t* vcap;
    const uint16_t* vcount;
    AdaptP2 ap;
    // byte path volumes (PV): pv[k] is path k of the eight in launch order (-> of pass +1, its three sweeps, pass -1's
    // three sweeps), cell c at pv[k] + c * pvs; the row path of pass +1 writes pv[0], the one of pass -1 sums them all
    uint8_t* pv[7] = {};
    int pvs = 0;            // byte stride between consecutive cells of a path volume
    const uint16_t* bagg_plus = nullptr;   // [H][cs] pass +1's border-column (x = W-1) costs from its replay
    const uint16_t* bagg_minus = nullptr;  // [H][cs] pass -1's border-column (x = 0) costs from its replay
};

// Census codes along a row. Lane l owns planes d = lo .. lo+2*NP-1 (lo = 2*NP*l) and keeps the window
// win[e] = cr(max(x - lo - e, 0)) of the right-image codes its planes compare with at the current x.
// - Moving to x+1 the window shifts by one code: win[e] <- win[e-1], and win[0] <- the left neighbour lane's
//   win[2*NP-1] = cr(max(x + 1 - lo, 0)) (DPP wave_shr), lane 0 taking cr(x+1) itself. So a step reads two
//   wave-uniform codes, cl(x) and cr(x+1), instead of 2*NP per lane.
// - Moving to x-1: win[e] <- win[e+1], and win[2*NP-1] <- the right neighbour lane's win[0] = cr(max(x - lo - 2*NP,
//   0)) (DPP wave_shl), lane 63 taking cr(max(x - 128*NP, 0)).
// The window is kept rotated instead of shifted: logical element e sits in win[(e + rot) % E], E = 2 * NP, so a
// step writes the one element that changes and moves none. A row path passes rot from its step index modulo E,
// which the unrolled step loop makes a constant (every step group starts at a multiple of E).
// Advance (x -> x+1): rot = (E - t % E) % E at step t; the new element takes the slot of the one dropped.
template <int NP>
__device__ __forceinline__ void census_advance_rot(uint64_t (&win)[2 * NP], uint64_t fresh, int rot)
{
    constexpr int E = 2 * NP;
    const int k = (rot + E - 1) % E;
    const uint64_t last = win[k];
    const uint32_t lo = shr1((uint32_t)last, (uint32_t)fresh), hi = shr1((uint32_t)(last >> 32), (uint32_t)(fresh >> 32));
    win[k] = ((uint64_t)hi << 32) | lo;
}
// Retreat (x -> x-1): rot = t % E at step t.
template <int NP>
__device__ __forceinline__ void census_retreat_rot(uint64_t (&win)[2 * NP], uint64_t fresh, int rot)
{
    const int k = rot;
    const uint64_t first = win[k];
    const uint32_t lo = shl1((uint32_t)first, (uint32_t)fresh), hi = shl1((uint32_t)(first >> 32), (uint32_t)(fresh >> 32));
    win[k] = ((uint64_t)hi << 32) | lo;
}
// Census Hamming costs popcount(cl(x) ^ cr(max(x-d, 0))) for d < D, 255 above (a plane left of the image
// repeats the last in-image plane, DenseStereo.cs:273-274), widened for the path step and packed as bytes.
// FULL: every lane's planes are below D (D = 128 * NP), so no plane takes 255.
// rot: the window is rotated (census_advance_rot / census_retreat_rot), element e in win[(e + rot) % (2 * NP)].
template <int NP, bool FULL = false>
__device__ __forceinline__ void census_cost(uint64_t cl, const uint64_t (&win)[2 * NP], int D, int lo, Cell<NP>& raw,
                                            uint32_t (&bytes)[NP / 2 + (NP == 1)], int rot = 0)
{
    uint32_t v[2 * NP];
#pragma unroll
    for (int e = 0; e < 2 * NP; e++)
        v[e] = (FULL || lo + e < D) ? (uint32_t)__popcll(cl ^ win[(e + rot) % (2 * NP)]) : 255u;
#pragma unroll
    for (int k = 0; k < NP; k++) raw.v[k] = v[2 * k] | (v[2 * k + 1] << 16);
    if constexpr (NP == 1) {
        bytes[0] = v[0] | (v[1] << 8);
    } else {
#pragma unroll
        for (int k = 0; k < NP / 2; k++) bytes[k] = v[4 * k] | (v[4 * k + 1] << 8) | (v[4 * k + 2] << 16) | (v[4 * k + 3] << 24);
    }
}

// Pass +1, path 0 (->): one wavefront per row, x = 0 .. W-1. The first aggregate writer (writes lc instead
// of adding; the border column W-1 gets 0 and is filled by the replay). COST: computes the census cost of
// each cell on the way and writes the raw volume for the later sweeps. Saves the W-2 column's states for
// the border replay. PV: the costs go to the byte path volume pv[0] instead of the aggregate.
template <int NP, bool MASKED, bool COST, bool VAR, bool ADP, bool PV = false>
__device__ __forceinline__ void row_fwd_wave(const RowArgs& a, int y, int lane)
{
    static_assert(!(COST && VAR), "per-pixel ranges take their raw volume from the cost kernel");
    constexpr int PF = 8;
    if (y >= a.H) return;
    const int W = a.W, cs = a.cs, lo = lane * 2 * NP;
    Cell<NP> mask;
    bool active;
    make_mask<NP>(lane, a.cap, mask, active);
    const bool store_lane = lo < cs;
    Cell<NP> state;
    set_init(state);
    uint32_t minprev = 2000;
    const size_t rowcell = (size_t)y * W;
    const uint64_t* crow = COST ? a.cr + rowcell : nullptr;
    const uint64_t* clrow = COST ? a.cl + rowcell : nullptr;
    const int lol = store_lane ? lo : 0;
    // COST: the lane's window of right codes, and the wave-uniform codes cl(x), cr(x+1) of 64 cells at a time in
    // VGPRs (lane k: cell x0 + k), read out per cell with v_readlane, the next 64 in flight. (Eight cells of codes
    // prefetched in SGPRs spill into VGPR lanes, and a spill waits for its scalar load every cell.)
    uint64_t win[2 * NP];
    uint64_t bl0 = 0, br0 = 0, bl1 = 0, br1 = 0;  // cells 64k .. 64k+63 in the registers of parity k & 1
    RawW<NP> rb[COST ? 1 : PF];
    Range from;
    // VAR / ADP: the row's range (start | cap << 16) and guidance words, 64 cells a VGPR (ChainWords)
    const auto rs_at = [&](int pos) {
        const size_t c = rowcell + min(pos, W - 1);
        return (uint32_t)a.vstart[c] | ((uint32_t)a.vcap[c] << 16);
    };
    const auto gq_at = [&](int pos) { return a.ap.guid[rowcell + min(pos, W - 1)]; };
    ChainWords rs;
    ChainP2 gq;
    if constexpr (VAR) rs.init(rs_at, lane);
    if constexpr (ADP) gq.init(gq_at, lane, a.ap, a.p2, 0u);
    if constexpr (COST) {
        const uint64_t c0 = crow[0];
#pragma unroll
        for (int e = 0; e < 2 * NP; e++) win[e] = c0;
        bl0 = clrow[min(lane, W - 1)];
        br0 = crow[min(lane + 1, W - 1)];
        bl1 = clrow[min(64 + lane, W - 1)];
        br1 = crow[min(65 + lane, W - 1)];
    }
#pragma unroll
    for (int j = 0; j < PF; j++) {
        const int xj = min(j, W - 1);
        if constexpr (!COST) ld_raw_w<NP>(a.raw + (rowcell + xj) * cs + lol, rb[j]);
    }
    // the step loop's stores do not depend on a condition but for padded lanes (see bstore_*); the column W-2
    // state for the border replay is kept and stored after the loop
    const bool st_raw = MASKED ? store_lane : true;
    Cell<NP> colv;
    uint32_t colm = 0;
    auto body = [&](int x, int j, uint64_t& bl, uint64_t& br) {
        Cell<NP> raw;
        const int xp = min(x + PF, W - 1);
        uint32_t el = kInit2, er = kInit2;
        if constexpr (VAR) {
            const uint32_t r = rs.at(x);
            rs.advance(x, rs_at, lane);
            reframe<NP, VAR>(state, from, (int)(r & 0xffffu), (int)(r >> 16), lane, el, er);
            make_mask<NP>(lane, (int)(r >> 16), mask, active);
        }
        if constexpr (COST) {  // the cost of this cell, and its raw bytes for the later sweeps
            const int q = x & 63;
            const auto rl64 = [&](uint64_t v) {
                return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), q) << 32) |
                       (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, q);
            };
            const uint64_t cl = rl64(bl), cr1 = rl64(br);
            uint32_t bytes[NP / 2 + (NP == 1)];
            // byte path volumes have no padding planes (cap = D), so an unmasked wave's planes are all below D
            // the window rotates with the step (x % (2 NP) == j % (2 NP): step groups start at multiples of PF)
            const int rot = (2 * NP - j % (2 * NP)) % (2 * NP);
            census_cost<NP, PV && !MASKED>(cl, win, a.D, lo, raw, bytes, rot);
            uint8_t* o = a.raw_out + (rowcell + x) * cs + lo;
            if (st_raw) {
                if constexpr (NP == 1) *(uint16_t*)o = (uint16_t)bytes[0];
                else {
#pragma unroll
                    for (int k = 0; k < NP / 2; k++) ((uint32_t*)o)[k] = bytes[k];
                }
            }
            census_advance_rot<NP>(win, cr1, rot);
        } else {
            widen<NP>(rb[j], raw);
            ld_raw_w<NP>(a.raw + (rowcell + xp) * cs + lol, rb[j]);
        }
        uint32_t p2 = a.p2;
        if constexpr (ADP) {
            p2 = gq.at(x);
            gq.advance(x, gq_at, lane, a.ap, a.p2);
        }
        Cell<NP> lc;
        minprev = path_step<NP, MASKED>(state, minprev, raw, a.p1x2, p2, mask, lc, el, er);
        Cell<NP> out;
#pragma unroll
        for (int k = 0; k < NP; k++) out.v[k] = x != W - 1 ? lc.v[k] : 0u;
        if constexpr (PV) store_bytes<NP>(a.pv[0] + (rowcell + x) * a.pvs + lo, MASKED ? active : true, out);
        else store_u16<NP>(a.agg + (rowcell + x) * cs + lo, MASKED ? active : true, out);
        if (x == W - 2) {
            colv = lc;
            colm = minprev;
        }
        state = lc;
    };
    int x0 = 0;
    if constexpr (COST) {
        // blocks of 64 cells alternate between the two register sets, so each step reads its block's codes from
        // registers fixed at compile time; a set is reloaded in place with the block after next when its block ends
        auto half = [&](int xa, uint64_t& bl, uint64_t& br) {
            int i = xa;
            const int ie = min(xa + 64, W);
            for (; i + PF <= ie; i += PF) {
#pragma unroll
                for (int j = 0; j < PF; j++) body(i + j, j, bl, br);
            }
#pragma unroll
            for (int j = 0; j < PF; j++)
                if (i + j < ie) body(i + j, j, bl, br);
            bl = clrow[min(xa + 128 + lane, W - 1)];
            br = crow[min(xa + 129 + lane, W - 1)];
        };
        for (; x0 < W; x0 += 128) {
            half(x0, bl0, br0);
            if (x0 + 64 < W) half(x0 + 64, bl1, br1);
        }
    } else {
        for (; x0 + PF <= W; x0 += PF) {
#pragma unroll
            for (int j = 0; j < PF; j++) body(x0 + j, j, bl0, br0);
        }
#pragma unroll
        for (int j = 0; j < PF; j++)
            if (x0 + j < W) body(x0 + j, j, bl0, br0);
    }
    if (W >= 2) {  // the column W-2 states for the border replay
        store_u16<NP>(a.cols + (size_t)y * cs + lo, store_lane, colv);
        if (lane == 0) a.colmins[y] = (uint16_t)colm;
    }
}
template <int NP, bool MASKED, bool COST, bool VAR, bool ADP>
__global__ __launch_bounds__(kBlock) void sgm_row_fwd(RowArgs a)
{
    row_fwd_wave<NP, MASKED, COST, VAR, ADP>(
        a, __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)), threadIdx.x & 63);
}
// Pass +1's row path of a request with byte path volumes: census cost fused, costs into pv[0].
template <int NP, bool MASKED>
__global__ __launch_bounds__(kBlock) void sgm_row_fwd_pv(RowArgs a)
{
    row_fwd_wave<NP, MASKED, true, false, false, true>(
        a, __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)), threadIdx.x & 63);
}

// Pass -1, path 0 (<-), the last aggregate contribution to every cell but the border column 0, with the
// winner-take-all fused: one wavefront per row, x = W-1 .. 1.
//  dL(x): first minimum over planes [0, D) of the final aggregate (DenseStereoSgm.cs:225-231), as the
//         minimum of the keys agg << 16 | d.
//  sub(x): parabola through agg(d-1), agg(d), agg(d+1) in float32 (correctly rounded division, no
//         contraction), no offset at the range ends (CrossViewCost.cs:197-202).
//  dR(xr) = first minimum over d of agg(xr+d, d): a shift register of keys along the planes. After cell x,
//         register plane p holds min over j >= p of key(x-p+j, j), i.e. dR candidates of xr = x-p, so
//         plane 0 completes xr = x. Each step shifts it down one plane (v_mov inside a lane, DPP wave_shl
//         across lanes) and folds in the cell's keys. Plane 1 after x = 1 is dR(0) without cell 0.
// Saves the column-1 states for the border replay. WTA = false: the path alone, its cost added to a.agg (the
// rest of the sum and the winner-take-all are sgm_finish's). PV: the final aggregate is the sum of the seven byte
// path volumes, this path's cost and, in column W-1, pass +1's border costs (bagg_plus); agg is only written (when
// asked for), never read.
template <int NP, bool MASKED, bool VAR, bool ADP, bool WTA, bool PV = false, int WAGG = -1>
__device__ __forceinline__ void row_bwd_wave(const RowArgs& a, int y, int lane)
{
    static_assert(!PV || (WTA && !VAR && !ADP), "byte path volumes: uniform constant-P2 requests");
    constexpr int PF = 8, E = 2 * NP, NV = PV ? 7 : 0;
    // byte path volumes: each cell's census cost is computed again from the codes, as the forward row path does,
    // instead of read from the raw volume (a window of right codes that moves one code a step, towards x = 0)
    constexpr bool COST = PV;
    if (y >= a.H) return;
    const int W = a.W, cs = a.cs, lo = lane * E;
    int D = a.D;
    Cell<NP> mask;
    bool active;
    make_mask<NP>(lane, a.cap, mask, active);
    Cell<NP> state;
    set_init(state);
    uint32_t minprev = 2000;
    uint32_t R[E];
#pragma unroll
    for (int e = 0; e < E; e++) R[e] = 0xffffffffu;
    const size_t rowcell = (size_t)y * W;
    const int lol = (VAR ? lo < cs : active) ? lo : 0;
    // the row's base addresses are wave-uniform (scalar registers) and each load adds a 32-bit lane offset below
    // W * cs, so the loads take their address as base + offset with no 64-bit arithmetic per lane
    const uint8_t* rawrow = COST ? nullptr : a.raw + rowcell * cs;
    const uint16_t* aggrow = a.agg + rowcell * cs;
    const uint8_t* pvrow[NV > 0 ? NV : 1];
#pragma unroll
    for (int v = 0; v < NV; v++) pvrow[v] = a.pv[v] + rowcell * a.pvs;
    RawW<NP> rb[COST ? 1 : PF];
    Cell<NP> ab[PF];
    RawW<NP> vb[PF][NV > 0 ? NV : 1];  // PV: the seven path volumes' bytes of the cells in flight
    // COST: the lane's window win[e] = cr(max(x - lo - e, 0)), and per step i (x = W-1-i) the wave-uniform codes
    // cl(x) and cr(max(x - 64*E, 0)) (the code lane 63 takes on the move to x-1), 64 steps a register (lane k: step
    // i0 + k) in two register sets that alternate between blocks of 64 steps, read out with v_readlane
    const uint64_t* crow = COST ? a.cr + rowcell : nullptr;
    const uint64_t* clrow = COST ? a.cl + rowcell : nullptr;
    uint64_t win[E];
    uint64_t bl0 = 0, bf0 = 0, bl1 = 0, bf1 = 0;
    if constexpr (COST) {
#pragma unroll
        for (int e = 0; e < E; e++) win[e] = crow[max(W - 1 - lo - e, 0)];
        bl0 = clrow[max(W - 1 - lane, 0)];
        bf0 = crow[max(W - 1 - lane - 64 * E, 0)];
        bl1 = clrow[max(W - 1 - (64 + lane), 0)];
        bf1 = crow[max(W - 1 - (64 + lane) - 64 * E, 0)];
    }
    const auto rl64 = [&](uint64_t v, int q) {
        return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), q) << 32) |
               (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, q);
    };
    Range from;
    // The step loop's stores do not depend on a condition (which would cost the prefetches their early issue, as
    // bstore_* explains): the aggregate is stored by every step or by none (WAGG, or a.write_agg: one branch for the
    // whole row), and the winner-take-all outputs leave once per block of 64 steps, after the block's loop.
    const bool wagg = WAGG >= 0 ? WAGG != 0 : (!WTA || a.write_agg);
    const __amdgpu_buffer_rsrc_t ragg = make_rsrc(wagg ? a.agg + rowcell * cs : nullptr, wagg ? (uint32_t)(W * cs * 2) : 0u);
    const __amdgpu_buffer_rsrc_t rsub = make_rsrc(WTA && a.sub ? a.sub + rowcell : nullptr, WTA && a.sub ? (uint32_t)(W * 4) : 0u);
    // PV: pass +1's border-column costs of this row, added at x = W-1 (loaded ahead of the loop)
    Cell<NP> eplus;
    if constexpr (PV) load_u16<NP>(a.bagg_plus + (size_t)y * cs + lo, lo < cs, eplus);
    // VAR / ADP: range (start | cap << 16), count and guidance words by position W-1-x, 64 a VGPR (ChainWords)
    const auto cell_at = [&](int pos) { return rowcell + max(W - 1 - pos, 0); };
    const auto rs_at = [&](int pos) {
        const size_t c = cell_at(pos);
        return (uint32_t)a.vstart[c] | ((uint32_t)a.vcap[c] << 16);
    };
    const auto rn_at = [&](int pos) { return (uint32_t)a.vcount[cell_at(pos)]; };
    const auto gq_at = [&](int pos) { return a.ap.guid[cell_at(pos)]; };
    ChainWords rs, rn;
    ChainP2 gq;
    if constexpr (VAR) {
        rs.init(rs_at, lane);
        rn.init(rn_at, lane);
    }
    if constexpr (ADP) gq.init(gq_at, lane, a.ap, a.p2, 0u);
#pragma unroll
    for (int j = 0; j < PF; j++) {
        const int xj = max(W - 1 - j, 0);
        if constexpr (!COST) ld_raw_w<NP>(rawrow + (uint32_t)(xj * cs + lol), rb[j]);
        if constexpr (PV) {
#pragma unroll
            for (int v = 0; v < NV; v++) ld_raw_w<NP>(pvrow[v] + (uint32_t)(xj * a.pvs + lol), vb[j][v]);
        } else {
            ld_u16<NP>(aggrow + (uint32_t)(xj * cs + lol), ab[j]);
        }
    }
    // lane x&63 keeps cell x's winner and the parabola's inputs; the float arithmetic of the subpixel runs once per
    // 64 cells, lane-parallel, when they are stored (the same operations on the same values as per cell)
    uint32_t myd = 0, myr = 0;
    int myl = 0, myc = 0, myrr = 0;
    bool mypar = false;
    auto body = [&](int x, int j, int xb, uint64_t bl, uint64_t bf) {
        Cell<NP> raw;
        if constexpr (COST) {
            const int q = (W - 1 - x) & 63;
            uint32_t bytes[NP / 2 + (NP == 1)];
            const int rot = j % (2 * NP);  // the step (W - 1 - x) is j modulo PF: blocks start at multiples of 64
            census_cost<NP, !MASKED>(rl64(bl, q), win, D, lo, raw, bytes, rot);  // COST is PV: D = 128 * NP unmasked
            census_retreat_rot<NP>(win, rl64(bf, q), rot);
        } else {
            widen<NP>(rb[j], raw);
        }
        Cell<NP> agg;
        if constexpr (PV) {  // the other seven paths' costs (and pass +1's border-column replay in column W-1)
#pragma unroll
            for (int k = 0; k < NP; k++) agg.v[k] = 0u;
#pragma unroll
            for (int v = 0; v < NV; v++) add_bytes<NP>(vb[j][v], agg);
            if (x == W - 1) {
#pragma unroll
                for (int k = 0; k < NP; k++) agg.v[k] += eplus.v[k];
            }
        } else {
            agg = ab[j];
        }
        const int xp = max(x - PF, 0);
        uint32_t el = kInit2, er = kInit2;
        int s0 = 0;
        if constexpr (VAR) {
            const uint32_t r = rs.at(W - 1 - x);
            D = (int)rn.at(W - 1 - x);
            rs.advance(W - 1 - x, rs_at, lane);
            rn.advance(W - 1 - x, rn_at, lane);
            s0 = (int)(r & 0xffffu);
            reframe<NP, VAR>(state, from, s0, (int)(r >> 16), lane, el, er);
            make_mask<NP>(lane, (int)(r >> 16), mask, active);
        }
        uint32_t p2 = a.p2;
        if constexpr (ADP) {
            p2 = gq.at(W - 1 - x);
            gq.advance(W - 1 - x, gq_at, lane, a.ap, a.p2);
        }
        if constexpr (!COST) ld_raw_w<NP>(rawrow + (uint32_t)(xp * cs + lol), rb[j]);
        if constexpr (PV) {
#pragma unroll
            for (int v = 0; v < NV; v++) ld_raw_w<NP>(pvrow[v] + (uint32_t)(xp * a.pvs + lol), vb[j][v]);
        } else {
            ld_u16<NP>(aggrow + (uint32_t)(xp * cs + lol), ab[j]);
        }
        Cell<NP> lc, fin;
        minprev = path_step<NP, MASKED>(state, minprev, raw, a.p1x2, p2, mask, lc, el, er);
        state = lc;  // at x = 1, the column-1 state the border replay reads (stored after the loop)
#pragma unroll
        for (int k = 0; k < NP; k++) fin.v[k] = padd(agg.v[k], lc.v[k]);
        if constexpr (WAGG == 1 || !WTA) bstore_cell<NP>(fin, ragg, (uint32_t)((x * cs + lo) * 2), active);
        else if constexpr (WAGG < 0) {
            if (wagg) bstore_cell<NP>(fin, ragg, (uint32_t)((x * cs + lo) * 2), active);
        }
        if constexpr (!WTA) return;
        // winner-take-all on the final aggregate of cell x
        uint32_t key[E];
        uint32_t best = 0xffffffffu;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int d = lo + e;
            const uint32_t v = (e & 1) ? (fin.v[e >> 1] >> 16) : (fin.v[e >> 1] & 0xffffu);
            key[e] = (PV && !MASKED) || d < D ? ((v << 16) | (uint32_t)d) : 0xffffffffu;  // PV unmasked: D = 128 * NP
            best = key[e] < best ? key[e] : best;
        }
        uint32_t rdone = 0;
        if constexpr (!VAR) {  // the right-image disparity needs one plane range for the whole row
            const uint32_t carry = shl1(R[0], 0xffffffffu);
#pragma unroll
            for (int e = 0; e < E - 1; e++) R[e] = R[e + 1];
            R[E - 1] = carry;
#pragma unroll
            for (int e = 0; e < E; e++) R[e] = key[e] < R[e] ? key[e] : R[e];
            rdone = (uint32_t)__builtin_amdgcn_readlane((int)R[0], 0);
        }
        best = wave_min(best);
        const int d = best == 0xffffffffu ? 0 : (int)(best & 0xffffu);  // no planes: index 0 (:227)
        const bool par = d > 0 && d < D - 1;
        const int l = (int)plane_of<NP>(fin, par ? d - 1 : 0), r = (int)plane_of<NP>(fin, par ? d + 1 : 0);
        if (lane == x - xb) {
            myd = (uint32_t)(d + s0);  // the plane index start + arg (:230)
            myr = rdone & 0xffffu;
            mypar = par;
            myl = l;
            myc = (int)(best >> 16);
            myrr = r;
        }
    };
    // the block's 64 pixels xb .. xb+63 (lane l: pixel xb + l; lanes [l0, l1] hold results), coalesced
    auto flush = [&](int xb, int l0, int l1) {
        if (lane < l0 || lane > l1) return;
        const size_t o = rowcell + xb + lane;
        a.disp[o] = (uint16_t)myd;
        float v = (float)(int)myd;  // the parabola through agg(d-1), agg(d), agg(d+1)
        const int den = myl - 2 * myc + myrr;
        if (mypar && den > 0) v = (float)(int)myd + __fdiv_rn(__fmul_rn(0.5f, (float)(myl - myrr)), (float)den);
        bstore_f32(v, rsub, (uint32_t)(xb + lane) * 4u);  // dropped when the caller wants no subpixel
        if (!VAR) a.dr[o] = (uint16_t)myr;
    };
    // x = W-1 .. 1: (W-1) steps, in blocks of 64 steps counted from x = W-1 (PF divides 64, so every block starts on
    // the same prefetch slot); a block's loop holds only the prefetch loads and, when asked, the aggregate stores
    const int n = W - 1;
    if constexpr (COST) {
        // block ib: steps ib .. min(ib + 64, n) - 1 with the codes of register set (bl, bf), which is then reloaded in
        // place with block ib + 128's codes
        auto block = [&](int ib, uint64_t& bl, uint64_t& bf) {
            const int xb = W - 64 - ib, ie = min(ib + 64, n);
            int i = ib;
            for (; i + PF <= ie; i += PF) {
#pragma unroll
                for (int j = 0; j < PF; j++) body(W - 1 - (i + j), j, xb, bl, bf);
            }
#pragma unroll
            for (int j = 0; j < PF; j++)
                if (i + j < ie) body(W - 1 - (i + j), j, xb, bl, bf);
            if (ie == ib + 64) flush(xb, 0, 63);
            else flush(xb, 1 - xb, W - 1 - ib - xb);  // pixels 1 .. W-1-ib
            bl = clrow[max(W - 1 - (ib + 128 + lane), 0)];
            bf = crow[max(W - 1 - (ib + 128 + lane) - 64 * E, 0)];
        };
        for (int ib = 0; ib < n; ib += 128) {
            block(ib, bl0, bf0);
            if (ib + 64 < n) block(ib + 64, bl1, bf1);
        }
    } else {
        int i0 = 0;
        if constexpr (WTA) {
            for (; i0 + 64 <= n; i0 += 64) {
                const int xb = W - 64 - i0;
                for (int i8 = i0; i8 < i0 + 64; i8 += PF) {
#pragma unroll
                    for (int j = 0; j < PF; j++) body(W - 1 - (i8 + j), j, xb, 0, 0);
                }
                flush(xb, 0, 63);
            }
        }
        // the rest (the last block's fewer than 64 steps, or the whole row without the winner-take-all)
        const int xb = W - 64 - i0, ilast = i0;
        for (; i0 + PF <= n; i0 += PF) {
#pragma unroll
            for (int j = 0; j < PF; j++) body(W - 1 - (i0 + j), j, xb, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < PF; j++)
            if (i0 + j < n) body(W - 1 - (i0 + j), j, xb, 0, 0);
        if constexpr (WTA) {
            if (n > ilast) flush(xb, 1 - xb, W - 1 - ilast - xb);  // pixels 1 .. W-1-ilast
        }
    }
    if (n > 0) {  // the column-1 states for the border replay
        store_u16<NP>(a.cols + (size_t)y * cs + lo, lo < cs, state);
        if (lane == 0) a.colmins[y] = (uint16_t)minprev;
    }
    if constexpr (WTA && !VAR) {
        const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)R[1], 0);
        if (lane == 0) a.dr0[y] = r1;
    }
}
template <int NP, bool MASKED, bool VAR, bool ADP>
__global__ __launch_bounds__(kBlock) void sgm_row_bwd_wta(RowArgs a)
{
    row_bwd_wave<NP, MASKED, VAR, ADP, true>(
        a, __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)), threadIdx.x & 63);
}
// Pass -1's row path of a request with byte path volumes: the eight paths' sum and the winner-take-all; WAGG: the
// final aggregate is stored too.
template <int NP, bool MASKED, int WAGG>
__global__ __launch_bounds__(kBlock) void sgm_row_bwd_wta_pv(RowArgs a)
{
    row_bwd_wave<NP, MASKED, false, false, true, true, WAGG>(
        a, __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)), threadIdx.x & 63);
}

// Column 0 (after the pass -1 replay) and the validity of every pixel of a row, one wavefront per row:
// valid(x) = x - dL >= 0 && |dL - dR(x - dL)| <= tol (RealStereo.cs:418-421).
template <int NP, bool VAR>
__device__ __forceinline__ void wta_finish_wave(const RowArgs& a, int tol, uint8_t* valid, int y, int lane)
{
    const size_t rowcell = (size_t)y * a.W;
    const int W = a.W, cs = a.cs, D = VAR ? (int)a.vcount[rowcell] : a.D, lo = lane * 2 * NP;
    const int s0 = VAR ? (int)a.vstart[rowcell] : 0;
    const bool active = lo < cs;
    Cell<NP> c;
    if (a.bagg_minus) {  // byte path volumes: column 0 is pass +1's four paths and pass -1's border replay
        load_u16<NP>(a.bagg_minus + (size_t)y * cs + lo, active, c);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            RawW<NP> w;
            load_raw_w<NP>(a.pv[v] + rowcell * a.pvs + lo, active, w);
            add_bytes<NP>(w, c);
        }
        if (a.write_agg) store_u16<NP>(a.agg + rowcell * cs + lo, active, c);
    } else {
        load_u16<NP>(a.agg + rowcell * cs + lo, active, c);
    }
    uint32_t best = 0xffffffffu;
#pragma unroll
    for (int e = 0; e < 2 * NP; e++) {
        const int d = lo + e;
        const uint32_t v = (e & 1) ? (c.v[e >> 1] >> 16) : (c.v[e >> 1] & 0xffffu);
        const uint32_t key = d < D ? ((v << 16) | (uint32_t)d) : 0xffffffffu;
        best = key < best ? key : best;
    }
    best = wave_min(best);
    const int d = best == 0xffffffffu ? 0 : (int)(best & 0xffffu);  // no planes: index 0 (:227)
    float v = (float)(d + s0);
    if (d > 0 && d < D - 1) {
        const int l = (int)plane_of<NP>(c, d - 1), cc = (int)(best >> 16), r = (int)plane_of<NP>(c, d + 1);
        const int den = l - 2 * cc + r;
        if (den > 0) v = (float)(d + s0) + __fdiv_rn(__fmul_rn(0.5f, (float)(l - r)), (float)den);
    }
    if (lane == 0) {
        a.disp[rowcell] = (uint16_t)(d + s0);
        if (a.sub) a.sub[rowcell] = v;
    }
    if constexpr (VAR) return;  // no right-image disparity with per-pixel ranges
    const uint32_t k00 = (plane_of<NP>(c, 0) << 16);
    const uint32_t part = a.dr0[y];
    if (lane == 0) a.dr[rowcell] = (uint16_t)((k00 < part ? k00 : part) & 0xffffu);
    __syncthreads();
    if (valid)
        for (int x = lane; x < W; x += 64) {
            const int dd = a.disp[rowcell + x], xr = x - dd;
            valid[rowcell + x] = (uint8_t)(xr >= 0 && abs(dd - (int)a.dr[rowcell + xr]) <= tol);
        }
}
template <int NP, bool VAR>
__global__ __launch_bounds__(64) void wta_finish(RowArgs a, int tol, uint8_t* valid)
{
    wta_finish_wave<NP, VAR>(a, tol, valid, blockIdx.x, threadIdx.x);
}

// Ranged layers after the fused passes: the final aggregate, agg + the three path volumes (saturating, so the
// order of the additions does not matter), stored when asked, and its first-minimum winner-take-all and subpixel
// (as sgm_row_bwd_wta and wta_finish compute them) for every cell, border columns included. One wavefront per 64
// cells of a row; the results are stored coalesced.
template <int NP>
__global__ __launch_bounds__(kBlock) void sgm_finish(RowArgs a, const uint16_t* __restrict__ v0,
                                                     const uint16_t* __restrict__ v1, const uint16_t* __restrict__ v2)
{
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * kBlock + threadIdx.x) >> 6));
    const int nb = (a.W + 63) >> 6;
    const int y = wv / nb;
    if (y >= a.H) return;
    const int xb = (wv - y * nb) << 6, n = min(64, a.W - xb);
    const size_t c0 = (size_t)y * a.W + xb;
    const int cs = a.cs, lo = lane * 2 * NP;
    // the 64 cells' ranges, one a lane, read out with v_readlane
    const uint32_t rs = lane < n ? ((uint32_t)a.vstart[c0 + lane] | ((uint32_t)a.vcap[c0 + lane] << 16)) : 0u;
    const uint32_t rn = lane < n ? (uint32_t)a.vcount[c0 + lane] : 0u;
    // lane j keeps cell j's winner and the parabola's inputs; the subpixel's float arithmetic runs once at the end,
    // lane-parallel (the same operations on the same values as per cell)
    uint32_t myd = 0;
    int myl = 0, myc = 0, myrr = 0;
    bool mypar = false;
    for (int j = 0; j < n; j++) {
        const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)rs, j);
        const int s0 = (int)(r & 0xffffu), cap = (int)(r >> 16), D = __builtin_amdgcn_readlane((int)rn, j);
        // lanes past the cell's cap read lane 0's planes (no bytes of their own fetched): none of them is stored
        // or reaches the minimum
        const size_t off = (c0 + j) * cs + (lo < cap ? lo : 0);
        Cell<NP> t, u0, u1, u2;
        ld_u16<NP>(a.agg + off, t);
        ld_u16<NP>(v0 + off, u0);
        ld_u16<NP>(v1 + off, u1);
        ld_u16<NP>(v2 + off, u2);
#pragma unroll
        for (int k = 0; k < NP; k++) t.v[k] = padd(padd(t.v[k], u0.v[k]), padd(u1.v[k], u2.v[k]));
        if (a.write_agg) store_u16<NP>(a.agg + (c0 + j) * cs + lo, lo < cap, t);
        uint32_t best = 0xffffffffu;
#pragma unroll
        for (int e = 0; e < 2 * NP; e++) {
            const int d = lo + e;
            const uint32_t v = (e & 1) ? (t.v[e >> 1] >> 16) : (t.v[e >> 1] & 0xffffu);
            const uint32_t key = d < D ? ((v << 16) | (uint32_t)d) : 0xffffffffu;
            best = key < best ? key : best;
        }
        best = wave_min(best);
        const int d = best == 0xffffffffu ? 0 : (int)(best & 0xffffu);  // no planes: index 0 (:227)
        const bool par = d > 0 && d < D - 1;
        const int l = (int)plane_of<NP>(t, par ? d - 1 : 0), rr = (int)plane_of<NP>(t, par ? d + 1 : 0);
        if (lane == j) {
            myd = (uint32_t)(d + s0);
            mypar = par;
            myl = l;
            myc = (int)(best >> 16);
            myrr = rr;
        }
    }
    if (lane < n) {
        a.disp[c0 + lane] = (uint16_t)myd;
        if (a.sub) {  // the parabola through agg(d-1), agg(d), agg(d+1)
            float v = (float)(int)myd;
            const int den = myl - 2 * myc + myrr;
            if (mypar && den > 0) v = (float)(int)myd + __fdiv_rn(__fmul_rn(0.5f, (float)(myl - myrr)), (float)den);
            a.sub[c0 + lane] = v;
        }
    }
}

// ------------------------------------------------------------------ segmented chains (narrow per-pixel ranges)
// With per-pixel ranges whose caps are at most 2*SEG planes (SEG = 8 or 16), one wavefront runs 64/SEG chains, SEG
// lanes each (one packed plane pair per lane, as the VAR kernels above). Everything the unsegmented kernels keep
// wave-uniform (the cell's range, minPrev, the chain position) is per segment here; the lane shifts of the path step
// and its minimum stay inside a segment (DPP row shifts and mirrors: a segment never straddles a row of 16 lanes),
// and realignment and plane reads use ds_bpermute inside the segment. The data layout is the VAR one, so the border
// replays and the column-0 finish are shared.
template <int SEG>
__device__ __forceinline__ uint32_t seg_shr1(uint32_t v, uint32_t edge, int li)  // lane i <- lane i-1; first: edge
{
    const uint32_t t = (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    return li == 0 ? edge : t;
}
template <int SEG>
__device__ __forceinline__ uint32_t seg_shl1(uint32_t v, uint32_t edge, int li)  // lane i <- lane i+1; last: edge
{
    const uint32_t t = (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)v, 0x101, 0xf, 0xf, false);  // row_shl:1
    return li == SEG - 1 ? edge : t;
}
// Minimum over the segment, in every lane of it: quads, then half-rows (8 lanes), then rows (16 lanes).
template <int SEG>
__device__ __forceinline__ uint32_t seg_min(uint32_t v)
{
    static_assert(SEG == 8 || SEG == 16, "segments of 8 or 16 lanes");
    uint32_t o = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xb1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
    v = v < o ? v : o;
    o = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4e, 0xf, 0xf, false);  // quad_perm [2,3,0,1]
    v = v < o ? v : o;
    o = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xf, 0xf, false);  // row_half_mirror
    v = v < o ? v : o;
    if constexpr (SEG == 16) {
        o = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xf, 0xf, false);  // row_mirror
        v = v < o ? v : o;
    }
    return v;
}
// Lane `src` (0 .. SEG-1) of this lane's segment.
template <int SEG>
__device__ __forceinline__ uint32_t seg_get(uint32_t v, int src, int lane)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute(((lane & ~(SEG - 1)) + (src & (SEG - 1))) << 2, (int)v);
}
// Plane r of the predecessor (start from.s, cap cp) in the successor's frame: 2000 outside [0, cp).
template <int SEG>
__device__ __forceinline__ uint32_t seg_rel_plane(uint32_t v, int r, int cp, int lane)
{
    const uint32_t w = seg_get<SEG>(v, r >> 1, lane);
    const uint32_t p = (r & 1) ? (w >> 16) : (w & 0xffffu);
    return (r >= 0 && r < cp) ? p : 2000u;
}
// realign() inside a segment, for every segment at once (delta may differ per segment; 0 is the identity).
template <int SEG>
__device__ __forceinline__ void seg_reframe(uint32_t& st, Range& from, int s, int c, int li, int lane, uint32_t& edge_l,
                                            uint32_t& edge_r)
{
    const int delta = s - from.s, cp = from.c;
    if (!__ballot(delta != 0)) {  // no segment moves its start: the identity (planes past cp already hold 2000)
        edge_l = edge_r = kInit2;
        from.c = c;
        return;
    }
    const uint32_t v = st;
    const uint32_t even = seg_get<SEG>(v, li + (delta >> 1), lane);
    const int sa = li + ((delta - 1) >> 1);
    const uint32_t a = seg_get<SEG>(v, sa, lane), b = seg_get<SEG>(v, sa + 1, lane);
    const uint32_t out = (delta & 1) ? __builtin_amdgcn_alignbit(b, a, 16) : even;
    const int r0 = 2 * li + delta;
    const bool ok0 = r0 >= 0 && r0 < cp, ok1 = r0 + 1 >= 0 && r0 + 1 < cp;
    st = (ok0 ? (out & 0xffffu) : 2000u) | (ok1 ? (out & 0xffff0000u) : (2000u << 16));
    const uint32_t el = seg_rel_plane<SEG>(v, delta - 1, cp, lane), er = seg_rel_plane<SEG>(v, 2 * SEG + delta, cp, lane);
    edge_l = el | (el << 16);
    edge_r = er | (er << 16);
    from.s = s;
    from.c = c;
}
// path_step<1, MASKED> inside a segment; returns the segment's new minPrev (in every lane of it).
template <int SEG>
__device__ __forceinline__ uint32_t seg_path_step(uint32_t prev, uint32_t minprev, uint32_t raw, uint32_t p1x2, uint32_t p2,
                                                  uint32_t mask, uint32_t& lc, uint32_t edge_l, uint32_t edge_r, int li)
{
    uint32_t p2p = p2 + minprev;
    p2p = p2p > 65535u ? 65535u : p2p;
    const uint32_t p2x2 = p2p | (p2p << 16), minx2 = minprev | (minprev << 16);
    const uint32_t left = seg_shr1<SEG>(prev, edge_l, li), right = seg_shl1<SEG>(prev, edge_r, li);
    const uint32_t lm = __builtin_amdgcn_alignbit(prev, left, 16), lp = __builtin_amdgcn_alignbit(right, prev, 16);
    uint32_t m = pmin(padd(pmin(lm, lp), p1x2), prev);
    m = pmin(m, p2x2);
    uint32_t t = psub(padd(raw, m), minx2);
    t = (t & mask) | (kInit2 & ~mask);
    const uint32_t run = t | ~mask;
    lc = t;
    const uint32_t lo = run & 0xffffu, hi = run >> 16;
    return seg_min<SEG>(lo < hi ? lo : hi);
}
__device__ __forceinline__ uint32_t seg_mask(int li, int cap)
{
    return (2 * li < cap ? 0xffffu : 0u) | (2 * li + 1 < cap ? 0xffff0000u : 0u);
}

// sgm_sweep<1, DIR, P, FIRST, true, true, false> with 64/SEG chains per wavefront.
// t: the lane's index in the path's grid (64/SEG chains per wavefront).
template <int SEG, int DIR, int P, bool FIRST>
__device__ __forceinline__ void sweep_seg_wave(const SweepArgs& a, int t, int lane)
{
    constexpr int DX = DIR * (P == 0 ? 1 : P == 1 ? 1 : P == 2 ? 0 : -1);
    constexpr int DY = DIR * (P == 0 ? 0 : 1);
    constexpr int PF = 8;
    const int li = lane & (SEG - 1);
    const int chain = t / SEG;
    const int W = a.W, H = a.H, cs = a.cs;
    const int nch = chain_count<DX, DY>(W, H);
    int x = 0, y = 0, n = 0;
    if (chain < nch) chain_start<DX, DY>(chain, W, H, x, y, n);
    int nmax = n;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
    if (nmax == 0) return;
    const int bx = DIR > 0 ? W - 1 : 0;
    const bool replays = P == 2 && x == bx;  // the border column's vertical chain is the replay's
    const int lo = li * 2;
    const int lol = lo < cs ? lo : 0;
    uint32_t state = kInit2, minprev = 2000;
    Range from;
    if (DIR < 0 && P != 0 && n > 0 && y == H - 1) {  // pass -1's first row reads half 0 as pass +1 left it
        const int sc = x - DX;
        if (sc >= 0 && sc < W) {
            state = lo < cs ? *(const uint32_t*)(a.seeds + ((size_t)(P - 1) * W + sc) * cs + lo) : 0u;
            from.s = a.seedstart[(P - 1) * W + sc];
            from.c = a.seedcap[(P - 1) * W + sc];
            const uint32_t m = seg_mask(li, from.c);
            state = (state & m) | (kInit2 & ~m);
            minprev = a.seedmins[(P - 1) * W + sc];
        }
    }
    const ptrdiff_t step = (ptrdiff_t)DY * W + DX;
    const size_t cell0 = n > 0 ? (size_t)y * W + x : 0;
    const int nl = n > 0 ? n - 1 : 0;
    uint32_t rb[PF], ab[PF], rs[PF];
#pragma unroll
    for (int j = 0; j < PF; j++) {
        const size_t cj = cell0 + (ptrdiff_t)min(j, nl) * step;
        rb[j] = (uint32_t)(*(const uint16_t*)(a.raw + cj * cs + lol));
        ab[j] = FIRST ? 0u : *(const uint32_t*)(a.agg + cj * cs + lol);
        rs[j] = (uint32_t)a.vstart[cj] | ((uint32_t)a.vcap[cj] << 16);
    }
    size_t cell = cell0;
    auto body = [&](int i, int j) {
        const bool live = i < n;
        const uint32_t raw = __builtin_amdgcn_perm(0u, rb[j], 0x0c010c00u);
        const uint32_t agg = ab[j];
        const int cs_ = (int)(rs[j] & 0xffffu), cc_ = (int)(rs[j] >> 16);
        uint32_t el, er;
        seg_reframe<SEG>(state, from, cs_, cc_, li, lane, el, er);
        const uint32_t mask = seg_mask(li, cc_);
        const bool active = 2 * li < cc_;
        const size_t cp = cell0 + (ptrdiff_t)min(i + PF, nl) * step;
        rb[j] = (uint32_t)(*(const uint16_t*)(a.raw + cp * cs + lol));
        if (!FIRST) ab[j] = *(const uint32_t*)(a.agg + cp * cs + lol);
        rs[j] = (uint32_t)a.vstart[cp] | ((uint32_t)a.vcap[cp] << 16);
        uint32_t lc;
        minprev = seg_path_step<SEG>(state, minprev, raw, a.p1x2, a.p2, mask, lc, el, er, li);
        if (live && !replays) {
            if (x != bx) {
                if (active) *(uint32_t*)(a.agg + cell * cs + lo) = FIRST ? lc : padd(agg, lc);
                if (DIR > 0 && P != 0 && y == a.seed_row) {
                    if (lo < cs) *(uint32_t*)(a.seeds + ((size_t)(P - 1) * W + x) * cs + lo) = lc;
                    if (li == 0) {
                        a.seedmins[(P - 1) * W + x] = (uint16_t)minprev;
                        a.seedstart[(P - 1) * W + x] = (uint16_t)cs_;
                        a.seedcap[(P - 1) * W + x] = (uint16_t)cc_;
                    }
                }
            } else if (FIRST && active) {
                *(uint32_t*)(a.agg + cell * cs + lo) = 0u;
            }
            if (P <= 1 && x == bx - DIR) {  // the border replay reads this column's path 0 / 1 states
                if (lo < cs) *(uint32_t*)(a.cols + ((size_t)P * H + y) * cs + lo) = lc;
                if (li == 0) a.colmins[P * H + y] = (uint16_t)minprev;
            }
        } else if (live && FIRST && lo < cs) {  // the replay's chain: its own volume still needs the zeros
            *(uint32_t*)(a.agg + cell * cs + lo) = 0u;
        }
        state = lc;
        cell += step;
        x += DX;
        y += DY;
    };
    int i0 = 0;
    for (; i0 + PF <= nmax; i0 += PF) {
#pragma unroll
        for (int j = 0; j < PF; j++) body(i0 + j, j);
    }
#pragma unroll
    for (int j = 0; j < PF; j++)
        if (i0 + j < nmax) body(i0 + j, j);
}
template <int SEG, int DIR, int P, bool FIRST>
__global__ __launch_bounds__(kBlock) void sgm_sweep_seg(SweepArgs a)
{
    sweep_seg_wave<SEG, DIR, P, FIRST>(a, (int)(blockIdx.x * kBlock + threadIdx.x), threadIdx.x & 63);
}

// sgm_row_fwd<1, true, false, true, false> with 64/SEG rows per wavefront.
template <int SEG>
__device__ __forceinline__ void row_fwd_seg_wave(const RowArgs& a, int t, int lane)
{
    constexpr int PF = 8;
    const int li = lane & (SEG - 1);
    const int yr = t / SEG;
    if (__builtin_amdgcn_readfirstlane(yr - lane / SEG) >= a.H) return;  // the wave's first row: whole waves only
    const bool row_ok = yr < a.H;
    const int y = row_ok ? yr : a.H - 1;  // idle segments repeat the last row and store nothing
    const int W = a.W, cs = a.cs, lo = li * 2;
    const int lol = lo < cs ? lo : 0;
    uint32_t state = kInit2, minprev = 2000;
    Range from;
    const size_t rowcell = (size_t)y * W;
    uint32_t rb[PF], rs[PF];
#pragma unroll
    for (int j = 0; j < PF; j++) {
        const int xj = min(j, W - 1);
        rb[j] = (uint32_t)(*(const uint16_t*)(a.raw + (rowcell + xj) * cs + lol));
        rs[j] = (uint32_t)a.vstart[rowcell + xj] | ((uint32_t)a.vcap[rowcell + xj] << 16);
    }
    auto body = [&](int x, int j) {
        const int xp = min(x + PF, W - 1);
        const int s_ = (int)(rs[j] & 0xffffu), c_ = (int)(rs[j] >> 16);
        uint32_t el, er;
        seg_reframe<SEG>(state, from, s_, c_, li, lane, el, er);
        const uint32_t mask = seg_mask(li, c_);
        const bool active = 2 * li < c_;
        const uint32_t raw = __builtin_amdgcn_perm(0u, rb[j], 0x0c010c00u);
        rb[j] = (uint32_t)(*(const uint16_t*)(a.raw + (rowcell + xp) * cs + lol));
        rs[j] = (uint32_t)a.vstart[rowcell + xp] | ((uint32_t)a.vcap[rowcell + xp] << 16);
        uint32_t lc;
        minprev = seg_path_step<SEG>(state, minprev, raw, a.p1x2, a.p2, mask, lc, el, er, li);
        if (row_ok) {
            if (active) *(uint32_t*)(a.agg + (rowcell + x) * cs + lo) = x != W - 1 ? lc : 0u;
            if (x == W - 2) {
                if (lo < cs) *(uint32_t*)(a.cols + (size_t)y * cs + lo) = lc;
                if (li == 0) a.colmins[y] = (uint16_t)minprev;
            }
        }
        state = lc;
    };
    int x0 = 0;
    for (; x0 + PF <= W; x0 += PF) {
#pragma unroll
        for (int j = 0; j < PF; j++) body(x0 + j, j);
    }
#pragma unroll
    for (int j = 0; j < PF; j++)
        if (x0 + j < W) body(x0 + j, j);
}
template <int SEG>
__global__ __launch_bounds__(kBlock) void sgm_row_fwd_seg(RowArgs a)
{
    row_fwd_seg_wave<SEG>(a, (int)(blockIdx.x * kBlock + threadIdx.x), threadIdx.x & 63);
}

// sgm_row_bwd_wta<1, true, true, false> with 64/SEG rows per wavefront: the last contribution and the
// winner-take-all of every cell but column 0 (the first minimum over the cell's count, the parabola).
// WTA = false: the path alone, its cost added to a.agg (sgm_finish takes the winner-take-all).
template <int SEG, bool WTA>
__device__ __forceinline__ void row_bwd_seg_wave(const RowArgs& a, int t, int lane)
{
    constexpr int PF = 8;
    const int li = lane & (SEG - 1);
    const int yr = t / SEG;
    if (__builtin_amdgcn_readfirstlane(yr - lane / SEG) >= a.H) return;  // the wave's first row: whole waves only
    const bool row_ok = yr < a.H;
    const int y = row_ok ? yr : a.H - 1;
    const int W = a.W, cs = a.cs, lo = li * 2;
    const int lol = lo < cs ? lo : 0;
    uint32_t state = kInit2, minprev = 2000;
    Range from;
    const size_t rowcell = (size_t)y * W;
    uint32_t rb[PF], ab[PF], rs[PF], rn[PF];
#pragma unroll
    for (int j = 0; j < PF; j++) {
        const int xj = max(W - 1 - j, 0);
        rb[j] = (uint32_t)(*(const uint16_t*)(a.raw + (rowcell + xj) * cs + lol));
        ab[j] = *(const uint32_t*)(a.agg + (rowcell + xj) * cs + lol);
        rs[j] = (uint32_t)a.vstart[rowcell + xj] | ((uint32_t)a.vcap[rowcell + xj] << 16);
        rn[j] = a.vcount[rowcell + xj];
    }
    auto body = [&](int x, int j) {
        const int xp = max(x - PF, 0);
        const int s0 = (int)(rs[j] & 0xffffu), c_ = (int)(rs[j] >> 16), D = (int)rn[j];
        uint32_t el, er;
        seg_reframe<SEG>(state, from, s0, c_, li, lane, el, er);
        const uint32_t mask = seg_mask(li, c_);
        const bool active = 2 * li < c_;
        const uint32_t raw = __builtin_amdgcn_perm(0u, rb[j], 0x0c010c00u);
        const uint32_t agg = ab[j];
        rb[j] = (uint32_t)(*(const uint16_t*)(a.raw + (rowcell + xp) * cs + lol));
        ab[j] = *(const uint32_t*)(a.agg + (rowcell + xp) * cs + lol);
        rs[j] = (uint32_t)a.vstart[rowcell + xp] | ((uint32_t)a.vcap[rowcell + xp] << 16);
        rn[j] = a.vcount[rowcell + xp];
        uint32_t lc;
        minprev = seg_path_step<SEG>(state, minprev, raw, a.p1x2, a.p2, mask, lc, el, er, li);
        const uint32_t fin = padd(agg, lc);
        if (row_ok) {
            if ((!WTA || a.write_agg) && active) *(uint32_t*)(a.agg + (rowcell + x) * cs + lo) = fin;
            if (x == 1) {
                if (lo < cs) *(uint32_t*)(a.cols + (size_t)y * cs + lo) = lc;
                if (li == 0) a.colmins[y] = (uint16_t)minprev;
            }
        }
        state = lc;
        if constexpr (!WTA) return;
        const uint32_t k0 = lo < D ? (((fin & 0xffffu) << 16) | (uint32_t)lo) : 0xffffffffu;
        const uint32_t k1 = lo + 1 < D ? ((fin & 0xffff0000u) | (uint32_t)(lo + 1)) : 0xffffffffu;
        const uint32_t best = seg_min<SEG>(k0 < k1 ? k0 : k1);
        const int d = best == 0xffffffffu ? 0 : (int)(best & 0xffffu);  // no planes: index 0 (:227)
        float v = (float)(d + s0);
        const uint32_t wl = seg_get<SEG>(fin, (d - 1) >> 1, lane), wr = seg_get<SEG>(fin, (d + 1) >> 1, lane);
        if (d > 0 && d < D - 1) {
            const int l = (int)(((d - 1) & 1) ? (wl >> 16) : (wl & 0xffffu)), cc = (int)(best >> 16);
            const int r = (int)(((d + 1) & 1) ? (wr >> 16) : (wr & 0xffffu));
            const int den = l - 2 * cc + r;
            if (den > 0) v = (float)(d + s0) + __fdiv_rn(__fmul_rn(0.5f, (float)(l - r)), (float)den);
        }
        if (row_ok && li == 0) {
            a.disp[rowcell + x] = (uint16_t)(d + s0);  // the plane index start + arg (:230)
            if (a.sub) a.sub[rowcell + x] = v;
        }
    };
    const int n = W - 1;  // x = W-1 .. 1
    int i0 = 0;
    for (; i0 + PF <= n; i0 += PF) {
#pragma unroll
        for (int j = 0; j < PF; j++) body(W - 1 - (i0 + j), j);
    }
#pragma unroll
    for (int j = 0; j < PF; j++)
        if (i0 + j < n) body(W - 1 - (i0 + j), j);
}
template <int SEG>
__global__ __launch_bounds__(kBlock) void sgm_row_bwd_wta_seg(RowArgs a)
{
    row_bwd_seg_wave<SEG, true>(a, (int)(blockIdx.x * kBlock + threadIdx.x), threadIdx.x & 63);
}

// ------------------------------------------------------------------ fused vertical + diagonal paths (bands)
// Paths 1, 2, 3 of one pass in one kernel, so each cell's raw cost is read once and its aggregate
// read-modify-written once for the three. Rows advance in bands of kBandRows; a band is one launch, with one
// workgroup per kStrip-column strip. A strip computes its three paths row by row from the states the previous
// band left at its last row (a ping-pong boundary buffer). The diagonal paths also run over halo columns,
// one fewer per row (kBandRows-1 at the band's first row), so a strip needs no other strip of its band.
// Waves own kBandG consecutive columns each and swap their edge columns' states through LDS once per row.
constexpr int kBandRows = 8, kStrip = 16, kBandWaves = 8, kBandG = 4;
static_assert(kStrip + 2 * kBandRows <= kBandWaves * kBandG, "band halo does not fit the waves");

struct BandArgs {
    const uint8_t* raw;
    uint16_t* agg;
    int W, H, cs, cap;
    uint32_t p1x2, p2;
    int split;                    // bands of kBandRows cover rows [0, split) then [split, H) (in pass order)
    int strips;                   // tiles per band
    const uint16_t* bin0;         // [3][W][cs] states of the row before the first band (null: the initial fill)
    const uint16_t* bin0min;      // [3][W] (u16, the seed table's)
    uint16_t* bnd;                // [2][3][W][cs] boundary ping-pong: band b writes slot b & 1
    uint32_t* bndmin;             // [2][3][W]
    uint16_t* extra;              // the last row of band `extra_band` also goes here (the pass +1 seed table)
    uint16_t* extramin;
    int extra_band;
    uint16_t* cols;               // [2][H][cs] of this pass: slot 1 takes path 1 of the column next to the border
    uint16_t* colmins;
    int* ticket;                  // tile order (zeroed before the launch)
    int* flags;                   // [bands][strips]: epoch when the tile's boundary is written
    int epoch;
    int* err;                     // set when a wait gives up
};

// Band b, in pass order: its first row (distance from the pass's first row) and its row count.
__device__ __forceinline__ void band_rows(int b, int split, int H, int& first, int& rows)
{
    const int nsplit = (split + kBandRows - 1) / kBandRows;
    if (b < nsplit) {
        first = b * kBandRows;
        rows = min(kBandRows, split - first);
    } else {
        first = split + (b - nsplit) * kBandRows;
        rows = min(kBandRows, H - first);
    }
}

template <int NP, int DIR, bool MASKED>
__global__ __launch_bounds__(kBandWaves * 64) void sgm_band(BandArgs a)
{
    constexpr int G = kBandG, NW = kBandWaves;
    // path j (0, 1, 2) is the reference's path j+1 of pass DIR; its predecessor column is x - DIR, x, x + DIR
    constexpr int JL = DIR > 0 ? 0 : 2;  // the path reading column x-1
    constexpr int JR = DIR > 0 ? 2 : 0;  // the path reading column x+1
    __shared__ uint32_t edge[2][NW][2][64 * NP];
    __shared__ uint32_t edgemin[2][NW][2];
    __shared__ int tile_s;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // tiles run in ticket order, band by band: every tile this one waits for took an earlier ticket, so it
    // has started and (by induction) finishes: no deadlock whatever the number of resident workgroups
    if (threadIdx.x == 0) tile_s = atomicAdd(a.ticket, 1);
    __syncthreads();
    const int tile = tile_s;
    const int band = tile / a.strips, strip = tile - band * a.strips;
    int first, rows;
    band_rows(band, a.split, a.H, first, rows);
    const int y0 = DIR > 0 ? first : a.H - 1 - first;
    const int W = a.W, cs = a.cs, H1 = rows - 1;
    const int xa = strip * kStrip, xb = min(W, xa + kStrip);
    const size_t bsz = (size_t)3 * W * cs, bmsz = (size_t)3 * W;
    const uint16_t* bin = band == 0 ? a.bin0 : a.bnd + ((band - 1) & 1) * bsz;
    const uint32_t* binmin = band == 0 ? nullptr : a.bndmin + ((band - 1) & 1) * bmsz;
    if (band > 0 && threadIdx.x == 0) {
        // wait for the previous band's tiles whose last row this tile's halo reads (strips s-1, s, s+1); their
        // boundary stores were drained before the flag, and are read below with agent-scope loads
        for (int d = max(strip - 1, 0); d <= min(strip + 1, a.strips - 1); d++) {
            const int* f = a.flags + (band - 1) * a.strips + d;
            int spins = 0;
            while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != a.epoch) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > (1 << 22)) {
                    atomicOr(a.err, 1);
                    break;
                }
            }
        }
    }
    __syncthreads();
    // columns [xa - kBandRows, ...): the strip, the halo, and one guard column each side that only ever holds
    // the boundary state its neighbour reads in the band's first row
    const int cbase = xa - kBandRows + wv * G;
    const int bx = DIR > 0 ? W - 1 : 0;
    Cell<NP> mask;
    bool active;
    make_mask<NP>(lane, a.cap, mask, active);
    const int lo = lane * 2 * NP, lol = active ? lo : 0;

    // the states of the row before the band ("row -1"): the boundary buffer, or the initial fill
    Cell<NP> st[G][3];
    uint32_t sm[G][3];
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int x = cbase + g;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            set_init(st[g][j]);
            sm[g][j] = 2000;
            if (bin && x >= 0 && x < W) {
                load_u16_agent<NP>(bin + ((size_t)j * W + x) * cs + lo, active, st[g][j]);
                if (MASKED) {
#pragma unroll
                    for (int q = 0; q < NP; q++) st[g][j].v[q] = (st[g][j].v[q] & mask.v[q]) | (kInit2 & ~mask.v[q]);
                }
                sm[g][j] = binmin ? __hip_atomic_load(binmin + j * W + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : (uint32_t)a.bin0min[j * W + x];
            }
        }
    }
    // the region of path j at band row k (L = halo width left at that row)
    auto need = [&](int x, int j, int k) -> bool {
        if (x < 0 || x >= W) return false;
        const int L = H1 - k;
        if (j == JL) return x >= xa - L && x < xb;
        if (j == JR) return x >= xa && x < xb + L;
        return x >= xa && x < xb && x != bx;
    };
    auto need_raw = [&](int x, int k) -> bool {
        const int L = H1 - k;
        return x >= 0 && x < W && x >= xa - L && x < xb + L;
    };
    auto load_row = [&](int k, RawW<NP> (&rb)[G], Cell<NP> (&ab)[G]) {
        const int y = y0 + DIR * k;
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int x = cbase + g;
            if (need_raw(x, k)) ld_raw_w<NP>(a.raw + ((size_t)y * W + x) * cs + lol, rb[g]);
            if (x >= xa && x < xb) ld_u16<NP>(a.agg + ((size_t)y * W + x) * cs + lol, ab[g]);
        }
    };
    auto do_row = [&](int k, RawW<NP> (&rb)[G], Cell<NP> (&ab)[G], RawW<NP> (&nrb)[G], Cell<NP> (&nab)[G]) {
        const int y = y0 + DIR * k;
        const int par = k & 1;
        // edge exchange: the previous row's states of this wave's boundary columns
#pragma unroll
        for (int q = 0; q < NP; q++) {
            edge[par][wv][0][lane * NP + q] = st[G - 1][JL].v[q];
            edge[par][wv][1][lane * NP + q] = st[0][JR].v[q];
        }
        if (lane == 0) {
            edgemin[par][wv][0] = sm[G - 1][JL];
            edgemin[par][wv][1] = sm[0][JR];
        }
        __syncthreads();
        if (k + 1 < rows) load_row(k + 1, nrb, nab);
        // path costs accumulate straight into the strip cells' prefetched aggregates ab[g]
        // the path reading x-1: columns right to left, so x-1 still holds the previous row
#pragma unroll
        for (int g = G - 1; g >= 0; g--) {
            const int x = cbase + g;
            if (need(x, JL, k)) {
                Cell<NP> prev;
                uint32_t pm;
                if (x - 1 < 0) { set_init(prev); pm = 2000; }
                else if (g > 0) { prev = st[g - 1][JL]; pm = sm[g - 1][JL]; }
                else {
#pragma unroll
                    for (int q = 0; q < NP; q++) prev.v[q] = edge[par][wv - 1][0][lane * NP + q];
                    pm = edgemin[par][wv - 1][0];
                }
                Cell<NP> lc, raw;
                widen<NP>(rb[g], raw);
                sm[g][JL] = path_step<NP, MASKED>(prev, pm, raw, a.p1x2, a.p2, mask, lc);
                st[g][JL] = lc;
                if (x >= xa && x < xb) {
#pragma unroll
                    for (int q = 0; q < NP; q++) ab[g].v[q] = padd(ab[g].v[q], lc.v[q]);
                }
            }
        }
        // the vertical path
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int x = cbase + g;
            if (need(x, 1, k)) {
                Cell<NP> prev;
                uint32_t pm;
                prev = st[g][1];
                pm = sm[g][1];
                Cell<NP> lc, raw;
                widen<NP>(rb[g], raw);
                sm[g][1] = path_step<NP, MASKED>(prev, pm, raw, a.p1x2, a.p2, mask, lc);
                st[g][1] = lc;
#pragma unroll
                for (int q = 0; q < NP; q++) ab[g].v[q] = padd(ab[g].v[q], lc.v[q]);
            }
        }
        // the path reading x+1: columns left to right
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int x = cbase + g;
            if (need(x, JR, k)) {
                Cell<NP> prev;
                uint32_t pm;
                if (x + 1 >= W) { set_init(prev); pm = 2000; }
                else if (g < G - 1) { prev = st[g + 1][JR]; pm = sm[g + 1][JR]; }
                else {
#pragma unroll
                    for (int q = 0; q < NP; q++) prev.v[q] = edge[par][wv + 1][1][lane * NP + q];
                    pm = edgemin[par][wv + 1][1];
                }
                Cell<NP> lc, raw;
                widen<NP>(rb[g], raw);
                sm[g][JR] = path_step<NP, MASKED>(prev, pm, raw, a.p1x2, a.p2, mask, lc);
                st[g][JR] = lc;
                if (x >= xa && x < xb) {
#pragma unroll
                    for (int q = 0; q < NP; q++) ab[g].v[q] = padd(ab[g].v[q], lc.v[q]);
                }
            }
        }
        // the strip's own cells: aggregate, seeds, the replay's column, the boundary for the next band
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int x = cbase + g;
            if (x >= xa && x < xb) {
                const size_t cell = (size_t)y * W + x;
                if (x != bx) store_u16<NP>(a.agg + cell * cs + lo, active, ab[g]);
                if (x == bx - DIR) {
                    store_u16<NP>(a.cols + ((size_t)a.H + y) * cs + lo, active, st[g][0]);
                    if (lane == 0) a.colmins[a.H + y] = (uint16_t)sm[g][0];
                }
            }
        }
    };
    // one row body: the next row's loads go out at the start of a row and are taken over at its end, a row
    // of work later
    RawW<NP> rb[G], nrb[G];
    Cell<NP> ab[G], nab[G];
    load_row(0, rb, ab);
    for (int k = 0; k < rows; k++) {
        do_row(k, rb, ab, nrb, nab);
#pragma unroll
        for (int g = 0; g < G; g++) {
            rb[g] = nrb[g];
            ab[g] = nab[g];
        }
    }
    // the band's last row: the boundary the next band starts from
    uint16_t* bout = a.bnd + (band & 1) * bsz;
    uint32_t* boutmin = a.bndmin + (band & 1) * bmsz;
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int x = cbase + g;
        if (x >= xa && x < xb) {
#pragma unroll
            for (int j = 0; j < 3; j++) {
                store_u16_agent<NP>(bout + ((size_t)j * W + x) * cs + lo, active, st[g][j]);
                if (lane == 0) __hip_atomic_store(boutmin + j * W + x, sm[g][j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (band == a.extra_band && x != bx) {  // the border cell's entries are the replays'
                    store_u16<NP>(a.extra + ((size_t)j * W + x) * cs + lo, active, st[g][j]);
                    if (lane == 0) a.extramin[j * W + x] = (uint16_t)sm[g][j];
                }
            }
        }
    }
    // publish: every wave drains its stores (the boundary's agent-scope stores are then visible device-wide),
    // then one flag store
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(a.flags + band * a.strips + strip, a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct ReplayArgs {
    const uint8_t* raw;
    uint16_t* agg;
    int W, H, cs, cap;
    uint32_t p1x2, p2;
    int dir, col;
    const int* visits;        // 4 ints per visit: y, half, path-0 label, path-1 label
    int nvis;
    const uint16_t* cols;     // [2][H][cs] of this pass
    const uint16_t* colmins;  // [2][H]
    uint16_t* seeds;          // [3][W][cs]
    uint16_t* seedmins;       // [3][W]
    uint16_t* partial;        // [nvis][cs]: paths 0 + 1 + 3 of each visit
    uint16_t* chainp;         // [nvis][cs]: path 2 of each visit (the chain's)
    const uint16_t* vstart;   // per-pixel ranges [H][W] (null: uniform)
    const uint16_t* vcap;
    uint16_t* seedstart;      // [3][W]
    uint16_t* seedcap;
    AdaptP2 ap;
    uint32_t* seedguid;       // [3][W] (adaptive P2)
    uint16_t* colagg = nullptr;  // byte path volumes: the border column's visit costs go here, [H][cs], instead of agg
};

// The state a path-0 / path-1 neighbour read returns, by plan label: -1 the initial fill, y >= 0 the saved
// neighbour-column state of row y, -2-c pass +1's half-0 leftover of cell c (path 1 only). `from` receives
// the range of the cell the state belongs to.
template <int NP, bool MASKED, bool VAR>
__device__ __forceinline__ void label_state(const ReplayArgs& a, int p, int l, int lo, const Cell<NP>& mask,
                                            Cell<NP>& prev, uint32_t& pm, Range& from, uint32_t& gp)
{
    set_init(prev);
    pm = 2000;
    gp = 0;
    from = Range();
    if (l >= 0) {
        load_u16<NP>(a.cols + ((size_t)p * a.H + l) * a.cs + lo, lo < a.cs, prev);
        pm = a.colmins[p * a.H + l];
        const int ncol = a.col - a.dir;  // the neighbour column (the plan labels only real columns)
        if (ncol >= 0 && ncol < a.W) {
            if (VAR) {
                from.s = a.vstart[(size_t)l * a.W + ncol];
                from.c = a.vcap[(size_t)l * a.W + ncol];
            }
            if (a.ap.on) gp = a.ap.guid[(size_t)l * a.W + ncol];
        }
    } else if (l <= -2) {
        const int c = -2 - l;
        load_u16<NP>(a.seeds + (size_t)c * a.cs + lo, lo < a.cs, prev);  // seed table of path 1
        pm = a.seedmins[c];
        if (VAR) {
            from.s = a.seedstart[c];
            from.c = a.seedcap[c];
        }
        if (a.ap.on) gp = a.seedguid[c];
    }
    if (VAR) {
        clip_to_cap<NP>(prev, from.c, lo / (2 * NP));
    } else if (MASKED) {
#pragma unroll
        for (int k = 0; k < NP; k++) prev.v[k] = (prev.v[k] & mask.v[k]) | (kInit2 & ~mask.v[k]);
    }
}

// Path step of a border visit from a state of another cell (range `from`) to the visited cell (start s, cap c).
template <int NP, bool MASKED, bool VAR>
__device__ __forceinline__ uint32_t step_from(Cell<NP> prev, Range from, uint32_t pm, int s, int c, int lane,
                                              const Cell<NP>& raw, const ReplayArgs& a, const Cell<NP>& mask, Cell<NP>& lc,
                                              uint32_t g, uint32_t gp)
{
    uint32_t el, er;
    reframe<NP, VAR>(prev, from, s, c, lane, el, er);
    return path_step<NP, MASKED>(prev, pm, raw, a.p1x2, p2_step(a.ap, a.p2, g, gp), mask, lc, el, er);
}

// Paths 0, 1 and 3 of every border visit, one wavefront per visit, all in parallel: their neighbours are
// fixed by the plan (saved neighbour-column states, pass +1 seeds, or the margin slot for path 3), so
// they do not chain. Each visit's sum of the three path costs goes to `partial`.
template <int NP, bool MASKED, bool VAR>
__device__ __forceinline__ void replay_paths_wave(const ReplayArgs& a, int i, int lane)
{
    if (i >= a.nvis) return;
    const int y = a.visits[4 * i], l0 = a.visits[4 * i + 2], l1 = a.visits[4 * i + 3];
    const size_t cell = (size_t)y * a.W + a.col;
    const int s = VAR ? (int)a.vstart[cell] : 0, c = VAR ? (int)a.vcap[cell] : a.cap;
    Cell<NP> mask;
    bool active;
    make_mask<NP>(lane, c, mask, active);
    const int lo = lane * 2 * NP;
    Cell<NP> raw, prev, lc0, lc1, lc3;
    load_raw<NP>(a.raw + cell * a.cs + lo, lo < a.cs, raw);
    const uint32_t g = a.ap.on ? a.ap.guid[cell] : 0u;
    uint32_t pm, gp;
    Range from;
    label_state<NP, MASKED, VAR>(a, 0, l0, lo, mask, prev, pm, from, gp);
    step_from<NP, MASKED, VAR>(prev, from, pm, s, c, lane, raw, a, mask, lc0, g, gp);
    label_state<NP, MASKED, VAR>(a, 1, l1, lo, mask, prev, pm, from, gp);
    step_from<NP, MASKED, VAR>(prev, from, pm, s, c, lane, raw, a, mask, lc1, g, gp);
    set_init(prev);
    path_step<NP, MASKED>(prev, 2000, raw, a.p1x2, p2_step(a.ap, a.p2, g, 0u), mask, lc3);  // the margin slot
    Cell<NP> sum;
#pragma unroll
    for (int k = 0; k < NP; k++) sum.v[k] = padd(padd(lc0.v[k], lc1.v[k]), lc3.v[k]);
    store_u16<NP>(a.partial + (size_t)i * a.cs + lo, active, sum);
}
template <int NP, bool MASKED, bool VAR>
__global__ __launch_bounds__(kBlock) void sgm_replay_paths(ReplayArgs a)
{
    replay_paths_wave<NP, MASKED, VAR>(a, (int)((blockIdx.x * kBlock + threadIdx.x) >> 6), threadIdx.x & 63);
}

// Path 2 of the border column: one chain through the column's visits in schedule order. Each visit reads
// the state the other line-buffer half holds for the border slot and leaves its own in the current half
// (:219, :265). The chain writes its path cost to `chainp` (the aggregate is updated afterwards, per cell, by
// sgm_replay_add). It reads only the cost volume, the ranges and guidance, its visit list and, in pass -1, the
// border cell's path-2 seed (pass +1's half-0 leftover), so it prefetches deeply and runs beside the sweeps.
// Pass +1 ends by writing the border cell's path-2 seed-table entry (half 0 as the pass leaves it).
template <int NP, bool MASKED, bool VAR, bool ADP>
__device__ __forceinline__ void replay_chain_wave(const ReplayArgs& a, const int* __restrict__ vis, int lane)
{
    constexpr int PF = 8;
    const int cs = a.cs, lo = lane * 2 * NP;
    Cell<NP> mask;
    bool active;
    make_mask<NP>(lane, a.cap, mask, active);
    Cell<NP> h0, h1;
    uint32_t m0 = 2000, m1 = 2000;
    // VAR: the cells whose states the two halves hold, as start | cap << 16 (kept in registers)
    uint32_t q0 = 0, q1 = 0;
    uint32_t g0 = 0, g1 = 0;  // guidance of the cells whose states the two halves hold (adaptive P2)
    set_init(h0);
    set_init(h1);
    if (a.dir < 0) {
        load_u16<NP>(a.seeds + ((size_t)1 * a.W + a.col) * cs + lo, lo < cs, h0);  // path 2's seed
        if (VAR) {
            q0 = (uint32_t)a.seedstart[1 * a.W + a.col] | ((uint32_t)a.seedcap[1 * a.W + a.col] << 16);
            clip_to_cap<NP>(h0, (int)(q0 >> 16), lane);
        } else if (MASKED) {
#pragma unroll
            for (int k = 0; k < NP; k++) h0.v[k] = (h0.v[k] & mask.v[k]) | (kInit2 & ~mask.v[k]);
        }
        m0 = a.seedmins[1 * a.W + a.col];
        if (ADP) g0 = a.seedguid[1 * a.W + a.col];
    }
    const int lol = lo < cs ? lo : 0;
    const int nv = a.nvis;
    RawW<NP> rb[PF];
    // the visit list's words 64 visits a VGPR (ChainWords): the rows of the visits being prefetched (PF ahead),
    // and the half, range and guidance of the visit being run
    const auto cell_of = [&](int pos) { return (size_t)vis[4 * min(pos, nv - 1)] * a.W + a.col; };
    const auto vrow = [&](int pos) { return (uint32_t)vis[4 * min(pos, nv - 1)]; };
    const auto vcur = [&](int pos) { return (uint32_t)vis[4 * min(pos, nv - 1) + 1]; };
    const auto vrs = [&](int pos) {
        const size_t c = cell_of(pos);
        return (uint32_t)a.vstart[c] | ((uint32_t)a.vcap[c] << 16);
    };
    const auto vg = [&](int pos) { return a.ap.guid[cell_of(pos)]; };
    ChainWords pr, cw, rw, gw;
    if (nv > 0) {
        pr.init(vrow, lane);
        cw.init(vcur, lane);
        if constexpr (VAR) rw.init(vrs, lane);
        if constexpr (ADP) gw.init(vg, lane);
#pragma unroll
        for (int j = 0; j < PF; j++) {
            const size_t cj = (size_t)pr.at(j) * a.W + a.col;
            ld_raw_w<NP>(a.raw + cj * cs + lol, rb[j]);
        }
    }
    auto body = [&](int i, int j) {
        const int cur = (int)cw.at(i);
        cw.advance(i, vcur, lane);
        Cell<NP> raw;
        widen<NP>(rb[j], raw);
        int s = 0, c = a.cap;
        if constexpr (VAR) {
            const uint32_t r = rw.at(i);
            rw.advance(i, vrs, lane);
            s = (int)(r & 0xffffu);
            c = (int)(r >> 16);
            make_mask<NP>(lane, c, mask, active);
        }
        const size_t cpn = (size_t)pr.at(i + PF) * a.W + a.col;
        pr.advance(i + PF, vrow, lane);
        ld_raw_w<NP>(a.raw + cpn * cs + lol, rb[j]);
        uint32_t g = 0;
        if constexpr (ADP) {
            g = gw.at(i);
            gw.advance(i, vg, lane);
        }
        Cell<NP> lc, prev;
        uint32_t mn, pm, q = 0;
#pragma unroll
        for (int k = 0; k < NP; k++) prev.v[k] = cur ? h0.v[k] : h1.v[k];  // the other half's state
        pm = cur ? m0 : m1;
        q = cur ? q0 : q1;
        const uint32_t gp = cur ? g0 : g1;
        Range from;
        from.s = (int)(q & 0xffffu);
        from.c = (int)(q >> 16);
        mn = step_from<NP, MASKED, VAR>(prev, from, pm, s, c, lane, raw, a, mask, lc, g, gp);
        store_u16<NP>(a.chainp + (size_t)i * cs + lo, active, lc);
        const uint32_t qn = VAR ? ((uint32_t)s | ((uint32_t)c << 16)) : 0u;
#pragma unroll
        for (int k = 0; k < NP; k++) {
            h1.v[k] = cur ? lc.v[k] : h1.v[k];
            h0.v[k] = cur ? h0.v[k] : lc.v[k];
        }
        m1 = cur ? mn : m1;
        m0 = cur ? m0 : mn;
        q1 = cur ? qn : q1;
        q0 = cur ? q0 : qn;
        g1 = cur ? g : g1;
        g0 = cur ? g0 : g;
    };
    int i0 = 0;
    for (; i0 + PF <= nv; i0 += PF) {
#pragma unroll
        for (int j = 0; j < PF; j++) body(i0 + j, j);
    }
#pragma unroll
    for (int j = 0; j < PF; j++)
        if (i0 + j < nv) body(i0 + j, j);
    if (a.dir > 0) {
        // path 2's seed: the chain's half-0 state (paths 1 and 3: sgm_replay_seeds)
        store_u16<NP>(a.seeds + ((size_t)1 * a.W + a.col) * cs + lo, lo < cs, h0);
        if (lane == 0) {
            a.seedmins[1 * a.W + a.col] = (uint16_t)m0;
            if (VAR) {
                a.seedstart[1 * a.W + a.col] = (uint16_t)(q0 & 0xffffu);
                a.seedcap[1 * a.W + a.col] = (uint16_t)(q0 >> 16);
            }
            if (ADP) a.seedguid[1 * a.W + a.col] = g0;
        }
    }
}
template <int NP, bool MASKED, bool VAR, bool ADP>
__global__ __launch_bounds__(64) void sgm_replay_chain(ReplayArgs a, const int* __restrict__ vis)
{
    replay_chain_wave<NP, MASKED, VAR, ADP>(a, vis, threadIdx.x);
}

// One pass's independent chains in one launch, for latency-bound (ranged) layers: the border chain (block 0,
// wave 0: the longest serial path starts first), the row path, then paths 2, 1 and 3. Pass +1's row path is
// the first aggregate writer and its paths 1-3 write a volume each (FIRST); pass -1 adds its row path to agg and
// its paths 1-3 to those volumes (the border columns, which each pass leaves to its replays, keep the other
// pass's costs), and sgm_finish sums them. One launch instead of one stream per path: streams beyond the
// device's hardware queues share a queue and run one after another.
struct PassArgs {
    SweepArgs s[3];   // paths 1, 2, 3
    RowArgs r;        // the row path
    ReplayArgs c;     // the border chain
    int n1, n2, nrow; // chains of paths 1, 2 and of the row path
};
template <int NP, int DIR, bool MASKED, bool VAR, bool ADP>
__global__ __launch_bounds__(kBlock) void sgm_pass(PassArgs p)
{
    const int lane = threadIdx.x & 63;
    if (blockIdx.x == 0) {
        if (threadIdx.x < 64) replay_chain_wave<NP, MASKED, VAR, ADP>(p.c, p.c.visits, lane);
        return;
    }
    int w = __builtin_amdgcn_readfirstlane((int)(((blockIdx.x - 1) * kBlock + threadIdx.x) >> 6));
    if (w < p.nrow) {
        if (DIR > 0) row_fwd_wave<NP, MASKED, false, VAR, ADP>(p.r, w, lane);
        else row_bwd_wave<NP, MASKED, VAR, ADP, false>(p.r, w, lane);  // into its own volume
        return;
    }
    w -= p.nrow;
    if (w < p.n2) {
        sweep_wave<NP, DIR, 2, (DIR > 0), MASKED, VAR, ADP>(p.s[1], w, lane);
        return;
    }
    w -= p.n2;
    if (w < p.n1) {
        sweep_wave<NP, DIR, 1, (DIR > 0), MASKED, VAR, ADP>(p.s[0], diag_order<DIR>(w, p.r.W, p.r.H), lane);
        return;
    }
    // returns past the last chain
    sweep_wave<NP, DIR, 3, (DIR > 0), MASKED, VAR, ADP>(p.s[2], diag_order<-DIR>(w - p.n1, p.r.W, p.r.H), lane);
}

// sgm_pass with segmented chains (64/SEG chains per wavefront; each wavefront runs chains of one path): the border
// chain in block 0, then the row path and paths 2, 1 and 3, as many wavefronts each as its chains fill.
template <int SEG, int DIR>
__global__ __launch_bounds__(kBlock) void sgm_pass_seg(PassArgs p)
{
    constexpr int CPW = 64 / SEG;
    const int lane = threadIdx.x & 63;
    if (blockIdx.x == 0) {
        if (threadIdx.x < 64) replay_chain_wave<1, true, true, false>(p.c, p.c.visits, lane);
        return;
    }
    int w = __builtin_amdgcn_readfirstlane((int)(((blockIdx.x - 1) * kBlock + threadIdx.x) >> 6));
    const int wr = (p.nrow + CPW - 1) / CPW, w2 = (p.n2 + CPW - 1) / CPW, w1 = (p.n1 + CPW - 1) / CPW;
    if (w < wr) {
        if (DIR > 0) row_fwd_seg_wave<SEG>(p.r, w * 64 + lane, lane);
        else row_bwd_seg_wave<SEG, false>(p.r, w * 64 + lane, lane);
        return;
    }
    w -= wr;
    if (w < w2) {
        sweep_seg_wave<SEG, DIR, 2, (DIR > 0)>(p.s[1], w * 64 + lane, lane);
        return;
    }
    w -= w2;
    if (w < w1) {
        sweep_seg_wave<SEG, DIR, 1, (DIR > 0)>(p.s[0], w * 64 + lane, lane);
        return;
    }
    sweep_seg_wave<SEG, DIR, 3, (DIR > 0)>(p.s[2], (w - w1) * 64 + lane, lane);  // returns past the last chain
}

// Pass +1's seed-table entries of paths 1 and 3 for the border cell: recomputed for the last visit that wrote
// half 0. Path 1 reads a saved neighbour-column state, so this runs after the sweeps (one wavefront).
template <int NP, bool MASKED, bool VAR, bool ADP>
__device__ __forceinline__ void replay_seeds_wave(const ReplayArgs& a, const int* __restrict__ vis, int lane)
{
    const int cs = a.cs, lo = lane * 2 * NP;
    Cell<NP> mask;
    bool active;
    make_mask<NP>(lane, a.cap, mask, active);
    int last = -1;
    for (int i = a.nvis - 1; i >= 0; i--)
        if (vis[4 * i + 1] == 0) { last = i; break; }
    Cell<NP> s1, s3;
    uint32_t sm1 = 2000, sm3 = 2000;
    Range rl;
    uint32_t gl = 0;
    set_init(s1);
    set_init(s3);
    if (last >= 0) {
        const int yl = vis[4 * last];
        const size_t cl = (size_t)yl * a.W + a.col;
        if (ADP) gl = a.ap.guid[cl];
        int s = 0, c = a.cap;
        if (VAR) {
            s = a.vstart[cl];
            c = a.vcap[cl];
            make_mask<NP>(lane, c, mask, active);
            rl.s = s;
            rl.c = c;
        }
        Cell<NP> r, prev;
        uint32_t pm, gp;
        Range from;
        load_raw<NP>(a.raw + cl * cs + lo, lo < cs, r);
        label_state<NP, MASKED, VAR>(a, 1, vis[4 * last + 3], lo, mask, prev, pm, from, gp);
        sm1 = step_from<NP, MASKED, VAR>(prev, from, pm, s, c, lane, r, a, mask, s1, gl, gp);
        set_init(prev);
        sm3 = path_step<NP, MASKED>(prev, 2000, r, a.p1x2, p2_step(a.ap, a.p2, gl, 0u), mask, s3);
    }
    store_u16<NP>(a.seeds + ((size_t)0 * a.W + a.col) * cs + lo, lo < cs, s1);
    store_u16<NP>(a.seeds + ((size_t)2 * a.W + a.col) * cs + lo, lo < cs, s3);
    if (lane == 0) {
        a.seedmins[0 * a.W + a.col] = (uint16_t)sm1;
        a.seedmins[2 * a.W + a.col] = (uint16_t)sm3;
        if (VAR) {
            a.seedstart[0 * a.W + a.col] = (uint16_t)rl.s;
            a.seedcap[0 * a.W + a.col] = (uint16_t)rl.c;
            a.seedstart[2 * a.W + a.col] = (uint16_t)rl.s;
            a.seedcap[2 * a.W + a.col] = (uint16_t)rl.c;
        }
        if (ADP) {
            a.seedguid[0 * a.W + a.col] = gl;
            a.seedguid[2 * a.W + a.col] = gl;
        }
    }
}
template <int NP, bool MASKED, bool VAR, bool ADP>
__global__ __launch_bounds__(64) void sgm_replay_seeds(ReplayArgs a, const int* __restrict__ vis)
{
    replay_seeds_wave<NP, MASKED, VAR, ADP>(a, vis, threadIdx.x);
}

// Adds every border visit's four path costs to its cell (one wavefront per row of the border column;
// `cellvis` lists each row's visits, at most two per pass).
template <int NP, bool VAR>
__device__ __forceinline__ void replay_add_wave(const ReplayArgs& a, const int* cellvis, int y, int lane)
{
    if (y >= a.H) return;
    const int lo = lane * 2 * NP;
    const bool active = lo < (VAR ? (int)a.vcap[(size_t)y * a.W + a.col] : a.cap);
    const size_t cell = ((size_t)y * a.W + a.col) * a.cs + lo;
    Cell<NP> agg, p;
    if (a.colagg) {  // the border column's costs of this pass alone, summed later with the path volumes
#pragma unroll
        for (int k = 0; k < NP; k++) agg.v[k] = 0u;
    } else {
        load_u16<NP>(a.agg + cell, active, agg);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int i = cellvis[2 * y + j];
        if (i >= 0) {
            load_u16<NP>(a.partial + (size_t)i * a.cs + lo, active, p);
#pragma unroll
            for (int k = 0; k < NP; k++) agg.v[k] = padd(agg.v[k], p.v[k]);
            load_u16<NP>(a.chainp + (size_t)i * a.cs + lo, active, p);
#pragma unroll
            for (int k = 0; k < NP; k++) agg.v[k] = padd(agg.v[k], p.v[k]);
        }
    }
    if (a.colagg) store_u16<NP>(a.colagg + (size_t)y * a.cs + lo, lo < a.cs, agg);
    else store_u16<NP>(a.agg + cell, active, agg);
}
template <int NP, bool VAR>
__global__ __launch_bounds__(kBlock) void sgm_replay_add(ReplayArgs a, const int* cellvis)
{
    replay_add_wave<NP, VAR>(a, cellvis, (int)((blockIdx.x * kBlock + threadIdx.x) >> 6), threadIdx.x & 63);
}

// Seeds of interior cells when pass +1 never writes half 0 (H = 1): the initial fill.
__global__ void fill_u16(uint16_t* p, size_t n, uint16_t v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_u32(uint32_t* p, size_t n, uint32_t v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
// out = sat(a0 + a1 + a2 + a3) on u16 pairs (pass +1's aggregate when its paths are still in their volumes).
__global__ void agg_sum4(uint32_t* out, const uint32_t* a0, const uint32_t* a1, const uint32_t* a2, const uint32_t* a3,
                         size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = padd(padd(a0[i], a1[i]), padd(a2[i], a3[i]));
}

// ------------------------------------------------------------------ census and cost volume
// The census code of pixel (x, y): bit k set where the k-th neighbour of the (2R+1)^2 window (row-major, the centre
// left out, coordinates clamped to the image) is darker than the centre.
template <typename T, int R>
__device__ __forceinline__ uint64_t census_at(const T* img, int W, int H, int stride, int x, int y)
{
    const int c = img[(size_t)y * stride + x];
    uint64_t code = 0;
    int bit = 0;
#pragma unroll
    for (int dy = -R; dy <= R; dy++) {
        const int yy = min(max(y + dy, 0), H - 1);
#pragma unroll
        for (int dx = -R; dx <= R; dx++) {
            if (dx == 0 && dy == 0) continue;
            const int xx = min(max(x + dx, 0), W - 1);
            const int v = img[(size_t)yy * stride + xx];
            code |= (uint64_t)(v < c) << bit;
            bit++;
        }
    }
    return code;
}
template <typename T, int R>
__global__ void census_kernel(const T* img, int W, int H, int stride, uint64_t* out)
{
    const int x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (x >= W || y >= H) return;
    out[(size_t)y * W + x] = census_at<T, R>(img, W, H, stride, x, y);
}

// Truncated 3x3 absolute-difference cost, four planes per thread (consecutive threads write consecutive
// 4-byte groups of a cell); blockIdx.y = row.
template <typename T>
__global__ __launch_bounds__(kBlock) void cost_ad_kernel(const T* L, const T* R, int W, int H, int stride, int D, int cs,
                                                          int shift, int adcap, uint8_t* raw)
{
    const int groups = cs >> 2;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < W * groups) {
        const int y = blockIdx.y;
        const int x = t / groups, d0 = (t - x * groups) * 4;
        int lv[9];
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++)
                lv[(dy + 1) * 3 + dx + 1] = L[(size_t)min(max(y + dy, 0), H - 1) * stride + min(max(x + dx, 0), W - 1)];
        uint32_t w = 0;
        for (int j = 0; j < 4; j++) {
            const int d = d0 + j;
            uint32_t v = 255u;
            if (d < D) {
                const int xr = max(x - d, 0);
                int s = 0;
#pragma unroll
                for (int dy = -1; dy <= 1; dy++)
#pragma unroll
                    for (int dx = -1; dx <= 1; dx++) {
                        const int b = R[(size_t)min(max(y + dy, 0), H - 1) * stride + min(max(xr + dx, 0), W - 1)];
                        const int df = abs(lv[(dy + 1) * 3 + dx + 1] - b) >> shift;
                        s += df < adcap ? df : adcap;
                    }
                v = (uint32_t)s;
            }
            w |= v << (8 * j);
        }
        ((uint32_t*)(raw + (size_t)y * W * cs))[t] = w;
    }
}

// ------------------------------------------------------------------ coarse-to-fine levels
// One level down: the rounded mean of each 2x2 block (the last row / column repeats on odd sizes).
template <typename T>
__global__ void downsample2_kernel(const T* src, int W, int H, int stride, T* dst, int W2, int H2)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W2 || y >= H2) return;
    const int x0 = min(2 * x, W - 1), x1 = min(2 * x + 1, W - 1), y0 = min(2 * y, H - 1), y1 = min(2 * y + 1, H - 1);
    const uint32_t sum = (uint32_t)src[(size_t)y0 * stride + x0] + src[(size_t)y0 * stride + x1] +
                         src[(size_t)y1 * stride + x0] + src[(size_t)y1 * stride + x1];
    dst[(size_t)y * W2 + x] = (T)((sum + 2) >> 2);
}

// The plane range of every pixel of a level from the disparity of the level below (half resolution, half
// the planes): start = clamp(2 d - r, 0, D - 2r), count = cap = 2r (a multiple of the alignment, so no
// range's padding reaches the end of its line-buffer segment).
[[maybe_unused]] __global__ void ranges_from_coarse_kernel(const uint16_t* dc, int Wc, int Hc, int W, int H, int r, int D, uint16_t* vstart,
                                          uint16_t* vcap, uint16_t* vcount)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)W * H) return;
    const int y = (int)(i / W), x = (int)(i - (size_t)y * W);
    const int d = dc[(size_t)min(y >> 1, Hc - 1) * Wc + min(x >> 1, Wc - 1)];
    const int st = max(0, min(2 * d - r, D - 2 * r));
    vstart[i] = (uint16_t)st;
    vcap[i] = (uint16_t)(2 * r);
    vcount[i] = (uint16_t)(2 * r);
}

// Census costs of per-pixel ranges: raw[y][x][j] = popcount(cl(x) ^ cr(max(x - (start + j), 0))) for j < count,
// 255 past it (DenseStereo.cs:380); one thread per byte of the cell-major volume.
[[maybe_unused]] __global__ void cost_census_ranges_kernel(const uint64_t* cl, const uint64_t* cr, int W, int H, int cs, const uint16_t* vstart,
                                          const uint16_t* vcount, uint8_t* raw)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t cells = (size_t)W * H;
    if (i >= cells * cs) return;
    const size_t c = i / cs;
    const int j = (int)(i - c * cs);
    const int y = (int)(c / W), x = (int)(c - (size_t)y * W);
    uint8_t v = 255;
    if (j < vcount[c]) {
        const int xr = max(x - ((int)vstart[c] + j), 0);
        v = (uint8_t)__popcll(cl[c] ^ cr[(size_t)y * W + xr]);
    }
    raw[i] = v;
}

// Jagged per-pixel arrays (cap[i] entries of pixel i at off[i], raster order: DenseLayer Raw/Agg) <-> the
// cell-major volume with stride cs; entries of the volume past a pixel's cap get `fill`.
template <typename T>
__global__ void unpack_kernel(const T* src, const uint64_t* off, const uint16_t* cap, T* dst, int cs, size_t cells, T fill)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells * cs) return;
    const size_t c = i / cs;
    const int j = (int)(i - c * cs);
    dst[i] = j < cap[c] ? src[off[c] + j] : fill;
}
template <typename T>
__global__ void pack_kernel(const T* src, const uint64_t* off, const uint16_t* cap, T* dst, int cs, size_t cells)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells * cs) return;
    const size_t c = i / cs;
    const int j = (int)(i - c * cs);
    if (j < cap[c]) dst[off[c] + j] = src[i];
}

template <typename T>
__global__ void restride_kernel(const T* src, int scs, T* dst, int dcs, size_t cells, int n, T fill)
{
    const size_t total = cells * (size_t)dcs;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t c = t / dcs;
        const int d = (int)(t - c * dcs);
        dst[t] = d < n ? src[c * scs + d] : fill;
    }
}

inline int grid_for(size_t total, int block) { return (int)std::min<size_t>((total + block - 1) / block, 1u << 20); }

// ------------------------------------------------------------------ launchers
// The argument blocks of one request's kernels: the sweeps' and row paths' (pass +1's border-column state slots) and
// pass +1's border replay. pv: the request runs with byte path volumes.
inline void sgm_base_args(const Geometry& g, const SgmBuffers& b, const SgmOutputs& o, int p1, int p2, bool pv,
                          SweepArgs& a, RowArgs& ra, ReplayArgs& r)
{
    a.raw = b.raw;
    a.agg = b.agg;
    a.W = g.W; a.H = g.H; a.cs = g.cs; a.cap = g.cap;
    a.p1x2 = (uint32_t)p1 | ((uint32_t)p1 << 16);
    a.p2 = (uint32_t)p2;
    a.seeds = b.seeds;
    a.seedmins = b.seedmins;
    a.seed_row = b.seed_row;
    a.vstart = b.vstart;
    a.vcap = b.vcap;
    a.seedstart = b.seedstart;
    a.seedcap = b.seedcap;
    a.ap = b.ap;
    a.seedguid = b.seedguid;
    ra.raw = b.raw; ra.raw_out = b.raw_out; ra.cl = b.cl; ra.cr = b.cr; ra.agg = b.agg;
    ra.W = g.W; ra.H = g.H; ra.cs = g.cs; ra.cap = g.cap; ra.D = g.D;
    ra.p1x2 = a.p1x2; ra.p2 = a.p2;
    ra.disp = o.disp; ra.sub = o.sub; ra.dr = b.dr; ra.dr0 = b.dr0; ra.write_agg = o.write_agg;
    ra.vstart = b.vstart; ra.vcap = b.vcap; ra.vcount = b.vcount;
    ra.ap = b.ap;
    if (pv) {
        for (int k = 0; k < 7; k++) ra.pv[k] = b.pv[k];
        ra.pvs = b.pvs;
        ra.bagg_plus = b.bagg[0];
    }
    r.raw = b.raw; r.agg = b.agg; r.W = g.W; r.H = g.H; r.cs = g.cs; r.cap = g.cap;
    r.p1x2 = a.p1x2; r.p2 = a.p2;
    r.seeds = b.seeds; r.seedmins = b.seedmins;
    r.dir = 1; r.col = g.W - 1; r.visits = b.visits_plus; r.nvis = b.n_plus;
    r.cols = b.cols; r.colmins = b.colmins; r.partial = b.partial; r.chainp = b.chainp;
    r.vstart = b.vstart; r.vcap = b.vcap; r.seedstart = b.seedstart; r.seedcap = b.seedcap;
    r.ap = b.ap; r.seedguid = b.seedguid;
    r.colagg = pv ? b.bagg[0] : nullptr;
    a.cols = b.cols; a.colmins = b.colmins;
    ra.cols = b.cols; ra.colmins = b.colmins;
}

template <int NP, bool MASKED, bool VAR, bool ADP, int SEG = 0>
hipError_t run_sgm(const Geometry& g, const SgmBuffers& b, const SgmOutputs& o, int p1, int p2, uint16_t* agg_pass1,
                   hipStream_t s, hipEvent_t* marks)
{
    static_assert(SEG == 0 || (NP == 1 && MASKED && VAR && !ADP), "segmented chains: narrow per-pixel ranges only");
    auto sblocks = [](int chains) { return (unsigned)(((size_t)chains * (SEG ? SEG : 64) + kBlock - 1) / kBlock); };
    int mk = 0;
    auto mark = [&]() { if (marks) (void)hipEventRecord(marks[mk++], s); };
    // byte path volumes: uniform constant-P2 census requests of at most 256 planes whose path costs fit a byte
    // (path_bytes in engine.hip decides and gives the volumes)
    constexpr bool kPv = !VAR && !ADP && SEG == 0 && NP <= 2;
    const bool pv = kPv && b.pv[0] != nullptr && b.cl != nullptr;
    SweepArgs a;
    RowArgs ra;
    ReplayArgs r;
    sgm_base_args(g, b, o, p1, p2, pv, a, ra, r);
    const size_t colsz = cols_half(g), colminsz = colmins_half(g);
    auto blocks = [&](int chains) { return (chains + 3) / 4; };
    // one chain a wavefront: one-wavefront workgroups when the request has the device to itself (SgmBuffers::solo)
    const unsigned wg = b.solo ? 64u : (unsigned)kBlock;
    auto grid = [&](int chains) { return b.solo ? (unsigned)chains : (unsigned)blocks(chains); };
    const int nh = g.H, nv = g.W, nd = g.W + g.H - 1;
    if (b.seed_row < 0) {  // pass +1 never writes half 0: every seed is the initial fill
        fill_u16<<<grid_for((size_t)3 * g.W * g.cs, 256), 256, 0, s>>>(b.seeds, (size_t)3 * g.W * g.cs, 2000);
        fill_u16<<<grid_for((size_t)3 * g.W, 256), 256, 0, s>>>(b.seedmins, (size_t)3 * g.W, 2000);
        if (VAR) fill_u16<<<grid_for((size_t)3 * g.W, 256), 256, 0, s>>>(b.seedcap, (size_t)3 * g.W, 0);  // empty ranges
        if (b.ap.on) fill_u32<<<grid_for((size_t)3 * g.W, 256), 256, 0, s>>>(b.seedguid, (size_t)3 * g.W, 0u);
    }
    // paths 1-3 of one pass: fused bands (up to 256 planes), else one sweep per path
    // (not with the census cost fused into the row path: the sweeps read the costs it writes)
    const size_t aggn = (size_t)g.W * g.H * g.cs / 2;  // u16 pairs of one volume (cs is a multiple of 4)
    auto vertical_diagonal = [&](auto dirc, uint16_t* cols, uint16_t* colmins) {
        constexpr int DIR = decltype(dirc)::value;
        if constexpr (kPv) {
            if (pv) {  // each path into its own byte volume: 1..3 (pass +1), 4..6 (pass -1)
                SweepArgs3 q;
                for (int k = 0; k < 3; k++) {
                    q.s[k] = a;
                    q.s[k].pvs = b.pvs;
                    q.s[k].pv = b.pv[(DIR > 0 ? 1 : 4) + k];
                }
                if (b.solo) {  // the device to itself: the three in one launch
                    sgm_sweeps_pv<NP, DIR, MASKED><<<grid(3 * std::max(nd, nv)), wg, 0, s>>>(q);
                    mark();  // the next two stages are empty
                    mark();
                    mark();
                    return;
                }
                sgm_sweep_pv<NP, DIR, 1, MASKED><<<grid(nd), wg, 0, s>>>(q.s[0]);
                mark();
                sgm_sweep_pv<NP, DIR, 2, MASKED><<<grid(nv), wg, 0, s>>>(q.s[1]);
                mark();
                sgm_sweep_pv<NP, DIR, 3, MASKED><<<grid(nd), wg, 0, s>>>(q.s[2]);
                mark();
                return;
            }
        }
        if constexpr (SEG > 0) {
            sgm_sweep_seg<SEG, DIR, 1, false><<<sblocks(nd), kBlock, 0, s>>>(a);
            mark();
            sgm_sweep_seg<SEG, DIR, 2, false><<<sblocks(nv), kBlock, 0, s>>>(a);
            mark();
            sgm_sweep_seg<SEG, DIR, 3, false><<<sblocks(nd), kBlock, 0, s>>>(a);
            mark();
            return;
        }
        if (NP > 2 || VAR || ADP || !o.bands) {  // the fused bands take neither ranges nor adaptive P2
            sgm_sweep<NP, DIR, 1, false, MASKED, VAR, ADP><<<grid(nd), wg, 0, s>>>(a);
            mark();
            sgm_sweep<NP, DIR, 2, false, MASKED, VAR, ADP><<<grid(nv), wg, 0, s>>>(a);
            mark();
            sgm_sweep<NP, DIR, 3, false, MASKED, VAR, ADP><<<grid(nd), wg, 0, s>>>(a);
            mark();
            return;
        }
        if constexpr (NP <= 2 && !VAR && !ADP) {
            BandArgs ba;
            ba.raw = b.raw; ba.agg = b.agg; ba.W = g.W; ba.H = g.H; ba.cs = g.cs; ba.cap = g.cap;
            ba.p1x2 = a.p1x2; ba.p2 = a.p2;
            ba.cols = cols; ba.colmins = colmins;
            ba.strips = (g.W + kStrip - 1) / kStrip;
            // pass +1 bands end at the seed row, whose states become the seed table (half 0 as pass +1 leaves
            // it for the interior cells): rows [0, seed_row] in bands of kBandRows, then the rest
            ba.split = DIR > 0 && b.seed_row >= 0 ? b.seed_row + 1 : g.H;
            const int nbands = (ba.split + kBandRows - 1) / kBandRows + (g.H - ba.split + kBandRows - 1) / kBandRows;
            ba.bin0 = DIR > 0 ? nullptr : b.seeds;
            ba.bin0min = DIR > 0 ? nullptr : b.seedmins;
            ba.bnd = b.bnd;
            ba.bndmin = b.bndmin;
            ba.extra = b.seeds;
            ba.extramin = b.seedmins;
            ba.extra_band = DIR > 0 && b.seed_row >= 0 ? (ba.split + kBandRows - 1) / kBandRows - 1 : -1;
            ba.ticket = b.ticket + (DIR > 0 ? 0 : 1);
            ba.flags = b.flags + (DIR > 0 ? 0 : (size_t)b.max_tiles);
            ba.epoch = b.epoch;
            ba.err = b.err;
            fill_u32<<<1, 1, 0, s>>>((uint32_t*)ba.ticket, 1, 0u);
            if (b.zero_flags)
                fill_u32<<<grid_for((size_t)b.max_tiles, 256), 256, 0, s>>>((uint32_t*)ba.flags, (size_t)b.max_tiles, 0u);
            sgm_band<NP, DIR, MASKED><<<nbands * ba.strips, kBandWaves * 64, 0, s>>>(ba);
            mark();  // one launch for the three paths: the next two stages are empty
            mark();
            mark();
        }
    };
    // The border chains (path 2 of the replayed column) run on b.chain when it is given: each forks from s once
    // its inputs are final (pass +1: the cost volume; pass -1: also pass +1's seeds) and s waits for it just
    // before the visits' costs are added.
    if (!b.chainp) return hipErrorInvalidValue;
    const bool side = b.chain && b.chain_fork && b.chain_join;
    bool chain_live = false;
    auto start_chain = [&](const ReplayArgs& rr) {
        hipStream_t cst = s;
        if (side) {
            (void)hipEventRecord(b.chain_fork, s);
            (void)hipStreamWaitEvent(b.chain, b.chain_fork, 0);
            cst = b.chain;
        }
        sgm_replay_chain<NP, MASKED, VAR, ADP><<<1, 64, 0, cst>>>(rr, rr.visits);
        if (side) (void)hipEventRecord(b.chain_join, b.chain);
        chain_live = side;
    };
    auto end_chain = [&]() {
        if (chain_live) (void)hipStreamWaitEvent(s, b.chain_join, 0);
        chain_live = false;
    };
    // Ranged layers given path volumes: each pass's border chain, row path and paths 1-3 in one launch (sgm_pass,
    // sgm_pass_seg); pass +1's paths 1-3 fill the volumes, pass -1 adds to them, and sgm_finish sums them with agg.
    constexpr bool kFusable = VAR;
    const bool fused = kFusable && b.aggp[0] && !o.bands && !b.cl;  // needs the volumes, not the side streams
    auto launch_pass = [&](auto dirc, const ReplayArgs& rr) {
        constexpr int DIR = decltype(dirc)::value;
        if constexpr (kFusable) {
            PassArgs p;
            for (int k = 0; k < 3; k++) {
                p.s[k] = a;
                p.s[k].agg = b.aggp[k];
            }
            p.r = ra;
            p.c = rr;
            p.n1 = nd;
            p.n2 = nv;
            p.nrow = nh;
            if constexpr (SEG > 0) {
                constexpr int CPW = 64 / (SEG > 0 ? SEG : 64);
                const int waves = (nh + CPW - 1) / CPW + (nv + CPW - 1) / CPW + 2 * ((nd + CPW - 1) / CPW);
                sgm_pass_seg<SEG, DIR><<<1 + (waves + 3) / 4, kBlock, 0, s>>>(p);
            } else {
                const int waves = p.nrow + nv + 2 * nd;
                sgm_pass<NP, DIR, MASKED, VAR, ADP><<<1 + (waves + 3) / 4, kBlock, 0, s>>>(p);
            }
        }
    };
    // pass +1: -> (with the census cost fused when codes are given), then down-right, down, down-left
    if (fused) {
        launch_pass(std::integral_constant<int, 1>(), r);
        mark();  // the row path and paths 1-3 in one launch: the next three stages are empty
        mark();
        mark();
        mark();
    } else {
        if (!b.cl) start_chain(r);  // else the row path writes the cost volume first
        if constexpr (SEG > 0) {
            sgm_row_fwd_seg<SEG><<<sblocks(nh), kBlock, 0, s>>>(ra);
        } else if constexpr (VAR) {
            sgm_row_fwd<NP, MASKED, false, true, ADP><<<grid(nh), wg, 0, s>>>(ra);
        } else if constexpr (kPv) {
            if (pv) sgm_row_fwd_pv<NP, MASKED><<<grid(nh), wg, 0, s>>>(ra);
            else if (b.cl) sgm_row_fwd<NP, MASKED, true, false, ADP><<<grid(nh), wg, 0, s>>>(ra);
            else sgm_row_fwd<NP, MASKED, false, false, ADP><<<grid(nh), wg, 0, s>>>(ra);
        } else {
            if (b.cl) sgm_row_fwd<NP, MASKED, true, false, ADP><<<grid(nh), wg, 0, s>>>(ra);
            else sgm_row_fwd<NP, MASKED, false, false, ADP><<<grid(nh), wg, 0, s>>>(ra);
        }
        if (b.cl) start_chain(r);
        mark();
        vertical_diagonal(std::integral_constant<int, 1>(), b.cols, b.colmins);
    }
    if (r.nvis > 0) sgm_replay_paths<NP, MASKED, VAR><<<blocks(r.nvis), kBlock, 0, s>>>(r);
    sgm_replay_seeds<NP, MASKED, VAR, ADP><<<1, 64, 0, s>>>(r, r.visits);
    end_chain();
    sgm_replay_add<NP, VAR><<<blocks(g.H), kBlock, 0, s>>>(r, b.cellvis_plus);
    mark();
    if (agg_pass1 && fused) {  // pass +1's paths 1-3 are still in their volumes
        agg_sum4<<<grid_for(aggn, 256), 256, 0, s>>>((uint32_t*)agg_pass1, (const uint32_t*)b.agg, (const uint32_t*)b.aggp[0],
                                                     (const uint32_t*)b.aggp[1], (const uint32_t*)b.aggp[2], aggn);
    } else if (agg_pass1) {
        hipError_t e = hipMemcpyAsync(agg_pass1, b.agg, (size_t)g.W * g.H * g.cs * 2, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
    }
    // pass -1: up-left, up, up-right (starting from pass +1's half-0 leftovers), then <- with the
    // winner-take-all fused (the last contribution)
    r.dir = -1; r.col = 0; r.visits = b.visits_minus; r.nvis = b.n_minus;
    r.colagg = pv ? b.bagg[1] : nullptr;
    r.cols = b.cols + colsz; r.colmins = b.colmins + colminsz;
    a.cols = b.cols + colsz; a.colmins = b.colmins + colminsz;
    ra.cols = b.cols + colsz; ra.colmins = b.colmins + colminsz;
    if (fused) {
        launch_pass(std::integral_constant<int, -1>(), r);
        mark();
        mark();
        mark();
    } else {
        start_chain(r);
        vertical_diagonal(std::integral_constant<int, -1>(), b.cols + colsz, b.colmins + colminsz);
    }
    if (fused) {
        mark();  // the row path ran in the pass launch
    } else {
        if constexpr (SEG > 0) {
            sgm_row_bwd_wta_seg<SEG><<<sblocks(nh), kBlock, 0, s>>>(ra);
        } else if constexpr (kPv) {
            if (pv && o.write_agg) sgm_row_bwd_wta_pv<NP, MASKED, 1><<<grid(nh), wg, 0, s>>>(ra);
            else if (pv) sgm_row_bwd_wta_pv<NP, MASKED, 0><<<grid(nh), wg, 0, s>>>(ra);
            else sgm_row_bwd_wta<NP, MASKED, VAR, ADP><<<grid(nh), wg, 0, s>>>(ra);
        } else {
            sgm_row_bwd_wta<NP, MASKED, VAR, ADP><<<grid(nh), wg, 0, s>>>(ra);
        }
        mark();
    }
    if (r.nvis > 0) sgm_replay_paths<NP, MASKED, VAR><<<blocks(r.nvis), kBlock, 0, s>>>(r);
    end_chain();
    sgm_replay_add<NP, VAR><<<blocks(g.H), kBlock, 0, s>>>(r, b.cellvis_minus);
    mark();
    if (fused) {  // the sum of pass -1's volumes and the winner-take-all of every cell
        const size_t waves = (size_t)g.H * ((g.W + 63) / 64);
        sgm_finish<NP><<<(unsigned)((waves + 3) / 4), kBlock, 0, s>>>(ra, b.aggp[0], b.aggp[1], b.aggp[2]);
    } else {
        if (pv) ra.bagg_minus = b.bagg[1];  // column 0 from the volumes and pass -1's border replay
        wta_finish<NP, VAR><<<g.H, 64, 0, s>>>(ra, o.tol, VAR ? nullptr : o.valid);
    }
    mark();
    return hipGetLastError();
}

}  // namespace

template <bool ADP>
hipError_t launch_sgm_t(const Geometry& g, const SgmBuffers& b, const SgmOutputs& o, int p1, int p2, uint16_t* agg_pass1,
                        hipStream_t s, hipEvent_t* m)
{
    const bool full = g.cap == 64 * g.dpl;
    if (b.vstart) {  // per-pixel ranges: 1, 2 or 4 packed pairs a lane (caps up to 128, 256, 512), per-cell masks
        if (g.dpl == 4) return run_sgm<2, true, true, ADP>(g, b, o, p1, p2, agg_pass1, s, m);
        if (g.dpl == 8) return run_sgm<4, true, true, ADP>(g, b, o, p1, p2, agg_pass1, s, m);
        if (g.dpl != 2) return hipErrorInvalidValue;
        if constexpr (!ADP) {  // narrow ranges: several chains per wavefront
            if (b.seg && g.cap <= 16) return run_sgm<1, true, true, false, 8>(g, b, o, p1, p2, agg_pass1, s, m);
            if (b.seg && g.cap <= 32) return run_sgm<1, true, true, false, 16>(g, b, o, p1, p2, agg_pass1, s, m);
        }
        return run_sgm<1, true, true, ADP>(g, b, o, p1, p2, agg_pass1, s, m);
    }
    switch (g.dpl) {
        case 2: return full ? run_sgm<1, false, false, ADP>(g, b, o, p1, p2, agg_pass1, s, m) : run_sgm<1, true, false, ADP>(g, b, o, p1, p2, agg_pass1, s, m);
        case 4: return full ? run_sgm<2, false, false, ADP>(g, b, o, p1, p2, agg_pass1, s, m) : run_sgm<2, true, false, ADP>(g, b, o, p1, p2, agg_pass1, s, m);
        case 8: return full ? run_sgm<4, false, false, ADP>(g, b, o, p1, p2, agg_pass1, s, m) : run_sgm<4, true, false, ADP>(g, b, o, p1, p2, agg_pass1, s, m);
        case 16: return full ? run_sgm<8, false, false, ADP>(g, b, o, p1, p2, agg_pass1, s, m) : run_sgm<8, true, false, ADP>(g, b, o, p1, p2, agg_pass1, s, m);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lux
