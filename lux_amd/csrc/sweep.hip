// The plane-sweep matching cost of the L16 dense stereo (Lux.Engine/Pipeline/Registration/DenseStereo.cs:217-381) on the
// device, and the entry points built on it.
//
// A layer pixel (x, y) looks at reference image pixel (min(scale*x + scale/2, W0-1), min(scale*y + scale/2, H0-1))
// (DenseStereoSgm.cs:195). For every plane of its range and every other camera, the pixel is projected with the
// camera's flow matrix; the 3x3 block around the projection (rows and pixels averaged with their successors at
// half-pixel positions) is compared with the reference's 3x3 block, channel by channel, with capped absolute
// differences; the weighted channel sums are clamped to 65535 and summed over the cameras in 16 bits; Normalise turns the
// sum into the raw byte. A plane whose projection leaves the camera's rect repeats the cost of the last plane that
// did not (the reference keeps a running `lastCost`); the first plane of a range is clamped into the rect instead.
//
// Device mapping: G lanes per pixel (G = 8 ... 64, from the widest range), one plane per lane, chunks of G planes. A
// lane evaluates its own plane; the carried cost is a segmented scan of "last lane that evaluated" within the G lanes,
// seeded by the previous chunk's last value. Every float operation is one rounded operation in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/lux_sgm.h"
#include "ctx.h"
#include "engine.h"
#include "kernels.h"
#include "rcp_dev.h"

namespace lux {

constexpr int kMaxViews = LUX_SGM_MAX_VIEWS;

// Kernel argument: the reference, the cameras and the planes, all device pointers.
struct SweepDev {
    const uint32_t* ref = nullptr;  // RGBA8 as u32, rw x rh
    int rw = 0, rh = 0;
    int nv = 0;
    const float* planes = nullptr;
    int n = 0;
    uint32_t cap_ev = 0, cap_od = 0;  // caps of channels {0, 2} and {1, 3} as u16 pairs
    float fnorm = 0.f;                // 0x3d17b426 / nv (Normalise, :378)
    const uint32_t* img[kMaxViews] = {};
    int w[kMaxViews] = {}, h[kMaxViews] = {};
    float m[kMaxViews][12] = {};      // flow matrix entries 0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14
    uint32_t wq[kMaxViews][4] = {};
};

namespace {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }

// float -> int as the reference's runtime converts: saturating, NaN -> 0.
__device__ __forceinline__ int f2i_sat(float f)
{
    return f != f ? 0 : (f >= 2147483648.0f ? INT_MAX : (f < -2147483648.0f ? INT_MIN : (int)f));
}

// pavgb on four bytes: (a + b + 1) >> 1 per byte = (a | b) - ((a ^ b) >> 1 per byte).
__device__ __forceinline__ uint32_t avg4(uint32_t a, uint32_t b) { return (a | b) - (((a ^ b) >> 1) & 0x7f7f7f7fu); }

// One camera's cost of one plane (:276-291): the 4x4 target block at (ix-1, iy-1); rows averaged with the next row
// for a half-pixel y, then pixels with the next pixel for a half-pixel x; capped per-channel differences against the
// reference block (refe/refo: channels {0, 2} / {1, 3} of its 9 pixels) summed over the 3x3 pixels, weighted, clamped.
__device__ __forceinline__ uint32_t block_cost(const uint32_t* img, int iw, int ix, int iy, int hx, int hy, const u16x2* refe,
                                               const u16x2* refo, u16x2 cape, u16x2 capo, const uint32_t* wq)
{
    const uint32_t* b = img + (size_t)(iy - 1) * iw + (ix - 1);
    uint32_t T[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int p = 0; p < 4; p++) T[q][p] = b[(size_t)q * iw + p];
    u16x2 se = {0, 0}, so = {0, 0};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        uint32_t a[4];
#pragma unroll
        for (int p = 0; p < 4; p++) a[p] = hy ? avg4(T[q][p], T[q + 1][p]) : T[q][p];
#pragma unroll
        for (int p = 0; p < 3; p++) {
            const uint32_t t = hx ? avg4(a[p], a[p + 1]) : a[p];
            const u16x2 te = as_u16x2(t & 0x00ff00ffu), to = as_u16x2((t >> 8) & 0x00ff00ffu);
            const u16x2 re = refe[q * 3 + p], ro = refo[q * 3 + p];
            se += __builtin_elementwise_min(__builtin_elementwise_max(te, re) - __builtin_elementwise_min(te, re), cape);
            so += __builtin_elementwise_min(__builtin_elementwise_max(to, ro) - __builtin_elementwise_min(to, ro), capo);
        }
    }
    const uint32_t cY = ((uint32_t)se.x * wq[0] + 16u) >> 5, cU = ((uint32_t)so.x * wq[1] + 16u) >> 5,
                   cV = ((uint32_t)se.y * wq[2] + 16u) >> 5, cA = ((uint32_t)so.y * wq[3] + 16u) >> 5;
    const uint32_t ck = (cY + cU) + (cV + cA);
    return ck < 65535u ? ck : 65535u;
}

// raw[cell*cs + j], j < cap: the normalised cost for j < count (0 for a skipped pixel), 255 past count (bytes past a
// pixel's cap are masked by the sweeps and left unwritten). vstart / vcount / vcap null: every pixel searches planes
// [0, ucount) with cap ucap.
// NV > 0: exactly NV other cameras, 1 to 4 (the unrolled view loop then keeps only their matrices live; unrolled over
// all kMaxViews, the sixteen matrices spill from SGPRs into VGPR lanes); NV = 0: a.nv of up to kMaxViews.
template <int G, int NV>
__global__ __launch_bounds__(256) void sweep_cost_kernel(const SweepDev a, int W, int H, int scale, const uint16_t* vstart,
                                                         const uint16_t* vcount, const uint16_t* vcap, int ucount, int ucap,
                                                         const uint8_t* skip, const uint32_t* list, const uint32_t* nlist,
                                                         uint8_t* raw, int cs)
{
    const int lane = threadIdx.x & 63, li = lane & (G - 1);
    const size_t seg = ((size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 / G) + (lane / G);
    const size_t cells = (size_t)W * H;
    // with a list, segment i takes the i-th computed cell; else cell i
    const size_t nseg = list ? (size_t)*nlist : cells;
    const bool live = seg < nseg;
    const size_t cell = live ? (list ? (size_t)list[seg] : seg) : 0;
    int start = 0, count = 0, cap = 0;
    bool skipped = false;
    int x = 0, y = 0;
    if (live) {
        start = vstart ? (int)vstart[cell] : 0;
        count = vcount ? (int)vcount[cell] : ucount;
        cap = vcap ? (int)vcap[cell] : ucap;
        skipped = skip && skip[cell];
        y = (int)(cell / W);
        x = (int)(cell - (size_t)y * W);
    }
    const int half = scale / 2;
    const int xi = min(scale * x + half, a.rw - 1), yi = min(scale * y + half, a.rh - 1);
    u16x2 refe[9], refo[9];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const uint32_t v = a.ref[(size_t)min(max(yi + r - 1, 0), a.rh - 1) * a.rw + min(max(xi + c - 1, 0), a.rw - 1)];
            refe[r * 3 + c] = as_u16x2(v & 0x00ff00ffu);
            refo[r * 3 + c] = as_u16x2((v >> 8) & 0x00ff00ffu);
        }
    const u16x2 cape = as_u16x2(a.cap_ev), capo = as_u16x2(a.cap_od);
    const float xf = (float)xi, yf = (float)yi;
    // every lane runs the wave's longest row of chunks, so the shuffles below see whole segments
    int need = live ? cap : 0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) need = max(need, __shfl_xor(need, o));
    const bool eval = live && !skipped;
    constexpr int KV = NV > 0 ? NV : kMaxViews;
    uint32_t carry[KV];
#pragma unroll
    for (int k = 0; k < KV; k++) carry[k] = 0;
    for (int c0 = 0; c0 < need; c0 += G) {
        const int d = c0 + li;
        const bool act = eval && d < count;
        const float z = act ? a.planes[start + d] : 1.0f;
        uint32_t tot = 0;
#pragma unroll
        for (int k = 0; k < KV; k++) {
            if (NV == 0 && k >= a.nv) continue;
            const float* m = a.m[k];
            const float xz = __fmul_rn(xf, z), yz = __fmul_rn(yf, z);
            const float px = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(z, m[6]), m[9]), __fmul_rn(xz, m[0])), __fmul_rn(yz, m[3]));
            const float py = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(z, m[7]), m[10]), __fmul_rn(xz, m[1])), __fmul_rn(yz, m[4]));
            const float pz = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(z, m[8]), m[11]), __fmul_rn(xz, m[2])), __fmul_rn(yz, m[5]));
            const float inv = __fdiv_rn(1.0f, pz);
            float fx = __fadd_rn(__fmul_rn(px, inv), 0.25f), fy = __fadd_rn(__fmul_rn(py, inv), 0.25f);
            const int iw = a.w[k], ih = a.h[k];
            const int x0 = 1, y0 = 1, x1 = iw - 2, y1 = ih - 2;  // rect + {1, 1, -2, -2}
            if (d == 0) {                                         // the first plane is clamped into the rect
                fx = fminf(fmaxf(fx, (float)x0), (float)(x1 - 1));
                fy = fminf(fmaxf(fy, (float)y0), (float)(y1 - 1));
            }
            const int ix = f2i_sat(fx), iy = f2i_sat(fy);
            const bool ok = act && (d == 0 || (ix >= x0 && ix < x1 && iy >= y0 && iy < y1));
            uint32_t cost = 0;
            if (ok) {
                const int hx = f2i_sat(__fadd_rn(fx, fx)) & 1, hy = f2i_sat(__fadd_rn(fy, fy)) & 1;
                cost = block_cost(a.img[k], iw, ix, iy, hx, hy, refe, refo, cape, capo, a.wq[k]);
            }
            // the cost of the last lane at or below this one that evaluated, else the carried one
            int v = ok ? li : -1;
            if constexpr (G == 16) {  // a segment is a DPP row: row_shr, lanes without a source taking -1
                v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x111, 0xf, 0xf, false));
                v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x112, 0xf, 0xf, false));
                v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x114, 0xf, 0xf, false));
                v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x118, 0xf, 0xf, false));
            } else {
#pragma unroll
                for (int o = 1; o < G; o <<= 1) {
                    const int t = __shfl_up(v, o, G);
                    if (li >= o) v = max(v, t);
                }
            }
            uint32_t got = (uint32_t)__shfl((int)cost, v < 0 ? 0 : v, G);
            got = v < 0 ? carry[k] : got;
            carry[k] = (uint32_t)__shfl((int)got, G - 1, G);
            tot += got;
        }
        if (live && d < cap) {
            uint8_t out = 255;
            if (d < count) out = skipped ? 0 : (uint8_t)f2i_sat(__fmul_rn((float)(tot & 0xffffu), a.fnorm));
            raw[cell * cs + d] = out;
        }
    }
}

// Guidance (DenseStereo.cs:162-175): the reference itself at its own size, else the pixel at
// (min(scale*x + scale/2, rw-1), min(scale*y + scale/2, rh-1)).
__global__ void guidance_kernel(const uint32_t* ref, int rw, int rh, int W, int H, int scale, uint32_t* out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    const int half = scale / 2;
    const int sx = (rw == W && rh == H) ? x : min(scale * x + half, rw - 1);
    const int sy = (rw == W && rh == H) ? y : min(scale * y + half, rh - 1);
    out[(size_t)y * W + x] = ref[(size_t)sy * rw + sx];
}

// RangeFromPrevious, first step (DenseStereoSgm.cs:95-104): the plane index of each previous pixel, the number of
// leading boundaries b[j] <= rcpss(depth). With non-decreasing boundaries (mono) that is a binary search.
__global__ void range_index_kernel(const float* prev, size_t np, const float* b, int n, int mono, uint16_t* idx)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float v = rcp_intel(prev[i]);
    int c = 0;
    if (mono) {
        int lo = 0, hi = n;  // first j with !(b[j] <= v)
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (b[mid] <= v) lo = mid + 1;
            else hi = mid;
        }
        c = lo;
    } else {
        while (c < n && b[c] <= v) c++;
    }
    idx[i] = (uint16_t)c;
}

// Statistics of a layer's ranges, as 64-bit words: 0 MinLo, 1 max hi, 2 max cap, 3 max count, 4 max(start + cap),
// 5 a range past the last plane, 6 sum of caps of computed pixels, 7 sum of caps of skipped pixels, 8 computed pixels.
constexpr int kStats = 9;
constexpr int kRangeBlock = 256;
// Reduces the per-block partials of range_window_kernel (min for word 0, max for 1-5, sums for 6-8) into st.
__global__ __launch_bounds__(1024) void stats_reduce_kernel(const unsigned long long* part, int nblocks, unsigned long long* st)
{
    __shared__ unsigned long long red[16][kStats];
    unsigned long long r[kStats];
    r[0] = ~0ull;
#pragma unroll
    for (int k = 1; k < kStats; k++) r[k] = 0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        const unsigned long long* q = part + (size_t)b * kStats;
        r[0] = min(r[0], q[0]);
#pragma unroll
        for (int k = 1; k < 6; k++) r[k] = max(r[k], q[k]);
#pragma unroll
        for (int k = 6; k < kStats; k++) r[k] += q[k];
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        r[0] = min(r[0], (unsigned long long)__shfl_xor((long long)r[0], o));
#pragma unroll
        for (int k = 1; k < 6; k++) r[k] = max(r[k], (unsigned long long)__shfl_xor((long long)r[k], o));
#pragma unroll
        for (int k = 6; k < kStats; k++) r[k] += (unsigned long long)__shfl_xor((long long)r[k], o);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < kStats; k++) red[wv][k] = r[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int w = 1; w < nw; w++) {
            r[0] = min(r[0], red[w][0]);
            for (int k = 1; k < 6; k++) r[k] = max(r[k], red[w][k]);
            for (int k = 6; k < kStats; k++) r[k] += red[w][k];
        }
        for (int k = 0; k < kStats; k++) st[k] = r[k];
    }
}

// RangeFromPrevious, the window (:105-126): min / max of the computed pixels' indices over x-1..x+2 (clamped),
// 0xffff / 0 when none is computed; then over y-1..y+2 of those. Packed min | max << 16.
__global__ void range_hwin_kernel(const uint16_t* idx, const uint8_t* pskip, int pw, int ph, uint32_t* hmm)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= pw) return;
    int mn = 0xffff, mx = 0;
#pragma unroll
    for (int k = -1; k <= 2; k++) {
        const size_t i = (size_t)y * pw + min(max(x + k, 0), pw - 1);
        if (pskip && pskip[i]) continue;
        mn = min(mn, (int)idx[i]);
        mx = max(mx, (int)idx[i]);
    }
    hmm[(size_t)y * pw + x] = (uint32_t)mn | ((uint32_t)mx << 16);
}
__global__ void range_vwin_kernel(const uint32_t* hmm, int pw, int ph, uint32_t* vmm)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= pw) return;
    int mn = 0xffff, mx = 0;
#pragma unroll
    for (int k = -1; k <= 2; k++) {
        const uint32_t v = hmm[(size_t)min(max(y + k, 0), ph - 1) * pw + x];
        mn = min(mn, (int)(v & 0xffffu));
        mx = max(mx, (int)(v >> 16));
    }
    vmm[(size_t)y * pw + x] = (uint32_t)mn | ((uint32_t)mx << 16);
}

// RangeFromPrevious, the ranges (:127-139): the window at ((int)(x * fx), (int)(y * fy)); start = max(0, min - margin),
// count = min(n - 1, max + 1) - start, in 16 bits as the reference stores them; cap = roundup(count, align).
__global__ __launch_bounds__(kRangeBlock) void range_window_kernel(const uint32_t* vmm, int pw, int ph, int W, int H, float fx,
                                                                   float fy, int margin, int n, int align, const uint8_t* skip,
                                                                   uint16_t* start, uint16_t* count, uint16_t* cap,
                                                                   unsigned long long* part)
{
    const int x0 = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    const bool live = x0 < W;
    const int x = live ? x0 : W - 1;  // lanes past the row repeat its last pixel for the reductions, write nothing
    const int py = f2i_sat(__fmul_rn((float)y, fy)), px = f2i_sat(__fmul_rn((float)x, fx));
    const uint32_t v = vmm[(size_t)py * pw + px];
    const int mn = (int)(v & 0xffffu), mx = (int)(v >> 16);
    const int lo = max(0, mn - margin), hi = min(n - 1, mx + 1);
    const int s = lo & 0xffff, c = (hi - lo) & 0xffff, cp = ((c + align - 1) / align) * align;
    const size_t o = (size_t)y * W + x;
    if (live) {
        start[o] = (uint16_t)s;
        count[o] = (uint16_t)c;
        cap[o] = (uint16_t)cp;
    }
    // one atomic per wave and statistic: reduce over the wave first
    const bool sk = skip && skip[o];
    int r0 = s, r1 = hi & 0xffff, r2 = cp, r3 = c, r4 = s + cp, r5 = (s + c > n || cp > 0xffff) ? 1 : 0;
    uint32_t r6 = live && !sk ? (uint32_t)cp : 0u, r7 = live && sk ? (uint32_t)cp : 0u, r8 = live && !sk ? 1u : 0u;
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) {
        r8 += (uint32_t)__shfl_xor((int)r8, k);
        r0 = min(r0, __shfl_xor(r0, k));
        r1 = max(r1, __shfl_xor(r1, k));
        r2 = max(r2, __shfl_xor(r2, k));
        r3 = max(r3, __shfl_xor(r3, k));
        r4 = max(r4, __shfl_xor(r4, k));
        r5 = max(r5, __shfl_xor(r5, k));
        r6 += (uint32_t)__shfl_xor((int)r6, k);
        r7 += (uint32_t)__shfl_xor((int)r7, k);
    }
    // the block's partial (no atomics: a second kernel reduces the partials)
    __shared__ uint32_t red[kRangeBlock / 64][kStats];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv][0] = (uint32_t)r0; red[wv][1] = (uint32_t)r1; red[wv][2] = (uint32_t)r2;
        red[wv][3] = (uint32_t)r3; red[wv][4] = (uint32_t)r4; red[wv][5] = (uint32_t)r5;
        red[wv][6] = r6; red[wv][7] = r7; red[wv][8] = r8;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long q[kStats];
        for (int k = 0; k < kStats; k++) q[k] = red[0][k];
        for (int w = 1; w < kRangeBlock / 64; w++) {
            q[0] = min(q[0], (unsigned long long)red[w][0]);
            for (int k = 1; k < 6; k++) q[k] = max(q[k], (unsigned long long)red[w][k]);
            for (int k = 6; k < kStats; k++) q[k] += red[w][k];
        }
        unsigned long long* o = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kStats;
        for (int k = 0; k < kStats; k++) o[k] = q[k];
    }
}

// Skipped pixels get their raw bytes here (0 over the planes, 255 over the padding, as Pixel uses them, :242); the
// computed ones are listed for the cost kernel (in no particular order: each is costed on its own). One wavefront
// per kSkipGroups x 64 cells: each skipped cell's bytes are written by the whole wavefront, four bytes a lane (cs is
// a multiple of 4; the cells' counts and caps are read 64 at a time and taken out with v_readlane), and the
// computed cells take one list reservation per wavefront.
constexpr int kSkipGroups = 4;
__global__ __launch_bounds__(256) void skip_list_kernel(const uint8_t* skip, const uint16_t* vcount, const uint16_t* vcap,
                                                        size_t cells, uint8_t* raw, int cs, uint32_t* list, uint32_t* nlist)
{
    const int lane = threadIdx.x & 63;
    const size_t c0 = (((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * (64 * kSkipGroups);
    if (c0 >= cells) return;
    uint64_t comp[kSkipGroups];
    uint32_t total = 0;
#pragma unroll
    for (int g = 0; g < kSkipGroups; g++) {
        const size_t i = c0 + (size_t)g * 64 + lane;
        const bool live = i < cells;
        const bool sk = live && skip[i];
        const int mycnt = live ? vcount[i] : 0, mycap = live ? vcap[i] : 0;
        comp[g] = __ballot(live && !sk);
        total += (uint32_t)__popcll(comp[g]);
        for (uint64_t sm = __ballot(sk); sm; sm &= sm - 1) {
            const int b = __builtin_ctzll(sm);
            const size_t c = c0 + (size_t)g * 64 + (size_t)b;
            const int cnt = __builtin_amdgcn_readlane(mycnt, b), cp = __builtin_amdgcn_readlane(mycap, b);
            for (int j = lane * 4; j < cp; j += 256) {
                uint32_t w = 0;
#pragma unroll
                for (int e = 0; e < 4; e++) w |= (j + e < cnt ? 0u : 0xffu) << (8 * e);
                if (j + 4 <= cp) {
                    *(uint32_t*)(raw + c * cs + j) = w;
                } else {
                    for (int e = 0; j + e < cp; e++) raw[c * cs + j + e] = (uint8_t)(w >> (8 * e));
                }
            }
        }
    }
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(nlist, total);
    base = (uint32_t)__shfl((int)base, 0);
#pragma unroll
    for (int g = 0; g < kSkipGroups; g++) {
        const uint64_t m = comp[g];
        if ((m >> lane) & 1ull) list[base + __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)(c0 + (size_t)g * 64 + lane);
        base += (uint32_t)__popcll(m);
    }
}

// The layer's depth (Compute :225-231): Planes[start + first minimum], the WTA index the sweeps leave.
__global__ void depth_kernel(const uint16_t* wta, const float* planes, int n, size_t cells, float* depth)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const int k = wta[i];
    depth[i] = k < n ? planes[k] : 0.0f;
}

}  // namespace

int check_sweep(const lux_sgm_sweep* s, std::string* err)
{
    if (!s || !s->ref || !s->views || !s->planes) {
        if (err) *err = "null sweep argument";
        return LUX_SGM_EINVAL;
    }
    if (s->n_views < 1 || s->n_views > kMaxViews) {
        if (err) *err = "1 to 16 other cameras";
        return LUX_SGM_EINVAL;
    }
    if (s->ref_width <= 0 || s->ref_height <= 0 || s->n_planes <= 0 || s->n_planes > 65535) {
        if (err) *err = "bad reference size or plane count";
        return LUX_SGM_EINVAL;
    }
    for (int k = 0; k < s->n_views; k++) {
        const lux_sgm_view& v = s->views[k];
        if (!v.rgba || v.width < 4 || v.height < 4) {
            if (err) *err = "every camera image needs at least 4 x 4 pixels";
            return LUX_SGM_EINVAL;
        }
    }
    return 0;
}

// Copies the sweep's images and planes into the workspace and fills the kernel argument.
int upload_sweep(Workspace& ws, const lux_sgm_sweep* s, bool device, hipStream_t st, SweepDev* out, std::string* err)
{
    int rc = check_sweep(s, err);
    if (rc) return rc;
    SweepDev a;
    a.rw = s->ref_width;
    a.rh = s->ref_height;
    a.nv = s->n_views;
    a.n = s->n_planes;
    a.cap_ev = (uint32_t)s->caps[0] | ((uint32_t)s->caps[2] << 16);
    a.cap_od = (uint32_t)s->caps[1] | ((uint32_t)s->caps[3] << 16);
    uint32_t fb = 0x3d17b426u;
    float f27;
    std::memcpy(&f27, &fb, 4);
    a.fnorm = f27 / (float)s->n_views;
    const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    auto put = [&](const char* name, const void* src, size_t bytes, const void** dst) -> int {
        if (device) {
            *dst = src;
            return 0;
        }
        void* p;
        int r = ws.get(name, bytes, &p, err);
        if (r) return r;
        hipError_t e = hipMemcpyAsync(p, src, bytes, kind, st);
        if (e != hipSuccess) return hip_status(e, "upload sweep", err);
        *dst = p;
        return 0;
    };
    const void* d;
    if ((rc = put("sw_ref", s->ref, (size_t)a.rw * a.rh * 4, &d))) return rc;
    a.ref = (const uint32_t*)d;
    if ((rc = put("sw_planes", s->planes, (size_t)a.n * 4, &d))) return rc;
    a.planes = (const float*)d;
    static const char* names[kMaxViews] = {"sw_img0", "sw_img1", "sw_img2",  "sw_img3",  "sw_img4",  "sw_img5",
                                           "sw_img6", "sw_img7", "sw_img8",  "sw_img9",  "sw_img10", "sw_img11",
                                           "sw_img12", "sw_img13", "sw_img14", "sw_img15"};
    static const int idx[12] = {0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14};
    for (int k = 0; k < a.nv; k++) {
        const lux_sgm_view& v = s->views[k];
        if ((rc = put(names[k], v.rgba, (size_t)v.width * v.height * 4, &d))) return rc;
        a.img[k] = (const uint32_t*)d;
        a.w[k] = v.width;
        a.h[k] = v.height;
        for (int j = 0; j < 12; j++) a.m[k][j] = v.warp[idx[j]];
        for (int j = 0; j < 4; j++) a.wq[k][j] = v.wq[j];
    }
    *out = a;
    return 0;
}

// G (lanes per pixel) follows the mean cap: wide enough that a typical pixel needs one chunk, narrow enough that
// the lanes of short ranges are not idle; a wave loops over its widest pixel's chunks.
hipError_t launch_sweep_cost(const SweepDev& a, int W, int H, int scale, const uint16_t* vstart, const uint16_t* vcount,
                             const uint16_t* vcap, int ucount, int ucap, double mean_cap, const uint8_t* skip, uint8_t* raw,
                             int cs, hipStream_t s, Workspace* ws = nullptr, uint64_t computed = 0)
{
    const size_t cells = (size_t)W * H;
    int G = 8;
    while (G < 64 && G < mean_cap) G <<= 1;
    // ranged layers with a skip mask: skipped pixels are written by a list kernel, the cost kernel runs over the
    // computed ones only (`computed` of them)
    const uint32_t* list = nullptr;
    const uint32_t* nlist = nullptr;
    size_t nseg = cells;
    if (ws && skip && vcount && vcap) {
        void *dl, *dn;
        std::string e2;
        if (ws->get("sw_list", cells * 4, &dl, &e2) || ws->get("sw_nlist", 256, &dn, &e2)) return hipErrorOutOfMemory;
        hipError_t e = launch_fill_u32((uint32_t*)dn, 1, 0u, s);
        if (e != hipSuccess) return e;
        const size_t waves = (cells + 64 * kSkipGroups - 1) / (64 * kSkipGroups);
        skip_list_kernel<<<(unsigned)((waves + 3) / 4), 256, 0, s>>>(skip, vcount, vcap, cells, raw, cs, (uint32_t*)dl,
                                                                     (uint32_t*)dn);
        list = (const uint32_t*)dl;
        nlist = (const uint32_t*)dn;
        nseg = computed;
        skip = nullptr;  // every listed cell is computed
    }
    const size_t per_block = 4 * (64 / G);
    const unsigned blocks = (unsigned)std::max<size_t>(1, (nseg + per_block - 1) / per_block);
    auto go = [&](auto gc, auto nvc) {
        sweep_cost_kernel<decltype(gc)::value, decltype(nvc)::value><<<blocks, 256, 0, s>>>(
            a, W, H, scale, vstart, vcount, vcap, ucount, ucap, skip, list, nlist, raw, cs);
    };
    auto by_views = [&](auto gc) {
        switch (a.nv) {
            case 1: go(gc, std::integral_constant<int, 1>()); break;
            case 2: go(gc, std::integral_constant<int, 2>()); break;
            case 3: go(gc, std::integral_constant<int, 3>()); break;
            case 4: go(gc, std::integral_constant<int, 4>()); break;
            default: go(gc, std::integral_constant<int, 0>()); break;
        }
    };
    switch (G) {
        case 8: by_views(std::integral_constant<int, 8>()); break;
        case 16: by_views(std::integral_constant<int, 16>()); break;
        case 32: by_views(std::integral_constant<int, 32>()); break;
        default: by_views(std::integral_constant<int, 64>()); break;
    }
    return hipGetLastError();
}

// The plane boundaries of RangeFromPrevious (:94-95) on the host: midpoints of neighbouring planes' rcpss, the last
// FLT_MAX; and whether they never decrease.
static std::vector<float> range_boundaries(const float* planes, int n, int* mono)
{
    std::vector<float> b(n);
    for (int i = 0; i + 1 < n; i++) b[i] = (rcp_intel_host(planes[i + 1]) + rcp_intel_host(planes[i])) * 0.5f;
    b[n - 1] = 3.40282347e38f;
    *mono = 1;
    for (int i = 0; i + 1 < n; i++)
        if (!(b[i] <= b[i + 1])) *mono = 0;
    return b;
}

// Host copies of a layer's range statistics.
struct RangeStats {
    int minlo = 0, maxhi = 0, maxcap = 0, maxcount = 0, maxend = 0, bad = 0;
    uint64_t cap_computed = 0, cap_skipped = 0, computed = 0;
};

// RangeFromPrevious on the device for a W x H layer from the previous layer's depth and skip (device), writing
// start/count/cap (device) and returning the statistics (synchronises the stream). `tag` keeps the buffers of
// two uses apart (the pyramid asks for align 8 and align 2 of one layer).
// Copies n floats into the workspace buffer `name` (at dst) unless it already holds them: the plane lists and their
// boundaries repeat across a pyramid's levels and its calls.
static hipError_t upload_once(Workspace& ws, const char* name, void* dst, const float* src, int n, hipStream_t s)
{
    std::vector<float>& held = ws.uploaded[name];
    if (ws.filled.count(name) && held.size() == (size_t)n && std::memcmp(held.data(), src, (size_t)n * 4) == 0)
        return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, (size_t)n * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    held.assign(src, src + n);
    ws.filled.insert(name);
    return hipSuccess;
}

int ranges_from_previous(Workspace& ws, const float* planes, int n, const float* d_prev, const uint8_t* d_pskip, int pw, int ph,
                         int W, int H, int margin, int align, const uint8_t* d_skip, uint16_t* d_start, uint16_t* d_count,
                         uint16_t* d_cap, RangeStats* out, hipStream_t s, std::string* err)
{
    if (n <= 0 || n > 65535) {
        if (err) *err = "plane count must be 1..65535";
        return LUX_SGM_EINVAL;
    }
    int mono = 0;
    std::vector<float> b = range_boundaries(planes, n, &mono);
    void *db, *didx, *dst, *dh, *dv, *dpart;
    int rc;
    const size_t np = (size_t)pw * ph;
    const int gx = (W + kRangeBlock - 1) / kRangeBlock, nblocks = gx * H;
    if ((rc = ws.get("rg_bounds", (size_t)n * 4, &db, err)) || (rc = ws.get("rg_idx", np * 2, &didx, err)) ||
        (rc = ws.get("rg_stats", kStats * 8, &dst, err)) || (rc = ws.get("rg_hwin", np * 4, &dh, err)) ||
        (rc = ws.get("rg_vwin", np * 4, &dv, err)) || (rc = ws.get("rg_part", (size_t)nblocks * kStats * 8, &dpart, err)))
        return rc;
    hipError_t e = upload_once(ws, "rg_bounds", db, b.data(), n, s);
    if (e != hipSuccess) return hip_status(e, "upload boundaries", err);
    range_index_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(d_prev, np, (const float*)db, n, mono, (uint16_t*)didx);
    const float fy = (float)(ph - 1) / (float)(H - 1), fx = (float)(pw - 1) / (float)(W - 1);
    // (int)(y * fy) stays inside the previous layer: y * fy <= (H-1) * (ph-1)/(H-1) up to rounding, checked here
    auto f2i = [](float f) { return f != f ? 0 : (f >= 2147483648.0f ? INT_MAX : (f < -2147483648.0f ? INT_MIN : (int)f)); };
    const int pyl = f2i((float)(H - 1) * fy), pxl = f2i((float)(W - 1) * fx);
    if (pyl > ph - 1 || pxl > pw - 1) {
        if (err) *err = "resampled position past the previous layer";
        return LUX_SGM_EINVAL;
    }
    range_hwin_kernel<<<dim3((pw + 127) / 128, ph), 128, 0, s>>>((const uint16_t*)didx, d_pskip, pw, ph, (uint32_t*)dh);
    range_vwin_kernel<<<dim3((pw + 127) / 128, ph), 128, 0, s>>>((const uint32_t*)dh, pw, ph, (uint32_t*)dv);
    range_window_kernel<<<dim3(gx, H), kRangeBlock, 0, s>>>((const uint32_t*)dv, pw, ph, W, H, fx, fy, margin, n, align, d_skip,
                                                            d_start, d_count, d_cap, (unsigned long long*)dpart);
    stats_reduce_kernel<<<1, 1024, 0, s>>>((const unsigned long long*)dpart, nblocks, (unsigned long long*)dst);
    unsigned long long st[kStats];
    void* hst;  // pinned: the copy is a DMA the stream wait covers, not a staged copy that blocks first
    if ((rc = ws.pinned(sizeof st, &hst, err))) return rc;
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hst, dst, sizeof st, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e, "ranges from previous", err);
    std::memcpy(st, hst, sizeof st);
    out->minlo = (int)st[0];
    out->maxhi = (int)st[1];
    out->maxcap = (int)st[2];
    out->maxcount = (int)st[3];
    out->maxend = (int)st[4];
    out->bad = (int)st[5];
    out->cap_computed = st[6];
    out->cap_skipped = st[7];
    out->computed = st[8];
    return 0;
}

// One DenseLayer.Compute on the device (DenseStereoSgm.cs:181-232) over device-resident ranges (start null: every
// pixel searches all planes) and skip mask: the plane-sweep cost straight into the sweeps' raw volume, the
// image-adaptive eight-path aggregation, the WTA and the depth. wta and depth are device outputs; the aggregates
// stay in w->agg (dense, stride w->g.cs) when keep_agg.
int dense_layer_device(Workspace& ws, const SweepDev& a, const lux_sgm_dense_params& p, const uint16_t* d_start,
                       const uint16_t* d_count, const uint16_t* d_cap, const RangeStats& rs, const uint8_t* d_skip,
                       bool keep_agg, uint16_t* agg_pass1, uint16_t* d_wta, float* d_depth, Prepared* w, hipStream_t s,
                       std::string* err, const Conc* conc = nullptr)
{
    const int W = p.width, H = p.height;
    const bool ranged = d_start != nullptr;
    auto refuse = [&](int code, const char* msg) {
        if (err) *err = msg;
        return code;
    };
    const int ucap = ((a.n + p.align - 1) / p.align) * p.align;
    bool spill = false;  // some range's padding passes its line-buffer segment (start - MinLo + cap > MaxHi)
    int maxhi_r = ucap, maxcap = ucap;
    if (ranged) {
        maxhi_r = ((rs.maxhi + p.align - 1) / p.align) * p.align;
        maxcap = rs.maxcap;
        if (rs.bad) return refuse(LUX_SGM_EINVAL, "a range passes the last plane");
        if (rs.maxcap == 0) return refuse(LUX_SGM_EINVAL, "every range is empty");
        if (rs.maxcap > 1024) return refuse(LUX_SGM_EUNSUPPORTED, "per-pixel ranges above 1024 lanes");
        spill = rs.maxend - rs.minlo > maxhi_r;
    }
    if (!ranged && ucap > 1024) return refuse(LUX_SGM_EUNSUPPORTED, "more than 1024 lanes per cell");
    // the path sweeps, or the line-buffer-literal path for spilling padding, caps of 513..1024 and odd tiles
    const bool literal = maxcap <= 1024 && (ws.force_literal || spill || (p.tile & 1) || (ranged && maxcap > 512));
    const Geometry g = geometry(W, H, ranged ? rs.maxcount : a.n, p.align);
    int rc = literal ? prepare_literal(ws, g, w, err) : prepare(ws, g, p.tile, 0, false, s, w, err);
    if (rc) return rc;
    if (ranged) {
        void *ss, *sc;
        if ((rc = ws.get("seedstart", (size_t)3 * W * 2, &ss, err)) || (rc = ws.get("seedcap", (size_t)3 * W * 2, &sc, err)))
            return rc;
        w->b.vstart = d_start;
        w->b.vcap = d_cap;
        w->b.vcount = d_count;
        w->b.seedstart = (uint16_t*)ss;
        w->b.seedcap = (uint16_t*)sc;
    }
    const double mean_cap = ranged ? (double)(rs.cap_computed + rs.cap_skipped) / ((double)W * H) : (double)ucap;
    hipError_t e = launch_sweep_cost(a, W, H, p.scale, d_start, d_count, d_cap, a.n, ucap, mean_cap, d_skip, w->raw, g.cs, s,
                                     &ws, rs.computed);
    if (e != hipSuccess) return hip_status(e, "sweep cost", err);
    void *gi, *sg;
    if ((rc = ws.get("guidance", (size_t)W * H * 4, &gi, err)) || (rc = ws.get("seedguid", (size_t)3 * W * 4, &sg, err)))
        return rc;
    guidance_kernel<<<dim3((W + 127) / 128, H), 128, 0, s>>>(a.ref, a.rw, a.rh, W, H, p.scale, (uint32_t*)gi);
    w->b.ap.guid = (const uint32_t*)gi;
    w->b.ap.w0 = p.guidance[0];
    w->b.ap.w1 = p.guidance[1];
    w->b.ap.w2 = p.guidance[2];
    w->b.ap.base = (float)p.p1 * p.p2_scale;  // p2Base (:253), one float multiply
    w->b.ap.on = 1;
    w->b.seedguid = (uint32_t*)sg;
    if (literal) {
        LiteralLayer l;
        l.W = W;
        l.H = H;
        l.T = p.tile;
        l.align = p.align;
        l.minlo = ranged ? rs.minlo : 0;
        l.maxhi = maxhi_r;
        l.start = d_start;
        l.cap = d_cap;
        l.count = d_count;
        l.ucap = ucap;
        l.ucount = a.n;
        l.raw = w->raw;
        l.agg = w->agg;
        l.cs = g.cs;
        l.ap = w->b.ap;
        l.p1 = p.p1;
        l.p2 = 0;
        if ((rc = run_literal(ws, l, maxcap, agg_pass1, d_wta, s, err))) return rc;
    } else {
        if (ranged) try_path_volumes(ws, w);  // ranged layers are latency-bound: one launch per pass
        attach_chain(w, conc);
        SgmOutputs o;
        o.disp = d_wta;
        o.tol = 0;
        o.write_agg = keep_agg;
        o.bands = 0;
        e = launch_sgm(g, w->b, o, p.p1, 0, agg_pass1, s);
        if (e != hipSuccess) return hip_status(e, "aggregation", err);
    }
    const size_t cells = (size_t)W * H;
    depth_kernel<<<(unsigned)((cells + 255) / 256), 256, 0, s>>>(d_wta, a.planes, a.n, cells, d_depth);
    e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_status(e, "depth", err);
}

int check_dense(const lux_sgm_sweep* sw, const lux_sgm_dense_params* p, std::string* err)
{
    int rc = check_sweep(sw, err);
    if (rc) return rc;
    auto bad = [&](const char* msg) {
        if (err) *err = msg;
        return LUX_SGM_EINVAL;
    };
    if (!p) return bad("null params");
    if (p->scale <= 0 || p->align <= 0 || p->tile < 2 || p->p1 < 0 || p->p1 > 65535) return bad("bad layer parameters");
    if (p->width != (sw->ref_width + p->scale - 1) / p->scale || p->height != (sw->ref_height + p->scale - 1) / p->scale)
        return bad("the layer size is (ref + scale - 1) / scale");
    if (p->width > 10240) {
        if (err) *err = "width above 10240";
        return LUX_SGM_EUNSUPPORTED;
    }
    return 0;
}

}  // namespace lux

using lux::fail;
using lux::hip_fail;

extern "C" {

int lux_sgm_sweep_cost(lux_sgm_ctx* c, const lux_sgm_sweep* sw, int32_t width, int32_t height, int32_t scale, int32_t align,
                       const uint16_t* start, const uint16_t* count, const uint8_t* skip, uint8_t* raw)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (width <= 0 || height <= 0 || scale <= 0 || align <= 0 || !raw) return fail(c, LUX_SGM_EINVAL, "bad layer arguments");
    if ((start == nullptr) != (count == nullptr)) return fail(c, LUX_SGM_EINVAL, "start and count go together");
    int rc = lux::check_sweep(sw, &c->err);
    if (rc) return rc;
    const size_t cells = (size_t)width * height;
    const int n = sw->n_planes;
    std::vector<uint16_t> cap(cells);
    std::vector<uint64_t> off(cells);
    uint64_t total = 0;
    int cs = 1;
    for (size_t i = 0; i < cells; i++) {
        const int st = start ? start[i] : 0, cn = count ? count[i] : n;
        if (st + cn > n) return fail(c, LUX_SGM_EINVAL, "a range passes the last plane");
        const int cp = ((cn + align - 1) / align) * align;
        if (cp > 65535) return fail(c, LUX_SGM_EINVAL, "a range's cap exceeds 16 bits");
        cap[i] = (uint16_t)cp;
        off[i] = total;
        total += (uint64_t)cp;
        cs = std::max(cs, cp);
    }
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    lux::Workspace& ws = c->ws;
    lux::SweepDev a;
    if ((rc = lux::upload_sweep(ws, sw, false, s, &a, &c->err))) return rc;
    void *vs = nullptr, *vn = nullptr, *vc, *vo, *sk = nullptr, *dense, *packed;
    if ((rc = ws.get("sw_dense", cells * cs, &dense, &c->err)) || (rc = ws.get("sw_packed", std::max<uint64_t>(total, 1), &packed, &c->err)) ||
        (rc = ws.get("vcap", cells * 2, &vc, &c->err)) || (rc = ws.get("voff", cells * 8, &vo, &c->err)))
        return rc;
    hipError_t e = hipMemcpyAsync(vc, cap.data(), cells * 2, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(vo, off.data(), cells * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && start) {
        if ((rc = ws.get("vstart", cells * 2, &vs, &c->err)) || (rc = ws.get("vcount", cells * 2, &vn, &c->err))) return rc;
        e = hipMemcpyAsync(vs, start, cells * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(vn, count, cells * 2, hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess && skip) {
        if ((rc = ws.get("skip", cells, &sk, &c->err))) return rc;
        e = hipMemcpyAsync(sk, skip, cells, hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess)
        e = lux::launch_sweep_cost(a, width, height, scale, (const uint16_t*)vs, (const uint16_t*)vn,
                                   start ? (const uint16_t*)vc : nullptr, n, ((n + align - 1) / align) * align,
                                   (double)total / (double)cells, (const uint8_t*)sk, (uint8_t*)dense, cs, s);
    if (e == hipSuccess && total)
        e = lux::launch_pack_u8((const uint8_t*)dense, (const uint64_t*)vo, (const uint16_t*)vc, (uint8_t*)packed, cs, cells, s);
    if (e == hipSuccess && total) e = hipMemcpyAsync(raw, packed, total, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : hip_fail(c, e, "sweep cost");
}

int lux_sgm_range_from_previous(lux_sgm_ctx* c, const float* planes, int32_t n, const float* prev_depth, const uint8_t* prev_skip,
                                int32_t pw, int32_t ph, int32_t width, int32_t height, int32_t margin, uint16_t* start,
                                uint16_t* count, int32_t* min_lo, int32_t* max_hi)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!planes || !prev_depth || !start || !count || n <= 0 || pw <= 0 || ph <= 0 || width <= 0 || height <= 0)
        return fail(c, LUX_SGM_EINVAL, "bad arguments");
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    lux::Workspace& ws = c->ws;
    const size_t np = (size_t)pw * ph, cells = (size_t)width * height;
    void *dp, *dk = nullptr, *ds, *dc, *dcap;
    int rc;
    if ((rc = ws.get("rg_prev", np * 4, &dp, &c->err)) || (rc = ws.get("vstart", cells * 2, &ds, &c->err)) ||
        (rc = ws.get("vcount", cells * 2, &dc, &c->err)) || (rc = ws.get("vcap", cells * 2, &dcap, &c->err)))
        return rc;
    hipError_t e = hipMemcpyAsync(dp, prev_depth, np * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && prev_skip) {
        if ((rc = ws.get("rg_pskip", np, &dk, &c->err))) return rc;
        e = hipMemcpyAsync(dk, prev_skip, np, hipMemcpyHostToDevice, s);
    }
    if (e != hipSuccess) return hip_fail(c, e, "upload previous layer");
    lux::RangeStats rs;
    if ((rc = lux::ranges_from_previous(ws, planes, n, (const float*)dp, (const uint8_t*)dk, pw, ph, width, height, margin, 1,
                                        nullptr, (uint16_t*)ds, (uint16_t*)dc, (uint16_t*)dcap, &rs, s, &c->err)))
        return rc;
    e = hipMemcpyAsync(start, ds, cells * 2, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(count, dc, cells * 2, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "download ranges");
    if (min_lo) *min_lo = rs.minlo;
    if (max_hi) *max_hi = rs.maxhi;
    return 0;
}

int lux_sgm_dense_layer(lux_sgm_ctx* c, const lux_sgm_sweep* sw, const lux_sgm_dense_params* p, const uint16_t* start,
                        const uint16_t* count, const uint8_t* skip, float* depth, uint16_t* agg, uint16_t* agg_pass1,
                        uint16_t* wta)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = lux::check_dense(sw, p, &c->err);
    if (rc) return rc;
    if ((start == nullptr) != (count == nullptr)) return fail(c, LUX_SGM_EINVAL, "start and count go together");
    if (!depth) return fail(c, LUX_SGM_EINVAL, "null depth");
    const int W = p->width, H = p->height, align = p->align, n = sw->n_planes;
    const size_t cells = (size_t)W * H;
    // the host side of the ranges: caps, offsets of the packed aggregates and the statistics the device checks
    lux::RangeStats rs;
    std::vector<uint16_t> cap(cells);
    std::vector<uint64_t> off(cells);
    uint64_t total = 0;
    rs.minlo = INT_MAX;
    for (size_t i = 0; i < cells; i++) {
        const int st = start ? start[i] : 0, cn = count ? count[i] : n;
        if (st + cn > n) return fail(c, LUX_SGM_EINVAL, "a range passes the last plane");
        const int cp = ((cn + align - 1) / align) * align;
        cap[i] = (uint16_t)cp;
        off[i] = total;
        total += (uint64_t)cp;
        rs.minlo = std::min(rs.minlo, st);
        rs.maxhi = std::max(rs.maxhi, st + cn);
        rs.maxcap = std::max(rs.maxcap, cp);
        rs.maxcount = std::max(rs.maxcount, cn);
        rs.maxend = std::max(rs.maxend, st + cp);
        rs.cap_computed += (uint64_t)cp;
        rs.computed += (skip && skip[i]) ? 0u : 1u;
    }
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    lux::Workspace& ws = c->ws;
    lux::SweepDev a;
    if ((rc = lux::upload_sweep(ws, sw, false, s, &a, &c->err))) return rc;
    void *vs = nullptr, *vn = nullptr, *vc = nullptr, *vo = nullptr, *sk = nullptr, *dw, *dd, *a1 = nullptr, *stage = nullptr;
    if ((rc = ws.get("layer_wta", cells * 2, &dw, &c->err)) || (rc = ws.get("dl_depth", cells * 4, &dd, &c->err))) return rc;
    hipError_t e = hipSuccess;
    if (start) {
        if ((rc = ws.get("vstart", cells * 2, &vs, &c->err)) || (rc = ws.get("vcount", cells * 2, &vn, &c->err)) ||
            (rc = ws.get("vcap", cells * 2, &vc, &c->err)))
            return rc;
        e = hipMemcpyAsync(vs, start, cells * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(vn, count, cells * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(vc, cap.data(), cells * 2, hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess && skip) {
        if ((rc = ws.get("skip", cells, &sk, &c->err))) return rc;
        e = hipMemcpyAsync(sk, skip, cells, hipMemcpyHostToDevice, s);
    }
    if (e != hipSuccess) return hip_fail(c, e, "upload layer");
    const lux::Geometry g = lux::geometry(W, H, start ? rs.maxcount : n, align);
    if (agg_pass1 && (rc = ws.get("layer_pass1", cells * g.cs * 2, &a1, &c->err))) return rc;
    lux::Prepared w;
    const lux::Conc conc = lux::conc_of(c);
    if ((rc = lux::dense_layer_device(ws, a, *p, (const uint16_t*)vs, (const uint16_t*)vn, (const uint16_t*)vc, rs,
                                      (const uint8_t*)sk, agg != nullptr, (uint16_t*)a1, (uint16_t*)dw, (float*)dd, &w, s,
                                      &c->err, &conc)))
        return rc;
    e = hipMemcpyAsync(depth, dd, cells * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && wta) e = hipMemcpyAsync(wta, dw, cells * 2, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(c, e, "download depth");
    uint16_t* outs[2] = {agg, agg_pass1};
    const uint16_t* srcs[2] = {w.agg, (const uint16_t*)a1};
    for (int k = 0; k < 2; k++) {
        if (!outs[k] || !total) continue;
        if (!stage && ((rc = ws.get("layer_stage", total * 2, &stage, &c->err)) || (rc = ws.get("voff", cells * 8, &vo, &c->err))))
            return rc;
        if (!vc && (rc = ws.get("vcap", cells * 2, &vc, &c->err))) return rc;
        e = hipMemcpyAsync(vo, off.data(), cells * 8, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(vc, cap.data(), cells * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = lux::launch_pack_u16(srcs[k], (const uint64_t*)vo, (const uint16_t*)vc, (uint16_t*)stage, g.cs, cells, s);
        if (e == hipSuccess) e = hipMemcpyAsync(outs[k], stage, total * 2, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(c, e, "download aggregate");
    }
    e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : hip_fail(c, e, "dense layer");
}

void lux_sgm_dense_pyramid_defaults(lux_sgm_dense_pyramid_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    static const int scales[6] = {32, 16, 8, 4, 2, 1}, tiles[6] = {4, 8, 16, 32, 64, 64};
    p->levels = 6;
    for (int i = 0; i < 6; i++) {
        p->scale[i] = scales[i];
        p->tile[i] = tiles[i];
        p->skip_mode[i] = i >= 4 ? 2 : 0;
    }
    p->p1 = 1;
    p->p2_scale = 500.0f;
    uint32_t cb = 0x3ef6384fu;
    float cg;
    std::memcpy(&cg, &cb, 4);
    p->guidance[0] = cg / 6.0f;
    p->guidance[1] = cg / 16.0f;
    p->guidance[2] = cg / 16.0f;
    p->guidance[3] = 0.0f;
    p->margin = 1;
    p->caps[0] = 2;
    p->caps[1] = 6;
    p->caps[2] = 6;
    p->caps[3] = 0;
    p->budget = 0x40000000ull;
}

}  // extern "C"

namespace {

// DenseStereoPyramid.Run; `device`: images and depths are device pointers, else host pointers.
int dense_pyramid(lux_sgm_ctx* c, const lux_sgm_dense_pyramid_params* pp, int32_t n_images, const uint8_t* const* images,
                  const int32_t* widths, const int32_t* heights, const uint8_t* gray, const float* warps, const float* planes8,
                  int32_t n8, const float* planes2, int32_t n2, float* const* depths, int32_t* aligns, bool device)
{
    if (!pp || pp->levels < 1 || pp->levels > 8) return fail(c, LUX_SGM_EINVAL, "1 to 8 levels");
    if (n_images < 2 || n_images > lux::kMaxViews + 1 || !images || !widths || !heights || !warps)
        return fail(c, LUX_SGM_EINVAL, "a reference and 1 to 16 other cameras");
    if (!planes8 || n8 <= 0 || n8 > 65535 || (planes2 && (n2 <= 0 || n2 > 65535)))
        return fail(c, LUX_SGM_EINVAL, "bad plane lists");
    for (int k = 0; k < n_images; k++)
        if (!images[k] || widths[k] < 4 || heights[k] < 4) return fail(c, LUX_SGM_EINVAL, "every image needs 4 x 4 pixels");
    for (int i = 0; i < pp->levels; i++)
        if (pp->scale[i] <= 0 || pp->tile[i] < 2 || (pp->skip_mode[i] != 0 && pp->skip_mode[i] != 2))
            return fail(c, LUX_SGM_EINVAL, "bad level parameters");
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    lux::Workspace& ws = c->ws;
    int rc;
    // the box-filtered images, shared by all layers (DenseStereoSgm.cs:336)
    static const char* names[lux::kMaxViews + 1] = {"py_f0", "py_f1", "py_f2", "py_f3", "py_f4", "py_f5",
                                                    "py_f6", "py_f7", "py_f8", "py_f9", "py_f10", "py_f11",
                                                    "py_f12", "py_f13", "py_f14", "py_f15", "py_f16"};
    std::vector<void*> filt(n_images);
    uint64_t imgs = 0;
    size_t maxbytes = 0;
    for (int k = 0; k < n_images; k++) maxbytes = std::max(maxbytes, (size_t)widths[k] * heights[k] * 4);
    void* src;  // one staging buffer at the largest size, so no reallocation happens under queued filters
    if ((rc = ws.get("py_src", maxbytes, &src, &c->err))) return rc;
    for (int k = 0; k < n_images; k++) {
        const size_t bytes = (size_t)widths[k] * heights[k] * 4;
        if ((rc = ws.get(names[k], bytes, &filt[k], &c->err))) return rc;
        hipError_t e = hipSuccess;
        const void* in = images[k];
        if (!device) {
            e = hipMemcpyAsync(src, images[k], bytes, hipMemcpyHostToDevice, s);
            in = src;
        }
        if (e == hipSuccess) e = lux::launch_box3(in, false, widths[k], heights[k], widths[k], 4, 255, filt[k], s);
        if (e != hipSuccess) return hip_fail(c, e, "box filter");
        imgs += (uint64_t)heights[k] * widths[k] * 4;
    }
    // the cost context of each alignment's planes (BuildContext, :190-200)
    std::vector<lux_sgm_view> views(n_images - 1);
    for (int k = 1; k < n_images; k++) {
        lux_sgm_view& v = views[k - 1];
        v.rgba = (const uint8_t*)filt[k];
        v.width = widths[k];
        v.height = heights[k];
        std::memcpy(v.warp, warps + (size_t)(k - 1) * 16, sizeof v.warp);
        const bool g = gray && gray[k];
        v.wq[0] = g ? 12240 : 8160;
        v.wq[1] = g ? 0 : 680;
        v.wq[2] = g ? 0 : 680;
        v.wq[3] = 0;
    }
    void *dp8, *dp2 = nullptr;
    if ((rc = ws.get("py_planes8", (size_t)n8 * 4, &dp8, &c->err))) return rc;
    hipError_t e = upload_once(ws, "py_planes8", dp8, planes8, n8, s);
    if (e == hipSuccess && planes2) {
        if ((rc = ws.get("py_planes2", (size_t)n2 * 4, &dp2, &c->err))) return rc;
        e = upload_once(ws, "py_planes2", dp2, planes2, n2, s);
    }
    if (e != hipSuccess) return hip_fail(c, e, "upload planes");
    lux_sgm_sweep sw{};
    sw.ref = (const uint8_t*)filt[0];
    sw.ref_width = widths[0];
    sw.ref_height = heights[0];
    sw.n_views = n_images - 1;
    sw.views = views.data();
    std::memcpy(sw.caps, pp->caps, 4);
    lux::SweepDev a8, a2;
    sw.planes = (const float*)dp8;
    sw.n_planes = n8;
    if ((rc = lux::upload_sweep(ws, &sw, true, s, &a8, &c->err))) return rc;
    if (planes2) {
        sw.planes = (const float*)dp2;
        sw.n_planes = n2;
        if ((rc = lux::upload_sweep(ws, &sw, true, s, &a2, &c->err))) return rc;
    }
    // per level: ranges (from the previous level), the memory mode, the layer; two buffer sets alternate so the
    // previous level's depth and skip mask survive into the next
    int pw = 0, ph = 0;
    const float* d_prev = nullptr;
    const uint8_t* d_pskip = nullptr;
    for (int i = 0; i < pp->levels; i++) {
        const char* tag = (i & 1) ? "b" : "a";
        auto nm = [&](const char* base) { return std::string("py_") + base + tag; };
        lux_sgm_dense_params p{};
        p.scale = pp->scale[i];
        p.tile = pp->tile[i];
        p.width = (widths[0] + p.scale - 1) / p.scale;
        p.height = (heights[0] + p.scale - 1) / p.scale;
        p.p1 = pp->p1;
        p.p2_scale = pp->p2_scale;
        std::memcpy(p.guidance, pp->guidance, sizeof p.guidance);
        const int W = p.width, H = p.height;
        const size_t cells = (size_t)W * H;
        if (W > 10240) return fail(c, LUX_SGM_EUNSUPPORTED, "width above 10240");
        // the skip mask of a mode and size never changes: made once per context and kept on the device
        const std::string skname = "py_skip" + std::to_string(pp->skip_mode[i]) + "_" + std::to_string(W) + "x" +
                                   std::to_string(H);
        void *dsk, *dd, *dw, *ds8, *dc8, *dcap8;
        if ((rc = ws.get(skname.c_str(), cells, &dsk, &c->err)) || (rc = ws.get(nm("depth").c_str(), cells * 4, &dd, &c->err)) ||
            (rc = ws.get("py_wta", cells * 2, &dw, &c->err)) || (rc = ws.get("py_start8", cells * 2, &ds8, &c->err)) ||
            (rc = ws.get("py_count8", cells * 2, &dc8, &c->err)) || (rc = ws.get("py_cap8", cells * 2, &dcap8, &c->err)))
            return rc;
        if (!ws.filled.count(skname)) {
            std::vector<uint8_t> skip(cells);
            if ((rc = lux_sgm_skip_mask(pp->skip_mode[i], W, H, skip.data()))) return fail(c, rc, "skip mask");
            e = hipMemcpy(dsk, skip.data(), cells, hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(c, e, "upload skip mask");
            ws.filled.insert(skname);
        }
        lux::RangeStats rs;
        const lux::SweepDev* a = &a8;
        const uint16_t *d_start = nullptr, *d_count = nullptr, *d_cap = nullptr;
        int align = 8;
        if (d_prev) {
            // mode 1 (align 8) if the estimate fits the budget, else mode 2 (align 2) (:340-343)
            auto estimate = [&](const lux::RangeStats& r, int al) {
                const uint64_t w = (uint64_t)W, h = (uint64_t)H;
                const uint64_t maxhi = (uint64_t)(((r.maxhi + al - 1) / al) * al);
                const uint64_t cost = 8 * w * h + 3 * r.cap_computed + 2 * r.cap_skipped;
                return cost + w * h * 4 + 20 * (2 * w + 4) + 2 * ((4 * maxhi + 5) * (2 * w + 4) + 8 * w + 16) + imgs + w * h * 8;
            };
            if ((rc = lux::ranges_from_previous(ws, planes8, n8, d_prev, d_pskip, pw, ph, W, H, pp->margin, 8,
                                                (const uint8_t*)dsk, (uint16_t*)ds8, (uint16_t*)dc8, (uint16_t*)dcap8, &rs, s,
                                                &c->err)))
                return rc;
            d_start = (const uint16_t*)ds8;
            d_count = (const uint16_t*)dc8;
            d_cap = (const uint16_t*)dcap8;
            if (estimate(rs, 8) >= pp->budget) {
                if (!planes2) return fail(c, LUX_SGM_EUNSUPPORTED, "the layer needs the align-2 memory mode and no align-2 planes were given");
                void *ds2, *dc2, *dcap2;
                if ((rc = ws.get("py_start2", cells * 2, &ds2, &c->err)) || (rc = ws.get("py_count2", cells * 2, &dc2, &c->err)) ||
                    (rc = ws.get("py_cap2", cells * 2, &dcap2, &c->err)))
                    return rc;
                if ((rc = lux::ranges_from_previous(ws, planes2, n2, d_prev, d_pskip, pw, ph, W, H, pp->margin, 2,
                                                    (const uint8_t*)dsk, (uint16_t*)ds2, (uint16_t*)dc2, (uint16_t*)dcap2, &rs,
                                                    s, &c->err)))
                    return rc;
                if (estimate(rs, 2) >= pp->budget)
                    return fail(c, LUX_SGM_EUNSUPPORTED, "StereoLayer memory budget exceeded (the upsample-only mode)");
                align = 2;
                a = &a2;
                d_start = (const uint16_t*)ds2;
                d_count = (const uint16_t*)dc2;
                d_cap = (const uint16_t*)dcap2;
            }
        }
        p.align = align;
        lux::Prepared w;
        const lux::Conc conc = lux::conc_of(c);
        rc = lux::dense_layer_device(ws, *a, p, d_start, d_count, d_cap, rs, (const uint8_t*)dsk, false, nullptr,
                                     (uint16_t*)dw, (float*)dd, &w, s, &c->err, &conc);
        if (rc) {
            c->err = "level " + std::to_string(i) + ": " + c->err;
            return rc;
        }
        if (aligns) aligns[i] = align;
        if (depths && depths[i]) {
            e = hipMemcpyAsync(depths[i], dd, cells * 4, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) return hip_fail(c, e, "download depth");
        }
        d_prev = (const float*)dd;
        d_pskip = (const uint8_t*)dsk;
        pw = W;
        ph = H;
    }
    // one wait at the end: each level's range statistics already wait for the level before
    e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : hip_fail(c, e, "dense pyramid");
}

}  // namespace

extern "C" {

int lux_sgm_dense_pyramid(lux_sgm_ctx* c, const lux_sgm_dense_pyramid_params* pp, int32_t n_images,
                          const uint8_t* const* images, const int32_t* widths, const int32_t* heights, const uint8_t* gray,
                          const float* warps, const float* planes8, int32_t n8, const float* planes2, int32_t n2,
                          float* const* depths, int32_t* aligns)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    return dense_pyramid(c, pp, n_images, images, widths, heights, gray, warps, planes8, n8, planes2, n2, depths, aligns,
                         false);
}

int lux_sgm_dense_pyramid_device(lux_sgm_ctx* c, const lux_sgm_dense_pyramid_params* pp, int32_t n_images,
                                 const uint8_t* const* d_images, const int32_t* widths, const int32_t* heights,
                                 const uint8_t* gray, const float* warps, const float* planes8, int32_t n8,
                                 const float* planes2, int32_t n2, float* const* d_depths, int32_t* aligns)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    return dense_pyramid(c, pp, n_images, d_images, widths, heights, gray, warps, planes8, n8, planes2, n2, d_depths, aligns,
                         true);
}

}  // extern "C"
