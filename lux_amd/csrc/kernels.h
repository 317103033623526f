// Internal interface between the C ABI (api.hip) and the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lux {

// Device layout of one uniform-range layer: cell-major volumes raw u8 / agg u16 [H][W][cs], where
// cs = roundup(cap, max(dpl, 4)) and dpl (planes per lane, a power of two >= 2) is chosen so that one
// 64-lane wavefront covers a cell: 64 * dpl >= cap.
struct Geometry {
    int W = 0, H = 0, D = 0, cap = 0, dpl = 0, cs = 0;
};
Geometry geometry(int W, int H, int D, int align);

// Per-pass buffers (neighbour-column states and minima, border replay partial sums and chain costs, border-column
// sums) keep pass -1's half this many u16 after pass +1's: whole 256-byte blocks, so no cache line holds both halves.
// (The persistent executor reads one pass's half while the other's is still to be written, and a line cached by
// such a read would be stale in the other half.)
inline size_t pass_half_u16(size_t n) { return (n + 127) & ~(size_t)127; }
inline size_t cols_half(const Geometry& g) { return pass_half_u16((size_t)2 * g.H * g.cs); }
inline size_t colmins_half(const Geometry& g) { return pass_half_u16((size_t)2 * g.H); }
inline size_t border_half(const Geometry& g) { return pass_half_u16((size_t)g.H * g.cs); }
inline size_t replay_half(const Geometry& g, size_t maxvis) { return pass_half_u16(maxvis * g.cs); }

// Image-adaptive P2 (DenseStereoSgm.cs:267-274): per path step, P2 = (int)(base * fastexp2(-S)) with
// S = (|d0| + |d2|) + |d1|, d_c = (g_c - q_c) * w_c, g the cell's guidance and q the guidance of the cell whose
// state the step reads (0 for the initial fill). guid: [H][W] RGBA8 as u32 (never null: a one-word dummy
// when off, so loads need no select).
struct AdaptP2 {
    const uint32_t* guid = nullptr;
    float w0 = 0.f, w1 = 0.f, w2 = 0.f;
    float base = 0.f;  // (float)P1 * P2Scale
    int on = 0;
};

struct SgmBuffers {
    const uint8_t* raw = nullptr;   // [H][W][cs]
    uint8_t* raw_out = nullptr;     // same buffer, written by the first sweep when the census cost is fused
    const uint64_t* cl = nullptr;   // census codes [H][W]; null: raw is given
    const uint64_t* cr = nullptr;
    uint16_t* dr = nullptr;         // [H][W] right-image disparity (scratch)
    uint32_t* dr0 = nullptr;        // [H]
    uint16_t* agg = nullptr;        // [H][W][cs]
    uint16_t* seeds = nullptr;      // [3][W][cs]  pass +1 half-0 leftovers of paths 1..3
    uint16_t* seedmins = nullptr;   // [3][W]
    uint16_t* cols = nullptr;       // [2 passes][2 paths][H][cs]  neighbour-column states for the replays
    uint16_t* colmins = nullptr;    // [2][2][H]
    uint16_t* partial = nullptr;    // [max(visits)][cs]  replay scratch: paths 0, 1 and 3 of each border visit
    uint16_t* chainp = nullptr;     // [max(visits)][cs]  path 2 of each border visit (the border chain's)
    uint16_t* bnd = nullptr;        // [2][3][W][cs]  band boundary states (ping-pong)
    uint32_t* bndmin = nullptr;     // [2][3][W]
    int* ticket = nullptr;          // [2] tile tickets of the two band launches
    int* flags = nullptr;           // [2][max_tiles] tile completion epochs
    int max_tiles = 0;
    int epoch = 0;                  // > 0, distinct per request
    int* err = nullptr;             // band wait gave up
    const int* visits_plus = nullptr;
    int n_plus = 0;
    const int* visits_minus = nullptr;
    int n_minus = 0;
    const int* cellvis_plus = nullptr;   // [H][2] visit indices of each border-column row (-1: none)
    const int* cellvis_minus = nullptr;
    int seed_row = -1;
    uint16_t* disp_tmp = nullptr;   // [H][W] disparity when the caller wants none
    // per-pixel plane ranges (null: every pixel has planes [0, cap)): start, cap and count of every cell, and
    // the start / cap of the cell behind each seed-table entry. Needs cap <= 512 (dpl 2, 4 or 8); no fused bands.
    const uint16_t* vstart = nullptr;
    const uint16_t* vcap = nullptr;
    const uint16_t* vcount = nullptr;
    uint16_t* seedstart = nullptr;  // [3][W]
    uint16_t* seedcap = nullptr;
    AdaptP2 ap;                     // adaptive P2 (ap.on = 0: every step uses the constant P2)
    uint32_t* seedguid = nullptr;   // [3][W] guidance of the cell behind each seed-table entry (adaptive P2)
    int seg = 1;                    // narrow per-pixel ranges (caps <= 32) may run several chains per wavefront
    // Path volumes [H][W][cs] for latency-bound (ranged) layers (set all or none): each pass then runs as one
    // launch, pass +1's paths 1-3 into the volumes, pass -1's added to them, all summed with agg once.
    uint16_t* aggp[3] = {};
    // Byte path volumes, cell c of volume k at pv[k] + c * pvs (set all seven or none; uniform constant-P2 census
    // requests whose every path cost fits a byte, PathBytes in engine.hip): each path's costs go to a volume of its
    // own, and pass -1's row path sums
    // them with its own into the final aggregate, which it then writes only when asked (agg may be null). bagg holds
    // the two border columns' replayed costs, [2][H][cs] (x = W-1 of pass +1, x = 0 of pass -1).
    uint8_t* pv[7] = {};
    int pvs = 0;                    // byte stride between consecutive cells of a path volume (7 * cs: interleaved)
    uint16_t* bagg[2] = {};
    int zero_flags = 0;             // clear the band flags by a kernel before each band launch (graph replays)
    // The request has the device to itself: each chain's wavefront is a workgroup of its own, so the chains spread
    // over the CUs to within one wavefront (four-wavefront workgroups left some CUs a whole workgroup more). With
    // other requests running beside it (0), four-wavefront workgroups.
    int solo = 1;
    // The border chains on a side stream (set all or none; none: they run in order on the request's stream). They
    // read only the cost volume, the ranges, the guidance and pass +1's seeds, so they run beside the sweeps.
    hipStream_t chain = nullptr;
    hipEvent_t chain_fork = nullptr, chain_join = nullptr;
};

// Census codes of one image (T = uint8_t or uint16_t pixels).
hipError_t launch_census(const void* img, bool img16, int W, int H, int stride, int radius, uint64_t* codes,
                         hipStream_t s);
// raw[y][x][d] for d < cs (truncated AD; 255 for d >= D). The census cost is computed inside the first sweep.
hipError_t launch_cost_ad(const void* L, const void* R, bool img16, int bits, int stride, int adcap, const Geometry& g,
                          uint8_t* raw, hipStream_t s);
struct SgmOutputs {
    uint16_t* disp = nullptr;       // [H][W] required (WTA plane index)
    float* sub = nullptr;           // may be null
    uint8_t* valid = nullptr;       // may be null
    int tol = 1;
    int write_agg = 0;              // leave the final aggregate in the agg volume
    int bands = 0;                  // paths 1-3 of a pass in fused band tiles (planes <= 256) instead of sweeps
};
// The eight path sweeps (the census cost fused into the first when b.cl is set, the winner-take-all into
// the last), the two border replays and the column-0 / validity finish. agg_pass1 (if not null) receives
// agg after pass +1. marks (if not null, kSgmMarks events) are recorded after each stage, in launch order.
constexpr int kSgmMarks = 11;
hipError_t launch_sgm(const Geometry& g, const SgmBuffers& b, const SgmOutputs& o, int p1, int p2, uint16_t* agg_pass1,
                      hipStream_t s, hipEvent_t* marks = nullptr);
// Line-buffer-literal aggregation (literal.hip): DenseLayer.Compute's two passes over the reference's own line
// buffers, word for word, for the layers the path sweeps do not model (padding that spills past its segment,
// odd tiles, caps up to 1024), then the first-minimum WTA.
struct LiteralLayer {
    int W = 0, H = 0, T = 0, align = 0, xsw = 0;
    int minlo = 0, maxhi = 0;          // MinLo, MaxHi rounded to align
    const uint16_t* start = nullptr;   // null: every cell [0, ucap) with ucount planes
    const uint16_t* cap = nullptr;
    const uint16_t* count = nullptr;
    int ucap = 0, ucount = 0;
    const uint8_t* raw = nullptr;      // [H][W][cs], skipped cells' bytes already 0 / 255
    uint16_t* agg = nullptr;           // [H][W][cs]
    int cs = 0;
    int npl = 2;                       // planes per lane: 2 (caps <= 128), 4 (<= 256), 0 (chunks, <= 1024)
    AdaptP2 ap;
    int p1 = 0, p2 = 0;
};
size_t literal_scratch_bytes(int W, int H, int maxhi);
// Rows whose wavefronts must be resident at once for every wait to end, and how many the device holds.
int literal_depth(int W, int T);
int literal_resident(int npl, bool adp);
hipError_t launch_literal(const LiteralLayer& l, void* scratch, uint16_t* agg_pass1, uint16_t* wta, hipStream_t s);
// Synchronises s and reads the error word (non-zero: a wait gave up).
hipError_t literal_error(void* scratch, int W, int H, int maxhi, hipStream_t s, uint32_t* err);

// Fill n u32 with v (a kernel, not a memset: safe inside captured graphs).
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
// Band tiles of a W x H layer (the flag array size per pass).
int band_tiles(int W, int H);
// Repack [H][W][cs] <-> [H][W][cap] on the device.
hipError_t launch_restride_u16(const uint16_t* src, int src_cs, uint16_t* dst, int dst_cs, int cells, int n, hipStream_t s);
// Coarse-to-fine levels: 2x2 mean downsampling, per-pixel ranges from the coarser disparity, range costs.
hipError_t launch_downsample2(const void* src, bool img16, int W, int H, int stride, void* dst, int W2, int H2, hipStream_t s);
hipError_t launch_ranges_from_coarse(const uint16_t* dc, int Wc, int Hc, int W, int H, int r, int D, uint16_t* vstart,
                                     uint16_t* vcap, uint16_t* vcount, hipStream_t s);
hipError_t launch_cost_census_ranges(const uint64_t* cl, const uint64_t* cr, int W, int H, int cs, const uint16_t* vstart,
                                     const uint16_t* vcount, uint8_t* raw, hipStream_t s);
// 3x3 box filter (1 or 4 interleaved channels; stride in pixels; output dense W*ch per row).
hipError_t launch_box3(const void* src, bool img16, int W, int H, int stride, int ch, int maxval, void* dst, hipStream_t s);
// Skipped pixels: raw 0 over their count (vcount[c], or `count` when null), 255 past it.
hipError_t launch_apply_skip(uint8_t* raw, const uint8_t* skip, const uint16_t* vcount, int count, int cs, size_t cells,
                             hipStream_t s);
// Jagged per-pixel arrays (cap[i] entries at off[i]) <-> the cell-major volume of stride cs (255 past a cap).
hipError_t launch_unpack_u8(const uint8_t* src, const uint64_t* off, const uint16_t* cap, uint8_t* dst, int cs, size_t cells,
                            hipStream_t s);
hipError_t launch_pack_u8(const uint8_t* src, const uint64_t* off, const uint16_t* cap, uint8_t* dst, int cs, size_t cells,
                          hipStream_t s);
hipError_t launch_pack_u16(const uint16_t* src, const uint64_t* off, const uint16_t* cap, uint16_t* dst, int cs, size_t cells,
                           hipStream_t s);
hipError_t launch_restride_u8(const uint8_t* src, int src_cs, uint8_t* dst, int dst_cs, int cells, int n, uint8_t fill,
                              hipStream_t s);

}  // namespace lux
