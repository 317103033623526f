// Device side of the persistent executor (exec.hip: batches; dloop.hip: the device-resident request loop): the
// stages of one rectified request with byte path volumes, the work-item functions that run them, and the flush
// epochs that carry one stage's results to the next across XCDs within a launch. Included by the two translation
// units (everything here is internal to each).
#pragma once
#include "exec.h"
#include "kernels_impl.h"

namespace lux {

namespace {


// The stages of one request, numbered so that every prerequisite has a smaller number.
enum XStage : int {
    X_IN = 0,      // host images into the request's device buffers (the device-resident loop, host memory)
    X_CENSUS,      // census codes: blocks of kXCensusRows rows of one image (left blocks, then right blocks)
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
    X_OUT,         // the outputs into host memory (the device-resident loop, host memory)
    X_NSTAGES
};
constexpr size_t kXChunk = 64 << 10;  // bytes one staging item copies
constexpr int kXCensusRows = 8;
constexpr uint32_t xbit(int s) { return 1u << s; }

// The stages whose items must all be complete before an item of stage s starts (same request): what the plain
// launch sequence orders by its stream. A stage completes only after its own prerequisites did, so direct ones
// suffice.
__host__ __device__ constexpr uint32_t exec_prereq(int s)
{
    switch (s) {
        case X_CENSUS: return xbit(X_IN);
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
        case X_OUT: return xbit(X_FINISH);
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
    // host staging (device-resident loop with host memory; else null): the images [left | right] of imgb bytes
    // each into din, and the outputs disp (2 bytes a pixel), sub (4), valid (1) from dout into [disp | sub | valid]
    const uint8_t* hin;
    uint8_t* hout;
    uint8_t* din[2];
    const uint8_t* dout[3];  // null: not asked for
    uint64_t imgb, px;
};

// Flush epochs (see the file comment): req counts the epochs raised, ack[x] is the last epoch XCD x has written its
// L2 back for, live[x] its running wavefronts (an XCD with none holds nothing unwritten: each leaves after a
// write-back).
constexpr int kXcds = 8;
struct XSync {
    uint32_t req;
    uint32_t ack[kXcds];
    int32_t live[kXcds];
    uint32_t claim[kXcds];  // the last epoch a wavefront of XCD x took on to write back (one writes back per epoch)
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

// This XCD's share of the protocol: write the L2 back if an epoch was raised that it has not acknowledged. One
// wavefront of the XCD takes each epoch on (the others would only repeat the write-back). True if this one did.
__device__ __forceinline__ bool xsync_poll(XSync* y, int x, int lane)
{
    const uint32_t r = ald(&y->req);
    if (ald(&y->ack[x]) >= r || ald(&y->claim[x]) >= r) return false;
    uint32_t prev = r;
    if (lane == 0) prev = __hip_atomic_fetch_max(&y->claim[x], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)__builtin_amdgcn_readfirstlane((int)prev) >= r) return false;  // another one took it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // buffer_wbl2: every dirty line of this XCD's L2 to memory
    if (lane == 0) __hip_atomic_fetch_max(&y->ack[x], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Whether epoch e has reached memory from every XCD.
__device__ __forceinline__ bool xsync_flushed(const XSync* y, uint32_t e)
{
    for (int x = 0; x < kXcds; x++)
        if (ald(&y->ack[x]) < e && (int)ald((const uint32_t*)&y->live[x]) > 0) return false;
    return true;
}

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

// A chunk of n bytes from src to dst, 16 bytes a lane (four loads in flight: one side may be host memory, where
// each load waits a bus round trip) when both are 16-byte aligned, the rest byte by byte.
__device__ __forceinline__ void copy_chunk(uint8_t* dst, const uint8_t* src, size_t n, int lane)
{
    const size_t n16 = (((uintptr_t)dst | (uintptr_t)src) & 15) ? 0 : n / 16;
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    size_t k = (size_t)lane;
    for (; k + 192 < n16; k += 256) {
        const uint4 a = s4[k], b = s4[k + 64], c = s4[k + 128], d = s4[k + 192];
        d4[k] = a;
        d4[k + 64] = b;
        d4[k + 128] = c;
        d4[k + 192] = d;
    }
    for (; k < n16; k += 64) d4[k] = s4[k];
    for (size_t j = n16 * 16 + (size_t)lane; j < n; j += 64) dst[j] = src[j];
}

// Staging item i: chunks of each image in turn (X_IN), of each output in turn (X_OUT).
__device__ __forceinline__ void x_in(const ExecJob* J, int i, int lane)
{
    const uint64_t imgb = uload(&J->imgb);
    const uint64_t per = (imgb + kXChunk - 1) / kXChunk;
    const int img = (uint64_t)i >= per ? 1 : 0;
    const uint64_t off = ((uint64_t)i - (uint64_t)img * per) * kXChunk;
    const uint8_t* hin = uload(&J->hin);
    uint8_t* d = uload(&J->din[img]);
    copy_chunk(d + off, hin + (uint64_t)img * imgb + off, (size_t)min<uint64_t>(kXChunk, imgb - off), lane);
}
__device__ __forceinline__ void x_out(const ExecJob* J, int i, int lane)
{
    const uint64_t px = uload(&J->px);
    const uint64_t bytes[3] = {2 * px, 4 * px, px}, base[3] = {0, 2 * px, 6 * px};
    uint64_t k = (uint64_t)i;
    for (int b = 0; b < 3; b++) {
        const uint64_t per = (bytes[b] + kXChunk - 1) / kXChunk;
        if (k < per) {
            const uint8_t* src = uload(&J->dout[b]);
            const uint64_t off = k * kXChunk;
            if (src) copy_chunk(uload(&J->hout) + base[b] + off, src + off, (size_t)min<uint64_t>(kXChunk, bytes[b] - off), lane);
            break;
        }
        k -= per;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");  // the host sees these bytes before it sees the request done
}

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


// The job descriptor of one prepared request (device pointers of its workspace, images and outputs; no staging).
inline ExecJob make_job(const ExecPair& q)
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
    const size_t colsz = cols_half(g), colminsz = colmins_half(g);
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
    r.partial = b.partial + replay_half(g, maxvis);
    r.chainp = b.chainp + replay_half(g, maxvis);
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

// Item `item` of stage `stage` of job J.
template <uint32_t SM>
__device__ __forceinline__ void run_item(const ExecJob* J, int stage, int item, int lane)
{
    if ((SM & xbit(X_IN)) && stage == X_IN) {
        x_in(J, item, lane);
    } else if ((SM & xbit(X_OUT)) && stage == X_OUT) {
        x_out(J, item, lane);
    } else if ((SM & xbit(X_CENSUS)) && stage == X_CENSUS) {
        census_item(J, item, lane);
    } else {
        switch (__builtin_amdgcn_readfirstlane(J->variant)) {
            case 0: chain_item<SM, 1, false>(J, stage, item, lane); break;
            case 1: chain_item<SM, 2, false>(J, stage, item, lane); break;
            case 2: chain_item<SM, 1, true>(J, stage, item, lane); break;
            default: chain_item<SM, 2, true>(J, stage, item, lane); break;
        }
    }
}

}  // namespace

}  // namespace lux
