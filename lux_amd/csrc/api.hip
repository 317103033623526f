// C ABI of liblux_sgm.so (declared in include/lux_sgm.h): contexts, pooled device buffers, the schedule-plan
// cache and the entry points that replace DenseLayer.Compute (DenseStereoSgm.cs:181-232).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>

#include "../../include/lux_sgm.h"
#include "kernels.h"
#include "plan.h"

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
};

thread_local std::string g_create_error;

}  // namespace

struct lux_sgm_ctx {
    int device = 0;
    uint64_t budget = 0;
    std::string err;
    std::mutex mu;
    hipStream_t stream = nullptr;
    std::map<std::string, DevBuf> bufs;
    std::map<std::tuple<int, int, int, int>, std::shared_ptr<lux::Plan>> plans;
    int opt_bands = 0;
    // band-tile synchronisation: flags hold the epoch of the request that wrote them
    int epoch = 0;
    void* sync_zeroed = nullptr;
    // stage timing (lux_sgm_set_profiling): events recorded between the launches of one request
    bool profiling = false;
    hipEvent_t ev[16] = {};
    int nev = 0;

    uint64_t held() const
    {
        uint64_t s = 0;
        for (auto& kv : bufs) s += kv.second.n;
        return s;
    }
};

namespace {

int fail(lux_sgm_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

int hip_fail(lux_sgm_ctx* c, hipError_t e, const char* where)
{
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    return fail(c, e == hipErrorOutOfMemory ? LUX_SGM_ENOMEM : LUX_SGM_EDEVICE, m);
}

// Grow-only pooled device buffer (the ScratchPool pattern, Lux.Engine/Pipeline/ScratchPool.cs:6-15).
int get_buf(lux_sgm_ctx* c, const char* name, size_t bytes, void** out)
{
    DevBuf& b = c->bufs[name];
    if (b.n < bytes) {
        uint64_t other = c->held() - b.n;
        if (c->budget && other + bytes > c->budget)
            return fail(c, LUX_SGM_ENOMEM, std::string("HBM budget exceeded allocating ") + name);
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) {
            b.p = nullptr;
            return hip_fail(c, e, "hipMalloc");
        }
        b.n = bytes;
    }
    *out = b.p;
    return 0;
}

int get_plan(lux_sgm_ctx* c, int W, int H, int T, int xsw, std::shared_ptr<lux::Plan>* out)
{
    auto key = std::make_tuple(W, H, T, xsw);
    auto it = c->plans.find(key);
    if (it == c->plans.end()) {
        auto p = std::make_shared<lux::Plan>(lux::make_plan(W, H, T, xsw));
        if (p->status != 0) return fail(c, p->status == -1 ? LUX_SGM_EINVAL : LUX_SGM_EUNSUPPORTED, p->error);
        it = c->plans.emplace(key, p).first;
    }
    *out = it->second;
    return 0;
}

struct Work {
    lux::Geometry g;
    lux::SgmBuffers b;
    uint8_t* raw = nullptr;
    uint16_t* agg = nullptr;
};

// Device buffers for one uniform-range layer, plus the plan's visit lists uploaded.
int prepare(lux_sgm_ctx* c, const lux::Geometry& g, int T, int xsw, hipStream_t s, Work* w)
{
    std::shared_ptr<lux::Plan> plan;
    int rc = get_plan(c, g.W, g.H, T, xsw, &plan);
    if (rc) return rc;
    const size_t cells = (size_t)g.W * g.H;
    void *raw, *agg, *seeds, *seedmins, *cols, *colmins, *vis, *partial, *drb, *dr0, *bnd, *bndmin;
    if ((rc = get_buf(c, "raw", cells * g.cs, &raw))) return rc;
    if ((rc = get_buf(c, "agg", cells * g.cs * 2, &agg))) return rc;
    if ((rc = get_buf(c, "seeds", (size_t)3 * g.W * g.cs * 2, &seeds))) return rc;
    if ((rc = get_buf(c, "seedmins", (size_t)3 * g.W * 2, &seedmins))) return rc;
    if ((rc = get_buf(c, "cols", (size_t)4 * g.H * g.cs * 2, &cols))) return rc;
    if ((rc = get_buf(c, "colmins", (size_t)4 * g.H * 2, &colmins))) return rc;
    const size_t nv = plan->plus.size() + plan->minus.size() + 4 * (size_t)g.H;
    const size_t maxvis = std::max(plan->plus.size(), plan->minus.size()) / 4 + 1;
    if ((rc = get_buf(c, "replay_partial", maxvis * g.cs * 2, &partial))) return rc;
    if ((rc = get_buf(c, "wta_dr", cells * 2, &drb))) return rc;
    if ((rc = get_buf(c, "band_boundary", (size_t)6 * g.W * g.cs * 2, &bnd))) return rc;
    const int tiles = lux::band_tiles(g.W, g.H);
    void* sync;
    if ((rc = get_buf(c, "band_sync", (size_t)(4 + 2 * tiles) * sizeof(int), &sync))) return rc;
    if (sync != c->sync_zeroed) {  // a new buffer: no flag may hold a live epoch
        hipError_t ez = lux::launch_fill_u32((uint32_t*)sync, 4 + 2 * (size_t)tiles, 0u, s);
        if (ez != hipSuccess) return hip_fail(c, ez, "zero band flags");
        c->sync_zeroed = sync;
    }
    if (++c->epoch <= 0) c->epoch = 1;
    if ((rc = get_buf(c, "band_boundary_min", (size_t)6 * g.W * 4, &bndmin))) return rc;
    if ((rc = get_buf(c, "wta_dr0", (size_t)g.H * 4, &dr0))) return rc;
    if ((rc = get_buf(c, "visits", (nv + 4) * sizeof(int), &vis))) return rc;
    hipError_t e = hipMemcpyAsync(vis, plan->plus.data(), plan->plus.size() * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync((int*)vis + plan->plus.size(), plan->minus.data(), plan->minus.size() * sizeof(int),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync((int*)vis + plan->plus.size() + plan->minus.size(), plan->cellvis_plus.data(),
                           plan->cellvis_plus.size() * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync((int*)vis + plan->plus.size() + plan->minus.size() + 2 * (size_t)g.H,
                           plan->cellvis_minus.data(), plan->cellvis_minus.size() * sizeof(int), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(c, e, "upload plan");
    w->g = g;
    w->raw = (uint8_t*)raw;
    w->agg = (uint16_t*)agg;
    w->b.raw = w->raw;
    w->b.raw_out = w->raw;
    w->b.agg = w->agg;
    w->b.dr = (uint16_t*)drb;
    w->b.dr0 = (uint32_t*)dr0;
    w->b.bnd = (uint16_t*)bnd;
    w->b.bndmin = (uint32_t*)bndmin;
    w->b.err = (int*)sync;
    w->b.ticket = (int*)sync + 2;
    w->b.flags = (int*)sync + 4;
    w->b.max_tiles = tiles;
    w->b.epoch = c->epoch;
    w->b.seeds = (uint16_t*)seeds;
    w->b.seedmins = (uint16_t*)seedmins;
    w->b.cols = (uint16_t*)cols;
    w->b.colmins = (uint16_t*)colmins;
    w->b.partial = (uint16_t*)partial;
    w->b.visits_plus = (const int*)vis;
    w->b.n_plus = (int)plan->plus.size() / 4;
    w->b.visits_minus = (const int*)vis + plan->plus.size();
    w->b.n_minus = (int)plan->minus.size() / 4;
    w->b.cellvis_plus = (const int*)vis + plan->plus.size() + plan->minus.size();
    w->b.cellvis_minus = w->b.cellvis_plus + 2 * (size_t)g.H;
    w->b.seed_row = plan->seed_row;
    return 0;
}

// After the stream has drained: a band tile that gave up waiting leaves the error word set.
int check_band_error(lux_sgm_ctx* c, hipStream_t s)
{
    auto it = c->bufs.find("band_sync");
    if (it == c->bufs.end() || !it->second.p) return 0;
    int err = 0;
    hipError_t e = hipMemcpyAsync(&err, it->second.p, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "read band error word");
    if (err) {
        int zero = 0;
        (void)hipMemcpyAsync(it->second.p, &zero, sizeof(int), hipMemcpyHostToDevice, s);
        (void)hipStreamSynchronize(s);
        return fail(c, LUX_SGM_EDEVICE, "a band tile timed out waiting for its neighbours");
    }
    return 0;
}

int check_rect(lux_sgm_ctx* c, const lux_sgm_rect_params* p)
{
    if (!p) return fail(c, LUX_SGM_EINVAL, "null params");
    if (p->width <= 0 || p->height <= 0 || p->stride < p->width) return fail(c, LUX_SGM_EINVAL, "bad image size/stride");
    if (p->width > 10240) return fail(c, LUX_SGM_EUNSUPPORTED, "width above 10240");
    if (!(p->bits == 8 || (p->bits >= 9 && p->bits <= 16))) return fail(c, LUX_SGM_EINVAL, "bits must be 8 or 9..16");
    if (p->cost != LUX_SGM_COST_CENSUS && p->cost != LUX_SGM_COST_AD) return fail(c, LUX_SGM_EINVAL, "unknown cost");
    if (p->cost == LUX_SGM_COST_CENSUS && (p->census_radius < 1 || p->census_radius > 3))
        return fail(c, LUX_SGM_EINVAL, "census radius must be 1..3");
    if (p->cost == LUX_SGM_COST_AD && (p->ad_cap < 0 || p->ad_cap > 28)) return fail(c, LUX_SGM_EINVAL, "ad_cap must be 0..28");
    if (p->disparities <= 0 || p->align <= 0) return fail(c, LUX_SGM_EINVAL, "disparities/align must be positive");
    if (((p->disparities + p->align - 1) / p->align) * p->align > 1024)
        return fail(c, LUX_SGM_EUNSUPPORTED, "more than 1024 lanes per cell");
    if (p->p1 < 0 || p->p1 > 65535 || p->p2 < 0 || p->p2 > 65535) return fail(c, LUX_SGM_EINVAL, "penalties must be 0..65535");
    if (p->tile < 2) return fail(c, LUX_SGM_EINVAL, "tile must be >= 2");
    if (p->lr_tolerance < 0) return fail(c, LUX_SGM_EINVAL, "negative L/R tolerance");
    return 0;
}

int rect_device(lux_sgm_ctx* c, const lux_sgm_rect_params* p, const void* dl, const void* dr, uint16_t* disp, float* sub,
                uint8_t* valid, uint16_t* agg_out, hipStream_t s)
{
    lux::Geometry g = lux::geometry(p->width, p->height, p->disparities, p->align);
    Work w;
    int rc = prepare(c, g, p->tile, p->xs_clamp_w, s, &w);
    if (rc) return rc;
    const bool img16 = p->bits > 8;
    hipError_t e;
    if (c->profiling) (void)hipEventRecord(c->ev[0], s);
    if (p->cost == LUX_SGM_COST_CENSUS) {  // the codes here; the cost volume inside the first sweep
        void *cl, *cr;
        const size_t n = (size_t)g.W * g.H * 8;
        if ((rc = get_buf(c, "census_l", n, &cl)) || (rc = get_buf(c, "census_r", n, &cr))) return rc;
        e = lux::launch_census(dl, img16, g.W, g.H, p->stride, p->census_radius, (uint64_t*)cl, s);
        if (e == hipSuccess) e = lux::launch_census(dr, img16, g.W, g.H, p->stride, p->census_radius, (uint64_t*)cr, s);
        w.b.cl = (const uint64_t*)cl;
        w.b.cr = (const uint64_t*)cr;
    } else {
        e = lux::launch_cost_ad(dl, dr, img16, p->bits, p->stride, p->ad_cap, g, w.raw, s);
    }
    if (e != hipSuccess) return hip_fail(c, e, "cost volume");
    if (c->profiling) (void)hipEventRecord(c->ev[1], s);
    lux::SgmOutputs o;
    void* td = disp;
    if (!disp && (rc = get_buf(c, "disp_tmp", (size_t)g.W * g.H * 2, &td))) return rc;
    o.disp = (uint16_t*)td;
    o.sub = sub;
    o.valid = valid;
    o.tol = p->lr_tolerance;
    o.write_agg = agg_out != nullptr;
    o.bands = c->opt_bands;
    e = lux::launch_sgm(g, w.b, o, p->p1, p->p2, nullptr, s, c->profiling ? c->ev + 2 : nullptr);
    if (e != hipSuccess) return hip_fail(c, e, "aggregation");
    if (c->profiling) c->nev = 2 + lux::kSgmMarks;
    if (agg_out) {
        e = lux::launch_restride_u16(w.agg, g.cs, agg_out, g.cap, g.W * g.H, g.cap, s);
        if (e != hipSuccess) return hip_fail(c, e, "agg copy");
    }
    return 0;
}

}  // namespace

extern "C" {

int lux_sgm_version(void) { return 100; }

int lux_sgm_create(int device, uint64_t hbm_budget, lux_sgm_ctx** out)
{
    if (!out) {
        g_create_error = "null out pointer";
        return LUX_SGM_EINVAL;
    }
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        g_create_error = std::string("no HIP device: ") + hipGetErrorString(e);
        return LUX_SGM_EDEVICE;
    }
    if (device < 0 || device >= n) {
        g_create_error = "device index out of range";
        return LUX_SGM_EINVAL;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        g_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return LUX_SGM_EDEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_create_error = std::string("built for gfx950, device is ") + prop.gcnArchName;
        return LUX_SGM_EUNSUPPORTED;
    }
    auto* c = new lux_sgm_ctx();
    c->device = device;
    c->budget = hbm_budget;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_create_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        delete c;
        return LUX_SGM_EDEVICE;
    }
    *out = c;
    return LUX_SGM_OK;
}

int lux_sgm_destroy(lux_sgm_ctx* c)
{
    if (!c) return LUX_SGM_EINVAL;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LUX_SGM_OK;
}

const char* lux_sgm_last_error(const lux_sgm_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int lux_sgm_rect_bytes(const lux_sgm_rect_params* p, uint64_t* bytes)
{
    if (!p || !bytes || p->width <= 0 || p->height <= 0 || p->disparities <= 0 || p->align <= 0) return LUX_SGM_EINVAL;
    lux::Geometry g = lux::geometry(p->width, p->height, p->disparities, p->align);
    const uint64_t cells = (uint64_t)g.W * g.H;
    uint64_t b = cells * g.cs * 3                      // raw u8 + agg u16
                 + cells * 16                          // census codes
                 + (uint64_t)3 * g.W * g.cs * 2 + (uint64_t)4 * g.H * g.cs * 2  // seeds, replay columns
                 + cells * 2 * (p->bits > 8 ? 2 : 1)   // images
                 + cells * 7;                          // outputs
    *bytes = b;
    return LUX_SGM_OK;
}

int lux_sgm_rectified_device(lux_sgm_ctx* c, const lux_sgm_rect_params* p, const void* d_left, const void* d_right,
                             uint16_t* d_disp, float* d_sub, uint8_t* d_valid, uint16_t* d_agg, void* stream)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_rect(c, p);
    if (rc) return rc;
    if (!d_left || !d_right) return fail(c, LUX_SGM_EINVAL, "null image");
    (void)hipSetDevice(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    return rect_device(c, p, d_left, d_right, d_disp, d_sub, d_valid, d_agg, s);
}

int lux_sgm_rectified(lux_sgm_ctx* c, const lux_sgm_rect_params* p, const void* left, const void* right, uint16_t* disp,
                      float* sub, uint8_t* valid, uint16_t* agg)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_rect(c, p);
    if (rc) return rc;
    if (!left || !right) return fail(c, LUX_SGM_EINVAL, "null image");
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    const size_t px = (size_t)p->width * p->height;
    const size_t esz = p->bits > 8 ? 2 : 1;
    const size_t imgb = (size_t)p->height * p->stride * esz;
    lux::Geometry g = lux::geometry(p->width, p->height, p->disparities, p->align);
    void *dl, *dr, *dd, *ds, *dv, *da = nullptr;
    if ((rc = get_buf(c, "img_l", imgb, &dl)) || (rc = get_buf(c, "img_r", imgb, &dr))) return rc;
    if ((rc = get_buf(c, "out_disp", px * 2, &dd)) || (rc = get_buf(c, "out_sub", px * 4, &ds)) ||
        (rc = get_buf(c, "out_valid", px, &dv)))
        return rc;
    if (agg && (rc = get_buf(c, "out_agg", px * g.cap * 2, &da))) return rc;
    hipError_t e = hipMemcpyAsync(dl, left, imgb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, right, imgb, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(c, e, "upload images");
    rc = rect_device(c, p, dl, dr, disp ? (uint16_t*)dd : nullptr, sub ? (float*)ds : nullptr,
                     valid ? (uint8_t*)dv : nullptr, agg ? (uint16_t*)da : nullptr, s);
    if (rc) return rc;
    if (disp) e = hipMemcpyAsync(disp, dd, px * 2, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && sub) e = hipMemcpyAsync(sub, ds, px * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && valid) e = hipMemcpyAsync(valid, dv, px, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && agg) e = hipMemcpyAsync(agg, da, px * g.cap * 2, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "download results");
    return check_band_error(c, s);
}

int lux_sgm_compute_layer(lux_sgm_ctx* c, const lux_sgm_layer_params* p, const uint8_t* raw, uint16_t* agg,
                          uint16_t* agg_pass1, uint16_t* wta)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!p || !raw) return fail(c, LUX_SGM_EINVAL, "null argument");
    if (p->width <= 0 || p->height <= 0 || p->planes <= 0 || p->align <= 0 || p->tile < 2)
        return fail(c, LUX_SGM_EINVAL, "bad layer parameters");
    if (p->width > 10240) return fail(c, LUX_SGM_EUNSUPPORTED, "width above 10240");
    if (((p->planes + p->align - 1) / p->align) * p->align > 1024)
        return fail(c, LUX_SGM_EUNSUPPORTED, "more than 1024 lanes per cell");
    if (p->p1 < 0 || p->p1 > 65535 || p->p2 < 0 || p->p2 > 65535) return fail(c, LUX_SGM_EINVAL, "penalties must be 0..65535");
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    lux::Geometry g = lux::geometry(p->width, p->height, p->planes, p->align);
    Work w;
    int rc = prepare(c, g, p->tile, p->xs_clamp_w, s, &w);
    if (rc) return rc;
    const size_t cells = (size_t)g.W * g.H;
    void *stage, *a1 = nullptr, *dw = nullptr;
    if ((rc = get_buf(c, "layer_stage", cells * g.cap * 2, &stage))) return rc;
    if (agg_pass1 && (rc = get_buf(c, "layer_pass1", cells * g.cs * 2, &a1))) return rc;
    if ((rc = get_buf(c, "layer_wta", cells * 2, &dw))) return rc;
    hipError_t e = hipMemcpyAsync(stage, raw, cells * g.cap, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = lux::launch_restride_u8((const uint8_t*)stage, g.cap, w.raw, g.cs, (int)cells, g.cap, 255, s);
    if (e != hipSuccess) return hip_fail(c, e, "upload raw");
    lux::SgmOutputs o;
    o.disp = (uint16_t*)dw;
    o.tol = 0;
    o.write_agg = agg != nullptr;
    o.bands = c->opt_bands;
    e = lux::launch_sgm(g, w.b, o, p->p1, p->p2, (uint16_t*)a1, s);
    if (e != hipSuccess) return hip_fail(c, e, "aggregation");
    if (wta) {
        e = hipMemcpyAsync(wta, dw, cells * 2, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return hip_fail(c, e, "wta");
    }
    uint16_t* outs[2] = {agg, agg_pass1};
    const uint16_t* srcs[2] = {w.agg, (const uint16_t*)a1};
    for (int k = 0; k < 2; k++) {
        if (!outs[k]) continue;
        e = lux::launch_restride_u16(srcs[k], g.cs, (uint16_t*)stage, g.cap, (int)cells, g.cap, s);
        if (e == hipSuccess) e = hipMemcpyAsync(outs[k], stage, cells * g.cap * 2, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(c, e, "download aggregate");
    }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "compute_layer");
    return check_band_error(c, s);
}

int lux_sgm_set_option(lux_sgm_ctx* c, int32_t option, int32_t value)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    switch (option) {
        case LUX_SGM_OPT_FUSED_BANDS: c->opt_bands = value != 0; return LUX_SGM_OK;
        default: return fail(c, LUX_SGM_EINVAL, "unknown option");
    }
}

int lux_sgm_set_profiling(lux_sgm_ctx* c, int on)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    if (on && !c->ev[0])
        for (auto& e : c->ev) {
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return hip_fail(c, r, "hipEventCreate");
        }
    c->profiling = on != 0;
    c->nev = 0;
    return LUX_SGM_OK;
}

int lux_sgm_stage_times(lux_sgm_ctx* c, float* ms, int32_t cap)
{
    if (!c || !ms) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->nev < 2) return fail(c, LUX_SGM_EINVAL, "no profiled request");
    if (cap < c->nev - 1) return fail(c, LUX_SGM_EINVAL, "stage array too small");
    hipError_t e = hipEventSynchronize(c->ev[c->nev - 1]);
    if (e != hipSuccess) return hip_fail(c, e, "hipEventSynchronize");
    for (int i = 0; i + 1 < c->nev; i++) {
        e = hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]);
        if (e != hipSuccess) return hip_fail(c, e, "hipEventElapsedTime");
    }
    return c->nev - 1;
}

int lux_sgm_plan_info(int32_t width, int32_t height, int32_t tile, int32_t xs_clamp_w, int32_t* visits_plus,
                      int32_t* visits_minus, int32_t* seed_row)
{
    lux::Plan p = lux::make_plan(width, height, tile, xs_clamp_w);
    if (p.status) return p.status == -1 ? LUX_SGM_EINVAL : LUX_SGM_EUNSUPPORTED;
    if (visits_plus) *visits_plus = (int32_t)p.plus.size() / 4;
    if (visits_minus) *visits_minus = (int32_t)p.minus.size() / 4;
    if (seed_row) *seed_row = p.seed_row;
    return LUX_SGM_OK;
}

int lux_sgm_plan_visits(int32_t width, int32_t height, int32_t tile, int32_t xs_clamp_w, int32_t* plus, int32_t cap_plus,
                        int32_t* minus, int32_t cap_minus)
{
    lux::Plan p = lux::make_plan(width, height, tile, xs_clamp_w);
    if (p.status) return p.status == -1 ? LUX_SGM_EINVAL : LUX_SGM_EUNSUPPORTED;
    if ((int)p.plus.size() / 4 > cap_plus || (int)p.minus.size() / 4 > cap_minus) return LUX_SGM_EINVAL;
    if (plus) std::memcpy(plus, p.plus.data(), p.plus.size() * sizeof(int));
    if (minus) std::memcpy(minus, p.minus.data(), p.minus.size() * sizeof(int));
    return LUX_SGM_OK;
}

}  // extern "C"
