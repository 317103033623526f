// Request engine: workspaces, HBM budget, plan cache, and the enqueue of one request (declared in engine.h).
#include "engine.h"

#include <algorithm>
#include <cstring>
#include <tuple>

namespace lux {

bool Budget::grow(uint64_t old_n, uint64_t new_n)
{
    std::lock_guard<std::mutex> lk(mu);
    if (limit && used - old_n + new_n > limit) return false;
    used = used - old_n + new_n;
    peak = std::max(peak, used);
    return true;
}

void Budget::release(uint64_t n)
{
    std::lock_guard<std::mutex> lk(mu);
    used -= std::min(used, n);
}

int hip_status(hipError_t e, const char* where, std::string* err)
{
    if (err) *err = std::string(where) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? LUX_SGM_ENOMEM : LUX_SGM_EDEVICE;
}

static int fail(int code, const std::string& msg, std::string* err)
{
    if (err) *err = msg;
    return code;
}

int Workspace::get(const char* name, size_t bytes, void** out, std::string* err)
{
    DevBuf& b = bufs[name];
    if (arena) {
        if (b.n < bytes) {
            const size_t at = (arena_used + 255) & ~(size_t)255;
            if (at > arena_cap || bytes > arena_cap - at)
                return fail(LUX_SGM_ENOMEM, std::string("scratch exhausted placing ") + name, err);
            b.p = arena + at;
            b.n = bytes;
            arena_used = at + bytes;
            filled.erase(name);
        }
        *out = b.p;
        return 0;
    }
    if (b.n < bytes) {
        if (budget && !budget->grow(b.n, bytes))
            return fail(LUX_SGM_ENOMEM, std::string("HBM budget exceeded allocating ") + name, err);
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
        filled.erase(name);
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) {
            b.p = nullptr;
            if (budget) budget->release(bytes);
            return hip_status(e, "hipMalloc", err);
        }
        b.n = bytes;
    }
    *out = b.p;
    return 0;
}

uint64_t Workspace::held() const
{
    uint64_t s = 0;
    for (auto& kv : bufs) s += kv.second.n;
    return s;
}

int Workspace::pinned(size_t bytes, void** out, std::string* err)
{
    if (host_pinned_n < bytes) {
        if (host_pinned) (void)hipHostFree(host_pinned);
        host_pinned = nullptr;
        host_pinned_n = 0;
        hipError_t e = hipHostMalloc(&host_pinned, bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            host_pinned = nullptr;
            return hip_status(e, "hipHostMalloc", err);
        }
        host_pinned_n = bytes;
    }
    *out = host_pinned;
    return 0;
}

void Workspace::release()
{
    if (arena) {  // the arena is not this workspace's to free
        bufs.clear();
        filled.clear();
        sync_zeroed = nullptr;
        arena_used = 0;
        return;
    }
    if (!bufs.empty()) (void)hipSetDevice(device);
    for (auto& kv : bufs) {
        if (kv.second.p) (void)hipFree(kv.second.p);
        if (budget) budget->release(kv.second.n);
    }
    bufs.clear();
    filled.clear();
    sync_zeroed = nullptr;
    if (host_pinned) (void)hipHostFree(host_pinned);
    host_pinned = nullptr;
    host_pinned_n = 0;
}

int check_rect(const lux_sgm_rect_params* p, std::string* err)
{
    if (!p) return fail(LUX_SGM_EINVAL, "null params", err);
    if (p->width <= 0 || p->height <= 0 || p->stride < p->width) return fail(LUX_SGM_EINVAL, "bad image size/stride", err);
    if (p->width > 10240) return fail(LUX_SGM_EUNSUPPORTED, "width above 10240", err);
    if (!(p->bits == 8 || (p->bits >= 9 && p->bits <= 16))) return fail(LUX_SGM_EINVAL, "bits must be 8 or 9..16", err);
    if (p->cost != LUX_SGM_COST_CENSUS && p->cost != LUX_SGM_COST_AD) return fail(LUX_SGM_EINVAL, "unknown cost", err);
    if (p->cost == LUX_SGM_COST_CENSUS && (p->census_radius < 1 || p->census_radius > 3))
        return fail(LUX_SGM_EINVAL, "census radius must be 1..3", err);
    if (p->cost == LUX_SGM_COST_AD && (p->ad_cap < 0 || p->ad_cap > 28))
        return fail(LUX_SGM_EINVAL, "ad_cap must be 0..28", err);
    if (p->disparities <= 0 || p->align <= 0) return fail(LUX_SGM_EINVAL, "disparities/align must be positive", err);
    if (p->disparities > 65535 || p->align > 65535)  // the uint16 range of the API; keeps the rounding below in int
        return fail(LUX_SGM_EUNSUPPORTED, "more than 1024 lanes per cell", err);
    if (((p->disparities + p->align - 1) / p->align) * p->align > 1024)
        return fail(LUX_SGM_EUNSUPPORTED, "more than 1024 lanes per cell", err);
    if (p->p1 < 0 || p->p1 > 65535 || p->p2 < 0 || p->p2 > 65535)
        return fail(LUX_SGM_EINVAL, "penalties must be 0..65535", err);
    if (p->tile < 2) return fail(LUX_SGM_EINVAL, "tile must be >= 2", err);
    if (p->lr_tolerance < 0) return fail(LUX_SGM_EINVAL, "negative L/R tolerance", err);
    return 0;
}

int check_layer(const lux_sgm_layer_params* p, std::string* err)
{
    if (!p) return fail(LUX_SGM_EINVAL, "null params", err);
    if (p->width <= 0 || p->height <= 0 || p->planes <= 0 || p->align <= 0 || p->tile < 2)
        return fail(LUX_SGM_EINVAL, "bad layer parameters", err);
    if (p->width > 10240) return fail(LUX_SGM_EUNSUPPORTED, "width above 10240", err);
    if (p->planes > 65535 || p->align > 65535)  // the uint16 range of the API; keeps the rounding below in int
        return fail(LUX_SGM_EUNSUPPORTED, "more than 1024 lanes per cell", err);
    if (((p->planes + p->align - 1) / p->align) * p->align > 1024)
        return fail(LUX_SGM_EUNSUPPORTED, "more than 1024 lanes per cell", err);
    if (p->p1 < 0 || p->p1 > 65535 || p->p2 < 0 || p->p2 > 65535)
        return fail(LUX_SGM_EINVAL, "penalties must be 0..65535", err);
    return 0;
}

// Each plan's visit lists have a buffer of their own, uploaded once: a pyramid's levels (or a server's request
// sizes) each keep theirs instead of re-uploading them per request.
static std::string visits_name(const Plan& plan)
{
    return "visits " + std::to_string(plan.W) + "x" + std::to_string(plan.H) + " T" + std::to_string(plan.T) +
           (plan.xsw ? " xsw" : "");
}

bool path_bytes(const lux_sgm_rect_params* p)
{
    // Every path cost lc = satsub(sat(raw + m), minPrev) is at most raw + P2 (m <= minPrev + P2, DenseStereoSgm.cs:288-298),
    // and the first step of a chain gives lc = raw (the line buffers hold 2000 and minPrev is 2000). A census cost of
    // radius r is at most (2r+1)^2 - 1, so with no padding planes (cap = count, whose raw would be 255) every lc fits a
    // byte when that maximum plus P2 does; the eight-path sum then stays below 65535 and needs no saturation.
    if (p->cost != LUX_SGM_COST_CENSUS) return false;
    const Geometry g = geometry(p->width, p->height, p->disparities, p->align);
    if (g.cap != p->disparities || g.dpl > 4) return false;
    const int r = p->census_radius, cmax = (2 * r + 1) * (2 * r + 1) - 1;
    return cmax + p->p2 <= 255;
}

std::vector<BufferNeed> request_buffers(const Geometry& g, const Plan& plan, bool census, size_t host_img_bytes, bool pv8)
{
    const size_t cells = (size_t)g.W * g.H;
    const size_t maxvis = std::max(plan.plus.size(), plan.minus.size()) / 4 + 1;
    const size_t nv = plan.plus.size() + plan.minus.size() + 4 * (size_t)g.H;
    std::vector<BufferNeed> v = {
        {"raw", cells * g.cs},
        {"agg", pv8 ? 0 : cells * g.cs * 2},  // with byte path volumes only when the caller asks for the aggregate
        {"seeds", (size_t)3 * g.W * g.cs * 2},
        {"seedmins", (size_t)3 * g.W * 2},
        {"cols", 2 * cols_half(g) * 2},
        {"colmins", 2 * colmins_half(g) * 2},
        // one half per pass: the persistent executor runs the two passes' replays in one launch, where a buffer
        // rewritten after another XCD read it could leave that XCD's cache stale
        {"replay_partial", 2 * replay_half(g, maxvis) * 2},
        {"replay_chain", 2 * replay_half(g, maxvis) * 2},
        {"wta_dr", cells * 2},
        {"wta_dr0", (size_t)g.H * 4},
        {"band_boundary", (size_t)6 * g.W * g.cs * 2},
        {"band_boundary_min", (size_t)6 * g.W * 4},
        {"band_sync", (size_t)(4 + 2 * band_tiles(g.W, g.H)) * sizeof(int)},
        {visits_name(plan), (nv + 4) * sizeof(int)},
        {"disp_tmp", cells * 2},
    };
    if (census) {
        v.push_back({"census_l", cells * 8});
        v.push_back({"census_r", cells * 8});
    }
    if (pv8) {
        v.push_back({"path_bytes", 7 * cells * g.cs});
        v.push_back({"border_agg", 2 * border_half(g) * 2});
    }
    if (host_img_bytes) {
        v.push_back({"img_l", host_img_bytes});
        v.push_back({"img_r", host_img_bytes});
        v.push_back({"out_disp", cells * 2});
        v.push_back({"out_sub", cells * 4});
        v.push_back({"out_valid", cells});
    }
    return v;
}

uint64_t growth(const Workspace& ws, const std::vector<BufferNeed>& need)
{
    uint64_t g = 0;
    for (const BufferNeed& b : need) {
        auto it = ws.bufs.find(b.name);
        const size_t have = it == ws.bufs.end() ? 0 : it->second.n;
        if (b.bytes > have) g += b.bytes - have;
    }
    return g;
}

uint64_t rect_workspace_bytes(const lux_sgm_rect_params* p)
{
    Geometry g = geometry(p->width, p->height, p->disparities, p->align);
    std::shared_ptr<Plan> plan;
    if (plan_for(g.W, g.H, p->tile, p->xs_clamp_w, &plan, nullptr)) return 0;
    uint64_t s = 0;
    for (const BufferNeed& b : request_buffers(g, *plan, p->cost == LUX_SGM_COST_CENSUS, 0, path_bytes(p))) s += b.bytes;
    return s;
}

int plan_for(int W, int H, int T, int xsw, std::shared_ptr<Plan>* out, std::string* err)
{
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, int>, std::shared_ptr<Plan>> cache;
    auto key = std::make_tuple(W, H, T, xsw);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            *out = it->second;
            return 0;
        }
    }
    auto p = std::make_shared<Plan>(make_plan(W, H, T, xsw));
    if (p->status != 0) return fail(p->status == -1 ? LUX_SGM_EINVAL : LUX_SGM_EUNSUPPORTED, p->error, err);
    std::lock_guard<std::mutex> lk(mu);
    *out = cache.emplace(key, p).first->second;
    return 0;
}

int prepare_layout(Workspace& ws, const Geometry& g, const Plan& plan, bool census, bool pv8, size_t host_img_bytes,
                   const int* visits, Prepared* w, std::map<std::string, void*>* named, std::string* err)
{
    std::map<std::string, void*> ptr;
    for (const BufferNeed& nb : request_buffers(g, plan, census, host_img_bytes, pv8)) {
        if (visits && nb.name == visits_name(plan)) continue;  // the caller keeps the visit lists elsewhere
        void* q;
        int rc = ws.get(nb.name.c_str(), nb.bytes, &q, err);
        if (rc) return rc;
        ptr[nb.name] = q;
    }
    void *raw = ptr["raw"], *agg = ptr["agg"], *seeds = ptr["seeds"], *seedmins = ptr["seedmins"], *cols = ptr["cols"],
         *colmins = ptr["colmins"], *partial = ptr["replay_partial"], *drb = ptr["wta_dr"], *dr0 = ptr["wta_dr0"],
         *bnd = ptr["band_boundary"], *bndmin = ptr["band_boundary_min"], *sync = ptr["band_sync"];
    const int* vis = visits ? visits : (const int*)ptr[visits_name(plan)];
    if (census) {
        w->b.cl = (const uint64_t*)ptr["census_l"];
        w->b.cr = (const uint64_t*)ptr["census_r"];
    }
    void* dtmp = ptr["disp_tmp"];
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
    w->b.max_tiles = band_tiles(g.W, g.H);
    w->b.epoch = ws.epoch;
    w->b.seeds = (uint16_t*)seeds;
    w->b.seedmins = (uint16_t*)seedmins;
    w->b.cols = (uint16_t*)cols;
    w->b.colmins = (uint16_t*)colmins;
    w->b.partial = (uint16_t*)partial;
    w->b.chainp = (uint16_t*)ptr["replay_chain"];
    w->b.visits_plus = vis;
    w->b.n_plus = (int)plan.plus.size() / 4;
    w->b.visits_minus = vis + plan.plus.size();
    w->b.n_minus = (int)plan.minus.size() / 4;
    w->b.cellvis_plus = vis + plan.plus.size() + plan.minus.size();
    w->b.cellvis_minus = w->b.cellvis_plus + 2 * (size_t)g.H;
    w->b.seed_row = plan.seed_row;
    w->b.disp_tmp = (uint16_t*)dtmp;
    // constant P2: the kernels' guidance loads read word 0 of an allocation that always exists
    w->b.ap = AdaptP2();
    w->b.ap.guid = (const uint32_t*)sync;
    if (pv8) {
        // seven volumes one after another (a per-cell interleave of the seven, 7 * cs bytes a cell, measured the
        // same for the last row path and slower for the sweeps)
        const size_t vol = (size_t)g.W * g.H * g.cs;
        for (int k = 0; k < 7; k++) w->b.pv[k] = (uint8_t*)ptr["path_bytes"] + k * vol;
        w->b.pvs = g.cs;
        w->b.bagg[0] = (uint16_t*)ptr["border_agg"];
        w->b.bagg[1] = w->b.bagg[0] + border_half(g);
    }
    if (named) *named = ptr;
    return 0;
}

int upload_plan(const Plan& plan, int* vis, hipStream_t s, std::string* err)
{
    const size_t np = plan.plus.size(), nm = plan.minus.size();
    hipError_t e = hipMemcpyAsync(vis, plan.plus.data(), np * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(vis + np, plan.minus.data(), nm * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(vis + np + nm, plan.cellvis_plus.data(), plan.cellvis_plus.size() * sizeof(int),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(vis + np + nm + plan.cellvis_plus.size(), plan.cellvis_minus.data(),
                           plan.cellvis_minus.size() * sizeof(int), hipMemcpyHostToDevice, s);
    return e == hipSuccess ? 0 : hip_status(e, "upload plan", err);
}

size_t plan_ints(const Plan& plan)
{
    return plan.plus.size() + plan.minus.size() + plan.cellvis_plus.size() + plan.cellvis_minus.size();
}

int prepare(Workspace& ws, const Geometry& g, int T, int xsw, bool census, hipStream_t s, Prepared* w, std::string* err,
            bool pv8)
{
    std::shared_ptr<Plan> plan;
    int rc = plan_for(g.W, g.H, T, xsw, &plan, err);
    if (rc) return rc;
    std::map<std::string, void*> ptr;
    if ((rc = prepare_layout(ws, g, *plan, census, pv8, 0, nullptr, w, &ptr, err))) return rc;
    void* sync = ptr["band_sync"];
    if (sync != ws.sync_zeroed) {  // a new buffer: no flag may hold a live epoch
        hipError_t ez = launch_fill_u32((uint32_t*)sync, 4 + 2 * (size_t)w->b.max_tiles, 0u, s);
        if (ez != hipSuccess) return hip_status(ez, "zero band flags", err);
        ws.sync_zeroed = sync;
    }
    if (++ws.epoch <= 0) ws.epoch = 1;
    w->b.epoch = ws.epoch;
    if (!ws.filled.count(visits_name(*plan))) {  // upload the visit lists once per plan
        if ((rc = upload_plan(*plan, (int*)ptr[visits_name(*plan)], s, err))) return rc;
        ws.filled.insert(visits_name(*plan));
    }
    return 0;
}

int enqueue_rect(Workspace& ws, Prepared& w, const lux_sgm_rect_params* p, const void* dl, const void* dr, const RectOut& o,
                 const EngineOpts& opt, hipStream_t s, hipEvent_t* ev, std::string* err)
{
    (void)ws;
    const Geometry& g = w.g;
    const bool img16 = p->bits > 8;
    hipError_t e;
    if (ev) (void)hipEventRecord(ev[0], s);
    if (p->cost == LUX_SGM_COST_CENSUS) {  // the codes here; the cost volume inside the first sweep
        e = launch_census(dl, img16, g.W, g.H, p->stride, p->census_radius, (uint64_t*)w.b.cl, s);
        if (e == hipSuccess) e = launch_census(dr, img16, g.W, g.H, p->stride, p->census_radius, (uint64_t*)w.b.cr, s);
    } else {
        w.b.cl = w.b.cr = nullptr;
        e = launch_cost_ad(dl, dr, img16, p->bits, p->stride, p->ad_cap, g, w.raw, s);
    }
    if (e != hipSuccess) return hip_status(e, "cost volume", err);
    if (ev) (void)hipEventRecord(ev[1], s);
    SgmOutputs so;
    so.disp = o.disp ? o.disp : w.b.disp_tmp;
    so.sub = o.sub;
    so.valid = o.valid;
    so.tol = p->lr_tolerance;
    so.write_agg = o.agg != nullptr;
    so.bands = opt.bands;
    w.b.zero_flags = opt.replayable;
    w.b.solo = !opt.shared;
    if (w.b.pv[0] && o.agg) {  // byte path volumes hold no aggregate: the last row path writes it here when asked
        void* a;
        int rc = ws.get("agg", (size_t)g.W * g.H * g.cs * 2, &a, err);
        if (rc) return rc;
        w.agg = w.b.agg = (uint16_t*)a;
    }
    e = launch_sgm(g, w.b, so, p->p1, p->p2, nullptr, s, ev ? ev + 2 : nullptr);
    if (e != hipSuccess) return hip_status(e, "aggregation", err);
    if (o.agg) {
        e = launch_restride_u16(w.agg, g.cs, o.agg, g.cap, g.W * g.H, g.cap, s);
        if (e != hipSuccess) return hip_status(e, "agg copy", err);
    }
    return 0;
}

// ------------------------------------------------------------------ coarse-to-fine
static const char* kPyrImg[2][4] = {{"pyr_l0", "pyr_l1", "pyr_l2", "pyr_l3"}, {"pyr_r0", "pyr_r1", "pyr_r2", "pyr_r3"}};
static const char* kPyrDisp[4] = {"pyr_d0", "pyr_d1", "pyr_d2", "pyr_d3"};

std::vector<PyrLevel> pyramid_levels(const lux_sgm_pyramid_params* p)
{
    const int L = p->levels;
    std::vector<PyrLevel> v(L);
    int W = p->base.width, H = p->base.height;
    for (int k = L - 1; k >= 0; k--) {
        v[k].W = W;
        v[k].H = H;
        v[k].stride = k == L - 1 ? p->base.stride : W;
        v[k].D = p->base.disparities >> (L - 1 - k);
        v[k].radius = k == 0 ? 0 : p->radius[k - 1];
        W = (W + 1) / 2;
        H = (H + 1) / 2;
    }
    return v;
}

static lux_sgm_rect_params level_params(const lux_sgm_pyramid_params* p, const PyrLevel& l)
{
    lux_sgm_rect_params q = p->base;
    q.width = l.W;
    q.height = l.H;
    q.stride = l.stride;
    q.disparities = l.radius ? 2 * l.radius : l.D;
    return q;
}

int check_pyramid(const lux_sgm_pyramid_params* p, std::string* err)
{
    if (!p) return fail(LUX_SGM_EINVAL, "null params", err);
    int rc = check_rect(&p->base, err);
    if (rc) return rc;
    if (p->levels < 1 || p->levels > 4) return fail(LUX_SGM_EINVAL, "levels must be 1..4", err);
    if (p->base.cost != LUX_SGM_COST_CENSUS) return fail(LUX_SGM_EUNSUPPORTED, "coarse-to-fine takes the census cost", err);
    const auto lv = pyramid_levels(p);
    if (lv[0].W < 2 || lv[0].H < 2 || lv[0].D < 1) return fail(LUX_SGM_EINVAL, "the coarsest level is empty", err);
    for (int k = 1; k < p->levels; k++) {
        const int r = lv[k].radius;
        if (r <= 0 || (2 * r) % p->base.align || 2 * r > 128 || 2 * r > lv[k].D)
            return fail(LUX_SGM_EINVAL, "each radius needs 2*radius a multiple of align, <= 128 and <= the level's planes", err);
    }
    return 0;
}

int pyramid_buffers(const lux_sgm_pyramid_params* p, size_t host_img_bytes, std::vector<BufferNeed>* out, std::string* err)
{
    int rc = check_pyramid(p, err);
    if (rc) return rc;
    const auto lv = pyramid_levels(p);
    std::map<std::string, size_t> need;
    auto put = [&](const std::string& name, size_t n) {
        size_t& v = need[name];
        v = std::max(v, n);
    };
    const size_t esz = p->base.bits > 8 ? 2 : 1;
    for (int k = 0; k < p->levels; k++) {
        const PyrLevel& l = lv[k];
        const lux_sgm_rect_params q = level_params(p, l);
        Geometry g = geometry(l.W, l.H, q.disparities, q.align);
        std::shared_ptr<Plan> plan;
        if ((rc = plan_for(l.W, l.H, q.tile, q.xs_clamp_w, &plan, err))) return rc;
        for (const BufferNeed& b : request_buffers(g, *plan, true)) put(b.name, b.bytes);
        const size_t cells = (size_t)l.W * l.H;
        if (l.radius) {
            put("vstart", cells * 2);
            put("vcap", cells * 2);
            put("vcount", cells * 2);
            put("seedstart", (size_t)3 * l.W * 2);
            put("seedcap", (size_t)3 * l.W * 2);
        }
        if (k < p->levels - 1) {
            put(kPyrImg[0][k], cells * esz);
            put(kPyrImg[1][k], cells * esz);
            put(kPyrDisp[k], cells * 2);
        }
    }
    if (host_img_bytes) {
        const size_t cells = (size_t)p->base.width * p->base.height;
        put("img_l", host_img_bytes);
        put("img_r", host_img_bytes);
        put("out_disp", cells * 2);
        put("out_sub", cells * 4);
    }
    out->clear();
    for (auto& kv : need) out->push_back({kv.first, kv.second});
    return 0;
}

int enqueue_pyramid(Workspace& ws, const lux_sgm_pyramid_params* p, const void* dl, const void* dr, uint16_t* disp,
                    float* sub, hipStream_t s, std::string* err, const Conc* conc)
{
    std::vector<BufferNeed> need;
    int rc = pyramid_buffers(p, 0, &need, err);
    if (rc) return rc;
    // every buffer at its largest size first: nothing is reallocated once work is queued
    std::map<std::string, void*> ptr;
    for (const BufferNeed& b : need) {
        void* q;
        if ((rc = ws.get(b.name.c_str(), b.bytes, &q, err))) return rc;
        ptr[b.name] = q;
    }
    bool volumes = true;
    {  // the path volumes of the ranged levels (one launch per pass), at the largest level's size
        size_t mx = 0;
        for (const PyrLevel& l : pyramid_levels(p))
            if (l.radius) {
                const lux_sgm_rect_params q = level_params(p, l);
                const Geometry g = geometry(l.W, l.H, q.disparities, q.align);
                mx = std::max(mx, (size_t)g.W * g.H * g.cs * 2);
            }
        std::string e2;
        void* q;
        for (const char* nm : {"aggp0", "aggp1", "aggp2"})
            if (mx && ws.get(nm, mx, &q, &e2)) volumes = false;  // over the budget: the paths run one after another
    }
    const auto lv = pyramid_levels(p);
    const int L = p->levels;
    const bool img16 = p->base.bits > 8;
    const void* imgs[2][4];
    imgs[0][L - 1] = dl;
    imgs[1][L - 1] = dr;
    for (int k = L - 2; k >= 0; k--)
        for (int e = 0; e < 2; e++) {
            imgs[e][k] = ptr[kPyrImg[e][k]];
            hipError_t h = launch_downsample2(imgs[e][k + 1], img16, lv[k + 1].W, lv[k + 1].H, lv[k + 1].stride,
                                              (void*)imgs[e][k], lv[k].W, lv[k].H, s);
            if (h != hipSuccess) return hip_status(h, "downsample", err);
        }
    for (int k = 0; k < L; k++) {
        const PyrLevel& l = lv[k];
        const lux_sgm_rect_params q = level_params(p, l);
        Geometry g = geometry(l.W, l.H, q.disparities, q.align);
        Prepared w;
        if ((rc = prepare(ws, g, q.tile, q.xs_clamp_w, true, s, &w, err))) return rc;
        uint16_t* ldisp = k == L - 1 ? disp : (uint16_t*)ptr[kPyrDisp[k]];
        float* lsub = k == L - 1 ? sub : nullptr;
        if (!l.radius) {  // the coarsest level: the whole plane range, as a rectified pair
            RectOut o;
            o.disp = ldisp;
            o.sub = lsub;
            EngineOpts opt;
            if ((rc = enqueue_rect(ws, w, &q, imgs[0][k], imgs[1][k], o, opt, s, nullptr, err))) return rc;
            continue;
        }
        uint16_t *vs = (uint16_t*)ptr["vstart"], *vc = (uint16_t*)ptr["vcap"], *vn = (uint16_t*)ptr["vcount"];
        hipError_t h = launch_ranges_from_coarse((const uint16_t*)ptr[kPyrDisp[k - 1]], lv[k - 1].W, lv[k - 1].H, l.W, l.H,
                                                 l.radius, l.D, vs, vc, vn, s);
        if (h == hipSuccess) h = launch_census(imgs[0][k], img16, l.W, l.H, l.stride, q.census_radius, (uint64_t*)w.b.cl, s);
        if (h == hipSuccess) h = launch_census(imgs[1][k], img16, l.W, l.H, l.stride, q.census_radius, (uint64_t*)w.b.cr, s);
        if (h == hipSuccess) h = launch_cost_census_ranges(w.b.cl, w.b.cr, l.W, l.H, g.cs, vs, vn, w.raw, s);
        if (h != hipSuccess) return hip_status(h, "level ranges and costs", err);
        w.b.cl = w.b.cr = nullptr;
        w.b.vstart = vs;
        w.b.vcap = vc;
        w.b.vcount = vn;
        w.b.seedstart = (uint16_t*)ptr["seedstart"];
        w.b.seedcap = (uint16_t*)ptr["seedcap"];
        if (volumes) try_path_volumes(ws, &w);  // sized above: no reallocation under queued work
        attach_chain(&w, conc);
        SgmOutputs so;
        so.disp = ldisp ? ldisp : w.b.disp_tmp;
        so.sub = lsub;
        so.tol = 0;
        h = launch_sgm(g, w.b, so, q.p1, q.p2, nullptr, s);
        if (h != hipSuccess) return hip_status(h, "level aggregation", err);
    }
    return 0;
}

int check_band_error(Workspace& ws, hipStream_t s, std::string* err)
{
    auto it = ws.bufs.find("band_sync");
    if (it == ws.bufs.end() || !it->second.p) return 0;
    int word = 0;
    hipError_t e = hipMemcpyAsync(&word, it->second.p, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e, "read band error word", err);
    if (word) {
        (void)launch_fill_u32((uint32_t*)it->second.p, 1, 0u, s);
        (void)hipStreamSynchronize(s);
        return fail(LUX_SGM_EDEVICE, "a band tile timed out waiting for its neighbours", err);
    }
    return 0;
}

}  // namespace lux

namespace lux {

bool try_path_volumes(Workspace& ws, Prepared* w)
{
    const size_t bytes = (size_t)w->g.W * w->g.H * w->g.cs * 2;
    static const char* names[3] = {"aggp0", "aggp1", "aggp2"};
    void* p[3];
    std::string err;
    for (int k = 0; k < 3; k++)
        if (ws.get(names[k], bytes, &p[k], &err)) return false;
    for (int k = 0; k < 3; k++) w->b.aggp[k] = (uint16_t*)p[k];
    return true;
}

void attach_chain(Prepared* w, const Conc* c)
{
    if (!c || !c->chain || !c->chain_fork || !c->chain_join) return;
    w->b.chain = c->chain;
    w->b.chain_fork = c->chain_fork;
    w->b.chain_join = c->chain_join;
}

int prepare_literal(Workspace& ws, const Geometry& g, Prepared* w, std::string* err)
{
    const size_t cells = (size_t)g.W * g.H;
    void *raw, *agg;
    int rc;
    if ((rc = ws.get("raw", cells * g.cs, &raw, err)) || (rc = ws.get("agg", cells * g.cs * 2, &agg, err))) return rc;
    w->g = g;
    w->raw = (uint8_t*)raw;
    w->agg = (uint16_t*)agg;
    w->b = SgmBuffers{};
    w->b.raw = w->raw;
    w->b.raw_out = w->raw;
    w->b.agg = w->agg;
    return 0;
}

int run_literal(Workspace& ws, LiteralLayer l, int maxcap, uint16_t* agg_pass1, uint16_t* wta, hipStream_t s,
                std::string* err)
{
    if (maxcap > 1024) return fail(LUX_SGM_EUNSUPPORTED, "the line-buffer-literal path takes at most 1024 lanes per cell", err);
    l.npl = maxcap <= 128 ? 2 : maxcap <= 256 ? 4 : 0;
    int& res = ws.lit_resident[l.npl == 2 ? 0 : l.npl == 4 ? 1 : 2][l.ap.on != 0];  // per context, so per device
    if (res < 0) res = literal_resident(l.npl, l.ap.on != 0);
    if (res < std::min(literal_depth(l.W, l.T), l.H))
        return fail(LUX_SGM_EUNSUPPORTED, "the device holds too few wavefronts for the line-buffer-literal schedule", err);
    void* scr;
    int rc = ws.get("literal", literal_scratch_bytes(l.W, l.H, l.maxhi), &scr, err);
    if (rc) return rc;
    hipError_t e = launch_literal(l, scr, agg_pass1, wta, s);
    uint32_t bad = 0;
    if (e == hipSuccess) e = literal_error(scr, l.W, l.H, l.maxhi, s, &bad);
    if (e != hipSuccess) return hip_status(e, "line-buffer-literal layer", err);
    if (bad) return fail(LUX_SGM_EDEVICE, "a line-buffer-literal wait gave up", err);
    return 0;
}

}  // namespace lux
