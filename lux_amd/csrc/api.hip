// C ABI of liblux_sgm.so (declared in include/lux_sgm.h): contexts and the entry points that replace
// DenseLayer.Compute (DenseStereoSgm.cs:181-232), the rectified front end, and the batched form.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lux_sgm.h"
#include "ctx.h"
#include "engine.h"
#include "exec.h"

namespace {

thread_local std::string g_create_error;

using lux::fail;
using lux::GraphEntry;
using lux::hip_fail;
using lux::kMaxBranches;

int rect_device(lux_sgm_ctx* c, const lux_sgm_rect_params* p, const void* dl, const void* dr, const lux::RectOut& o,
                hipStream_t s)
{
    lux::Geometry g = lux::geometry(p->width, p->height, p->disparities, p->align);
    lux::Prepared w;
    int rc = lux::prepare(c->ws, g, p->tile, p->xs_clamp_w, p->cost == LUX_SGM_COST_CENSUS, s, &w, &c->err,
                          lux::path_bytes(p));
    if (rc) return rc;
    const lux::Conc conc = lux::conc_of(c);
    lux::attach_chain(&w, &conc);
    lux::EngineOpts opt;
    opt.bands = c->opt_bands;
    rc = lux::enqueue_rect(c->ws, w, p, dl, dr, o, opt, s, c->profiling ? c->ev : nullptr, &c->err);
    if (!rc && c->profiling) c->nev = 2 + lux::kSgmMarks;
    return rc;
}

void destroy_graph(GraphEntry& g)
{
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    lux::exec_destroy(g.xb);
    g = GraphEntry();
}

void destroy_graphs(lux_sgm_ctx* c)
{
    for (auto& kv : c->graphs) destroy_graph(kv.second);
    c->graphs.clear();
}

std::string shape_key(int n, const lux_sgm_pair* pairs)
{
    std::string k((const char*)&n, sizeof n);
    for (int i = 0; i < n; i++) k.append((const char*)&pairs[i].params, sizeof(lux_sgm_rect_params));
    return k;
}

lux_sgm_schedule default_schedule(int n)
{
    lux_sgm_schedule s;
    s.bands = 0;
    s.branches = std::min(n, 4);
    s.ms = 0.f;
    s.executor = 0;
    s.order = 0;
    return s;
}

// Longest pair first onto the least loaded branch.
std::vector<int> assign_branches(int n, const lux_sgm_pair* pairs, int branches)
{
    std::vector<int> order(n), br(n);
    for (int i = 0; i < n; i++) order[i] = i;
    auto cells = [&](int i) {
        const lux_sgm_rect_params& p = pairs[i].params;
        return (double)p.width * p.height * p.disparities;
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells(a) > cells(b); });
    std::vector<double> load(branches, 0.0);
    for (int i : order) {
        int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        br[i] = k;
        load[k] += cells(i);
    }
    return br;
}

int ensure_batch_streams(lux_sgm_ctx* c) { return lux::ensure_side_streams(c); }

// Prepare every pair's workspace (allocation and plan upload happen here, outside any capture).
int prepare_batch(lux_sgm_ctx* c, int n, const lux_sgm_pair* pairs, std::vector<lux::Prepared>* prep)
{
    while ((int)c->bws.size() < n) {
        c->bws.emplace_back(new lux::Workspace());
        c->bws.back()->device = c->device;
        c->bws.back()->budget = &c->budget;
    }
    prep->assign(n, lux::Prepared());
    for (int i = 0; i < n; i++) {
        const lux_sgm_rect_params* p = &pairs[i].params;
        int rc = lux::check_rect(p, &c->err);
        if (rc) return rc;
        if (!pairs[i].left || !pairs[i].right) return fail(c, LUX_SGM_EINVAL, "null image in batch");
        lux::Workspace& ws = *c->bws[i];
        std::map<std::string, void*> before;
        for (auto& kv : ws.bufs) before[kv.first] = kv.second.p;
        lux::Geometry g = lux::geometry(p->width, p->height, p->disparities, p->align);
        rc = lux::prepare(ws, g, p->tile, p->xs_clamp_w, p->cost == LUX_SGM_COST_CENSUS, c->stream, &(*prep)[i], &c->err,
                          lux::path_bytes(p));
        if (rc) return rc;
        for (auto& kv : ws.bufs) {
            auto it = before.find(kv.first);
            if (it == before.end() || it->second != kv.second.p) {
                c->gen++;
                break;
            }
        }
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? 0 : hip_fail(c, e, "prepare batch");
}

// Capture the executor's launches of the batch into a graph (the work lists are uploaded first, outside the capture).
int capture_exec(lux_sgm_ctx* c, int n, const lux_sgm_pair* pairs, std::vector<lux::Prepared>& prep,
                 const lux_sgm_schedule& sch, GraphEntry* out)
{
    std::vector<lux::ExecPair> xp(n);
    for (int i = 0; i < n; i++) {
        if (!lux::exec_supported(&pairs[i].params, prep[i]))
            return fail(c, LUX_SGM_EUNSUPPORTED, "the executor takes census pairs whose path costs fit a byte");
        xp[i].p = &pairs[i].params;
        xp[i].w = &prep[i];
        xp[i].left = pairs[i].left;
        xp[i].right = pairs[i].right;
        xp[i].disp = pairs[i].disp;
        xp[i].sub = pairs[i].sub;
        xp[i].valid = pairs[i].valid;
    }
    lux::ExecBatch* xb = nullptr;
    int rc = lux::exec_create(xp.data(), n, sch.executor, sch.order, c->device, &xb, &c->err);
    if (rc) return rc;
    hipStream_t cs = c->stream;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t el = lux::exec_enqueue(xb, 0, cs);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(cs, &graph);
        if (e == hipSuccess) e = el;
        if (e == hipSuccess) e = hipGraphInstantiate(&out->exec, graph, nullptr, nullptr, 0);
        out->graph = graph;
    }
    out->xb = xb;
    out->gen = c->gen;
    if (e != hipSuccess) {
        destroy_graph(*out);
        return hip_fail(c, e, "capture executor batch");
    }
    return 0;
}

// Capture the batch into a graph: fork from the capture stream onto the branches, one request per pair.
int capture_batch(lux_sgm_ctx* c, int n, const lux_sgm_pair* pairs, std::vector<lux::Prepared>& prep,
                  const lux_sgm_schedule& sch, GraphEntry* out)
{
    if (sch.executor) return capture_exec(c, n, pairs, prep, sch, out);
    const int nb = std::max(1, std::min({sch.branches, n, kMaxBranches}));
    std::vector<int> br = assign_branches(n, pairs, nb);
    lux::EngineOpts opt;
    opt.bands = sch.bands;
    opt.replayable = true;
    opt.shared = true;
    hipStream_t cs = c->stream;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamBeginCapture");
    int rc = 0;
    e = hipEventRecord(c->fork, cs);
    for (int k = 0; e == hipSuccess && k < nb; k++) e = hipStreamWaitEvent(c->side[k], c->fork, 0);
    for (int i = 0; e == hipSuccess && !rc && i < n; i++) {
        lux::RectOut o;
        o.disp = pairs[i].disp;
        o.sub = pairs[i].sub;
        o.valid = pairs[i].valid;
        rc = lux::enqueue_rect(*c->bws[i], prep[i], &pairs[i].params, pairs[i].left, pairs[i].right, o, opt, c->side[br[i]],
                               nullptr, &c->err);
    }
    for (int k = 0; e == hipSuccess && k < nb; k++) {
        e = hipEventRecord(c->join[k], c->side[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(cs, c->join[k], 0);
    }
    hipGraph_t graph = nullptr;
    hipError_t ee = hipStreamEndCapture(cs, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e == hipSuccess) e = ee;
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(c, e, "capture batch");
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return hip_fail(c, e, "hipGraphInstantiate");
    }
    out->graph = graph;
    out->exec = exec;
    out->gen = c->gen;
    return 0;
}

// The instantiated graph of this batch under schedule sch (built on first use, cached by descriptor).
int batch_graph(lux_sgm_ctx* c, int n, const lux_sgm_pair* pairs, const lux_sgm_schedule& sch, hipGraphExec_t* exec,
                GraphEntry** entry = nullptr)
{
    int rc = ensure_batch_streams(c);
    if (rc) return rc;
    std::vector<lux::Prepared> prep;
    if ((rc = prepare_batch(c, n, pairs, &prep))) return rc;
    std::string key((const char*)&sch.bands, sizeof sch.bands);
    key.append((const char*)&sch.branches, sizeof sch.branches);
    key.append((const char*)&sch.executor, sizeof sch.executor);
    key.append((const char*)&sch.order, sizeof sch.order);
    key.append((const char*)&n, sizeof n);
    key.append((const char*)pairs, sizeof(lux_sgm_pair) * (size_t)n);
    auto it = c->graphs.find(key);
    if (it != c->graphs.end() && it->second.gen != c->gen) {  // a workspace moved since: rebuild
        destroy_graph(it->second);
        c->graphs.erase(it);
        it = c->graphs.end();
    }
    if (it == c->graphs.end()) {
        if (c->graphs.size() >= 16) destroy_graphs(c);
        GraphEntry ge;
        if ((rc = capture_batch(c, n, pairs, prep, sch, &ge))) return rc;
        it = c->graphs.emplace(key, ge).first;
    }
    *exec = it->second.exec;
    if (entry) *entry = &it->second;
    return 0;
}

// After the stream has drained: the executor's error word of this batch's graph (a stage wait that gave up).
int batch_exec_check(lux_sgm_ctx* c, GraphEntry* ge, hipStream_t s)
{
    return ge && ge->xb ? lux::exec_check(ge->xb, s, &c->err) : 0;
}

}  // namespace

extern "C" {

int lux_sgm_version(void) { return 200; }

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
    if (lux::noalloc_setup(device, &g_create_error)) return LUX_SGM_EDEVICE;
    auto* c = new lux_sgm_ctx();
    c->device = device;
    c->budget.limit = hbm_budget;
    c->ws.device = device;
    c->ws.budget = &c->budget;
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
    for (auto s : c->side)
        if (s) (void)hipStreamSynchronize(s);
    destroy_graphs(c);
    c->ws.release();
    c->bws.clear();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (int k = 0; k < kMaxBranches; k++) {
        if (c->join[k]) (void)hipEventDestroy(c->join[k]);
        if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
    }
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->chain_fork) (void)hipEventDestroy(c->chain_fork);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LUX_SGM_OK;
}

const char* lux_sgm_last_error(const lux_sgm_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int lux_sgm_rect_bytes(const lux_sgm_rect_params* p, uint64_t* bytes)
{
    if (!bytes) return LUX_SGM_EINVAL;
    std::string err;
    const int rc = lux::check_rect(p, &err);
    if (rc) return rc;
    const uint64_t px = (uint64_t)p->width * p->height;
    *bytes = lux::rect_workspace_bytes(p) + px * 2 * (p->bits > 8 ? 2 : 1)  // images
             + px * 7;                                                     // outputs
    return LUX_SGM_OK;
}

int lux_sgm_rectified_device(lux_sgm_ctx* c, const lux_sgm_rect_params* p, const void* d_left, const void* d_right,
                             uint16_t* d_disp, float* d_sub, uint8_t* d_valid, uint16_t* d_agg, void* stream)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = lux::check_rect(p, &c->err);
    if (rc) return rc;
    if (!d_left || !d_right) return fail(c, LUX_SGM_EINVAL, "null image");
    (void)hipSetDevice(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    lux::RectOut o;
    o.disp = d_disp;
    o.sub = d_sub;
    o.valid = d_valid;
    o.agg = d_agg;
    return rect_device(c, p, d_left, d_right, o, s);
}

int lux_sgm_rectified(lux_sgm_ctx* c, const lux_sgm_rect_params* p, const void* left, const void* right, uint16_t* disp,
                      float* sub, uint8_t* valid, uint16_t* agg)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = lux::check_rect(p, &c->err);
    if (rc) return rc;
    if (!left || !right) return fail(c, LUX_SGM_EINVAL, "null image");
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    const size_t px = (size_t)p->width * p->height;
    const size_t esz = p->bits > 8 ? 2 : 1;
    const size_t imgb = (size_t)p->height * p->stride * esz;
    lux::Geometry g = lux::geometry(p->width, p->height, p->disparities, p->align);
    void *dl, *dr, *dd, *ds, *dv, *da = nullptr;
    lux::Workspace& ws = c->ws;
    if ((rc = ws.get("img_l", imgb, &dl, &c->err)) || (rc = ws.get("img_r", imgb, &dr, &c->err))) return rc;
    if ((rc = ws.get("out_disp", px * 2, &dd, &c->err)) || (rc = ws.get("out_sub", px * 4, &ds, &c->err)) ||
        (rc = ws.get("out_valid", px, &dv, &c->err)))
        return rc;
    if (agg && (rc = ws.get("out_agg", px * g.cap * 2, &da, &c->err))) return rc;
    hipError_t e = hipMemcpyAsync(dl, left, imgb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, right, imgb, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(c, e, "upload images");
    lux::RectOut o;
    o.disp = disp ? (uint16_t*)dd : nullptr;
    o.sub = sub ? (float*)ds : nullptr;
    o.valid = valid ? (uint8_t*)dv : nullptr;
    o.agg = agg ? (uint16_t*)da : nullptr;
    rc = rect_device(c, p, dl, dr, o, s);
    if (rc) return rc;
    if (disp) e = hipMemcpyAsync(disp, dd, px * 2, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && sub) e = hipMemcpyAsync(sub, ds, px * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && valid) e = hipMemcpyAsync(valid, dv, px, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && agg) e = hipMemcpyAsync(agg, da, px * g.cap * 2, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "download results");
    return lux::check_band_error(ws, s, &c->err);
}

int lux_sgm_box3_device(lux_sgm_ctx* c, const void* d_src, int32_t bits, int32_t channels, int32_t width, int32_t height,
                        int32_t stride, void* d_dst, void* stream)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!d_src || !d_dst || width <= 0 || height <= 0 || stride < width) return fail(c, LUX_SGM_EINVAL, "bad image");
    if (!(bits == 8 || (bits >= 9 && bits <= 16))) return fail(c, LUX_SGM_EINVAL, "bits must be 8 or 9..16");
    if (channels != 1 && channels != 4) return fail(c, LUX_SGM_EINVAL, "channels must be 1 or 4");
    if (channels == 4 && bits != 8) return fail(c, LUX_SGM_EINVAL, "RGBA images are 8-bit");
    (void)hipSetDevice(c->device);
    hipError_t e = lux::launch_box3(d_src, bits > 8, width, height, stride, channels, (1 << bits) - 1, d_dst,
                                    stream ? (hipStream_t)stream : c->stream);
    return e == hipSuccess ? 0 : hip_fail(c, e, "box3");
}

int lux_sgm_box3(lux_sgm_ctx* c, const void* src, int32_t bits, int32_t channels, int32_t width, int32_t height,
                 int32_t stride, void* dst)
{
    if (!c) return LUX_SGM_EINVAL;
    if (!src || !dst || width <= 0 || height <= 0 || stride < width) return fail(c, LUX_SGM_EINVAL, "bad image");
    const size_t esz = bits > 8 ? 2 : 1, inb = (size_t)height * stride * channels * esz,
                 outb = (size_t)height * width * channels * esz;
    void *di, *dd;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        int rc;
        if ((rc = c->ws.get("box_in", inb, &di, &c->err)) || (rc = c->ws.get("box_out", outb, &dd, &c->err))) return rc;
        (void)hipSetDevice(c->device);
        hipError_t e = hipMemcpyAsync(di, src, inb, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "upload image");
    }
    int rc = lux_sgm_box3_device(c, di, bits, channels, width, height, stride, dd, nullptr);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    hipError_t e = hipMemcpyAsync(dst, dd, outb, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? 0 : hip_fail(c, e, "download image");
}

static int check_upsample(lux_sgm_ctx* c, int pw, int ph, int W, int H, int setW, int setH)
{
    if (pw <= 0 || ph <= 0 || W <= 0 || H <= 0) return fail(c, LUX_SGM_EINVAL, "bad sizes");
    if (W != setW * 2 || H != setH * 2) return fail(c, LUX_SGM_EINVAL, "Incorrect output depth size!");  // (:46-47)
    if ((int64_t)W * 65536 > INT32_MAX || (int64_t)H * 65536 > INT32_MAX) return fail(c, LUX_SGM_EUNSUPPORTED, "image too large");
    return 0;
}

int lux_sgm_upsample_layer_device(lux_sgm_ctx* c, const float* d_prev_depth, const float* d_prev_conf, int32_t pw, int32_t ph,
                                  const uint8_t* d_guide, int32_t width, int32_t height, int32_t set_w, int32_t set_h,
                                  float sigma, float* d_depth, float* d_conf, void* stream)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_upsample(c, pw, ph, width, height, set_w, set_h);
    if (rc) return rc;
    if (!d_prev_depth || !d_guide || !d_depth) return fail(c, LUX_SGM_EINVAL, "null argument");
    if (d_prev_conf && !d_conf) return fail(c, LUX_SGM_EINVAL, "a previous confidence needs a confidence output");
    (void)hipSetDevice(c->device);
    return lux::enqueue_upsample(c->ws, d_prev_depth, d_prev_conf, pw, ph, d_guide, width, height, sigma, d_depth, d_conf,
                                 stream ? (hipStream_t)stream : c->stream, &c->err);
}

int lux_sgm_upsample_layer(lux_sgm_ctx* c, const float* prev_depth, const float* prev_conf, int32_t pw, int32_t ph,
                           const uint8_t* guide, int32_t width, int32_t height, int32_t set_w, int32_t set_h, float sigma,
                           float* depth, float* conf)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_upsample(c, pw, ph, width, height, set_w, set_h);
    if (rc) return rc;
    if (!prev_depth || !guide || !depth) return fail(c, LUX_SGM_EINVAL, "null argument");
    if (prev_conf && !conf) return fail(c, LUX_SGM_EINVAL, "a previous confidence needs a confidence output");
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    const size_t n = (size_t)pw * ph, m = (size_t)width * height;
    void *dp, *dg, *dd, *dpc = nullptr, *dc = nullptr;
    if ((rc = c->ws.get("ups_prev", n * 4, &dp, &c->err)) || (rc = c->ws.get("ups_guide", m * 4, &dg, &c->err)) ||
        (rc = c->ws.get("ups_out", m * 4, &dd, &c->err)))
        return rc;
    if (prev_conf && ((rc = c->ws.get("ups_prev_conf", n * 4, &dpc, &c->err)) || (rc = c->ws.get("ups_conf", m * 4, &dc, &c->err))))
        return rc;
    hipError_t e = hipMemcpyAsync(dp, prev_depth, n * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dg, guide, m * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && prev_conf) e = hipMemcpyAsync(dpc, prev_conf, n * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(c, e, "upload");
    if ((rc = lux::enqueue_upsample(c->ws, (const float*)dp, (const float*)dpc, pw, ph, (const uint8_t*)dg, width, height, sigma,
                                    (float*)dd, (float*)dc, s, &c->err)))
        return rc;
    e = hipMemcpyAsync(depth, dd, m * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && prev_conf) e = hipMemcpyAsync(conf, dc, m * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : hip_fail(c, e, "download depth");
}

int lux_sgm_pyramid_bytes(const lux_sgm_pyramid_params* p, uint64_t* bytes)
{
    if (!p || !bytes) return LUX_SGM_EINVAL;
    std::vector<lux::BufferNeed> need;
    const size_t imgb = (size_t)p->base.height * p->base.stride * (p->base.bits > 8 ? 2 : 1);
    int rc = lux::pyramid_buffers(p, imgb, &need, nullptr);
    if (rc) return rc;
    uint64_t t = 0;
    for (auto& b : need) t += b.bytes;
    *bytes = t;
    return LUX_SGM_OK;
}

int lux_sgm_pyramid_device(lux_sgm_ctx* c, const lux_sgm_pyramid_params* p, const void* d_left, const void* d_right,
                           uint16_t* d_disp, float* d_sub, void* stream)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!d_left || !d_right) return fail(c, LUX_SGM_EINVAL, "null image");
    (void)hipSetDevice(c->device);
    const lux::Conc conc = lux::conc_of(c);
    return lux::enqueue_pyramid(c->ws, p, d_left, d_right, d_disp, d_sub, stream ? (hipStream_t)stream : c->stream,
                                &c->err, &conc);
}

int lux_sgm_pyramid(lux_sgm_ctx* c, const lux_sgm_pyramid_params* p, const void* left, const void* right, uint16_t* disp,
                    float* sub)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = lux::check_pyramid(p, &c->err);
    if (rc) return rc;
    if (!left || !right) return fail(c, LUX_SGM_EINVAL, "null image");
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    const size_t px = (size_t)p->base.width * p->base.height;
    const size_t imgb = (size_t)p->base.height * p->base.stride * (p->base.bits > 8 ? 2 : 1);
    lux::Workspace& ws = c->ws;
    void *dl, *dr, *dd, *ds;
    if ((rc = ws.get("img_l", imgb, &dl, &c->err)) || (rc = ws.get("img_r", imgb, &dr, &c->err)) ||
        (rc = ws.get("out_disp", px * 2, &dd, &c->err)) || (rc = ws.get("out_sub", px * 4, &ds, &c->err)))
        return rc;
    hipError_t e = hipMemcpyAsync(dl, left, imgb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, right, imgb, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(c, e, "upload images");
    const lux::Conc conc = lux::conc_of(c);
    if ((rc = lux::enqueue_pyramid(ws, p, dl, dr, (uint16_t*)dd, sub ? (float*)ds : nullptr, s, &c->err, &conc))) return rc;
    if (disp) e = hipMemcpyAsync(disp, dd, px * 2, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && sub) e = hipMemcpyAsync(sub, ds, px * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : hip_fail(c, e, "download results");
}

int lux_sgm_compute_layer(lux_sgm_ctx* c, const lux_sgm_layer_params* p, const uint8_t* raw, uint16_t* agg,
                          uint16_t* agg_pass1, uint16_t* wta)
{
    return lux_sgm_compute_layer_ex(c, p, nullptr, nullptr, nullptr, nullptr, raw, agg, agg_pass1, wta);
}

int lux_sgm_compute_layer_ranges(lux_sgm_ctx* c, const lux_sgm_layer_params* p, const uint16_t* start,
                                 const uint16_t* count, const uint8_t* raw, uint16_t* agg, uint16_t* agg_pass1,
                                 uint16_t* wta)
{
    if (!start || !count) return fail(c, LUX_SGM_EINVAL, "null ranges");
    return lux_sgm_compute_layer_ex(c, p, start, count, nullptr, nullptr, raw, agg, agg_pass1, wta);
}

int lux_sgm_compute_layer_ex(lux_sgm_ctx* c, const lux_sgm_layer_params* p, const uint16_t* start, const uint16_t* count,
                             const uint8_t* skip, const lux_sgm_guidance* gd, const uint8_t* raw, uint16_t* agg,
                             uint16_t* agg_pass1, uint16_t* wta)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!raw) return fail(c, LUX_SGM_EINVAL, "null argument");
    if ((start == nullptr) != (count == nullptr)) return fail(c, LUX_SGM_EINVAL, "start and count go together");
    if (gd && !gd->guidance) return fail(c, LUX_SGM_EINVAL, "null guidance image");
    int rc = lux::check_layer(p, &c->err);
    if (rc) return rc;
    const bool ranged = start != nullptr;
    const int W = p->width, H = p->height, align = p->align;
    const size_t cells = (size_t)W * H;
    // ranges as DenseLayer.Init leaves them (:53-58): cap = roundup(count, align); MinLo = min start,
    // MaxHi = roundup(max(start + count), align) (:46, :51, :136). Uniform: start 0, count = planes.
    std::vector<uint16_t> cap;
    std::vector<uint64_t> off;
    uint64_t total = 0;
    int maxcount = p->planes;
    bool spill = false;  // some range's padding passes its line-buffer segment (start - MinLo + cap > MaxHi)
    int lminlo = 0, lmaxhi = 0, lmaxcap = 0;
    if (ranged) {
        cap.resize(cells);
        off.resize(cells);
        int minlo = 1 << 30, maxhi = 0, maxcap = 0;
        maxcount = 0;
        for (size_t i = 0; i < cells; i++) {
            if ((int)start[i] + (int)count[i] > p->planes) return fail(c, LUX_SGM_EINVAL, "a range passes the last plane");
            const int cp = ((count[i] + align - 1) / align) * align;
            cap[i] = (uint16_t)cp;
            off[i] = total;
            total += (uint64_t)cp;
            minlo = std::min(minlo, (int)start[i]);
            maxhi = std::max(maxhi, (int)start[i] + (int)count[i]);
            maxcap = std::max(maxcap, cp);
            maxcount = std::max(maxcount, (int)count[i]);
        }
        maxhi = ((maxhi + align - 1) / align) * align;
        if (maxcap == 0) return fail(c, LUX_SGM_EINVAL, "every range is empty");
        if (maxcap > 1024) return fail(c, LUX_SGM_EUNSUPPORTED, "per-pixel ranges above 1024 lanes");
        for (size_t i = 0; i < cells; i++)
            if ((int)start[i] - minlo + (int)cap[i] > maxhi) spill = true;
        lminlo = minlo;
        lmaxhi = maxhi;
        lmaxcap = maxcap;
    }
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    lux::Geometry g = lux::geometry(W, H, maxcount, align);
    if (!ranged) {
        total = cells * g.cap;
        lmaxhi = lmaxcap = g.cap;
    }
    lux::Workspace& ws = c->ws;
    // The path sweeps take even tiles, caps up to 512 with ranges and ranges whose padding stays inside its
    // line-buffer segment; the line-buffer-literal path takes the rest up to 1024 lanes (and everything when forced).
    const bool literal = lmaxcap <= 1024 && (ws.force_literal || spill || (p->tile & 1) || (ranged && lmaxcap > 512));
    lux::Prepared w;
    if (literal) rc = lux::prepare_literal(ws, g, &w, &c->err);
    else rc = lux::prepare(ws, g, p->tile, p->xs_clamp_w, false, s, &w, &c->err);
    if (rc) return rc;
    void *stage, *a1 = nullptr, *dw;
    if ((rc = ws.get("layer_stage", std::max<size_t>(total * 2, 2), &stage, &c->err)) ||
        (rc = ws.get("layer_wta", cells * 2, &dw, &c->err)))
        return rc;
    if (agg_pass1 && (rc = ws.get("layer_pass1", cells * g.cs * 2, &a1, &c->err))) return rc;
    hipError_t e = hipSuccess;
    void* vo = nullptr;
    if (ranged) {
        void *vs, *vc, *vn, *ss, *sc;
        if ((rc = ws.get("vstart", cells * 2, &vs, &c->err)) || (rc = ws.get("vcap", cells * 2, &vc, &c->err)) ||
            (rc = ws.get("vcount", cells * 2, &vn, &c->err)) || (rc = ws.get("seedstart", (size_t)3 * W * 2, &ss, &c->err)) ||
            (rc = ws.get("seedcap", (size_t)3 * W * 2, &sc, &c->err)) || (rc = ws.get("voff", cells * 8, &vo, &c->err)))
            return rc;
        e = hipMemcpyAsync(vs, start, cells * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(vc, cap.data(), cells * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(vn, count, cells * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(vo, off.data(), cells * 8, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && total) e = hipMemcpyAsync(stage, raw, total, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = lux::launch_unpack_u8((const uint8_t*)stage, (const uint64_t*)vo, (const uint16_t*)vc, w.raw, g.cs, cells, s);
        w.b.vstart = (const uint16_t*)vs;
        w.b.vcap = (const uint16_t*)vc;
        w.b.vcount = (const uint16_t*)vn;
        w.b.seedstart = (uint16_t*)ss;
        w.b.seedcap = (uint16_t*)sc;
    } else {
        e = hipMemcpyAsync(stage, raw, cells * g.cap, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = lux::launch_restride_u8((const uint8_t*)stage, g.cap, w.raw, g.cs, (int)cells, g.cap, 255, s);
    }
    if (e == hipSuccess && skip) {  // skipped pixels: raw 0 in range, 255 padding, whatever the caller's bytes
        void* sk;
        if ((rc = ws.get("skip", cells, &sk, &c->err))) return rc;
        e = hipMemcpyAsync(sk, skip, cells, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = lux::launch_apply_skip(w.raw, (const uint8_t*)sk, w.b.vcount, maxcount, g.cs, cells, s);
    }
    if (e != hipSuccess) return hip_fail(c, e, "upload raw");
    if (gd) {  // adaptive P2: the guidance image and the seed table's guidance
        void *gi, *sg;
        if ((rc = ws.get("guidance", cells * 4, &gi, &c->err)) ||
            (rc = ws.get("seedguid", (size_t)3 * W * 4, &sg, &c->err)))
            return rc;
        e = hipMemcpyAsync(gi, gd->guidance, cells * 4, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_fail(c, e, "upload guidance");
        w.b.ap.guid = (const uint32_t*)gi;
        w.b.ap.w0 = gd->weights[0];
        w.b.ap.w1 = gd->weights[1];
        w.b.ap.w2 = gd->weights[2];
        w.b.ap.base = (float)p->p1 * gd->p2_scale;  // p2Base (:253), one float multiply
        w.b.ap.on = 1;
        w.b.seedguid = (uint32_t*)sg;
    }
    if (literal) {
        lux::LiteralLayer l;
        l.W = W;
        l.H = H;
        l.T = p->tile;
        l.align = align;
        l.xsw = p->xs_clamp_w;
        l.minlo = lminlo;
        l.maxhi = lmaxhi;
        l.start = w.b.vstart;
        l.cap = w.b.vcap;
        l.count = w.b.vcount;
        l.ucap = g.cap;
        l.ucount = p->planes;
        l.raw = w.raw;
        l.agg = w.agg;
        l.cs = g.cs;
        l.ap = w.b.ap;
        l.p1 = p->p1;
        l.p2 = p->p2;
        if ((rc = lux::run_literal(ws, l, lmaxcap, (uint16_t*)a1, (uint16_t*)dw, s, &c->err))) return rc;
    } else {
        const lux::Conc conc = lux::conc_of(c);
        if (ranged) lux::try_path_volumes(ws, &w);  // ranged layers are latency-bound: one launch per pass
        lux::attach_chain(&w, &conc);
        lux::SgmOutputs o;
        o.disp = (uint16_t*)dw;
        o.tol = 0;
        o.write_agg = agg != nullptr;
        o.bands = ranged || gd ? 0 : c->opt_bands;
        e = lux::launch_sgm(g, w.b, o, p->p1, p->p2, (uint16_t*)a1, s);
        if (e != hipSuccess) return hip_fail(c, e, "aggregation");
    }
    if (wta) {
        e = hipMemcpyAsync(wta, dw, cells * 2, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return hip_fail(c, e, "wta");
    }
    uint16_t* outs[2] = {agg, agg_pass1};
    const uint16_t* srcs[2] = {w.agg, (const uint16_t*)a1};
    for (int k = 0; k < 2; k++) {
        if (!outs[k] || !total) continue;
        if (ranged) e = lux::launch_pack_u16(srcs[k], (const uint64_t*)vo, w.b.vcap, (uint16_t*)stage, g.cs, cells, s);
        else e = lux::launch_restride_u16(srcs[k], g.cs, (uint16_t*)stage, g.cap, (int)cells, g.cap, s);
        if (e == hipSuccess) e = hipMemcpyAsync(outs[k], stage, total * 2, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(c, e, "download aggregate");
    }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "compute_layer");
    return literal ? 0 : lux::check_band_error(ws, s, &c->err);
}

int lux_sgm_set_option(lux_sgm_ctx* c, int32_t option, int32_t value)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    switch (option) {
        case LUX_SGM_OPT_FUSED_BANDS: c->opt_bands = value != 0; return LUX_SGM_OK;
        case LUX_SGM_OPT_LITERAL: c->ws.force_literal = value != 0; return LUX_SGM_OK;
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

int lux_sgm_batch(lux_sgm_ctx* c, int32_t n, const lux_sgm_pair* pairs, void* stream)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (n <= 0 || !pairs) return fail(c, LUX_SGM_EINVAL, "empty batch");
    (void)hipSetDevice(c->device);
    auto t = c->tuned.find(shape_key(n, pairs));
    const lux_sgm_schedule sch = t != c->tuned.end() ? t->second : default_schedule(n);
    hipGraphExec_t exec;
    int rc = batch_graph(c, n, pairs, sch, &exec);
    if (rc) return rc;
    hipError_t e = hipGraphLaunch(exec, stream ? (hipStream_t)stream : c->stream);
    return e == hipSuccess ? 0 : hip_fail(c, e, "hipGraphLaunch");
}

int lux_sgm_batch_set_schedule(lux_sgm_ctx* c, int32_t n, const lux_sgm_pair* pairs, const lux_sgm_schedule* sch)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (n <= 0 || !pairs || !sch) return fail(c, LUX_SGM_EINVAL, "empty batch or null schedule");
    if (sch->executor != 0 && sch->executor != 1 && sch->executor != 7)
        return fail(c, LUX_SGM_EINVAL, "executor must be 0, 1 or 7");
    if (sch->executor)
        for (int i = 0; i < n; i++) {
            int rc = lux::check_rect(&pairs[i].params, &c->err);
            if (rc) return rc;
            if (!lux::exec_params_supported(&pairs[i].params))
                return fail(c, LUX_SGM_EUNSUPPORTED, "the executor takes census pairs whose path costs fit a byte");
        }
    c->tuned[shape_key(n, pairs)] = *sch;
    return 0;
}

int lux_sgm_batch_tune(lux_sgm_ctx* c, int32_t n, const lux_sgm_pair* pairs, void* stream, int32_t reps,
                       lux_sgm_schedule* best, lux_sgm_schedule* tried, int32_t cap)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (n <= 0 || !pairs) return fail(c, LUX_SGM_EINVAL, "empty batch");
    if (reps < 1) reps = 1;
    (void)hipSetDevice(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    bool bands_ok = true;  // the fused bands take up to 256 planes
    for (int i = 0; i < n; i++)
        if (lux::geometry(pairs[i].params.width, pairs[i].params.height, pairs[i].params.disparities,
                          pairs[i].params.align).dpl > 4)
            bands_ok = false;
    std::vector<lux_sgm_schedule> cand;
    for (int bands = 0; bands <= (bands_ok ? 1 : 0); bands++)
        for (int b : {1, 2, 3, 4, 6, 8}) {
            if (b > n) continue;
            lux_sgm_schedule sc = default_schedule(n);
            sc.bands = bands;
            sc.branches = b;
            cand.push_back(sc);
        }
    // the persistent executor, when it takes every pair: one launch or one per pass group, two item orders
    bool exec_ok = true;
    for (int i = 0; i < n && exec_ok; i++) exec_ok = lux::exec_params_supported(&pairs[i].params);
    if (exec_ok)
        for (int groups : {1, 7})
            for (int order : {0, 1}) {
                lux_sgm_schedule sc = default_schedule(n);
                sc.executor = groups;
                sc.order = order;
                cand.push_back(sc);
            }
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) return hip_fail(c, e, "hipEventCreate");
    int rc = 0, ntried = 0;
    lux_sgm_schedule win = cand[0];
    win.ms = -1.f;
    for (auto& sc : cand) {
        hipGraphExec_t exec;
        GraphEntry* ge = nullptr;
        if ((rc = batch_graph(c, n, pairs, sc, &exec, &ge))) break;
        e = hipGraphLaunch(exec, s);  // warm
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; e == hipSuccess && r < reps; r++) e = hipGraphLaunch(exec, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            rc = hip_fail(c, e, "tune batch");
            break;
        }
        for (int i = 0; i < n && !rc; i++) rc = lux::check_band_error(*c->bws[i], s, &c->err);
        if (!rc) rc = batch_exec_check(c, ge, s);
        if (rc) break;
        sc.ms = ms / reps;
        if (tried && ntried < cap) tried[ntried] = sc;
        ntried++;
        if (win.ms < 0.f || sc.ms < win.ms) win = sc;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    c->tuned[shape_key(n, pairs)] = win;
    if (best) *best = win;
    return ntried;
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
