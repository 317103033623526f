// The rectified request without device allocation (lux_sgm_rectified_device_noalloc): for callers that cannot call
// hipMalloc (code under graph capture, serving loops that forbid it). Its buffers are carved out of one file-scope
// device array of kScratchBytes instead of the context's pooled workspace; the kernels and the enqueue are the ones
// lux_sgm_rectified_device runs (in their u16-aggregate form), so the results are the same bit for bit.
//
// The array exists once per device the library runs on, for the life of the process. Every call on a device uses
// that device's copy and leaves its contents behind: what has to start at zero (the band flags) is cleared by a fill
// kernel in each call, and the visit lists of the border replays are uploaded in each call. Calls that overlap on one
// device race on the array; a caller serialises them (one stream, or one call at a time).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/lux_sgm.h"
#include "ctx.h"
#include "engine.h"

namespace {

constexpr size_t kScratchBytes = (size_t)96 << 20;  // 96 MiB: a 640x480 D=64 census request takes about 68 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

namespace lux {

// The device address of this device's scratch (the symbol's name is a host-side stand-in, not a device address).
// Called with `device` current: at context creation, and again by a call on a device whose address is not known yet.
int noalloc_setup(int device, std::string* err)
{
    if (device < 0 || device >= kMaxDevices) {
        if (err) *err = "device index above 63";
        return LUX_SGM_EUNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hip_status(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address", err);
    g_scratch[device] = ptr;
    return 0;
}

}  // namespace lux

int lux_sgm_rectified_device_noalloc(lux_sgm_ctx* c, const lux_sgm_rect_params* p, const void* d_left, const void* d_right,
                                     uint16_t* d_disp, float* d_sub, uint8_t* d_valid, uint16_t* d_agg, void* stream)
{
    if (!c) return LUX_SGM_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = lux::check_rect(p, &c->err);
    if (rc) return rc;
    if (!d_left || !d_right) return lux::fail(c, LUX_SGM_EINVAL, "null image");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return lux::hip_fail(c, e, "hipSetDevice");
    if ((rc = lux::noalloc_setup(c->device, &c->err))) return rc;
    void* base;
    {
        std::lock_guard<std::mutex> g(g_mu);
        base = g_scratch[c->device];
    }
    // what the request's buffers take, each placed on a 256-byte boundary as the arena places them
    const lux::Geometry g = lux::geometry(p->width, p->height, p->disparities, p->align);
    std::shared_ptr<lux::Plan> plan;
    if ((rc = lux::plan_for(g.W, g.H, p->tile, p->xs_clamp_w, &plan, &c->err))) return rc;
    // the u16-aggregate form (3 bytes a cell), not the byte path volumes (8): the scratch holds more requests that way,
    // and both give the same results
    const bool census = p->cost == LUX_SGM_COST_CENSUS, pv8 = false;
    std::vector<lux::BufferNeed> bufs = lux::request_buffers(g, *plan, census, 0, pv8);
    size_t need = 0;
    for (const lux::BufferNeed& b : bufs) need = ((need + 255) & ~(size_t)255) + b.bytes;
    if (need > kScratchBytes)
        return lux::fail(c, LUX_SGM_ENOMEM,
                         "request needs " + std::to_string(need) + " bytes of scratch; the no-allocation entry point holds " +
                             std::to_string(kScratchBytes));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    lux::Workspace ws;  // a fresh arena workspace: the visit lists are uploaded and the band flags cleared below
    ws.device = c->device;
    ws.arena = (uint8_t*)base;
    ws.arena_cap = kScratchBytes;
    lux::Prepared w;
    if ((rc = lux::prepare(ws, g, p->tile, p->xs_clamp_w, census, s, &w, &c->err, pv8))) return rc;
    // no side stream: creating one is not this entry point's to do; the border chains run in order on s
    lux::RectOut o;
    o.disp = d_disp;
    o.sub = d_sub;
    o.valid = d_valid;
    o.agg = d_agg;
    lux::EngineOpts opt;
    opt.bands = c->opt_bands;
    rc = lux::enqueue_rect(ws, w, p, d_left, d_right, o, opt, s, nullptr, &c->err);
    if (rc) return rc;
    // the visit lists were copied from host memory the plan cache owns; the copies are complete once s reaches here,
    // which the caller's synchronisation of s covers (the cache keeps the plan alive for the process)
    return LUX_SGM_OK;
}
