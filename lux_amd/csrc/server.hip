// Served form of liblux_sgm.so (declared in include/lux_sgm.h): a request loop over concurrent device slots.
//
// One dispatcher thread admits queued requests in arrival order while a slot is free and the request's
// workspace fits in what the other slots leave of the HBM budget. Like the reference's job queue
// (Lux.Library/Core/JobQueue.cs:23-29), a heavy request that does not fit holds back every heavy request
// behind it, a light one (workspace <= 1.5 GB) may pass it, and with nothing running the head always starts.
// Each slot is a non-blocking stream with a pooled workspace; a request's copies and kernels go onto its
// slot's stream and end in a host callback that records the completion time and wakes the dispatcher.
// The dispatcher never holds the server lock across a HIP call, so the callbacks can always run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lux_sgm.h"
#include "dloop.h"
#include "engine.h"
#include "pinned.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint64_t kLightBytes = 1536ull << 20;  // 1.5 GB: a light request may pass a heavy one that waits
using lux::PinnedPool;

struct Request {
    uint64_t ticket = 0;
    lux_sgm_rect_params p{};       // the pair (for a coarse-to-fine request: its full-resolution level)
    bool pyramid = false;
    lux_sgm_pyramid_params pp{};
    int mem = LUX_SGM_MEM_HOST;
    const void* left = nullptr;   // device mode
    const void* right = nullptr;
    uint16_t* disp = nullptr;     // caller's outputs (host or device per mem)
    float* sub = nullptr;
    uint8_t* valid = nullptr;
    void* stage_in = nullptr;     // host mode: pinned [left | right]
    size_t stage_in_n = 0;
    void* stage_out = nullptr;    // host mode: pinned [disp | sub | valid]
    size_t stage_out_n = 0;
    std::vector<lux::BufferNeed> bufs;  // the workspace buffers it uses
    uint64_t need = 0;                  // their total bytes
    int state = LUX_SGM_REQ_PENDING;
    int status = 0;
    std::string err;
    int slot = -1;
    Clock::time_point t_submit, t_admit, t_done;
};

struct Slot {
    hipStream_t stream = nullptr;
    lux::Workspace ws;
    std::shared_ptr<Request> req;  // running request (null: free)
    bool completed = false;        // set by the host callback
    Clock::time_point t_done;
    uint64_t held = 0;             // ws.held(), kept under the server lock for readers off the dispatcher
};

struct CallbackArg;

}  // namespace

struct lux_sgm_server {
    int device = 0;
    lux::DLoop* dloop = nullptr;  // LUX_SGM_SERVER_DEVICE_LOOP: the device-resident loop serves every call
    lux::Budget budget;
    std::vector<std::unique_ptr<Slot>> slots;
    std::deque<std::shared_ptr<Request>> pending;
    std::map<uint64_t, std::shared_ptr<Request>> reqs;
    std::mutex mu;
    std::condition_variable wake, done;
    std::thread thr;
    bool stop = false;
    uint64_t next_ticket = 1;
    uint64_t submitted = 0, completed = 0, failed = 0;
    PinnedPool pinned;
    std::mutex err_mu;
    std::string err;
};

namespace {

thread_local std::string g_server_create_error;

struct CallbackArg {
    lux_sgm_server* srv;
    int slot;
};

void set_err(lux_sgm_server* s, const std::string& m)
{
    std::lock_guard<std::mutex> lk(s->err_mu);
    s->err = m;
}

int fail(lux_sgm_server* s, int code, const std::string& m)
{
    set_err(s, m);
    return code;
}

void on_complete(void* arg)
{
    auto* a = (CallbackArg*)arg;
    lux_sgm_server* s = a->srv;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        Slot& sl = *s->slots[a->slot];
        sl.completed = true;
        sl.t_done = Clock::now();
    }
    s->wake.notify_all();
    delete a;
}

size_t image_bytes(const lux_sgm_rect_params& p) { return (size_t)p.height * p.stride * (p.bits > 8 ? 2 : 1); }

// Enqueue request r on slot k (called without the server lock). Returns a status; on success the slot's
// stream ends in the completion callback.
int launch(lux_sgm_server* s, int k, Request& r)
{
    Slot& sl = *s->slots[k];
    (void)hipSetDevice(s->device);
    hipStream_t st = sl.stream;
    const lux_sgm_rect_params& p = r.p;
    const size_t px = (size_t)p.width * p.height, imgb = image_bytes(p);
    lux::Workspace& ws = sl.ws;
    std::string& err = r.err;
    const void *dl = r.left, *dr = r.right;
    lux::RectOut o;
    void *od = nullptr, *os = nullptr, *ov = nullptr;
    int rc = 0;
    if (r.mem == LUX_SGM_MEM_HOST) {
        void *il, *ir;
        if ((rc = ws.get("img_l", imgb, &il, &err)) || (rc = ws.get("img_r", imgb, &ir, &err))) return rc;
        if ((rc = ws.get("out_disp", px * 2, &od, &err)) || (rc = ws.get("out_sub", px * 4, &os, &err)) ||
            (rc = ws.get("out_valid", px, &ov, &err)))
            return rc;
        hipError_t e = hipMemcpyAsync(il, r.stage_in, imgb, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(ir, (char*)r.stage_in + imgb, imgb, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return lux::hip_status(e, "upload images", &err);
        dl = il;
        dr = ir;
        o.disp = r.disp ? (uint16_t*)od : nullptr;
        o.sub = r.sub ? (float*)os : nullptr;
        o.valid = r.valid ? (uint8_t*)ov : nullptr;
    } else {
        o.disp = r.disp;
        o.sub = r.sub;
        o.valid = r.valid;
    }
    if (r.pyramid) {
        if ((rc = lux::enqueue_pyramid(ws, &r.pp, dl, dr, o.disp, o.sub, st, &err))) return rc;
    } else {
        lux::Geometry g = lux::geometry(p.width, p.height, p.disparities, p.align);
        lux::Prepared w;
        if ((rc = lux::prepare(ws, g, p.tile, p.xs_clamp_w, p.cost == LUX_SGM_COST_CENSUS, st, &w, &err,
                               lux::path_bytes(&p))))
            return rc;
        lux::EngineOpts opt;
        opt.shared = true;  // the other slots run beside this request
        if ((rc = lux::enqueue_rect(ws, w, &p, dl, dr, o, opt, st, nullptr, &err))) return rc;
    }
    if (r.mem == LUX_SGM_MEM_HOST) {
        char* so = (char*)r.stage_out;
        hipError_t e = hipSuccess;
        if (r.disp) e = hipMemcpyAsync(so, od, px * 2, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && r.sub) e = hipMemcpyAsync(so + px * 2, os, px * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && r.valid && !r.pyramid) e = hipMemcpyAsync(so + px * 6, ov, px, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return lux::hip_status(e, "download results", &err);
    }
    auto* a = new CallbackArg{s, k};
    hipError_t e = hipLaunchHostFunc(st, on_complete, a);
    if (e != hipSuccess) {
        delete a;
        return lux::hip_status(e, "hipLaunchHostFunc", &err);
    }
    return 0;
}

// Pick a slot for request r (dispatcher thread, under the lock): the free slot whose workspace needs the least
// growth to hold r's buffers, provided the budget covers that growth once idle workspaces are released (the
// largest first, and the chosen slot's own last). Returns the slot and, in evict, the slots to release first;
// -1 if r does not fit now.
int pick_slot(lux_sgm_server* s, const Request& r, std::vector<int>* evict)
{
    evict->clear();
    int best = -1;
    uint64_t bg = 0;
    for (int k = 0; k < (int)s->slots.size(); k++) {
        Slot& sl = *s->slots[k];
        if (sl.req) continue;
        const uint64_t g = lux::growth(sl.ws, r.bufs);
        if (best < 0 || g < bg) {
            best = k;
            bg = g;
        }
    }
    if (best < 0 || !s->budget.limit) return best;
    uint64_t used = 0;
    for (auto& sl : s->slots) used += sl->ws.held();
    if (used + bg <= s->budget.limit) return best;
    std::vector<std::pair<uint64_t, int>> idle;
    for (int k = 0; k < (int)s->slots.size(); k++)
        if (k != best && !s->slots[k]->req && s->slots[k]->ws.held()) idle.push_back({s->slots[k]->ws.held(), k});
    std::sort(idle.rbegin(), idle.rend());
    for (auto& kv : idle) {
        used -= kv.first;
        evict->push_back(kv.second);
        if (used + bg <= s->budget.limit) return best;
    }
    // release the chosen slot's own buffers too: it then grows from nothing to r's total
    if (used - s->slots[best]->ws.held() + r.need <= s->budget.limit) {
        evict->push_back(best);
        return best;
    }
    evict->clear();
    return -1;
}

void dispatcher(lux_sgm_server* s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // completions
        for (auto& slp : s->slots) {
            Slot& sl = *slp;
            if (!sl.req || !sl.completed) continue;
            Request& r = *sl.req;
            r.t_done = sl.t_done;
            r.state = LUX_SGM_REQ_DONE;
            s->completed++;
            sl.req.reset();
            sl.completed = false;
            s->done.notify_all();
        }
        if (s->stop && s->pending.empty()) {
            bool busy = false;
            for (auto& sl : s->slots) busy |= (bool)sl->req;
            if (!busy) break;
        }
        // admission, in arrival order; light requests may pass a heavy one that does not fit
        bool launched = false;
        bool heavy_blocked = false;
        for (auto it = s->pending.begin(); it != s->pending.end();) {
            std::shared_ptr<Request> r = *it;
            const bool heavy = r->need > kLightBytes;
            if (heavy && heavy_blocked) {
                ++it;
                continue;
            }
            std::vector<int> evict;
            int k = pick_slot(s, *r, &evict);
            if (k < 0) {
                if (heavy) heavy_blocked = true;
                ++it;
                continue;
            }
            it = s->pending.erase(it);
            Slot& sl = *s->slots[k];
            sl.req = r;
            sl.completed = false;
            r->slot = k;
            r->state = LUX_SGM_REQ_RUNNING;
            r->t_admit = Clock::now();
            std::vector<Slot*> ev;
            for (int e : evict) ev.push_back(s->slots[e].get());
            for (Slot* e : ev)
                if (e != &sl) e->req = std::make_shared<Request>();  // keep slots being released out of use
            lk.unlock();
            for (Slot* e : ev) e->ws.release();
            int rc = launch(s, k, *r);
            lk.lock();
            for (Slot* e : ev) {
                e->held = e->ws.held();
                if (e != &sl) e->req.reset();
            }
            sl.held = sl.ws.held();
            if (rc) {
                // nothing of the request runs on the device beyond what was enqueued: drain the stream
                lk.unlock();
                (void)hipStreamSynchronize(sl.stream);
                lk.lock();
                r->status = rc;
                r->state = LUX_SGM_REQ_DONE;
                r->t_done = Clock::now();
                s->failed++;
                sl.req.reset();
                s->done.notify_all();
            }
            launched = true;
            it = s->pending.begin();  // the lock was released: start again from the head
            heavy_blocked = false;
        }
        if (launched) continue;
        s->wake.wait(lk);
    }
}

int submit(lux_sgm_server* s, std::shared_ptr<Request> r, const void* left, const void* right, uint64_t* ticket);

}  // namespace

extern "C" {

int lux_sgm_server_create_ex(int device, uint64_t hbm_budget, int32_t nslots, uint32_t flags, lux_sgm_server** out)
{
    if (!(flags & LUX_SGM_SERVER_DEVICE_LOOP)) return lux_sgm_server_create(device, hbm_budget, nslots, out);
    if (!out) {
        g_server_create_error = "null out pointer";
        return LUX_SGM_EINVAL;
    }
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        g_server_create_error = std::string("no HIP device: ") + hipGetErrorString(e);
        return LUX_SGM_EDEVICE;
    }
    if (device < 0 || device >= n) {
        g_server_create_error = "device index out of range";
        return LUX_SGM_EINVAL;
    }
    lux::DLoop* d = nullptr;
    int rc = lux::dloop_create(device, hbm_budget, nslots, &d, &g_server_create_error);
    if (rc) return rc;
    auto* s = new lux_sgm_server();
    s->device = device;
    s->budget.limit = hbm_budget;
    s->dloop = d;
    *out = s;
    return LUX_SGM_OK;
}

int lux_sgm_server_create(int device, uint64_t hbm_budget, int32_t nslots, lux_sgm_server** out)
{
    if (!out) {
        g_server_create_error = "null out pointer";
        return LUX_SGM_EINVAL;
    }
    *out = nullptr;
    if (nslots < 1 || nslots > 16) {
        g_server_create_error = "slots must be 1..16";
        return LUX_SGM_EINVAL;
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        g_server_create_error = std::string("no HIP device: ") + hipGetErrorString(e);
        return LUX_SGM_EDEVICE;
    }
    if (device < 0 || device >= n) {
        g_server_create_error = "device index out of range";
        return LUX_SGM_EINVAL;
    }
    if ((e = hipSetDevice(device)) != hipSuccess) {
        g_server_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return LUX_SGM_EDEVICE;
    }
    auto* s = new lux_sgm_server();
    s->device = device;
    s->budget.limit = hbm_budget;
    for (int k = 0; k < nslots; k++) {
        auto sl = std::make_unique<Slot>();
        sl->ws.device = device;
        sl->ws.budget = &s->budget;
        if ((e = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking)) != hipSuccess) {
            g_server_create_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
            for (auto& x : s->slots) (void)hipStreamDestroy(x->stream);
            delete s;
            return LUX_SGM_EDEVICE;
        }
        s->slots.push_back(std::move(sl));
    }
    s->thr = std::thread(dispatcher, s);
    *out = s;
    return LUX_SGM_OK;
}

int lux_sgm_server_destroy(lux_sgm_server* s)
{
    if (!s) return LUX_SGM_EINVAL;
    if (s->dloop) {
        lux::dloop_destroy(s->dloop);
        delete s;
        return LUX_SGM_OK;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto& r : s->pending) {
            r->status = LUX_SGM_ECANCELED;
            r->state = LUX_SGM_REQ_DONE;
        }
        s->pending.clear();
        s->stop = true;
    }
    s->wake.notify_all();
    s->done.notify_all();
    if (s->thr.joinable()) s->thr.join();
    (void)hipSetDevice(s->device);
    for (auto& sl : s->slots) {
        (void)hipStreamSynchronize(sl->stream);
        sl->ws.release();
        (void)hipStreamDestroy(sl->stream);
    }
    for (auto& kv : s->reqs) {
        s->pinned.put(kv.second->stage_in, kv.second->stage_in_n);
        s->pinned.put(kv.second->stage_out, kv.second->stage_out_n);
    }
    s->pinned.clear();
    delete s;
    return LUX_SGM_OK;
}

const char* lux_sgm_server_last_error(const lux_sgm_server* s)
{
    if (!s) return g_server_create_error.c_str();
    return s->err.c_str();
}

int lux_sgm_server_submit(lux_sgm_server* s, const lux_sgm_rect_params* p, int32_t mem, const void* left,
                          const void* right, uint16_t* disp, float* sub, uint8_t* valid, uint64_t* ticket)
{
    if (!s || !ticket) return LUX_SGM_EINVAL;
    std::string err;
    if (s->dloop) {
        int rc = lux::dloop_submit(s->dloop, p, mem, left, right, disp, sub, valid, ticket, &err);
        return rc ? fail(s, rc, err) : LUX_SGM_OK;
    }
    int rc = lux::check_rect(p, &err);
    if (rc) return fail(s, rc, err);
    if (mem != LUX_SGM_MEM_HOST && mem != LUX_SGM_MEM_DEVICE) return fail(s, LUX_SGM_EINVAL, "unknown memory kind");
    auto r = std::make_shared<Request>();
    r->p = *p;
    r->mem = mem;
    r->disp = disp;
    r->sub = sub;
    r->valid = valid;
    const size_t imgb = image_bytes(*p);
    lux::Geometry g = lux::geometry(p->width, p->height, p->disparities, p->align);
    std::shared_ptr<lux::Plan> plan;
    if ((rc = lux::plan_for(g.W, g.H, p->tile, p->xs_clamp_w, &plan, &err))) return fail(s, rc, err);
    r->bufs = lux::request_buffers(g, *plan, p->cost == LUX_SGM_COST_CENSUS, mem == LUX_SGM_MEM_HOST ? imgb : 0,
                                   lux::path_bytes(p));
    return submit(s, r, left, right, ticket);
}

int lux_sgm_server_submit_pyramid(lux_sgm_server* s, const lux_sgm_pyramid_params* p, int32_t mem, const void* left,
                                  const void* right, uint16_t* disp, float* sub, uint64_t* ticket)
{
    if (!s || !ticket) return LUX_SGM_EINVAL;
    std::string err;
    if (s->dloop) return fail(s, LUX_SGM_EUNSUPPORTED, "coarse-to-fine requests are the host dispatcher's");
    int rc = lux::check_pyramid(p, &err);
    if (rc) return fail(s, rc, err);
    if (mem != LUX_SGM_MEM_HOST && mem != LUX_SGM_MEM_DEVICE) return fail(s, LUX_SGM_EINVAL, "unknown memory kind");
    auto r = std::make_shared<Request>();
    r->p = p->base;
    r->pyramid = true;
    r->pp = *p;
    r->mem = mem;
    r->disp = disp;
    r->sub = sub;
    if ((rc = lux::pyramid_buffers(p, mem == LUX_SGM_MEM_HOST ? image_bytes(p->base) : 0, &r->bufs, &err)))
        return fail(s, rc, err);
    if (mem == LUX_SGM_MEM_HOST) {  // the host form stages outputs in out_disp / out_sub (valid unused)
        const size_t px = (size_t)p->base.width * p->base.height;
        r->bufs.push_back({"out_valid", px});
    }
    return submit(s, r, left, right, ticket);
}

}  // extern "C"

namespace {

int submit(lux_sgm_server* s, std::shared_ptr<Request> r, const void* left, const void* right, uint64_t* ticket)
{
    int rc = 0;
    const lux_sgm_rect_params* p = &r->p;
    const int mem = r->mem;
    if (!left || !right) return fail(s, LUX_SGM_EINVAL, "null image");
    const size_t px = (size_t)p->width * p->height, imgb = image_bytes(*p);
    for (const auto& b : r->bufs) r->need += b.bytes;
    if (s->budget.limit && r->need > s->budget.limit)
        return fail(s, LUX_SGM_ENOMEM, "request larger than the server's HBM budget");
    if (mem == LUX_SGM_MEM_HOST) {
        (void)hipSetDevice(s->device);
        r->stage_in = s->pinned.get(2 * imgb, &r->stage_in_n);
        r->stage_out = s->pinned.get(px * 7, &r->stage_out_n);
        if (!r->stage_in || !r->stage_out) {
            s->pinned.put(r->stage_in, r->stage_in_n);
            s->pinned.put(r->stage_out, r->stage_out_n);
            return fail(s, LUX_SGM_ENOMEM, "pinned staging");
        }
        std::memcpy(r->stage_in, left, imgb);
        std::memcpy((char*)r->stage_in + imgb, right, imgb);
    } else {
        r->left = left;
        r->right = right;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return LUX_SGM_ECANCELED;
        r->ticket = s->next_ticket++;
        r->t_submit = Clock::now();
        s->reqs[r->ticket] = r;
        s->pending.push_back(r);
        s->submitted++;
        *ticket = r->ticket;
    }
    s->wake.notify_all();
    (void)rc;
    return LUX_SGM_OK;
}

}  // namespace

extern "C" {

int lux_sgm_server_poll(lux_sgm_server* s, uint64_t ticket, int32_t* state)
{
    if (!s || !state) return LUX_SGM_EINVAL;
    if (s->dloop) {
        std::string err;
        int rc = lux::dloop_poll(s->dloop, ticket, state, &err);
        return rc ? fail(s, rc, err) : LUX_SGM_OK;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(s, LUX_SGM_EINVAL, "unknown ticket");
    *state = it->second->state;
    return LUX_SGM_OK;
}

int lux_sgm_server_wait(lux_sgm_server* s, uint64_t ticket, int32_t timeout_ms, lux_sgm_request_stats* stats)
{
    if (!s) return LUX_SGM_EINVAL;
    if (s->dloop) {
        std::string err;
        int rc = lux::dloop_wait(s->dloop, ticket, timeout_ms, stats, &err);
        return rc ? fail(s, rc, err) : LUX_SGM_OK;
    }
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return fail(s, LUX_SGM_EINVAL, "unknown ticket");
        r = it->second;
        auto ready = [&] { return r->state == LUX_SGM_REQ_DONE; };
        if (timeout_ms < 0) s->done.wait(lk, ready);
        else if (!s->done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
            return fail(s, LUX_SGM_ECANCELED, "timed out waiting for the request");
        s->reqs.erase(ticket);
    }
    const lux_sgm_rect_params& p = r->p;
    const size_t px = (size_t)p.width * p.height;
    if (r->status == 0 && r->mem == LUX_SGM_MEM_HOST) {
        const char* so = (const char*)r->stage_out;
        if (r->disp) std::memcpy(r->disp, so, px * 2);
        if (r->sub) std::memcpy(r->sub, so + px * 2, px * 4);
        if (r->valid) std::memcpy(r->valid, so + px * 6, px);
    }
    s->pinned.put(r->stage_in, r->stage_in_n);
    s->pinned.put(r->stage_out, r->stage_out_n);
    r->stage_in = r->stage_out = nullptr;
    if (stats) {
        auto ms = [](Clock::time_point a, Clock::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        const bool admitted = r->slot >= 0;
        stats->queued_ms = admitted ? ms(r->t_submit, r->t_admit) : ms(r->t_submit, r->t_done);
        stats->run_ms = admitted ? ms(r->t_admit, r->t_done) : 0.0;
        stats->total_ms = ms(r->t_submit, r->t_done);
        stats->slot = r->slot;
        stats->status = r->status;
    }
    if (r->status) set_err(s, r->err.empty() ? "request failed" : r->err);
    return r->status;
}

int lux_sgm_server_cancel(lux_sgm_server* s, uint64_t ticket)
{
    if (!s) return LUX_SGM_EINVAL;
    if (s->dloop) return fail(s, LUX_SGM_EUNSUPPORTED, "the device-resident loop admits requests as they come");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(s, LUX_SGM_EINVAL, "unknown ticket");
    auto r = it->second;
    if (r->state != LUX_SGM_REQ_PENDING) return fail(s, LUX_SGM_EINVAL, "the request is already running or done");
    s->pending.erase(std::find(s->pending.begin(), s->pending.end(), r));
    r->status = LUX_SGM_ECANCELED;
    r->err = "cancelled";
    r->state = LUX_SGM_REQ_DONE;
    r->t_done = Clock::now();
    s->done.notify_all();
    return LUX_SGM_OK;
}

int lux_sgm_server_info_get(lux_sgm_server* s, lux_sgm_server_info* info)
{
    if (!s || !info) return LUX_SGM_EINVAL;
    if (s->dloop) return lux::dloop_info(s->dloop, info);
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(info, 0, sizeof *info);
    info->submitted = s->submitted;
    info->completed = s->completed;
    info->failed = s->failed;
    info->pending = (int32_t)s->pending.size();
    for (auto& sl : s->slots) {
        info->running += sl->req ? 1 : 0;
        info->bytes_held += sl->held;
    }
    info->slots = (int32_t)s->slots.size();
    info->budget = s->budget.limit;
    {
        std::lock_guard<std::mutex> lb(s->budget.mu);
        info->bytes_peak = s->budget.peak;
    }
    return LUX_SGM_OK;
}

}  // extern "C"
