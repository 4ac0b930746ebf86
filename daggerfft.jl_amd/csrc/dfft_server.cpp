// dfft_server.cpp -- the serving loop: a stream of independent 3D transform
// requests with plans and device buffers kept resident.
//
// Threads: clients call submit / wait / info / release from any thread; one
// scheduler thread per server owns every device launch of the requests.
//
//   submit     looks the plan up in the server's plan cache (building it at
//              the first submission of a shape, on the caller's thread), sizes
//              the request's buffers and queues it.
//   scheduler  retires requests whose completion events fired, then admits
//              queued requests in arrival order while in-flight slots and
//              memory under the budget allow.  A request that does not fit may
//              be overtaken by later ones that do, but the oldest such request
//              keeps its bytes reserved, so overtaking never starves it.  An
//              admitted request takes buffers from the arena, and its launches (input
//              fill when the caller gave no input, the plan's passes, an event)
//              go to the lane (HIP stream) with the least outstanding work.
//   arena      device memory grows in chunks up to the budget and is carved
//              into request buffers; nothing is allocated or freed on the
//              device while requests run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/daggerfft_hip.h"
#include "dfft_plan.hpp"

namespace dfft {
size_t plan_resident_bytes(const dfft_plan *p);
}

namespace {

using clk = std::chrono::steady_clock;

// Buffers of at least LARGE bytes come from large chunks at 32 MiB granularity,
// smaller ones from small chunks at 2 MiB granularity: small buffers never
// split the holes that large ones need.
constexpr size_t LARGE = (size_t)256 << 20;
constexpr size_t GRAIN_S = (size_t)2 << 20, GRAIN_L = (size_t)32 << 20;
constexpr size_t CHUNK_S = (size_t)2 << 30, CHUNK_L = (size_t)8 << 30;

size_t round_up(size_t b) {
    const size_t g = b >= LARGE ? GRAIN_L : GRAIN_S;
    return (b + g - 1) / g * g;
}

struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    int chunk = -1;
};

// Device memory of the server: chunks allocated as the load grows (up to the
// budget) and kept until the server is destroyed, carved into request
// buffers by a best-fit free list that merges neighbours on release.  No
// device allocation or free happens while requests run, so no request ever
// waits for an implicit device synchronisation of hipFree.
struct Arena {
    struct Chunk {
        char *base;
        size_t size;
        bool large;
        std::map<size_t, size_t> free;  // offset -> length of each free range
    };
    size_t budget = 0;
    size_t held = 0;  // bytes of all chunks
    size_t busy = 0;  // bytes carved out
    int64_t allocs = 0, grows = 0, waits = 0;
    std::vector<Chunk> chunks;

    bool carve(size_t bytes, bool large, Buf &b) {
        int bc = -1;
        size_t boff = 0, blen = ~(size_t)0;
        for (size_t c = 0; c < chunks.size(); ++c) {
            if (chunks[c].large != large) continue;
            for (const auto &fr : chunks[c].free)
                if (fr.second >= bytes && fr.second < blen) {
                    bc = (int)c;
                    boff = fr.first;
                    blen = fr.second;
                }
        }
        if (bc < 0) return false;
        auto &fl = chunks[bc].free;
        fl.erase(boff);
        if (blen > bytes) fl.emplace(boff + bytes, blen - bytes);
        b.p = chunks[bc].base + boff;
        b.bytes = bytes;
        b.chunk = bc;
        busy += bytes;
        ++allocs;
        return true;
    }
    // a buffer of at least `bytes`; `other` = bytes the budget must also hold (plans)
    int acquire(size_t bytes, size_t other, Buf &b) {
        b = Buf{};
        if (bytes == 0) return DFFT_OK;
        bytes = round_up(bytes);
        const bool large = bytes >= LARGE;
        if (carve(bytes, large, b)) return DFFT_OK;
        const size_t g = large ? GRAIN_L : GRAIN_S;
        const size_t room = budget > held + other ? budget - held - other : 0;
        size_t want = std::min(room, std::max(bytes, large ? CHUNK_L : CHUNK_S)) / g * g;
        if (want < bytes) return DFFT_ERR_BUDGET;
        void *p = nullptr;
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            want = bytes;
            if (hipMalloc(&p, want) != hipSuccess) {
                (void)hipGetLastError();
                return DFFT_ERR_OOM;
            }
        }
        chunks.push_back(Chunk{static_cast<char *>(p), want, large, {{0, want}}});
        held += want;
        ++grows;
        return carve(bytes, large, b) ? DFFT_OK : DFFT_ERR_OOM;
    }
    void release(Buf &b) {
        if (!b.p) return;
        Chunk &ch = chunks[b.chunk];
        size_t off = (size_t)(static_cast<char *>(b.p) - ch.base), len = b.bytes;
        auto &fl = ch.free;
        auto next = fl.lower_bound(off);
        if (next != fl.end() && off + len == next->first) {
            len += next->second;
            next = fl.erase(next);
        }
        if (next != fl.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                len += prev->second;
                fl.erase(prev);
            }
        }
        fl.emplace(off, len);
        busy -= b.bytes;
        b = Buf{};
    }
    // frees every chunk; only when nothing is carved out (no request holds memory)
    void free_all() {
        for (auto &c : chunks) (void)hipFree(c.base);
        chunks.clear();
        held = 0;
    }
};

using PlanKey = std::tuple<int64_t, int64_t, int64_t, int, int, int, int, uint32_t>;

struct Req {
    int64_t ticket = 0;
    dfft_request r{};
    int state = DFFT_REQ_QUEUED;
    int status = DFFT_ERR_NOT_READY;
    int lane = -1;
    double t_arrival = 0, t_admit = 0, t_done = 0;
    const dfft_plan *plan = nullptr;
    size_t in_bytes = 0, out_bytes = 0, work_bytes = 0;
    size_t cls_in = 0, cls_out = 0, cls_work = 0;  // server buffer bytes this request takes
    Buf in, out, work;
    const void *result = nullptr;
    size_t result_bytes = 0;
    double cost = 0;  // algorithmic bytes
    bool waited = false;  // admission once deferred for want of a contiguous range
    hipEvent_t ev = nullptr;
    size_t need() const { return cls_in + cls_out + cls_work; }
};

}  // namespace

struct dfft_server {
    int device = 0;
    int max_inflight = 32;
    clk::time_point t0;
    std::vector<hipStream_t> lanes;
    std::vector<double> lane_load;

    std::mutex mu;                    // everything below
    std::condition_variable cv_sched;  // scheduler: new work, released memory, stop
    std::condition_variable cv_done;   // clients: a request finished
    bool stop = false;
    Arena pool;
    std::map<PlanKey, dfft_plan *> plans;
    size_t plan_bytes = 0;
    size_t plan_reserve = 0;  // budget kept out of the arena for plans built later
    std::unordered_map<int64_t, Req *> reqs;
    std::deque<Req *> queue;
    std::vector<Req *> running;
    std::deque<int64_t> finished;  // not yet returned by wait_any
    std::vector<hipEvent_t> spare_events;
    int64_t next_ticket = 1;
    int64_t submitted = 0, completed = 0, failed = 0, peak_running = 0, overtakes = 0;
    size_t peak_held = 0;
    std::mutex plan_mu;  // serialises plan construction
    std::thread sched;

    double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
    void loop();
    bool reap();
    bool admit();
    void launch(Req *q);
    void finish(Req *q, int status);
};

// Retires requests whose completion events fired.  Caller holds mu.
bool dfft_server::reap() {
    bool any = false;
    for (size_t i = 0; i < running.size();) {
        Req *q = running[i];
        const hipError_t e = hipEventQuery(q->ev);
        if (e == hipErrorNotReady) {
            ++i;
            continue;
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            if (q->status == DFFT_OK) q->status = DFFT_ERR_HIP;
        }
        running[i] = running.back();
        running.pop_back();
        lane_load[q->lane] -= q->cost;
        finish(q, q->status);
        any = true;
    }
    return any;
}

// Marks a request finished; returns every buffer that does not hold its result.
void dfft_server::finish(Req *q, int status) {
    q->t_done = now();
    if (q->ev) {
        spare_events.push_back(q->ev);
        q->ev = nullptr;
    }
    const bool ok = status == DFFT_OK;
    q->status = status;
    q->state = ok ? DFFT_REQ_DONE : DFFT_REQ_FAILED;
    // DFFT_REQ_DISCARD: nobody reads the result, so its server buffer goes
    // back to the arena now rather than at dfft_server_release
    const bool discard = (q->r.req_flags & DFFT_REQ_DISCARD) != 0;
    const void *keep = (ok && !discard) ? q->result : nullptr;
    if (q->in.p && q->in.p != keep) pool.release(q->in);
    if (q->work.p) pool.release(q->work);
    if (q->out.p && q->out.p != keep) pool.release(q->out);
    if (!ok || (discard && !q->r.d_out)) q->result = nullptr;
    if (ok) ++completed;
    else ++failed;
    finished.push_back(q->ticket);
    cv_done.notify_all();
}

// Enqueues a request's device work on its lane.  Caller holds mu.
void dfft_server::launch(Req *q) {
    hipStream_t st = lanes[q->lane];
    const dfft_request &r = q->r;
    const bool in_real = r.kind == DFFT_R2C;
    const int64_t nin = q->in_bytes / (size_t)((in_real ? 1 : 2) * (r.precision ? 8 : 4));
    // where the input lives and where the result goes
    const void *src = r.d_in;
    void *dst = r.d_out ? r.d_out : q->out.p;
    int rc = DFFT_OK;
    if (!src) {
        void *fill = q->in.p ? q->in.p : dst;  // C2C without separate input: fill in place
        rc = dfft_fill_uniform(fill, nin, in_real ? 0 : 1, r.precision, r.seed, st);
        src = fill;
    }
    if (rc == DFFT_OK) rc = dfft_execute(q->plan, src, dst, q->work.p, st);
    q->result = dst;
    q->result_bytes = q->out_bytes;
    if (spare_events.empty()) {
        hipEvent_t ev;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            ev = nullptr;
        }
        q->ev = ev;
    } else {
        q->ev = spare_events.back();
        spare_events.pop_back();
    }
    if (!q->ev || hipEventRecord(q->ev, st) != hipSuccess) {
        (void)hipGetLastError();
        // cannot observe completion: wait for the lane, then retire
        (void)hipStreamSynchronize(st);
        if (q->ev) spare_events.push_back(q->ev);
        q->ev = nullptr;
        finish(q, rc == DFFT_OK ? DFFT_ERR_HIP : rc);
        return;
    }
    q->status = rc;
    q->state = DFFT_REQ_RUNNING;
    lane_load[q->lane] += q->cost;
    running.push_back(q);
    peak_running = std::max<int64_t>(peak_running, (int64_t)running.size());
}

// Admits queued requests in arrival order.  Caller holds mu.
bool dfft_server::admit() {
    bool any = false;
    size_t reserved = 0;  // bytes kept for the oldest request waiting for memory
    bool blocked = false;
    for (auto it = queue.begin(); it != queue.end() && (int)running.size() < max_inflight;) {
        Req *q = *it;
        const size_t need = q->need();
        if (pool.busy + plan_bytes + reserved + need > pool.budget) {
            if (!blocked) {
                blocked = true;
                reserved = need;
            }
            ++it;
            continue;
        }
        int rc = DFFT_OK;
        for (int attempt = 0; attempt < 2; ++attempt) {
            const size_t other = std::max(plan_bytes, plan_reserve);
            rc = pool.acquire(q->cls_in, other, q->in);
            if (rc == DFFT_OK) rc = pool.acquire(q->cls_out, other, q->out);
            if (rc == DFFT_OK) rc = pool.acquire(q->cls_work, other, q->work);
            if (rc == DFFT_OK) break;
            pool.release(q->in);
            pool.release(q->out);
            pool.release(q->work);
            // no contiguous room although the budget has it: with nothing
            // carved out, return the chunks and lay the arena out afresh
            if (attempt == 0 && pool.busy == 0 && running.empty()) pool.free_all();
            else break;
        }
        if (rc != DFFT_OK) {
            if (!q->waited) ++pool.waits;
            q->waited = true;
            if (rc == DFFT_ERR_OOM && running.empty() && pool.busy == 0) {
                // nothing else holds device memory: it can never fit
                it = queue.erase(it);
                finish(q, DFFT_ERR_OOM);
                any = true;
                continue;
            }
            if (!blocked) {
                blocked = true;
                reserved = need;
            }
            ++it;
            continue;
        }
        if (blocked) ++overtakes;
        peak_held = std::max(peak_held, pool.held + plan_bytes);
        it = queue.erase(it);
        q->t_admit = now();
        int lane = 0;
        for (int l = 1; l < (int)lanes.size(); ++l)
            if (lane_load[l] < lane_load[lane]) lane = l;
        q->lane = lane;
        launch(q);
        any = true;
    }
    return any;
}

void dfft_server::loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        reap();
        const bool admitted = admit();
        if (stop && queue.empty() && running.empty()) break;
        if (running.empty()) {
            if (!admitted) cv_sched.wait(lk);  // woken by submit, release or stop
            continue;
        }
        // requests on the device: poll their events
        lk.unlock();
        std::this_thread::yield();
        lk.lock();
    }
}

static int validate(const dfft_request *r) {
    if (!r) return DFFT_ERR_ARG;
    if (r->req_flags & ~DFFT_REQ_DISCARD) return DFFT_ERR_ARG;
    if (r->kind != DFFT_C2C && r->kind != DFFT_R2C && r->kind != DFFT_C2R) return DFFT_ERR_UNSUPPORTED;
    if (r->sign != -1 && r->sign != 1) return DFFT_ERR_ARG;
    if (r->kind == DFFT_R2C && r->sign != -1) return DFFT_ERR_ARG;
    if (r->kind == DFFT_C2R && r->sign != 1) return DFFT_ERR_ARG;
    if (r->precision != DFFT_F32 && r->precision != DFFT_F64) return DFFT_ERR_ARG;
    if (r->decomp != DFFT_PENCIL && r->decomp != DFFT_SLAB) return DFFT_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        if (r->n[d] < 1) return DFFT_ERR_ARG;
    if (r->d_in && r->d_in == r->d_out && r->kind != DFFT_C2C) return DFFT_ERR_ARG;
    return DFFT_OK;
}

// The resident plan of a request's shape, built on first use.
static int get_plan(dfft_server *s, const dfft_request *r, const dfft_plan **out) {
    const PlanKey key{r->n[0], r->n[1], r->n[2], r->kind, r->sign, r->precision, r->decomp, r->flags};
    std::lock_guard<std::mutex> pg(s->plan_mu);
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->plans.find(key);
        if (it != s->plans.end()) {
            *out = it->second;
            return DFFT_OK;
        }
    }
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (prev != s->device) (void)hipSetDevice(s->device);
    const int dims[3] = {0, 1, 2};
    dfft_plan *p = nullptr;
    int rc = dfft_plan_create(&p, 3, r->n, r->kind, r->sign, nullptr, r->precision, r->decomp, dims, 3, r->flags);
    if (prev != s->device) (void)hipSetDevice(prev);
    if (rc != DFFT_OK) return rc;
    std::lock_guard<std::mutex> g(s->mu);
    s->plans.emplace(key, p);
    s->plan_bytes += dfft::plan_resident_bytes(p);

    *out = p;
    return DFFT_OK;
}

static void free_server(dfft_server *s) {
    for (auto &kv : s->reqs) {
        Req *q = kv.second;
        if (q->ev) (void)hipEventDestroy(q->ev);
        s->pool.release(q->in);
        s->pool.release(q->out);
        s->pool.release(q->work);
        delete q;
    }
    s->reqs.clear();
    s->pool.free_all();
    for (auto ev : s->spare_events) (void)hipEventDestroy(ev);
    for (auto &kv : s->plans) dfft_plan_destroy(kv.second);
    for (auto st : s->lanes) (void)hipStreamDestroy(st);
    delete s;
}

static bool finished_state(int st) { return st == DFFT_REQ_DONE || st == DFFT_REQ_FAILED; }

template <class Pred>
static bool wait_for(dfft_server *s, std::unique_lock<std::mutex> &lk, double timeout_s, Pred pred) {
    if (timeout_s < 0) {
        s->cv_done.wait(lk, pred);
        return true;
    }
    return s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
}

extern "C" {

int dfft_server_create(dfft_server **server, const dfft_server_config *cfg) {
    if (!server) return DFFT_ERR_ARG;
    *server = nullptr;
    dfft_server_config c{};
    if (cfg) c = *cfg;
    if (c.flags != 0 || c.max_inflight < 0 || c.lanes < 0 || c.lanes > 16) return DFFT_ERR_ARG;
    dfft_server *s = new (std::nothrow) dfft_server();
    if (!s) return DFFT_ERR_OOM;
    s->t0 = clk::now();
    if (hipError_t e = hipGetDevice(&s->device)) {
        delete s;
        return dfft::hip_fail(e, "hipGetDevice");
    }
    size_t free_b = 0, total_b = 0;
    if (hipError_t e = hipMemGetInfo(&free_b, &total_b)) {
        delete s;
        return dfft::hip_fail(e, "hipMemGetInfo");
    }
    s->pool.budget = c.hbm_budget ? c.hbm_budget : free_b / 10 * 9;
    s->plan_reserve = std::min<size_t>(s->pool.budget / 64, (size_t)64 << 20);
    s->max_inflight = c.max_inflight ? c.max_inflight : 32;
    const int nl = c.lanes ? c.lanes : 4;
    for (int l = 0; l < nl; ++l) {
        hipStream_t st;
        // non-blocking: the lanes never wait for the legacy default stream
        if (hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) {
            free_server(s);
            return dfft::hip_fail(e, "hipStreamCreateWithFlags");
        }
        s->lanes.push_back(st);
        s->lane_load.push_back(0.0);
    }
    try {
        s->sched = std::thread([s] { s->loop(); });
    } catch (...) {
        free_server(s);
        return DFFT_ERR_OOM;
    }
    *server = s;
    return DFFT_OK;
}

int dfft_server_clock(const dfft_server *s, double *now) {
    if (!s || !now) return DFFT_ERR_ARG;
    *now = s->now();
    return DFFT_OK;
}

int dfft_server_prepare(dfft_server *s, const dfft_request *reqs, int count) {
    if (!s || count < 0 || (count && !reqs)) return DFFT_ERR_ARG;
    for (int i = 0; i < count; ++i) {
        int rc = validate(&reqs[i]);
        const dfft_plan *p = nullptr;
        if (rc == DFFT_OK) rc = get_plan(s, &reqs[i], &p);
        if (rc != DFFT_OK) return rc;
    }
    return DFFT_OK;
}

int dfft_server_submit(dfft_server *s, const dfft_request *r, int64_t *ticket) {
    if (!s || !ticket) return DFFT_ERR_ARG;
    *ticket = 0;
    int rc = validate(r);
    if (rc != DFFT_OK) return rc;
    const double t_sub = s->now();
    const dfft_plan *p = nullptr;
    rc = get_plan(s, r, &p);
    if (rc != DFFT_OK) return rc;
    Req *q = new (std::nothrow) Req();
    if (!q) return DFFT_ERR_OOM;
    q->r = *r;
    q->plan = p;
    q->t_arrival = r->t_arrival > 0 ? r->t_arrival : t_sub;
    dfft_plan_io_size(p, &q->in_bytes, &q->out_bytes);
    dfft_plan_workspace_size(p, &q->work_bytes);
    dfft_plan_algorithmic_bytes(p, &q->cost);
    // server buffers: the input when the caller gave none (C2C fills its
    // output and runs in place), the output when the caller gave none
    const bool inplace_c2c = r->kind == DFFT_C2C && !r->d_in;
    if (!r->d_in && !inplace_c2c) q->cls_in = round_up(q->in_bytes);
    if (!r->d_out) q->cls_out = round_up(q->out_bytes);
    q->cls_work = round_up(q->work_bytes);
    std::lock_guard<std::mutex> g(s->mu);
    if (q->need() + s->plan_bytes > s->pool.budget) {
        delete q;
        return DFFT_ERR_BUDGET;
    }
    q->ticket = s->next_ticket++;
    s->reqs.emplace(q->ticket, q);
    s->queue.push_back(q);
    ++s->submitted;
    *ticket = q->ticket;
    s->cv_sched.notify_one();
    return DFFT_OK;
}

int dfft_server_wait(dfft_server *s, int64_t ticket, double timeout_s) {
    if (!s) return DFFT_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return DFFT_ERR_ARG;
    Req *q = it->second;
    return wait_for(s, lk, timeout_s, [q] { return finished_state(q->state); }) ? DFFT_OK : DFFT_ERR_NOT_READY;
}

int dfft_server_wait_any(dfft_server *s, double timeout_s, int64_t *ticket) {
    if (!s || !ticket) return DFFT_ERR_ARG;
    *ticket = 0;
    std::unique_lock<std::mutex> lk(s->mu);
    if (!wait_for(s, lk, timeout_s, [s] { return !s->finished.empty(); })) return DFFT_ERR_NOT_READY;
    *ticket = s->finished.front();
    s->finished.pop_front();
    return DFFT_OK;
}

int dfft_server_info(dfft_server *s, int64_t ticket, dfft_request_info *info) {
    if (!s || !info) return DFFT_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return DFFT_ERR_ARG;
    const Req *q = it->second;
    std::memset(info, 0, sizeof(*info));
    info->state = q->state;
    info->status = finished_state(q->state) ? q->status : DFFT_ERR_NOT_READY;
    info->lane = q->lane;
    info->t_arrival = q->t_arrival;
    info->t_admit = q->t_admit;
    info->t_done = q->t_done;
    info->d_result = q->state == DFFT_REQ_DONE ? q->result : nullptr;
    info->result_bytes = q->state == DFFT_REQ_DONE ? q->result_bytes : 0;
    info->algorithmic_bytes = q->cost;
    return DFFT_OK;
}

int dfft_server_copy_result(dfft_server *s, int64_t ticket, void *d_dst, size_t bytes, void *stream) {
    if (!s || !d_dst) return DFFT_ERR_ARG;
    const void *src;
    size_t nb;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return DFFT_ERR_ARG;
        const Req *q = it->second;
        if (q->state != DFFT_REQ_DONE) return q->state == DFFT_REQ_FAILED ? q->status : DFFT_ERR_NOT_READY;
        if (!q->result) return DFFT_ERR_STATE;  // a discarded result
        src = q->result;
        nb = q->result_bytes;
    }
    if (bytes != nb) return DFFT_ERR_ARG;
    hipError_t e = hipMemcpyAsync(d_dst, src, nb, hipMemcpyDeviceToDevice, (hipStream_t)stream);
    return e == hipSuccess ? DFFT_OK : dfft::hip_fail(e, "hipMemcpyAsync (result)");
}

int dfft_server_release(dfft_server *s, int64_t ticket) {
    if (!s) return DFFT_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return DFFT_ERR_ARG;
    Req *q = it->second;
    if (q->state == DFFT_REQ_RUNNING) return DFFT_ERR_STATE;
    if (q->state == DFFT_REQ_QUEUED) s->queue.erase(std::find(s->queue.begin(), s->queue.end(), q));
    s->pool.release(q->in);
    s->pool.release(q->out);
    s->pool.release(q->work);
    s->reqs.erase(it);
    auto f = std::find(s->finished.begin(), s->finished.end(), ticket);
    if (f != s->finished.end()) s->finished.erase(f);
    delete q;
    s->cv_sched.notify_one();
    return DFFT_OK;
}

int dfft_server_drain(dfft_server *s, double timeout_s) {
    if (!s) return DFFT_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    return wait_for(s, lk, timeout_s, [s] { return s->queue.empty() && s->running.empty(); }) ? DFFT_OK
                                                                                              : DFFT_ERR_NOT_READY;
}

int dfft_server_stats_get(dfft_server *s, dfft_server_stats *st) {
    if (!s || !st) return DFFT_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    st->queued = (int64_t)s->queue.size();
    st->running = (int64_t)s->running.size();
    st->peak_running = s->peak_running;
    st->plans = (int64_t)s->plans.size();
    st->overtakes = s->overtakes;
    st->buffer_allocs = s->pool.allocs;
    st->arena_chunks = s->pool.grows;
    st->arena_waits = s->pool.waits;
    st->budget = s->pool.budget;
    st->bytes_held = s->pool.held + s->plan_bytes;
    st->bytes_busy = s->pool.busy;
    st->peak_bytes_held = std::max(s->peak_held, s->pool.held + s->plan_bytes);
    return DFFT_OK;
}

int dfft_server_destroy(dfft_server *s) {
    if (!s) return DFFT_ERR_ARG;
    {
        std::lock_guard<std::mutex> g(s->mu);
        // queued requests are cancelled; running ones finish first
        for (Req *q : s->queue) {
            q->state = DFFT_REQ_FAILED;
            q->status = DFFT_ERR_STATE;
        }
        s->queue.clear();
        s->stop = true;
        s->cv_sched.notify_one();
    }
    if (s->sched.joinable()) s->sched.join();
    free_server(s);
    return DFFT_OK;
}

}  // extern "C"
