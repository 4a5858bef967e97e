// Continuous-batching tile server for the segmentNuclei chain.
//
// The reference scales out one tile per MPI process (nu-segment.cpp:274-433), each running the
// whole chain on its own tile.  Here one process drives one GPU: requests of any tile size are
// queued, admitted when a slot (host worker + HIP stream + device workspace) is free AND their
// device-memory footprint fits the remaining HBM budget, and run concurrently on their own
// streams, so the latency-bound phases of different tiles overlap on the device.  Intermediates
// never leave the device; a request's outputs are written into the caller's device buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/nscale_hip.h"
#include "common.hpp"

namespace nsc {
size_t workspace_bytes(int max_w, int max_h);
size_t workspace_arena_bytes(int max_w, int max_h);
bool workspace_fits(const nsc_workspace* ws, int w, int h);
}  // namespace nsc

namespace {

using Clock = std::chrono::steady_clock;

// Device bytes one in-flight request holds beyond its slot's workspace: the operators' stream-ordered
// scratch (watershed queues and tie pool, union-find parents, reconstruction tile lists) plus the
// host-input staging copy.  Measured peak of the chain's scratch is below 100 B/px.
constexpr size_t kScratchPerPx = 112;

struct Req {
    int64_t ticket;
    nsc_request r;
    nsc_seg_params params;
    size_t bytes;
    Clock::time_point t_submit, t_start, t_end;
    nsc_result res{};
    bool done = false;
};

struct Slot {
    std::thread th;
    hipStream_t stream = nullptr;
    nsc_workspace* ws = nullptr;
    uint8_t* staging = nullptr;  // device copy of host inputs
    size_t staging_bytes = 0;
};

}  // namespace

struct nsc_server {
    nsc_server_config cfg{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::unordered_map<int64_t, Req*> all;
    std::vector<Slot> slots;
    int64_t next_ticket = 1;
    size_t inflight_bytes = 0, budget_left = 0;
    bool stopping = false;
    Clock::time_point t0 = Clock::now();

    // first queued request that fits the budget (skip-ahead keeps small tiles flowing while a large one
    // waits for memory); nullptr if none
    Req* take_fitting() {
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            if ((*it)->bytes <= budget_left - inflight_bytes) {
                Req* q = *it;
                queue.erase(it);
                return q;
            }
        }
        return nullptr;
    }

    void run(Slot& S, Req* q) {
        const nsc_request& r = q->r;
        nsc_result& res = q->res;
        res.error = NSC_OK;
        const uint8_t* in = r.bgr;
        if (r.host_input) {
            const size_t b = (size_t)3 * r.w * r.h;
            if (b > S.staging_bytes) {
                if (S.staging) (void)hipFree(S.staging);
                S.staging = nullptr;
                S.staging_bytes = 0;
                if (hipMalloc(&S.staging, b) != hipSuccess) { res.error = NSC_ERR_NOMEM; return; }
                S.staging_bytes = b;
            }
            if (hipMemcpyAsync(S.staging, r.bgr, b, hipMemcpyHostToDevice, S.stream) != hipSuccess) {
                res.error = NSC_ERR_HIP;
                return;
            }
            in = S.staging;
        }
        int cc = 0;
        const int st = nsc_segment_nuclei(in, r.w, r.h, &q->params, r.labels, &cc, S.ws, nullptr, nullptr, S.stream);
        if (st <= NSC_ERR_ARG) { res.error = st; return; }
        res.status = st;
        res.compcount = cc;
        if (st != NSC_SEG_SUCCESS) return;
        if (r.canonical || r.bbox5 || r.area) {
            int n = 0;
            int rc = nsc_cc_relabel(r.labels, r.w, r.h, -1, &n, S.stream);
            if (rc != NSC_OK) { res.error = rc; return; }
            res.compcount = n;
            if (r.bbox5 && n > 0) {
                if (n > r.max_labels) { res.error = NSC_ERR_LIMIT; return; }
                rc = nsc_bounding_boxes(r.labels, r.w, r.h, n, r.bbox5, r.area, S.stream);
                if (rc == NSC_OK && nsc::stream_sync(S.stream) != hipSuccess) rc = NSC_ERR_HIP;
                if (rc != NSC_OK) { res.error = rc; return; }
            }
        }
    }

    void worker(int k) {
        Slot& S = slots[k];
        (void)hipSetDevice(device);
        nsc::set_thread_blocking_sync(true);
        for (;;) {
            Req* q = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stopping || (q = take_fitting()) != nullptr; });
                if (!q) return;  // stopping with nothing admissible
                inflight_bytes += q->bytes;
            }
            q->t_start = Clock::now();
            run(S, q);
            q->t_end = Clock::now();
            {
                std::lock_guard<std::mutex> lk(mu);
                inflight_bytes -= q->bytes;
                auto ms = [](Clock::time_point a, Clock::time_point b) {
                    return std::chrono::duration<double, std::milli>(b - a).count();
                };
                q->res.queue_ms = ms(q->t_submit, q->t_start);
                q->res.run_ms = ms(q->t_start, q->t_end);
                q->res.total_ms = ms(q->t_submit, q->t_end);
                q->res.start_ms = ms(t0, q->t_start);
                q->res.end_ms = ms(t0, q->t_end);
                q->res.slot = k;
                q->done = true;
            }
            cv_done.notify_all();
            cv_work.notify_all();  // memory freed: another queued request may fit now
        }
    }
};

using nsc::set_error;

static bool is_device_ptr(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // clear the sticky-free error of an unknown host pointer
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

extern "C" {

size_t nsc_request_bytes(int w, int h) { return (size_t)w * h * kScratchPerPx + ((size_t)1 << 20); }

int nsc_server_create(const nsc_server_config* cfg, nsc_server** out) {
    NSC_ARG(cfg && out && cfg->slots > 0 && cfg->slots <= 64 && cfg->max_w > 0 && cfg->max_h > 0, "server_create");
    NSC_CHECK(nsc::ensure_init());
    const size_t ws_bytes = nsc::workspace_bytes(cfg->max_w, cfg->max_h) + nsc::workspace_arena_bytes(cfg->max_w, cfg->max_h);
    const size_t fixed = ws_bytes * cfg->slots;
    if (cfg->hbm_budget < fixed + nsc_request_bytes(cfg->max_w, cfg->max_h)) {
        set_error("server_create: HBM budget %zu B below the %d slot workspaces (%zu B) plus one largest request",
                  cfg->hbm_budget, cfg->slots, fixed);
        return NSC_ERR_ARG;
    }
    auto* S = new nsc_server();
    S->cfg = *cfg;
    S->budget_left = cfg->hbm_budget - fixed;
    NSC_HIP(hipGetDevice(&S->device));
    S->slots.resize(cfg->slots);
    for (auto& sl : S->slots) {
        int rc = NSC_OK;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) rc = NSC_ERR_HIP;
        if (rc == NSC_OK) rc = nsc_workspace_create(cfg->max_w, cfg->max_h, &sl.ws);
        if (rc != NSC_OK) {
            for (auto& t : S->slots) {
                if (t.ws) nsc_workspace_destroy(t.ws);
                if (t.stream) (void)hipStreamDestroy(t.stream);
            }
            delete S;
            return rc;
        }
    }
    for (int k = 0; k < cfg->slots; ++k) S->slots[k].th = std::thread([S, k] { S->worker(k); });
    *out = S;
    return NSC_OK;
}

int nsc_server_submit(nsc_server* S, const nsc_request* r, int64_t* ticket) {
    NSC_ARG(S && r && ticket && r->bgr && r->labels && r->w > 0 && r->h > 0, "server_submit");
    // outputs (and a device input) must be device memory: kernels write them
    NSC_ARG(is_device_ptr(r->labels) && (!r->bbox5 || is_device_ptr(r->bbox5)) && (!r->area || is_device_ptr(r->area)),
            "server_submit: labels/bbox5/area must be device pointers");
    NSC_ARG(r->host_input || is_device_ptr(r->bgr), "server_submit: bgr must be a device pointer (or set host_input)");
    NSC_ARG(nsc::workspace_fits(S->slots[0].ws, r->w, r->h), "server_submit: tile larger than the slot workspace");
    NSC_ARG(!r->bbox5 || r->max_labels > 0, "server_submit: bbox5 needs max_labels");
    auto* q = new Req();
    q->r = *r;
    if (r->params) q->params = *r->params; else nsc_seg_params_default(&q->params);
    q->r.params = nullptr;
    q->bytes = nsc_request_bytes(r->w, r->h);
    if (q->bytes > S->budget_left) {
        delete q;
        set_error("server_submit: request needs %zu B, budget for requests is %zu B", nsc_request_bytes(r->w, r->h),
                  S->budget_left);
        return NSC_ERR_LIMIT;
    }
    q->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stopping) {
            delete q;
            set_error("server_submit: server is shutting down");
            return NSC_ERR_ARG;
        }
        q->ticket = S->next_ticket++;
        S->all[q->ticket] = q;
        S->queue.push_back(q);
    }
    S->cv_work.notify_one();
    *ticket = q->ticket;
    return NSC_OK;
}

static int finish(nsc_server* S, std::unique_lock<std::mutex>& lk, int64_t ticket, nsc_result* res, bool block) {
    auto it = S->all.find(ticket);
    if (it == S->all.end()) {
        set_error("server: unknown ticket %lld", (long long)ticket);
        return NSC_ERR_ARG;
    }
    Req* q = it->second;
    if (!q->done) {
        if (!block) return 1;
        S->cv_done.wait(lk, [&] { return q->done; });
    }
    if (res) *res = q->res;
    S->all.erase(it);
    const int err = q->res.error;
    delete q;
    return err;
}

int nsc_server_wait(nsc_server* S, int64_t ticket, nsc_result* res) {
    NSC_ARG(S, "server_wait");
    std::unique_lock<std::mutex> lk(S->mu);
    return finish(S, lk, ticket, res, true);
}

int nsc_server_poll(nsc_server* S, int64_t ticket, nsc_result* res) {
    NSC_ARG(S, "server_poll");
    std::unique_lock<std::mutex> lk(S->mu);
    return finish(S, lk, ticket, res, false);
}

int nsc_server_stats(nsc_server* S, int* queued, int* running, size_t* inflight_bytes, size_t* request_budget) {
    NSC_ARG(S, "server_stats");
    std::lock_guard<std::mutex> lk(S->mu);
    int run = 0;
    for (auto& kv : S->all) run += (!kv.second->done) ? 1 : 0;
    if (queued) *queued = (int)S->queue.size();
    if (running) *running = run - (int)S->queue.size();
    if (inflight_bytes) *inflight_bytes = S->inflight_bytes;
    if (request_budget) *request_budget = S->budget_left;
    return NSC_OK;
}

int nsc_server_destroy(nsc_server* S) {
    if (!S) return NSC_OK;
    {
        // drain: workers finish every queued request before they see `stopping` with an empty queue
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.wait(lk, [&] {
            for (auto& kv : S->all)
                if (!kv.second->done) return false;
            return true;
        });
        S->stopping = true;
    }
    S->cv_work.notify_all();
    for (auto& sl : S->slots)
        if (sl.th.joinable()) sl.th.join();
    for (auto& sl : S->slots) {
        if (sl.staging) (void)hipFree(sl.staging);
        nsc_workspace_destroy(sl.ws);
        (void)hipStreamDestroy(sl.stream);
    }
    for (auto& kv : S->all) delete kv.second;
    delete S;
    return NSC_OK;
}

}  // extern "C"
