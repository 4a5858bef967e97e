// Continuous-batching tile server for the segmentNuclei chain and its two task stages.
//
// The reference scales out one tile per MPI process (nu-segment.cpp:274-433), each running the
// whole chain on its own tile, and splits the chain into tasks S1 / A4 for pipelining
// (tasks/seg-tasks).  Here one process drives one GPU: requests of any tile size are queued and
// admitted when a slot (host worker + HIP stream + device memory) is free AND the device memory
// the slot must add for them -- a chain workspace sized for the tile, staging for host inputs --
// fits the remaining HBM budget.  Admitted requests run concurrently on their own streams, so the
// latency-bound phases of different tiles overlap on the device.  Intermediates never leave the
// device; a request's outputs are written into the caller's device buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
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

struct Req {
    int64_t ticket;
    nsc_request r;
    nsc_seg_params params;
    Clock::time_point t_submit, t_start, t_end;
    nsc_result res{};
    bool done = false;
};

// A slot: worker thread, non-blocking stream, and the device memory it holds -- a chain workspace
// sized for the largest tile it has served (grown on demand) and a staging buffer for host inputs.
struct Slot {
    std::thread th;
    hipStream_t stream = nullptr;
    nsc_workspace* ws = nullptr;
    int ws_w = 0, ws_h = 0;
    size_t ws_bytes = 0;
    uint8_t* staging = nullptr;
    size_t staging_bytes = 0;
};

size_t ws_total(int w, int h) { return nsc::workspace_bytes(w, h) + nsc::workspace_arena_bytes(w, h); }

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
    size_t used = 0;  // device bytes the slots hold (workspaces + staging), reserved before allocation
    bool stopping = false;
    Clock::time_point t0 = Clock::now();

    // Device bytes slot S must add to serve q (negative when a smaller workspace replaces one that
    // does not fit q's tile): a workspace for q's tile when its own does not fit it, staging for a
    // host input larger than its buffer.
    long long extra_for(const Slot& S, const Req* q, size_t* ws_new, size_t* stage_new) const {
        const nsc_request& r = q->r;
        *ws_new = nsc::workspace_fits(S.ws, r.w, r.h) ? S.ws_bytes : ws_total(r.w, r.h);
        const size_t sb = r.host_input ? (size_t)3 * r.w * r.h : 0;
        *stage_new = std::max(sb, S.staging_bytes);
        return ((long long)*ws_new - (long long)S.ws_bytes) + ((long long)*stage_new - (long long)S.staging_bytes);
    }

    // first queued request that slot S can serve within the budget (skip-ahead keeps small tiles
    // flowing while a large one waits for memory); reserves the memory it adds.  nullptr if none.
    Req* take_fitting(Slot& S, size_t* ws_new, size_t* stage_new) {
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            const long long add = extra_for(S, *it, ws_new, stage_new);
            if ((long long)used + add <= (long long)cfg.hbm_budget) {
                Req* q = *it;
                queue.erase(it);
                used = (size_t)((long long)used + add);
                return q;
            }
        }
        return nullptr;
    }

    // device memory for q in slot S, as reserved by take_fitting; on failure the slot's fields say
    // what it still holds
    int provision(Slot& S, const Req* q, size_t ws_new, size_t stage_new) {
        const nsc_request& r = q->r;
        if (!S.ws || ws_new != S.ws_bytes) {
            if (S.ws) nsc_workspace_destroy(S.ws);
            S.ws = nullptr;
            S.ws_bytes = 0;
            S.ws_w = S.ws_h = 0;
            const int rc = nsc_workspace_create(r.w, r.h, &S.ws);
            if (rc != NSC_OK) {
                S.ws = nullptr;
                return rc;
            }
            S.ws_w = r.w;
            S.ws_h = r.h;
            S.ws_bytes = ws_new;
        }
        if (stage_new > S.staging_bytes) {
            if (S.staging) (void)hipFree(S.staging);
            S.staging = nullptr;
            S.staging_bytes = 0;
            if (hipMalloc(&S.staging, stage_new) != hipSuccess) {
                (void)hipGetLastError();
                S.staging = nullptr;
                return NSC_ERR_NOMEM;
            }
            S.staging_bytes = stage_new;
        }
        return NSC_OK;
    }

    void run(Slot& S, Req* q) {
        const nsc_request& r = q->r;
        nsc_result& res = q->res;
        res.error = NSC_OK;
        const uint8_t* in = r.bgr;
        if (r.host_input) {
            const size_t b = (size_t)3 * r.w * r.h;
            if (hipMemcpyAsync(S.staging, r.bgr, b, hipMemcpyHostToDevice, S.stream) != hipSuccess) {
                res.error = NSC_ERR_HIP;
                return;
            }
            in = S.staging;
        }
        if (r.stage == NSC_STAGE_FIND_CANDIDATES) {  // S1 task
            const int st = nsc_pl_find_nucleus_candidates(in, 0, r.w, r.h, &q->params, r.mask_out, 0, S.ws, S.stream);
            if (st <= NSC_ERR_ARG) res.error = st; else res.status = st;
            return;
        }
        if (r.stage == NSC_STAGE_SEPARATE_NUCLEI) {  // A4 task
            const int st = nsc_pl_separate_nuclei(in, 0, r.mask_in, 0, r.w, r.h, q->params.minSizePl,
                                                  q->params.watershedConnectivity, r.mask_out, 0, S.ws, S.stream);
            if (st <= NSC_ERR_ARG) res.error = st; else res.status = st;
            return;
        }
        nsc_seg_io io;
        memset(&io, 0, sizeof(io));
        io.canonical = (r.canonical || r.bbox5 || r.area) ? 1 : 0;
        io.bbox_mode = r.bbox5 ? NSC_BBOX_ALL : NSC_BBOX_NONE;
        io.bbox5 = r.bbox5;
        io.area = r.area;
        io.max_boxes = r.max_labels;
        const int st = nsc_segment_nuclei_io(in, r.w, r.h, &q->params, r.labels, &io, S.ws, nullptr, nullptr, S.stream);
        if (st <= NSC_ERR_ARG) { res.error = st; return; }
        res.status = st;
        res.compcount = io.nuclei;
    }

    void worker(int k) {
        Slot& S = slots[k];
        (void)hipSetDevice(device);
        nsc::set_thread_blocking_sync(true);
        (void)nsc::set_thread_pass_bridge(false);
        nsc::thread_serving() = true;
        for (;;) {
            Req* q = nullptr;
            size_t ws_new = 0, stage_new = 0;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if ((q = take_fitting(S, &ws_new, &stage_new)) != nullptr) break;
                    if (stopping && queue.empty()) return;
                    // nothing this slot could serve fits: under memory pressure an idle slot gives
                    // its memory back, so the budget eventually admits every request (submit
                    // rejects any request larger than the whole budget)
                    if (!queue.empty() && (S.ws || S.staging)) {
                        const size_t held = S.ws_bytes + S.staging_bytes;
                        lk.unlock();
                        if (S.ws) nsc_workspace_destroy(S.ws);
                        if (S.staging) (void)hipFree(S.staging);
                        S.ws = nullptr;
                        S.staging = nullptr;
                        S.ws_bytes = S.staging_bytes = 0;
                        S.ws_w = S.ws_h = 0;
                        lk.lock();
                        used -= held;
                        cv_work.notify_all();
                        continue;
                    }
                    cv_work.wait(lk);
                }
            }
            q->t_start = Clock::now();
            const int prc = provision(S, q, ws_new, stage_new);
            if (prc != NSC_OK) {
                q->res.error = prc;
                // the reservation assumed ws_new + stage_new; the slot holds what its fields say
                std::lock_guard<std::mutex> lk(mu);
                used = used - ws_new - stage_new + S.ws_bytes + S.staging_bytes;
            } else {
                run(S, q);
            }
            q->t_end = Clock::now();
            {
                std::lock_guard<std::mutex> lk(mu);
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
            cv_work.notify_all();
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

size_t nsc_request_bytes(int w, int h) { return ws_total(w, h); }

int nsc_server_create(const nsc_server_config* cfg, nsc_server** out) {
    NSC_ARG(cfg && out && cfg->slots > 0 && cfg->slots <= 64 && cfg->max_w > 0 && cfg->max_h > 0, "server_create");
    NSC_LABEL_SIZE(cfg->max_w, cfg->max_h);
    NSC_CHECK(nsc::ensure_init());
    if (cfg->hbm_budget < ws_total(cfg->max_w, cfg->max_h)) {
        set_error("server_create: HBM budget %zu B below one largest request (%zu B)", cfg->hbm_budget,
                  ws_total(cfg->max_w, cfg->max_h));
        return NSC_ERR_ARG;
    }
    auto* S = new nsc_server();
    S->cfg = *cfg;
    NSC_HIP(hipGetDevice(&S->device));
    S->slots.resize(cfg->slots);
    for (auto& sl : S->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) {
            for (auto& t : S->slots)
                if (t.stream) (void)hipStreamDestroy(t.stream);
            delete S;
            set_error("server_create: stream creation failed");
            return NSC_ERR_HIP;
        }
    }
    for (int k = 0; k < cfg->slots; ++k) S->slots[k].th = std::thread([S, k] { S->worker(k); });
    *out = S;
    return NSC_OK;
}

int nsc_server_submit(nsc_server* S, const nsc_request* r, int64_t* ticket) {
    NSC_ARG(S && r && ticket && r->bgr && r->w > 0 && r->h > 0, "server_submit");
    NSC_ARG(r->stage >= NSC_STAGE_CHAIN && r->stage <= NSC_STAGE_SEPARATE_NUCLEI, "server_submit: stage");
    NSC_ARG((size_t)r->w * r->h <= (size_t)S->cfg.max_w * S->cfg.max_h,
            "server_submit: tile has more pixels than the server's largest tile");
    // outputs (and a device input) must be device memory: kernels write them
    if (r->stage == NSC_STAGE_CHAIN) {
        NSC_ARG(r->labels && is_device_ptr(r->labels) && (!r->bbox5 || is_device_ptr(r->bbox5)) &&
                    (!r->area || is_device_ptr(r->area)),
                "server_submit: labels/bbox5/area must be device pointers");
        NSC_ARG(!r->bbox5 || r->max_labels > 0, "server_submit: bbox5 needs max_labels");
    } else {
        NSC_ARG(r->mask_out && is_device_ptr(r->mask_out), "server_submit: mask_out must be a device pointer");
        NSC_ARG(r->stage != NSC_STAGE_SEPARATE_NUCLEI || (r->mask_in && is_device_ptr(r->mask_in)),
                "server_submit: plSeparateNuclei needs mask_in (seg_open) on the device");
    }
    NSC_ARG(r->host_input || is_device_ptr(r->bgr), "server_submit: bgr must be a device pointer (or set host_input)");
    const size_t need = ws_total(r->w, r->h) + (r->host_input ? (size_t)3 * r->w * r->h : 0);
    if (need > S->cfg.hbm_budget) {
        set_error("server_submit: request needs %zu B, the budget is %zu B", need, S->cfg.hbm_budget);
        return NSC_ERR_LIMIT;
    }
    auto* q = new Req();
    q->r = *r;
    if (r->params) q->params = *r->params; else nsc_seg_params_default(&q->params);
    q->r.params = nullptr;
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
    S->cv_work.notify_all();  // the slot that can serve it within the budget may not be the first awake
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

int nsc_server_stats(nsc_server* S, int* queued, int* running, size_t* used_bytes, size_t* budget) {
    NSC_ARG(S, "server_stats");
    std::lock_guard<std::mutex> lk(S->mu);
    int run = 0;
    for (auto& kv : S->all) run += (!kv.second->done) ? 1 : 0;
    if (queued) *queued = (int)S->queue.size();
    if (running) *running = run - (int)S->queue.size();
    if (used_bytes) *used_bytes = S->used;
    if (budget) *budget = S->cfg.hbm_budget;
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
        if (sl.ws) nsc_workspace_destroy(sl.ws);
        (void)hipStreamDestroy(sl.stream);
    }
    for (auto& kv : S->all) delete kv.second;
    delete S;
    return NSC_OK;
}

}  // extern "C"
