// engine.hip -- serving engine of the GPU strategy: circuits arrive at any
// time, are admitted against an HBM budget, keep their states resident, and
// are advanced together.
//
// The reference executes one circuit per call (execute(), executors.mojo:
// 723-730).  Here a worker thread owns one HIP stream at default priority and
// issues "rounds": the next plan step of every small running request, with
// the tile passes of a round that share a tile size fused into one batched
// persistent launch (tile_batch_kernel).  Requests of n >= large_qubits run
// their whole plan on their own state stream and overlap with the rounds.
// New requests join at the next round boundary (continuous batching); a
// request's arithmetic is the same alone or in a round, so its result is too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"

using namespace bfly;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    int32_t n = 0, dtype = 0;
    int32_t status = BFLY_REQ_QUEUED;
    int32_t init_kind = BFLY_INIT_ZERO;
    uint64_t init_seed = 0;
    std::vector<uint8_t> init_re, init_im;
    bfly_plan* plan = nullptr;
    bfly_state* state = nullptr;
    hipEvent_t done_ev = nullptr;
    uint64_t bytes = 0;
    size_t next_step = 0;
    bool large = false;
    bool issued_all = false;
    bool cancel = false;
    std::string message;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

constexpr size_t kSlotItems = 4096;  // batched items per round and tile size (x4 per ring slot)

}  // namespace

struct bfly_engine {
    bfly_engine_opts opts{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::vector<uint64_t> running;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    uint64_t bytes_in_use = 0, peak_bytes = 0;
    int64_t rounds = 0, launches = 0, completed = 0, admitted = 0;
    // ring of item slots: host (pinned) -> device copy per round
    int nslots = 0;
    BatchItem* h_items = nullptr;
    BatchItem* d_items = nullptr;
    std::vector<hipEvent_t> slot_ev;
    std::vector<bool> slot_used;
    int slot = 0;
    std::thread worker;
    bool stop = false;
    Clock::time_point t0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

uint64_t state_charge(int n, int dtype) { return (dtype == BFLY_C64 ? 8ull : 16ull) << n; }

void finish(bfly_engine* e, Request* r, int status) {
    r->status = status;
    r->t_done = e->now();
    if (status == BFLY_REQ_DONE) e->completed += 1;
    e->cv_done.notify_all();
}

void free_request_resources(bfly_engine* e, Request* r) {
    if (r->state) {
        bfly_state_destroy(r->state);
        r->state = nullptr;
        e->bytes_in_use -= r->bytes;
    }
    if (r->plan) {
        bfly_plan_destroy(r->plan);
        r->plan = nullptr;
    }
    if (r->done_ev) {
        (void)hipEventDestroy(r->done_ev);
        r->done_ev = nullptr;
    }
}

// Allocate and initialise the state of an admitted request.  Errors stay
// with the request (FAILED) and never stop the engine.
void admit(bfly_engine* e, Request* r) {
    int rc = bfly_state_create(&r->state, r->n, r->dtype);
    if (rc == BFLY_OK) {
        e->bytes_in_use += r->bytes;
        e->peak_bytes = std::max(e->peak_bytes, e->bytes_in_use);
        if (r->init_kind == BFLY_INIT_HOST) {
            rc = bfly_state_upload(r->state, r->init_re.data(), r->init_im.data(), 1ull << r->n);
            std::vector<uint8_t>().swap(r->init_re);
            std::vector<uint8_t>().swap(r->init_im);
        } else if (r->init_kind == BFLY_INIT_RANDOM) {
            rc = bfly_state_fill_random(r->state, r->init_seed);
            if (rc == BFLY_OK) rc = bfly_state_synchronize(r->state);
        }
    }
    if (rc == BFLY_OK) rc = upload_plan(r->plan, e->device);
    if (rc == BFLY_OK && hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming) != hipSuccess)
        rc = fail(BFLY_EHIP, "hipEventCreate");
    if (rc != BFLY_OK) {
        r->message = bfly_last_error();
        free_request_resources(e, r);
        finish(e, r, BFLY_REQ_FAILED);
        return;
    }
    r->large = r->n >= e->opts.large_qubits;
    r->status = BFLY_REQ_RUNNING;
    r->t_admit = e->now();
    e->admitted += 1;
    e->running.push_back(r->id);
}

// First-fit admission over the pending queue; once the oldest request has
// waited 2 s, nothing may pass it, so large requests are not starved.
void admit_pending(bfly_engine* e) {
    const double now = e->now();
    bool head_blocked = false;
    for (auto it = e->pending.begin(); it != e->pending.end();) {
        Request* r = e->reqs[*it].get();
        if (r->cancel) {
            finish(e, r, BFLY_REQ_CANCELLED);
            it = e->pending.erase(it);
            continue;
        }
        if (e->bytes_in_use + r->bytes <= e->opts.budget_bytes) {
            it = e->pending.erase(it);
            admit(e, r);
            continue;
        }
        if (it == e->pending.begin() && now - r->t_submit > 2.0) head_blocked = true;
        if (head_blocked) break;
        ++it;
    }
}

// One round: the next step of every small running request (tile passes
// grouped by dtype and tile size into batched launches) and the whole
// remaining plan of every large one on its own stream.
int issue_round(bfly_engine* e) {
    struct Group { int key; std::vector<Request*> reqs; };
    std::vector<Group> groups;
    std::vector<Request*> finished_now;
    bool any = false;
    int count = 0;
    for (uint64_t id : e->running) {
        Request* r = e->reqs[id].get();
        if (r->issued_all) continue;
        auto& steps = r->plan->hp.steps;
        if (r->cancel) {
            // nothing more is issued; completion is ordered after what was
            (void)hipEventRecord(r->done_ev, r->large ? r->state->stream : e->stream);
            r->issued_all = true;
            continue;
        }
        if (r->large) {
            for (; r->next_step < steps.size(); ++r->next_step) {
                hipError_t err = launch_step(steps[r->next_step], r->state, r->plan, r->state->stream);
                if (err != hipSuccess) return hip_fail(err, "large request launch");
                e->launches += 1;
            }
            (void)hipEventRecord(r->done_ev, r->state->stream);
            r->issued_all = true;
            continue;
        }
        if (e->opts.max_round_requests > 0 && count >= e->opts.max_round_requests) continue;
        ++count;
        any = true;
        if (steps.empty()) {
            finished_now.push_back(r);
            continue;
        }
        const Step& st = steps[r->next_step];
        if (st.kind == S_TILE) {
            const int key = r->dtype * 1024 + st.tile.slots * 64 + st.tile.k;
            auto g = std::find_if(groups.begin(), groups.end(), [&](const Group& x) { return x.key == key; });
            if (g == groups.end()) groups.push_back({key, {r}});
            else g->reqs.push_back(r);
        } else {
            hipError_t err = launch_step(st, r->state, r->plan, e->stream);
            if (err != hipSuccess) return hip_fail(err, "round launch");
            e->launches += 1;
            if (++r->next_step == steps.size()) finished_now.push_back(r);
        }
    }
    if (!any) return BFLY_OK;
    if (!groups.empty()) {
        // reuse a ring slot only after the round that used it has finished
        const int s = e->slot;
        if (e->slot_used[s]) HIP_TRY(hipEventSynchronize(e->slot_ev[s]), "hipEventSynchronize");
        BatchItem* hi = e->h_items + static_cast<size_t>(s) * kSlotItems * 4;
        BatchItem* di = e->d_items + static_cast<size_t>(s) * kSlotItems * 4;
        size_t used = 0;
        struct L { int dtype, slots, k; size_t off; int n; uint64_t tiles; };
        std::vector<L> ls;
        for (auto& g : groups) {
            for (size_t base = 0; base < g.reqs.size(); base += kSlotItems) {
                const size_t cnt = std::min(kSlotItems, g.reqs.size() - base);
                if (used + cnt > kSlotItems * 4) return fail(BFLY_ENOMEM, "round too large for the item ring");
                uint64_t tiles = 0;
                for (size_t i = 0; i < cnt; ++i) {
                    Request* r = g.reqs[base + i];
                    const Step& st = r->plan->hp.steps[r->next_step];
                    BatchItem& it = hi[used + i];
                    it.re = r->state->re;
                    it.im = r->state->im;
                    it.sweeps = r->plan->d_sweeps;
                    it.instrs = r->plan->d_instrs;
                    it.tile_begin = tiles;
                    it.P = st.tile;
                    tiles += 1ull << (st.tile.n - st.tile.k);
                }
                ls.push_back({g.key / 1024, (g.key / 64) % 16, g.key % 64, used, static_cast<int>(cnt), tiles});
                used += cnt;
            }
        }
        HIP_TRY(hipMemcpyAsync(di, hi, used * sizeof(BatchItem), hipMemcpyHostToDevice, e->stream), "round items");
        for (const L& l : ls) {
            HIP_TRY(prepare_tile_kernels(l.dtype, l.k, l.slots), "tile kernel preparation");
            HIP_TRY(launch_tile_batch(l.dtype, l.k, l.slots, di + l.off, l.n, l.tiles, e->stream), "batched launch");
            e->launches += 1;
        }
        HIP_TRY(hipEventRecord(e->slot_ev[s], e->stream), "hipEventRecord");
        e->slot_used[s] = true;
        e->slot = (s + 1) % e->nslots;
        for (auto& g : groups)
            for (Request* r : g.reqs)
                if (++r->next_step == r->plan->hp.steps.size()) finished_now.push_back(r);
    }
    for (Request* r : finished_now) {
        HIP_TRY(hipEventRecord(r->done_ev, e->stream), "hipEventRecord");
        r->issued_all = true;
    }
    e->rounds += 1;
    return BFLY_OK;
}

void reap(bfly_engine* e) {
    for (auto it = e->running.begin(); it != e->running.end();) {
        Request* r = e->reqs[*it].get();
        if (r->issued_all && hipEventQuery(r->done_ev) == hipSuccess) {
            finish(e, r, r->cancel ? BFLY_REQ_CANCELLED : BFLY_REQ_DONE);
            if (r->cancel) free_request_resources(e, r);
            it = e->running.erase(it);
        } else {
            ++it;
        }
    }
}

void worker_loop(bfly_engine* e) {
    (void)hipSetDevice(e->device);
    std::unique_lock<std::mutex> lk(e->mu);
    while (true) {
        admit_pending(e);
        if (e->running.empty()) {
            if (e->stop && e->pending.empty()) break;
            // nothing runs: wait for a submission, or for finished results to
            // be fetched / released so that pending requests fit the budget
            e->cv_work.wait_for(lk, std::chrono::milliseconds(10));
            continue;
        }
        // bounded look-ahead: at most max_inflight_rounds rounds not yet finished
        const int ahead = e->opts.max_inflight_rounds;
        const int s_old = (e->slot + e->nslots - ahead) % e->nslots;
        if (e->slot_used[s_old] && hipEventQuery(e->slot_ev[s_old]) == hipErrorNotReady) {
            hipEvent_t ev = e->slot_ev[s_old];
            lk.unlock();
            (void)hipEventSynchronize(ev);
            lk.lock();
        }
        const int64_t before = e->rounds;
        int rc = issue_round(e);
        if (rc != BFLY_OK) {
            // fail every request that had work in flight in this round
            const std::string msg = bfly_last_error();
            (void)hipStreamSynchronize(e->stream);
            for (uint64_t id : e->running) {
                Request* r = e->reqs[id].get();
                if (r->large && r->issued_all) continue;
                r->message = msg;
                free_request_resources(e, r);
                finish(e, r, BFLY_REQ_FAILED);
            }
            e->running.erase(std::remove_if(e->running.begin(), e->running.end(),
                                            [&](uint64_t id) { return e->reqs[id]->status == BFLY_REQ_FAILED; }),
                             e->running.end());
        }
        reap(e);
        if (e->rounds == before && !e->running.empty()) {
            // everything issued: wait briefly for completions or new work
            e->cv_work.wait_for(lk, std::chrono::microseconds(50));
        }
    }
}

}  // namespace

extern "C" {

int bfly_engine_create(bfly_engine** out, const bfly_engine_opts* opts) {
    if (!out) return fail(BFLY_EINVAL, "null output pointer");
    *out = nullptr;
    auto* e = new bfly_engine();
    if (opts) e->opts = *opts;
    if (e->opts.max_inflight_rounds <= 0) e->opts.max_inflight_rounds = 3;
    if (e->opts.large_qubits <= 0) e->opts.large_qubits = 26;
    e->t0 = Clock::now();
    hipError_t err = hipGetDevice(&e->device);
    size_t free_b = 0, total_b = 0;
    if (err == hipSuccess) err = hipMemGetInfo(&free_b, &total_b);
    if (err == hipSuccess && e->opts.budget_bytes == 0) e->opts.budget_bytes = free_b / 10 * 9;
    e->nslots = e->opts.max_inflight_rounds + 2;
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    const size_t ring_bytes = static_cast<size_t>(e->nslots) * kSlotItems * 4 * sizeof(BatchItem);
    if (err == hipSuccess) err = hipHostMalloc(reinterpret_cast<void**>(&e->h_items), ring_bytes, hipHostMallocDefault);
    if (err == hipSuccess) err = hipMalloc(&e->d_items, ring_bytes);
    e->slot_ev.assign(e->nslots, nullptr);
    e->slot_used.assign(e->nslots, false);
    for (int i = 0; i < e->nslots && err == hipSuccess; ++i)
        err = hipEventCreateWithFlags(&e->slot_ev[i], hipEventDisableTiming);
    if (err != hipSuccess) {
        int rc = hip_fail(err, "engine creation");
        bfly_engine_destroy(e);
        return rc;
    }
    e->worker = std::thread(worker_loop, e);
    *out = e;
    return BFLY_OK;
}

int bfly_engine_submit(bfly_engine* e, const bfly_request* rq, uint64_t* id) {
    if (!e || !rq || !id) return fail(BFLY_EINVAL, "null argument");
    const uint64_t charge = state_charge(rq->num_qubits, rq->dtype);
    if (rq->num_qubits < 1 || rq->num_qubits > kMaxQubits) return fail(BFLY_EINVAL, "qubit count out of range");
    if (charge > e->opts.budget_bytes)
        return fail(BFLY_ENOMEM, "state of " + std::to_string(charge) + " bytes exceeds the engine budget of " +
                                     std::to_string(e->opts.budget_bytes));
    auto r = std::make_unique<Request>();
    r->n = rq->num_qubits;
    r->dtype = rq->dtype;
    r->bytes = charge;
    r->init_kind = rq->init_kind;
    r->init_seed = rq->init_seed;
    if (rq->init_kind == BFLY_INIT_HOST) {
        if (!rq->init_re || !rq->init_im) return fail(BFLY_EINVAL, "null initial state");
        const size_t bytes = (1ull << r->n) * (r->dtype == BFLY_C64 ? 4 : 8);
        r->init_re.assign(static_cast<const uint8_t*>(rq->init_re), static_cast<const uint8_t*>(rq->init_re) + bytes);
        r->init_im.assign(static_cast<const uint8_t*>(rq->init_im), static_cast<const uint8_t*>(rq->init_im) + bytes);
    } else if (rq->init_kind != BFLY_INIT_ZERO && rq->init_kind != BFLY_INIT_RANDOM) {
        return fail(BFLY_EINVAL, "unknown init kind");
    }
    int rc = bfly_plan_create(&r->plan, rq->num_qubits, rq->dtype, rq->ops, rq->nops, rq->udata, rq->udata_len,
                              &rq->exec);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(e->mu);
    r->id = e->next_id++;
    r->t_submit = e->now();
    *id = r->id;
    e->pending.push_back(r->id);
    e->reqs[r->id] = std::move(r);
    e->cv_work.notify_all();
    return BFLY_OK;
}

int bfly_engine_status(bfly_engine* e, uint64_t id, int32_t* status, double* t_submit, double* t_admit,
                       double* t_done) {
    if (!e) return fail(BFLY_EINVAL, "null engine");
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->reqs.find(id);
    if (it == e->reqs.end()) return fail(BFLY_EINVAL, "unknown request " + std::to_string(id));
    Request* r = it->second.get();
    if (status) *status = r->status;
    if (t_submit) *t_submit = r->t_submit;
    if (t_admit) *t_admit = r->t_admit;
    if (t_done) *t_done = r->t_done;
    return BFLY_OK;
}

int bfly_engine_wait(bfly_engine* e, uint64_t id, double timeout_s) {
    if (!e) return fail(BFLY_EINVAL, "null engine");
    std::unique_lock<std::mutex> lk(e->mu);
    auto it = e->reqs.find(id);
    if (it == e->reqs.end()) return fail(BFLY_EINVAL, "unknown request " + std::to_string(id));
    Request* r = it->second.get();
    auto finished = [&] { return r->status >= BFLY_REQ_DONE; };
    if (timeout_s < 0) {
        e->cv_done.wait(lk, finished);
    } else if (!e->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished)) {
        return fail(BFLY_ETIMEDOUT, "request " + std::to_string(id) + " not finished");
    }
    return BFLY_OK;
}

int bfly_engine_fetch(bfly_engine* e, uint64_t id, void* re, void* im, uint64_t len) {
    if (!e) return fail(BFLY_EINVAL, "null engine");
    std::unique_lock<std::mutex> lk(e->mu);
    auto it = e->reqs.find(id);
    if (it == e->reqs.end()) return fail(BFLY_EINVAL, "unknown request " + std::to_string(id));
    Request* r = it->second.get();
    if (r->status != BFLY_REQ_DONE) return fail(BFLY_ESTATE, "request " + std::to_string(id) + " is not done");
    int rc = bfly_state_download(r->state, re, im, len);
    if (rc) return rc;
    free_request_resources(e, r);
    e->reqs.erase(it);
    e->cv_work.notify_all();
    return BFLY_OK;
}

int bfly_engine_release(bfly_engine* e, uint64_t id) {
    if (!e) return fail(BFLY_EINVAL, "null engine");
    std::unique_lock<std::mutex> lk(e->mu);
    auto it = e->reqs.find(id);
    if (it == e->reqs.end()) return fail(BFLY_EINVAL, "unknown request " + std::to_string(id));
    Request* r = it->second.get();
    if (r->status < BFLY_REQ_DONE) return fail(BFLY_ESTATE, "request " + std::to_string(id) + " still in flight");
    free_request_resources(e, r);
    e->reqs.erase(it);
    e->cv_work.notify_all();
    return BFLY_OK;
}

int bfly_engine_cancel(bfly_engine* e, uint64_t id) {
    if (!e) return fail(BFLY_EINVAL, "null engine");
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->reqs.find(id);
    if (it == e->reqs.end()) return fail(BFLY_EINVAL, "unknown request " + std::to_string(id));
    Request* r = it->second.get();
    if (r->status >= BFLY_REQ_DONE) return fail(BFLY_ESTATE, "request already finished");
    r->cancel = true;
    e->cv_work.notify_all();
    return BFLY_OK;
}

int bfly_engine_stats(bfly_engine* e, int64_t* admitted, int64_t* completed, int64_t* rounds, int64_t* launches,
                      uint64_t* bytes_in_use, uint64_t* peak_bytes) {
    if (!e) return fail(BFLY_EINVAL, "null engine");
    std::lock_guard<std::mutex> lk(e->mu);
    if (admitted) *admitted = e->admitted;
    if (completed) *completed = e->completed;
    if (rounds) *rounds = e->rounds;
    if (launches) *launches = e->launches;
    if (bytes_in_use) *bytes_in_use = e->bytes_in_use;
    if (peak_bytes) *peak_bytes = e->peak_bytes;
    return BFLY_OK;
}

int bfly_engine_error(bfly_engine* e, uint64_t id, char* buf, uint64_t len) {
    if (!e || !buf || !len) return fail(BFLY_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->reqs.find(id);
    if (it == e->reqs.end()) return fail(BFLY_EINVAL, "unknown request " + std::to_string(id));
    std::strncpy(buf, it->second->message.c_str(), len - 1);
    buf[len - 1] = 0;
    return BFLY_OK;
}

int bfly_engine_destroy(bfly_engine* e) {
    if (!e) return BFLY_OK;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        e->stop = true;
        for (uint64_t id : e->pending) e->reqs[id]->cancel = true;
        for (uint64_t id : e->running) e->reqs[id]->cancel = true;
        e->cv_work.notify_all();
    }
    if (e->worker.joinable()) e->worker.join();
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto& kv : e->reqs) {
        Request* r = kv.second.get();
        if (r->state) (void)hipStreamSynchronize(r->state->stream);
        free_request_resources(e, r);
    }
    for (hipEvent_t ev : e->slot_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->d_items) (void)hipFree(e->d_items);
    if (e->h_items) (void)hipHostFree(e->h_items);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return BFLY_OK;
}

}  // extern "C"
