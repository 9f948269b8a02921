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
#include <map>
#include <set>
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
    std::vector<bfly_op> ops;      // until a planner thread has compiled them
    std::vector<double> udata;
    bfly_exec_opts exec{};
    bfly_plan* plan = nullptr;
    bfly_state* state = nullptr;
    void* plan_buf = nullptr;      // device records of the plan (pooled with the state)
    size_t plan_cap = 0;
    hipEvent_t done_ev = nullptr;
    uint64_t bytes = 0;
    size_t next_step = 0;
    bool large = false;
    bool issued_all = false;
    bool cancel = false;
    int64_t touched = -1;          // last issue attempt that enqueued work for this request
    std::string message;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

constexpr size_t kSlotItems = 4096;  // batched items per round and tile size (x4 per ring slot)
// Large requests run whole plans on their own streams, at most this many at
// once (more would only share the device's bandwidth and stretch every one).
constexpr int kLargeLanes = 2;
// Steps of every small request issued per round (host work amortised).
constexpr int kSubsteps = 16;

// A state kept for reuse after its request was released, with its plan buffer.
struct PoolEntry {
    bfly_state* state = nullptr;
    void* plan_buf = nullptr;
    size_t plan_cap = 0;
};

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
    // released states by (n, dtype): reused without hipMalloc / hipFree (both
    // synchronise the device); evicted only when a new shape needs the room
    std::map<std::pair<int, int>, std::vector<PoolEntry>> pool;
    uint64_t pool_bytes = 0;
    int64_t rounds = 0, launches = 0, completed = 0, admitted = 0;
    int64_t attempts = 0;        // issue_round calls (marks the requests each one touched)
    // ring of item slots: host (pinned) -> device copy per round
    int nslots = 0;
    int large_running = 0;
    BatchItem* h_items = nullptr;
    BatchItem* d_items = nullptr;
    OpItem* h_ops = nullptr;     // flip / diffuser items, same ring layout
    OpItem* d_ops = nullptr;
    std::vector<hipEvent_t> slot_ev;
    std::vector<bool> slot_used;
    int slot = 0;
    std::thread worker;
    std::vector<std::thread> planners;
    std::deque<uint64_t> to_plan;      // submitted, not yet compiled
    std::set<uint64_t> planning;       // ids submitted and not yet compiled (incl. in progress)
    std::condition_variable cv_plan;
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

// A released request's state (and plan buffer) goes back to the pool.  Its
// work is complete: release is only allowed once the request finished.
void free_request_resources(bfly_engine* e, Request* r) {
    if (r->state) {
        e->pool[{r->n, r->dtype}].push_back(PoolEntry{r->state, r->plan_buf, r->plan_cap});
        e->pool_bytes += r->bytes;
        e->bytes_in_use -= r->bytes;
        r->state = nullptr;
        r->plan_buf = nullptr;
        r->plan_cap = 0;
    }
    if (r->plan) {
        bfly_plan_destroy(r->plan);
        r->plan = nullptr;
    }
    if (r->done_ev) {
        (void)hipEventDestroy(r->done_ev);
        r->done_ev = nullptr;
    }
    std::vector<uint8_t>().swap(r->init_re);
    std::vector<uint8_t>().swap(r->init_im);
}

// A failed request's state may still be written by launches queued before
// the failure: wait for its stream, then free the state instead of pooling it.
void discard_request_resources(bfly_engine* e, Request* r) {
    if (r->state) {
        (void)hipStreamSynchronize(r->state->stream);
        bfly_state_destroy(r->state);
        if (r->plan_buf) (void)hipFree(r->plan_buf);
        e->bytes_in_use -= r->bytes;
        r->state = nullptr;
        r->plan_buf = nullptr;
        r->plan_cap = 0;
    }
    free_request_resources(e, r);
}

void destroy_entry(PoolEntry& pe) {
    if (pe.state) bfly_state_destroy(pe.state);
    if (pe.plan_buf) (void)hipFree(pe.plan_buf);
    pe = PoolEntry{};
}

// Frees pooled states (largest first) until `need` more bytes fit the budget.
void evict_pool(bfly_engine* e, uint64_t need) {
    while (e->bytes_in_use + e->pool_bytes + need > e->opts.budget_bytes && e->pool_bytes > 0) {
        auto best = e->pool.end();
        for (auto it = e->pool.begin(); it != e->pool.end(); ++it)
            if (!it->second.empty() && (best == e->pool.end() || it->first.first > best->first.first)) best = it;
        if (best == e->pool.end()) break;
        PoolEntry pe = best->second.back();
        best->second.pop_back();
        e->pool_bytes -= state_charge(best->first.first, best->first.second);
        destroy_entry(pe);
    }
}

// Allocate (or reuse) and initialise the state of an admitted request; all
// initialisation and the plan upload are enqueued on the engine stream, so
// admission does not wait for the device.  Errors stay with the request
// (FAILED) and never stop the engine.
void admit(bfly_engine* e, Request* r) {
    int rc = BFLY_OK;
    auto& free_list = e->pool[{r->n, r->dtype}];
    if (!free_list.empty()) {
        PoolEntry pe = free_list.back();
        free_list.pop_back();
        e->pool_bytes -= r->bytes;
        r->state = pe.state;
        r->plan_buf = pe.plan_buf;
        r->plan_cap = pe.plan_cap;
    } else {
        evict_pool(e, r->bytes);
        rc = bfly_state_create(&r->state, r->n, r->dtype);
        if (rc != BFLY_OK) r->state = nullptr;
    }
    if (rc == BFLY_OK) {
        e->bytes_in_use += r->bytes;
        e->peak_bytes = std::max(e->peak_bytes, e->bytes_in_use);
        bfly_state* st = r->state;
        const uint64_t size = 1ull << r->n;
        hipError_t err = hipSuccess;
        if (r->init_kind == BFLY_INIT_HOST) {
            const size_t b = size * (r->dtype == BFLY_C64 ? sizeof(float) : sizeof(double));
            err = hipMemcpyAsync(st->re, r->init_re.data(), b, hipMemcpyHostToDevice, e->stream);
            if (err == hipSuccess) err = hipMemcpyAsync(st->im, r->init_im.data(), b, hipMemcpyHostToDevice, e->stream);
        } else if (r->init_kind == BFLY_INIT_RANDOM) {
            rc = fill_random_on(st, r->init_seed, e->stream);
        } else {
            err = launch_reset(r->dtype, st->re, st->im, size, e->stream);
        }
        if (err != hipSuccess) rc = hip_fail(err, "state initialisation");
    }
    if (rc == BFLY_OK) {
        const size_t need = plan_device_bytes(r->plan);
        if (need > r->plan_cap) {
            if (r->plan_buf) (void)hipFree(r->plan_buf);
            r->plan_buf = nullptr;
            r->plan_cap = 0;
            const size_t cap = std::max<size_t>(need + need / 2, 64 * 1024);
            if (hipMalloc(&r->plan_buf, cap) != hipSuccess) {
                (void)hipGetLastError();
                r->plan_buf = nullptr;
                rc = fail(BFLY_ENOMEM, "plan buffer allocation failed");
            } else {
                r->plan_cap = cap;
            }
        }
        if (rc == BFLY_OK && need) rc = upload_plan_into(r->plan, e->device, r->plan_buf, e->stream);
        if (rc == BFLY_OK && !need) r->plan->device = e->device;
    }
    if (rc == BFLY_OK && hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming) != hipSuccess)
        rc = fail(BFLY_EHIP, "hipEventCreate");
    // a large request runs on its state's stream: after the initialisation
    if (rc == BFLY_OK && r->n >= e->opts.large_qubits) {
        if (hipEventRecord(r->done_ev, e->stream) != hipSuccess ||
            hipStreamWaitEvent(r->state->stream, r->done_ev, 0) != hipSuccess)
            rc = fail(BFLY_EHIP, "large request ordering");
    }
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

// First-fit admission over the pending queue (in submission order); no
// request passes one submitted earlier that is still being compiled, and
// once the oldest request has waited 2 s, nothing may pass it, so large
// requests are not starved.
void admit_pending(bfly_engine* e) {
    const double now = e->now();
    bool head_blocked = false;
    const uint64_t first_planning = e->planning.empty() ? UINT64_MAX : *e->planning.begin();
    for (auto it = e->pending.begin(); it != e->pending.end();) {
        Request* r = e->reqs[*it].get();
        if (r->id > first_planning) break;
        if (r->cancel) {
            finish(e, r, BFLY_REQ_CANCELLED);
            it = e->pending.erase(it);
            continue;
        }
        // large requests take one of kLargeLanes lanes, in arrival order
        const bool large = r->n >= e->opts.large_qubits;
        if (large && e->large_running >= kLargeLanes) {
            ++it;
            continue;
        }
        if (e->bytes_in_use + r->bytes <= e->opts.budget_bytes) {
            it = e->pending.erase(it);
            admit(e, r);
            if (r->large) e->large_running += 1;
            continue;
        }
        if (it == e->pending.begin() && now - r->t_submit > 2.0) head_blocked = true;
        if (head_blocked) break;
        ++it;
    }
}

// One round: up to kSubsteps next steps of every small running request (in
// each substep, tile passes grouped by dtype, sweep width and tile size,
// phase flips and whole-state diffusers by dtype, into batched launches;
// other steps launched alone), all items of the round copied to the device
// at once; and the whole remaining plan of every large request on its own
// stream.
int issue_round(bfly_engine* e) {
    std::vector<Request*> small;
    const size_t cap = kSlotItems * 4;  // items per ring slot (each kind)
    const int64_t attempt = ++e->attempts;
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
            r->touched = attempt;
            for (; r->next_step < steps.size(); ++r->next_step) {
                hipError_t err = launch_step(steps[r->next_step], r->state, r->plan, r->state->stream);
                if (err != hipSuccess) return hip_fail(err, "large request launch");
                e->launches += 1;
            }
            (void)hipEventRecord(r->done_ev, r->state->stream);
            r->issued_all = true;
            continue;
        }
        if (e->opts.max_round_requests > 0 && small.size() >= static_cast<size_t>(e->opts.max_round_requests)) continue;
        if (small.size() >= cap) continue;
        small.push_back(r);
    }
    if (small.empty()) return BFLY_OK;

    // reuse a ring slot only after the round that used it has finished
    const int s = e->slot;
    if (e->slot_used[s]) HIP_TRY(hipEventSynchronize(e->slot_ev[s]), "hipEventSynchronize");
    BatchItem* hi = e->h_items + static_cast<size_t>(s) * cap;
    BatchItem* di = e->d_items + static_cast<size_t>(s) * cap;
    OpItem* ho = e->h_ops + static_cast<size_t>(s) * cap;
    OpItem* dop = e->d_ops + static_cast<size_t>(s) * cap;
    enum ActKind { A_STEP, A_TILES, A_OPS, A_DONE };
    struct Act {
        ActKind kind;
        Request* r = nullptr;       // A_STEP, A_DONE
        const Step* st = nullptr;   // A_STEP
        int dtype = 0, slots = 0, k = 0, op = 0;
        size_t off = 0;
        int n = 0;
        uint64_t units = 0;
    };
    std::vector<Act> acts;
    size_t used = 0, oused = 0;
    struct Group { int key; std::vector<Request*> reqs; };
    for (int sub = 0; sub < kSubsteps; ++sub) {
        std::vector<Group> groups, op_groups;
        std::vector<Request*> stepped;
        size_t tile_need = 0, op_need = 0;
        for (Request* r : small) {
            const auto& steps = r->plan->hp.steps;
            if (r->next_step >= steps.size()) continue;
            const Step& st = steps[r->next_step];
            if (st.kind == S_TILE) ++tile_need;
            if (st.kind == S_FLIP || st.kind == S_DIFFUSE) ++op_need;
        }
        if (used + tile_need > cap || oused + op_need > cap) break;
        for (Request* r : small) {
            const auto& steps = r->plan->hp.steps;
            if (r->next_step >= steps.size()) continue;
            const Step& st = steps[r->next_step];
            const bool whole_diffuse = st.kind == S_DIFFUSE && st.mask == (1ull << r->n) - 1;
            if (st.kind == S_TILE) {
                const int key = r->dtype * 1024 + st.tile.slots * 64 + st.tile.k;
                auto g = std::find_if(groups.begin(), groups.end(), [&](const Group& x) { return x.key == key; });
                if (g == groups.end()) groups.push_back({key, {r}});
                else g->reqs.push_back(r);
            } else if (st.kind == S_FLIP || whole_diffuse) {
                const int key = st.kind * 16 + r->dtype;
                auto g = std::find_if(op_groups.begin(), op_groups.end(), [&](const Group& x) { return x.key == key; });
                if (g == op_groups.end()) op_groups.push_back({key, {r}});
                else g->reqs.push_back(r);
            } else {
                Act a{A_STEP};
                a.r = r;
                a.st = &st;
                acts.push_back(a);
            }
            stepped.push_back(r);
        }
        if (stepped.empty()) break;
        for (auto& g : op_groups) {
            uint64_t units = 0;
            const int kind = g.key / 16;
            for (size_t i = 0; i < g.reqs.size(); ++i) {
                Request* r = g.reqs[i];
                const Step& st = r->plan->hp.steps[r->next_step];
                bfly_state* sst = r->state;
                OpItem it = kind == S_FLIP ? flip_item(sst->re, sst->im, sst->n, st.mask, st.controls)
                                           : diffuse_item(sst->re, sst->im, sst->n, st.a, sst->scratch, st.chain,
                                                          st.fmask, st.fvalue);
                it.begin = units;
                units += it.count;
                ho[oused + i] = it;
            }
            Act a{A_OPS};
            a.op = kind;
            a.dtype = g.key % 16;
            a.off = oused;
            a.n = static_cast<int>(g.reqs.size());
            a.units = units;
            acts.push_back(a);
            oused += g.reqs.size();
        }
        for (auto& g : groups) {
            uint64_t tiles = 0;
            for (size_t i = 0; i < g.reqs.size(); ++i) {
                Request* r = g.reqs[i];
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
            Act a{A_TILES};
            a.dtype = g.key / 1024;
            a.slots = (g.key / 64) % 16;
            a.k = g.key % 64;
            a.off = used;
            a.n = static_cast<int>(g.reqs.size());
            a.units = tiles;
            acts.push_back(a);
            used += g.reqs.size();
        }
        for (Request* r : stepped) {
            if (++r->next_step == r->plan->hp.steps.size()) {
                Act a{A_DONE};
                a.r = r;
                acts.push_back(a);
            }
        }
    }
    // requests whose plan has no steps complete at once
    for (Request* r : small) {
        if (r->plan->hp.steps.empty()) {
            Act a{A_DONE};
            a.r = r;
            acts.push_back(a);
        }
    }
    for (Request* r : small) r->touched = attempt;   // this round's launches carry their steps
    if (oused)
        HIP_TRY(hipMemcpyAsync(dop, ho, oused * sizeof(OpItem), hipMemcpyHostToDevice, e->stream), "round op items");
    if (used)
        HIP_TRY(hipMemcpyAsync(di, hi, used * sizeof(BatchItem), hipMemcpyHostToDevice, e->stream), "round items");
    for (const Act& a : acts) {
        switch (a.kind) {
            case A_STEP: {
                hipError_t err = launch_step(*a.st, a.r->state, a.r->plan, e->stream);
                if (err != hipSuccess) return hip_fail(err, "round launch");
                e->launches += 1;
                break;
            }
            case A_OPS:
                if (a.op == S_FLIP)
                    HIP_TRY(launch_flip_items(a.dtype, OpItem{}, dop + a.off, a.n, a.units, e->stream), "batched flips");
                else
                    HIP_TRY(launch_diffuse_items(a.dtype, OpItem{}, dop + a.off, a.n, a.units, e->stream),
                            "batched diffusers");
                e->launches += 1;
                break;
            case A_TILES:
                HIP_TRY(prepare_tile_kernels(a.dtype, a.k, a.slots), "tile kernel preparation");
                HIP_TRY(launch_tile_batch(a.dtype, a.k, a.slots, di + a.off, a.n, a.units, e->stream), "batched launch");
                e->launches += 1;
                break;
            case A_DONE:
                HIP_TRY(hipEventRecord(a.r->done_ev, e->stream), "hipEventRecord");
                a.r->issued_all = true;
                break;
        }
    }
    HIP_TRY(hipEventRecord(e->slot_ev[s], e->stream), "hipEventRecord");
    e->slot_used[s] = true;
    e->slot = (s + 1) % e->nslots;
    e->rounds += 1;
    return BFLY_OK;
}

void reap(bfly_engine* e) {
    for (auto it = e->running.begin(); it != e->running.end();) {
        Request* r = e->reqs[*it].get();
        if (r->issued_all && hipEventQuery(r->done_ev) == hipSuccess) {
            if (r->large) e->large_running -= 1;
            finish(e, r, r->cancel ? BFLY_REQ_CANCELLED : BFLY_REQ_DONE);
            if (r->cancel) free_request_resources(e, r);
            it = e->running.erase(it);
        } else {
            ++it;
        }
    }
}

// Compiles submitted circuits off the submitting thread; a compiled request
// joins the admission queue.
void planner_loop(bfly_engine* e) {
    std::unique_lock<std::mutex> lk(e->mu);
    while (true) {
        e->cv_plan.wait(lk, [&] { return e->stop || !e->to_plan.empty(); });
        if (e->to_plan.empty()) {
            if (e->stop) return;
            continue;
        }
        const uint64_t id = e->to_plan.front();
        e->to_plan.pop_front();
        Request* r = e->reqs[id].get();
        lk.unlock();
        bfly_plan* plan = nullptr;
        int rc = BFLY_OK;
        std::string msg;
        if (!r->cancel) {
            rc = bfly_plan_create(&plan, r->n, r->dtype, r->ops.data(), r->ops.size(),
                                  r->udata.empty() ? nullptr : r->udata.data(), r->udata.size() / 2, &r->exec);
            if (rc != BFLY_OK) {
                msg = bfly_last_error();
            } else if (plan->hp.measures) {
                rc = BFLY_EUNSUPPORTED;
                msg = "measurement runs through bfly_plan_execute, not in the engine";
                bfly_plan_destroy(plan);
                plan = nullptr;
            }
        }
        std::vector<bfly_op>().swap(r->ops);
        std::vector<double>().swap(r->udata);
        lk.lock();
        e->planning.erase(id);
        r->plan = plan;
        if (r->cancel) {
            if (plan) bfly_plan_destroy(plan);
            r->plan = nullptr;
            finish(e, r, BFLY_REQ_CANCELLED);
        } else if (rc != BFLY_OK) {
            r->message = msg;
            finish(e, r, BFLY_REQ_FAILED);
        } else {
            // admission follows submission order: keep the queue sorted by id
            auto pos = e->pending.end();
            while (pos != e->pending.begin() && *(pos - 1) > id) --pos;
            e->pending.insert(pos, id);
        }
        e->cv_work.notify_all();
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
            // fail the requests the failed round enqueued work for (the others
            // keep running); their states are discarded, not pooled
            const std::string msg = bfly_last_error();
            (void)hipStreamSynchronize(e->stream);
            for (uint64_t id : e->running) {
                Request* r = e->reqs[id].get();
                if (r->touched != e->attempts) continue;
                if (r->large && r->issued_all) continue;
                r->message = msg;
                if (r->large) e->large_running -= 1;
                discard_request_resources(e, r);
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
    if (e->opts.large_qubits <= 0) e->opts.large_qubits = 21;
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
    const size_t ops_bytes = static_cast<size_t>(e->nslots) * kSlotItems * 4 * sizeof(OpItem);
    if (err == hipSuccess) err = hipHostMalloc(reinterpret_cast<void**>(&e->h_ops), ops_bytes, hipHostMallocDefault);
    if (err == hipSuccess) err = hipMalloc(&e->d_ops, ops_bytes);
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
    const int np = e->opts.planner_threads > 0 ? std::min(e->opts.planner_threads, 64) : 8;
    for (int i = 0; i < np; ++i) e->planners.emplace_back(planner_loop, e);
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
    if (rq->dtype != BFLY_C128 && rq->dtype != BFLY_C64) return fail(BFLY_EINVAL, "unknown dtype");
    if (rq->nops && !rq->ops) return fail(BFLY_EINVAL, "null op array");
    if (rq->exec.mode < BFLY_MODE_FUSED || rq->exec.mode > BFLY_MODE_AUTO)
        return fail(BFLY_EINVAL, "unknown execution mode");
    // invalid circuits are refused here; compilation runs on a planner thread
    std::string err;
    int rc = validate_ops(rq->num_qubits, rq->ops, rq->nops, rq->udata_len, err);
    if (rc) return fail(rc, err);
    r->ops.assign(rq->ops, rq->ops + rq->nops);
    if (rq->udata && rq->udata_len) r->udata.assign(rq->udata, rq->udata + 2 * rq->udata_len);
    r->exec = rq->exec;
    std::lock_guard<std::mutex> lk(e->mu);
    r->id = e->next_id++;
    r->t_submit = e->now();
    *id = r->id;
    e->to_plan.push_back(r->id);
    e->planning.insert(r->id);
    e->reqs[r->id] = std::move(r);
    e->cv_plan.notify_one();
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

int bfly_engine_fetch_range(bfly_engine* e, uint64_t id, uint64_t first, uint64_t count, void* re, void* im) {
    if (!e) return fail(BFLY_EINVAL, "null engine");
    if (count && (!re || !im)) return fail(BFLY_EINVAL, "null output buffer");
    std::unique_lock<std::mutex> lk(e->mu);
    auto it = e->reqs.find(id);
    if (it == e->reqs.end()) return fail(BFLY_EINVAL, "unknown request " + std::to_string(id));
    Request* r = it->second.get();
    if (r->status != BFLY_REQ_DONE) return fail(BFLY_ESTATE, "request " + std::to_string(id) + " is not done");
    const uint64_t size = 1ull << r->n;
    if (first > size || count > size - first)
        return fail(BFLY_EINVAL, "amplitude range [" + std::to_string(first) + ", +" + std::to_string(count) +
                                     ") outside a state of " + std::to_string(size));
    if (!count) return BFLY_OK;
    bfly_state* st = r->state;
    const size_t esz = st->dtype == BFLY_C64 ? sizeof(float) : sizeof(double);
    HIP_TRY(hipSetDevice(st->device), "hipSetDevice");
    HIP_TRY(hipMemcpyAsync(re, static_cast<char*>(st->re) + first * esz, count * esz, hipMemcpyDeviceToHost,
                           st->stream),
            "hipMemcpyAsync(range re)");
    HIP_TRY(hipMemcpyAsync(im, static_cast<char*>(st->im) + first * esz, count * esz, hipMemcpyDeviceToHost,
                           st->stream),
            "hipMemcpyAsync(range im)");
    HIP_TRY(hipStreamSynchronize(st->stream), "hipStreamSynchronize");
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
        for (uint64_t id : e->to_plan) e->reqs[id]->cancel = true;
        e->cv_work.notify_all();
        e->cv_plan.notify_all();
    }
    for (std::thread& t : e->planners)
        if (t.joinable()) t.join();
    if (e->worker.joinable()) e->worker.join();
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto& kv : e->reqs) {
        Request* r = kv.second.get();
        if (r->state) (void)hipStreamSynchronize(r->state->stream);
        free_request_resources(e, r);
    }
    for (auto& kv : e->pool)
        for (PoolEntry& pe : kv.second) destroy_entry(pe);
    e->pool.clear();
    for (hipEvent_t ev : e->slot_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->d_items) (void)hipFree(e->d_items);
    if (e->h_items) (void)hipHostFree(e->h_items);
    if (e->d_ops) (void)hipFree(e->d_ops);
    if (e->h_ops) (void)hipHostFree(e->h_ops);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return BFLY_OK;
}

}  // extern "C"
