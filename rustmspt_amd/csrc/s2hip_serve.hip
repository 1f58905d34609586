// Serving loop: S2 requests of mixed grid sizes, r_max and method (exact or
// FP32 FFT), admitted as device memory and slots free up.
//
// One dispatcher thread owns admission: requests wait in arrival order; the
// first waiting request whose estimated device footprint fits the remaining HBM
// budget is handed to a free slot (a request skipped because it does not fit
// counts one deferral).  Each slot is a worker thread with its own context and
// HIP stream, so admitted requests run concurrently and their kernels overlap on
// the device.  A request's volume is uploaded when it is admitted and stays
// resident in HBM until its result is finished.  Results are bit-identical to the
// same request run alone (all device reductions are integer or fixed-order).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "s2hip_internal.h"

namespace s2h {

void exact_plans_clear(s2h_ctx *ctx);

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t ticket;
    uint64_t tag;
    std::vector<uint64_t> words;
    uint64_t dims[3];
    uint32_t r_max;
    double pitch;
    int32_t method;
    uint64_t footprint;
    double t_submit = 0, t_start = 0, t_done = 0;
    int32_t status = -1;  // -1 pending, S2H_OK or an error code
    uint32_t deferrals = 0;
    std::string error;
    std::vector<double> s2;
};

// Device bytes one request holds while admitted: packed volume, shell plan
// (offsets, gather, per-offset hits and terms, dense table) and, unless they
// fit the slot's FFT arena (reserved apart from the admission budget), the
// pruned spectra -- FP32 for FFT32; an exact request whose FP64 spectra do
// not fit the arena counts directly and holds none.  Shell size bound: the
// ball of radius r_max/pitch + 1.
uint64_t footprint(const uint64_t dims[3], uint32_t r_max, double pitch, int32_t method, uint64_t arena) {
    const double Rb = (double)r_max / pitch + 0.5 / pitch + 1.0;
    double K = 4.18879 * Rb * Rb * Rb;
    const double box = (2.0 * dims[0] - 1) * (2.0 * dims[1] - 1) * (2.0 * dims[2] - 1);
    K = std::min(K, box);
    const uint64_t row_words = (dims[2] + 63) / 64 + 1;
    double bytes = (double)(dims[0] * dims[1] * row_words * 8 + kVolTailBytes);
    bytes += K * (12 + 4 + 8 + 8) + K * 8 / 2 + 65536.0 * 64;
    if (method == 1) {
        int R[3];
        for (int a = 0; a < 3; ++a) R[a] = (int)std::min<double>((double)dims[a] - 1, std::floor(Rb));
        const uint64_t spectra = fft_plan_buffer_bytes(dims, R, 4);
        if (!arena || spectra > arena) bytes += (double)spectra;
    }
    return (uint64_t)(bytes * 1.1) + (1ull << 20);
}

}  // namespace
}  // namespace s2h

struct s2h_server {
    int device = 0;
    uint64_t budget = 0;
    uint64_t in_use = 0;
    uint64_t peak = 0;
    uint64_t next_ticket = 1;
    uint64_t admitted = 0, deferred = 0, completed = 0, failed = 0;
    bool stop = false;
    s2h::Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_dispatch, cv_done;
    std::deque<s2h::Req *> pending;
    std::unordered_map<uint64_t, s2h::Req *> all;
    struct Slot {
        s2h_ctx *ctx = nullptr;
        s2h::Req *job = nullptr;
        std::thread th;
        std::condition_variable cv;
        uint64_t held = 0;  // device bytes the slot's context holds (in use + cached) after its last job
    };
    std::vector<Slot *> slots;
    std::thread dispatcher;
    uint64_t slot_plan_cap = 8ull << 30;
    uint64_t arena_bytes = 2ull << 30;  // each slot's FFT arena
    uint64_t admit_budget = 0;  // budget minus the slots' plan-cache allowances and arenas
    uint64_t over_budget = 0;   // jobs whose context held more device memory than its share
};

namespace s2h {
namespace {

double now_s(s2h_server *sv) {
    return std::chrono::duration<double>(Clock::now() - sv->t0).count();
}

// Runs one request on the slot's context.  Touches only the request's inputs
// and its result vector; the status is published by the caller under the lock.
s2h_status run_job(s2h_ctx *ctx, Req *r, std::string &err) {
    s2h_vol *vol = nullptr;
    s2h_status st = s2h_vol_from_packed(ctx, r->words.data(), r->words.size(), r->dims, &vol);
    if (st == S2H_OK) {
        r->s2.assign(r->r_max + 1, 0.0);
        st = r->method == 1 ? s2h_fft32_s2(ctx, vol, r->r_max, r->pitch, r->s2.data(), r->s2.size())
                            : s2h_exact_s2(ctx, vol, r->r_max, r->pitch, r->s2.data(), r->s2.size());
    }
    if (st != S2H_OK) err = s2h_last_error();
    if (vol) s2h_vol_free(ctx, vol);
    return st;
}

void worker_loop(s2h_server *sv, s2h_server::Slot *slot) {
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        slot->cv.wait(lk, [&] { return sv->stop || slot->job != nullptr; });
        if (!slot->job) return;
        Req *r = slot->job;
        r->t_start = now_s(sv);
        // the context may hold the request's admitted footprint on top of its
        // cached plans (at most the slot's plan allowance): its device memory,
        // cached blocks included, stays within that, so the slots together
        // stay within the server's budget
        const uint64_t keep = sv->slot_plan_cap + sv->arena_bytes;  // what an idle slot may hold
        const uint64_t job_budget = sv->budget ? r->footprint + keep : 0;
        lk.unlock();
        s2h_ctx *ctx = slot->ctx;
        ctx->budget = job_budget;
        ctx->held_peak = ctx->in_use + ctx->cached_bytes;
        std::string err;
        s2h_status st = run_job(ctx, r, err);
        if (st == S2H_ERR_BUDGET) {
            // the plans cached from earlier requests left too little room: drop
            // them, give their blocks back to the device and run once more
            exact_plans_clear(ctx);
            dev_cache_release(ctx);
            err.clear();
            st = run_job(ctx, r, err);
        }
        const double t_done = now_s(sv);
        // cached plans beyond the per-slot cap are dropped between requests,
        // and under a budget the idle context keeps at most the allowance
        // (its FFT arena included)
        if (ctx->in_use > keep) exact_plans_clear(ctx);
        if (sv->budget && ctx->in_use + ctx->cached_bytes > keep)
            dev_cache_trim(ctx, keep > ctx->in_use ? keep - ctx->in_use : 0);
        const uint64_t held = ctx->in_use + ctx->cached_bytes;
        const bool over = job_budget && ctx->held_peak > job_budget;
        lk.lock();
        slot->held = held;
        if (over) ++sv->over_budget;
        // publishing the status is the worker's last access to the request: a
        // poller may release it as soon as the lock is dropped
        sv->in_use -= r->footprint;
        if (st == S2H_OK) ++sv->completed;
        else ++sv->failed;
        r->words.clear();
        r->words.shrink_to_fit();
        r->error = err;
        r->t_done = t_done;
        r->status = st;
        slot->job = nullptr;
        sv->cv_dispatch.notify_one();
        sv->cv_done.notify_all();
    }
}

void dispatch_loop(s2h_server *sv) {
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        sv->cv_dispatch.wait(lk, [&] {
            if (sv->stop) return true;
            if (sv->pending.empty()) return false;
            for (auto *s : sv->slots)
                if (!s->job) return true;
            return false;
        });
        if (sv->stop) return;
        bool progress = true;
        while (progress && !sv->pending.empty()) {
            progress = false;
            s2h_server::Slot *free_slot = nullptr;
            for (auto *s : sv->slots)
                if (!s->job) {
                    free_slot = s;
                    break;
                }
            if (!free_slot) break;
            for (auto it = sv->pending.begin(); it != sv->pending.end(); ++it) {
                Req *r = *it;
                if (sv->budget && sv->in_use + r->footprint > sv->admit_budget) {
                    if (r->deferrals++ == 0) ++sv->deferred;  // requests ever passed over
                    continue;
                }
                sv->pending.erase(it);
                sv->in_use += r->footprint;
                sv->peak = std::max(sv->peak, sv->in_use);
                ++sv->admitted;
                free_slot->job = r;
                free_slot->cv.notify_one();
                progress = true;
                break;
            }
        }
        if (!sv->pending.empty()) {
            // nothing admissible right now: wait for a completion
            sv->cv_dispatch.wait_for(lk, std::chrono::milliseconds(50));
        }
    }
}

}  // namespace

void exact_plans_clear(s2h_ctx *ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->exact_plans) exact_plan_release(ctx, kv.second.get());
    ctx->exact_plans.clear();
    shell_devs_release(ctx);
    for (auto &kv : ctx->fft_plans) fft_plan_release(ctx, kv.second);
    ctx->fft_plans.clear();
    for (auto &kv : ctx->fftd_plans) fft_plan_release(ctx, kv.second);
    ctx->fftd_plans.clear();
}

}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_serve_create(int device, uint64_t hbm_budget_bytes, uint32_t nslots, s2h_server **out) {
    if (!out || nslots == 0 || nslots > 64) return fail(S2H_ERR_ARG, "need 1..64 slots");
    auto *sv = new s2h_server();
    sv->device = device;
    sv->budget = hbm_budget_bytes;
    sv->t0 = Clock::now();
    if (hbm_budget_bytes) {
        // a quarter of the budget is left to the slots' cached shell plans and
        // a quarter to their FFT arenas; requests are admitted against the rest
        sv->slot_plan_cap = std::max<uint64_t>(hbm_budget_bytes / (4 * nslots), 64ull << 20);
        sv->arena_bytes = hbm_budget_bytes / (4 * nslots);
        sv->admit_budget =
            hbm_budget_bytes - std::min(hbm_budget_bytes, (sv->slot_plan_cap + sv->arena_bytes) * nslots);
    }
    for (uint32_t i = 0; i < nslots; ++i) {
        auto *slot = new s2h_server::Slot();
        s2h_status st = s2h_create(device, 0, &slot->ctx);
        // a stream of mixed shapes sees most (grid, r_max) pairs once: take the
        // cost model's launch shape (or one timed earlier in the process)
        if (st == S2H_OK) st = s2h_set_autotune(slot->ctx, 0);
        // the FFT arena, reserved once: spectra that fit it are never
        // allocated or freed on the request path (a hipFree stalls every
        // slot); an exact request takes the FP64 FFT kernel only where its
        // spectra fit the arena
        if (st == S2H_OK) st = dev_alloc(slot->ctx, slot->ctx->fft_arena, sv->arena_bytes);
        if (st == S2H_OK) slot->ctx->fft_budget = sv->arena_bytes;
        if (st != S2H_OK) {
            if (slot->ctx) s2h_destroy(slot->ctx);
            delete slot;
            s2h_serve_destroy(sv);
            return st;
        }
        sv->slots.push_back(slot);
    }
    for (auto *slot : sv->slots) slot->th = std::thread(worker_loop, sv, slot);
    sv->dispatcher = std::thread(dispatch_loop, sv);
    *out = sv;
    return S2H_OK;
}

void s2h_serve_destroy(s2h_server *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        sv->cv_dispatch.notify_all();
        for (auto *s : sv->slots) s->cv.notify_all();
    }
    if (sv->dispatcher.joinable()) sv->dispatcher.join();
    for (auto *s : sv->slots) {
        if (s->th.joinable()) s->th.join();
        if (s->ctx) s2h_destroy(s->ctx);
        delete s;
    }
    for (auto &kv : sv->all) delete kv.second;
    delete sv;
}

s2h_status s2h_serve_submit(s2h_server *sv, const s2h_request *req, uint64_t *ticket) {
    if (!sv || !req || !ticket || !req->words) return fail(S2H_ERR_ARG, "null argument");
    if (req->method != 0 && req->method != 1) return fail(S2H_ERR_ARG, "method must be 0 (exact) or 1 (fft32)");
    if (!(req->pitch > 0.0)) return fail(S2H_ERR_ARG, "pitch must be > 0");
    for (int a = 0; a < 3; ++a)
        if (req->dims[a] == 0 || req->dims[a] > 65535) return fail(S2H_ERR_ARG, "grid dimension outside 1..65535");
    const uint64_t need = req->dims[0] * req->dims[1] * ((req->dims[2] + 63) / 64 + 1);
    if (req->nwords != need) return fail(S2H_ERR_SIZE, "packed volume length mismatch");
    auto *r = new Req();
    r->tag = req->tag;
    r->words.assign(req->words, req->words + req->nwords);
    for (int a = 0; a < 3; ++a) r->dims[a] = req->dims[a];
    r->r_max = req->r_max;
    r->pitch = req->pitch;
    r->method = req->method;
    r->footprint = footprint(r->dims, r->r_max, r->pitch, r->method, sv->arena_bytes);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->budget && r->footprint > sv->admit_budget) {
        delete r;
        return fail(S2H_ERR_BUDGET, "request footprint exceeds the server's device budget");
    }
    r->ticket = sv->next_ticket++;
    r->t_submit = now_s(sv);
    sv->all[r->ticket] = r;
    sv->pending.push_back(r);
    *ticket = r->ticket;
    sv->cv_dispatch.notify_one();
    return S2H_OK;
}

s2h_status s2h_serve_poll(s2h_server *sv, uint64_t ticket, int32_t wait, s2h_result *res, double *s2, uint64_t n) {
    if (!sv || !res) return fail(S2H_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->all.find(ticket);
    if (it == sv->all.end()) return fail(S2H_ERR_ARG, "unknown ticket");
    Req *r = it->second;
    if (wait) sv->cv_done.wait(lk, [&] { return r->status != -1; });
    res->ready = r->status != -1;
    res->tag = r->tag;
    res->status = r->status;
    res->deferrals = r->deferrals;
    res->footprint = r->footprint;
    res->t_submit = r->t_submit;
    res->t_start = r->t_start;
    res->t_done = r->t_done;
    if (!res->ready) return S2H_OK;
    if (r->status != S2H_OK) {
        std::string msg = r->error;
        sv->all.erase(it);
        delete r;
        return fail(S2H_ERR_HIP, "request failed: " + msg);
    }
    if (s2) {
        if (n != r->s2.size()) return fail(S2H_ERR_SIZE, "s2 buffer must hold r_max+1 values");
        std::memcpy(s2, r->s2.data(), n * 8);
    }
    sv->all.erase(it);
    delete r;
    return S2H_OK;
}

s2h_status s2h_serve_stats(s2h_server *sv, uint64_t *stats, uint64_t n) {
    if (!sv || !stats) return fail(S2H_ERR_ARG, "null argument");
    if (n != 8 && n != 10) return fail(S2H_ERR_SIZE, "stats buffer must hold 8 or 10 values");
    std::lock_guard<std::mutex> lk(sv->mu);
    stats[0] = sv->admitted;
    stats[1] = sv->deferred;
    stats[2] = sv->completed;
    stats[3] = sv->failed;
    stats[4] = sv->pending.size();
    stats[5] = sv->in_use;
    stats[6] = sv->peak;
    stats[7] = sv->budget;
    if (n == 10) {
        uint64_t held = 0;
        for (auto *s : sv->slots) held += s->held;
        stats[8] = held;
        stats[9] = sv->over_budget;
    }
    return S2H_OK;
}

}  // extern "C"
