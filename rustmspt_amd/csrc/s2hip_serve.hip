// Serving loop: S2 requests of mixed grid sizes, r_max and method (exact or
// FP32 FFT), admitted as device memory and slots free up.
//
// The queue, admission and slot dispatch are the device-independent core
// (s2hip_serve_core.h); this file is its device backend.  Each slot is a
// worker thread with its own context and HIP stream, so admitted requests run
// concurrently and their kernels overlap on the device.  A request's volume is
// uploaded when it is admitted and stays resident in HBM until its result is
// finished.  Results are bit-identical to the same request run alone (all
// device reductions are integer or fixed-order).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "s2hip_internal.h"
#include "s2hip_serve_core.h"

namespace s2h {

void exact_plans_clear(s2h_ctx *ctx);

namespace {

using serve::JobOutcome;
using serve::Req;

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

// Runs one request on the slot's context.  Touches only the request's inputs
// and its result vector.
s2h_status run_job(s2h_ctx *ctx, Req &r, std::string &err) {
    s2h_vol *vol = nullptr;
    s2h_status st = s2h_vol_from_packed(ctx, r.words.data(), r.words.size(), r.dims, &vol);
    if (st == S2H_OK) {
        r.s2.assign(r.r_max + 1, 0.0);
        st = r.method == 1 ? s2h_fft32_s2(ctx, vol, r.r_max, r.pitch, r.s2.data(), r.s2.size())
                           : s2h_exact_s2(ctx, vol, r.r_max, r.pitch, r.s2.data(), r.s2.size());
    }
    if (st != S2H_OK) err = s2h_last_error();
    if (vol) s2h_vol_free(ctx, vol);
    return st;
}

// The device backend: the request runs under a context budget of its share
// (its admitted footprint plus what an idle slot keeps: the plan allowance and
// the FFT arena); after it, plans beyond the allowance are dropped and cached
// blocks trimmed, so the slots together stay within the server's budget.
struct HipBackend {
    using Ctx = s2h_ctx;
    uint64_t budget = 0, keep = 0;
    JobOutcome run(s2h_ctx *ctx, Req &r, uint64_t share) {
        JobOutcome out;
        ctx->budget = share;
        ctx->held_peak = ctx->in_use + ctx->cached_bytes;
        out.status = run_job(ctx, r, out.error);
        if (out.status == S2H_ERR_BUDGET) {
            // the plans cached from earlier requests left too little room: drop
            // them, give their blocks back to the device and run once more
            exact_plans_clear(ctx);
            dev_cache_release(ctx);
            out.error.clear();
            out.status = run_job(ctx, r, out.error);
        }
        if (ctx->in_use > keep) exact_plans_clear(ctx);
        if (budget && ctx->in_use + ctx->cached_bytes > keep)
            dev_cache_trim(ctx, keep > ctx->in_use ? keep - ctx->in_use : 0);
        out.held = ctx->in_use + ctx->cached_bytes;
        out.over = share && ctx->held_peak > share;
        return out;
    }
};

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

struct s2h_server {
    int device = 0;
    uint64_t budget = 0;
    uint64_t slot_plan_cap = 8ull << 30;  // cached shell plans a slot keeps
    uint64_t arena_bytes = 2ull << 30;    // each slot's FFT arena
    s2h::serve::Core<s2h::HipBackend> *core = nullptr;
};

using namespace s2h;

extern "C" {

s2h_status s2h_serve_create(int device, uint64_t hbm_budget_bytes, uint32_t nslots, s2h_server **out) {
    if (!out || nslots == 0 || nslots > 64) return fail(S2H_ERR_ARG, "need 1..64 slots");
    auto *sv = new s2h_server();
    sv->device = device;
    sv->budget = hbm_budget_bytes;
    uint64_t admit = 0;
    if (hbm_budget_bytes) {
        // a quarter of the budget is left to the slots' cached shell plans and
        // a quarter to their FFT arenas; requests are admitted against the rest
        sv->slot_plan_cap = std::max<uint64_t>(hbm_budget_bytes / (4 * nslots), 64ull << 20);
        sv->arena_bytes = hbm_budget_bytes / (4 * nslots);
        admit = hbm_budget_bytes - std::min(hbm_budget_bytes, (sv->slot_plan_cap + sv->arena_bytes) * nslots);
    }
    std::vector<s2h_ctx *> ctxs;
    s2h_status st = S2H_OK;
    for (uint32_t i = 0; i < nslots && st == S2H_OK; ++i) {
        s2h_ctx *ctx = nullptr;
        st = s2h_create(device, 0, &ctx);
        // a stream of mixed shapes sees most (grid, r_max) pairs once: take the
        // cost model's launch shape (or one timed earlier in the process)
        if (st == S2H_OK) st = s2h_set_autotune(ctx, 0);
        // the FFT arena, reserved once: spectra that fit it are never
        // allocated or freed on the request path (a hipFree stalls every
        // slot); an exact request takes the FP64 FFT kernel only where its
        // spectra fit the arena
        if (st == S2H_OK) st = dev_alloc(ctx, ctx->fft_arena, sv->arena_bytes);
        if (st == S2H_OK) ctx->fft_budget = sv->arena_bytes;
        if (ctx) ctxs.push_back(ctx);
    }
    if (st != S2H_OK) {
        for (s2h_ctx *c : ctxs) s2h_destroy(c);
        delete sv;
        return st;
    }
    HipBackend be;
    be.budget = hbm_budget_bytes;
    be.keep = sv->slot_plan_cap + sv->arena_bytes;
    sv->core = new serve::Core<HipBackend>(be, hbm_budget_bytes, admit, be.keep);
    sv->core->start(ctxs);
    *out = sv;
    return S2H_OK;
}

void s2h_serve_destroy(s2h_server *sv) {
    if (!sv) return;
    if (sv->core) {
        for (s2h_ctx *c : sv->core->stop()) s2h_destroy(c);
        delete sv->core;
    }
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
    const s2h_status st = sv->core->submit(r, ticket);
    if (st == S2H_ERR_BUDGET) return fail(st, "request footprint exceeds the server's device budget");
    return st;
}

s2h_status s2h_serve_poll(s2h_server *sv, uint64_t ticket, int32_t wait, s2h_result *res, double *s2, uint64_t n) {
    if (!sv || !res) return fail(S2H_ERR_ARG, "null argument");
    std::string err;
    const s2h_status st = sv->core->poll(ticket, wait != 0, res, s2, n, err);
    return st == S2H_OK ? st : fail(st, err);
}

s2h_status s2h_serve_stats(s2h_server *sv, uint64_t *stats, uint64_t n) {
    if (!sv || !stats) return fail(S2H_ERR_ARG, "null argument");
    if (n != 8 && n != 10) return fail(S2H_ERR_SIZE, "stats buffer must hold 8 or 10 values");
    uint64_t all[10];
    sv->core->stats(all);
    std::memcpy(stats, all, n * 8);
    return S2H_OK;
}

}  // extern "C"
