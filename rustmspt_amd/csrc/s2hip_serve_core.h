// The serving loop's request queue, admission and slot dispatch, independent of
// the device: a Backend runs one admitted request on a slot's context.  The
// HIP library instantiates it with its device backend (s2hip_serve.hip); the
// sanitizer stress test instantiates it with a host-only backend, so the
// threading (submitters, one dispatcher, one worker per slot, pollers) is
// checked under ThreadSanitizer without a GPU.
//
// One dispatcher thread owns admission: requests wait in arrival order; the
// first waiting request whose estimated footprint fits the remaining budget is
// handed to a free slot (a request skipped because it does not fit counts one
// deferral).  Each slot is a worker thread with its own context.  A request
// lives in `all` from submission until a poll returns it finished; the worker's
// last access to it is publishing its status under the lock.
#pragma once
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/s2hip.h"

namespace s2h {
namespace serve {

using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t ticket = 0;
    uint64_t tag = 0;
    std::vector<uint64_t> words;
    uint64_t dims[3] = {0, 0, 0};
    uint32_t r_max = 0;
    double pitch = 1.0;
    int32_t method = 0;
    uint64_t footprint = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int32_t status = -1;  // -1 pending, S2H_OK or an error code
    uint32_t deferrals = 0;
    std::string error;
    std::vector<double> s2;
};

// What a backend reports after running one request on a slot's context.
struct JobOutcome {
    s2h_status status = S2H_OK;
    std::string error;
    uint64_t held = 0;  // device bytes the context holds afterwards
    bool over = false;  // the context exceeded the request's share while it ran
};

// Backend: `using Ctx = ...;` and
//   JobOutcome run(Ctx *ctx, Req &r, uint64_t budget_share)
// touching only the request's inputs and its s2 vector.
template <class Backend>
class Core {
  public:
    using Ctx = typename Backend::Ctx;

    Core(Backend backend, uint64_t budget, uint64_t admit_budget, uint64_t keep_per_slot)
        : backend_(std::move(backend)), budget_(budget), admit_budget_(admit_budget), keep_(keep_per_slot),
          t0_(Clock::now()) {}

    // Takes ownership of the contexts; starts the slot workers and the dispatcher.
    void start(const std::vector<Ctx *> &ctxs) {
        for (Ctx *c : ctxs) {
            auto *s = new Slot();
            s->ctx = c;
            slots_.push_back(s);
        }
        for (Slot *s : slots_) s->th = std::thread(&Core::worker_loop, this, s);
        dispatcher_ = std::thread(&Core::dispatch_loop, this);
    }

    // Stops the threads; returns the slots' contexts for the caller to release.
    std::vector<Ctx *> stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            cv_dispatch_.notify_all();
            for (Slot *s : slots_) s->cv.notify_all();
        }
        if (dispatcher_.joinable()) dispatcher_.join();
        std::vector<Ctx *> ctxs;
        for (Slot *s : slots_) {
            if (s->th.joinable()) s->th.join();
            ctxs.push_back(s->ctx);
            delete s;
        }
        slots_.clear();
        for (auto &kv : all_) delete kv.second;
        all_.clear();
        pending_.clear();
        return ctxs;
    }

    // Queues r (taking ownership).  S2H_ERR_BUDGET (r deleted) when its
    // footprint exceeds what the budget can ever admit.
    s2h_status submit(Req *r, uint64_t *ticket) {
        std::lock_guard<std::mutex> lk(mu_);
        if (budget_ && r->footprint > admit_budget_) {
            delete r;
            return S2H_ERR_BUDGET;
        }
        r->ticket = next_ticket_++;
        r->t_submit = now_s();
        all_[r->ticket] = r;
        pending_.push_back(r);
        *ticket = r->ticket;
        cv_dispatch_.notify_one();
        return S2H_OK;
    }

    // The request's state; once finished, its curve (s2 may be null) and the
    // request released.  S2H_ERR_ARG: unknown ticket; S2H_ERR_SIZE: s2 holds
    // the wrong count (the request stays); the request's own error status
    // otherwise, with its message in err.
    s2h_status poll(uint64_t ticket, bool wait, s2h_result *res, double *s2, uint64_t n, std::string &err) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = all_.find(ticket);
        if (it == all_.end()) {
            err = "unknown ticket";
            return S2H_ERR_ARG;
        }
        Req *r = it->second;
        if (wait) cv_done_.wait(lk, [&] { return r->status != -1; });
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
            err = "request failed: " + r->error;
            all_.erase(it);
            delete r;
            return S2H_ERR_HIP;
        }
        if (s2) {
            if (n != r->s2.size()) {
                err = "s2 buffer must hold r_max+1 values";
                return S2H_ERR_SIZE;
            }
            std::memcpy(s2, r->s2.data(), n * 8);
        }
        all_.erase(it);
        delete r;
        return S2H_OK;
    }

    // admitted, deferred, completed, failed, waiting, bytes in use (estimates),
    // peak, budget, device bytes the slots hold, requests over their share.
    void stats(uint64_t out[10]) {
        std::lock_guard<std::mutex> lk(mu_);
        out[0] = admitted_;
        out[1] = deferred_;
        out[2] = completed_;
        out[3] = failed_;
        out[4] = pending_.size();
        out[5] = in_use_;
        out[6] = peak_;
        out[7] = budget_;
        uint64_t held = 0;
        for (Slot *s : slots_) held += s->held;
        out[8] = held;
        out[9] = over_budget_;
    }

  private:
    struct Slot {
        Ctx *ctx = nullptr;
        Req *job = nullptr;
        std::thread th;
        std::condition_variable cv;
        uint64_t held = 0;  // device bytes the slot's context holds after its last job
    };

    double now_s() const { return std::chrono::duration<double>(Clock::now() - t0_).count(); }

    void worker_loop(Slot *slot) {
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            slot->cv.wait(lk, [&] { return stop_ || slot->job != nullptr; });
            if (!slot->job) return;
            Req *r = slot->job;
            r->t_start = now_s();
            // the context may hold the request's admitted footprint on top of
            // what an idle slot keeps, so the slots together stay within the budget
            const uint64_t share = budget_ ? r->footprint + keep_ : 0;
            lk.unlock();
            JobOutcome out = backend_.run(slot->ctx, *r, share);
            const double t_done = now_s();
            lk.lock();
            slot->held = out.held;
            if (out.over) ++over_budget_;
            in_use_ -= r->footprint;
            if (out.status == S2H_OK) ++completed_;
            else ++failed_;
            r->words.clear();
            r->words.shrink_to_fit();
            r->error = out.error;
            r->t_done = t_done;
            // publishing the status is the worker's last access to the request:
            // a poller may release it as soon as the lock is dropped
            r->status = out.status;
            slot->job = nullptr;
            cv_dispatch_.notify_one();
            cv_done_.notify_all();
        }
    }

    void dispatch_loop() {
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            cv_dispatch_.wait(lk, [&] {
                if (stop_) return true;
                if (pending_.empty()) return false;
                for (Slot *s : slots_)
                    if (!s->job) return true;
                return false;
            });
            if (stop_) return;
            bool progress = true;
            while (progress && !pending_.empty()) {
                progress = false;
                Slot *free_slot = nullptr;
                for (Slot *s : slots_)
                    if (!s->job) {
                        free_slot = s;
                        break;
                    }
                if (!free_slot) break;
                for (auto it = pending_.begin(); it != pending_.end(); ++it) {
                    Req *r = *it;
                    if (budget_ && in_use_ + r->footprint > admit_budget_) {
                        if (r->deferrals++ == 0) ++deferred_;  // requests ever passed over
                        continue;
                    }
                    pending_.erase(it);
                    in_use_ += r->footprint;
                    peak_ = std::max(peak_, in_use_);
                    ++admitted_;
                    free_slot->job = r;
                    free_slot->cv.notify_one();
                    progress = true;
                    break;
                }
            }
            // nothing admissible right now: wait for a completion
            if (!pending_.empty()) cv_dispatch_.wait_for(lk, std::chrono::milliseconds(50));
        }
    }

    Backend backend_;
    const uint64_t budget_, admit_budget_, keep_;
    const Clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable cv_dispatch_, cv_done_;
    std::deque<Req *> pending_;
    std::unordered_map<uint64_t, Req *> all_;
    std::vector<Slot *> slots_;
    std::thread dispatcher_;
    uint64_t in_use_ = 0, peak_ = 0, next_ticket_ = 1;
    uint64_t admitted_ = 0, deferred_ = 0, completed_ = 0, failed_ = 0, over_budget_ = 0;
    bool stop_ = false;
};

}  // namespace serve
}  // namespace s2h
