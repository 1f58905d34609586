// Stress of the serving core (rustmspt_amd/csrc/s2hip_serve_core.h) with a
// host-only backend, for the ThreadSanitizer and AddressSanitizer builds.
// Submitter threads queue requests of varied footprints under a budget that
// defers some; poller threads poll every ticket without waiting in tight loops
// (the pattern in which a finished request is released while its worker could
// still touch it) and with waiting; one round stops the server with requests
// still queued.  Every result must be the backend's deterministic curve of its
// request.  Exit status 0 and no sanitizer report: pass.
#include <atomic>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "s2hip_serve_core.h"

using s2h::serve::Core;
using s2h::serve::JobOutcome;
using s2h::serve::Req;

namespace {

struct FakeCtx {
    uint64_t jobs = 0;  // touched only by the slot's worker
};

double expected(const Req &r, uint32_t k) {
    uint64_t h = r.tag * 1000003ull + k;
    for (uint64_t w : r.words) h = (h ^ w) * 1099511628211ull;
    return (double)(h % 1000003ull) / 1000003.0;
}

struct FakeBackend {
    using Ctx = FakeCtx;
    JobOutcome run(FakeCtx *ctx, Req &r, uint64_t share) {
        JobOutcome out;
        ++ctx->jobs;
        std::this_thread::sleep_for(std::chrono::microseconds((r.tag * 37) % 200));
        if (r.tag % 97 == 13) {  // a failing request
            out.status = S2H_ERR_HIP;
            out.error = "backend failure";
            return out;
        }
        r.s2.resize(r.r_max + 1);
        for (uint32_t k = 0; k <= r.r_max; ++k) r.s2[k] = expected(r, k);
        out.held = 1000 + ctx->jobs;
        out.over = share && r.footprint > share;
        return out;
    }
};

int round_trip(int round, int submitters, int per, int pollers, bool stop_early) {
    const uint64_t budget = 10000, admit = 6000, keep = 500;
    Core<FakeBackend> core(FakeBackend{}, budget, admit, keep);
    std::vector<FakeCtx *> ctxs;
    for (int i = 0; i < 4; ++i) ctxs.push_back(new FakeCtx());
    core.start(ctxs);
    const int total = submitters * per;
    std::vector<std::atomic<uint64_t>> tickets(total);
    for (auto &t : tickets) t = 0;
    std::vector<std::vector<uint64_t>> words(total);
    std::atomic<int> bad{0}, done{0}, failed{0};
    std::vector<std::thread> th;
    for (int s = 0; s < submitters; ++s)
        th.emplace_back([&, s] {
            std::mt19937_64 g(round * 1000 + s);
            for (int i = 0; i < per; ++i) {
                const int idx = s * per + i;
                auto *r = new Req();
                r->tag = (uint64_t)idx;
                r->words.assign(1 + g() % 64, 0);
                for (uint64_t &w : r->words) w = g();
                words[idx] = r->words;
                r->r_max = 1 + (uint32_t)(g() % 20);
                r->footprint = 500 + g() % 2500;  // two or three fit the admission budget
                uint64_t t = 0;
                if (core.submit(r, &t) != S2H_OK) {
                    bad++;
                    continue;
                }
                tickets[idx] = t;
                if (g() % 8 == 0) std::this_thread::yield();
            }
        });
    for (int p = 0; p < pollers; ++p)
        th.emplace_back([&, p] {
            std::mt19937_64 g(round * 7919 + p);
            // poller p owns the tickets idx % pollers == p
            std::vector<int> mine;
            for (int idx = p; idx < total; idx += pollers) mine.push_back(idx);
            while (!mine.empty()) {
                for (size_t k = 0; k < mine.size();) {
                    const int idx = mine[k];
                    const uint64_t t = tickets[idx];
                    if (!t) {
                        ++k;
                        continue;
                    }
                    s2h_result res{};
                    std::string err;
                    const bool wait = g() % 16 == 0;
                    s2h_status st = core.poll(t, wait, &res, nullptr, 0, err);
                    if (st == S2H_ERR_HIP) {
                        failed++;
                        mine.erase(mine.begin() + k);
                        continue;
                    }
                    if (st != S2H_OK) {
                        bad++;
                        mine.erase(mine.begin() + k);
                        continue;
                    }
                    if (res.ready) {
                        done++;
                        mine.erase(mine.begin() + k);
                        continue;
                    }
                    ++k;
                }
                if (stop_early && done > total / 3) return;
            }
        });
    for (auto &t : th) t.join();
    uint64_t st[10];
    core.stats(st);
    for (FakeCtx *c : core.stop()) delete c;
    if (bad) {
        std::fprintf(stderr, "round %d: %d unexpected statuses\n", round, (int)bad);
        return 1;
    }
    if (!stop_early && (int)(done + failed) != total) {
        std::fprintf(stderr, "round %d: %d done + %d failed of %d\n", round, (int)done, (int)failed, total);
        return 1;
    }
    if (!stop_early && st[1] == 0) {
        std::fprintf(stderr, "round %d: nothing deferred\n", round);
        return 1;
    }
    return 0;
}

// Results: the curve returned equals the backend's function of the request.
int check_values() {
    Core<FakeBackend> core(FakeBackend{}, 0, 0, 0);
    core.start({new FakeCtx(), new FakeCtx()});
    std::vector<uint64_t> tickets;
    std::vector<Req> copies;
    for (int i = 0; i < 64; ++i) {
        auto *r = new Req();
        r->tag = (uint64_t)i + 1;
        r->words = {(uint64_t)i * 3 + 1, (uint64_t)i * 7 + 2};
        r->r_max = 5 + i % 7;
        copies.push_back(*r);
        uint64_t t = 0;
        core.submit(r, &t);
        tickets.push_back(t);
    }
    int bad = 0;
    for (int i = 0; i < 64; ++i) {
        s2h_result res{};
        std::string err;
        std::vector<double> s2(copies[i].r_max + 1);
        const s2h_status st = core.poll(tickets[i], true, &res, s2.data(), s2.size(), err);
        if (copies[i].tag % 97 == 13) {
            bad += st != S2H_ERR_HIP;
            continue;
        }
        if (st != S2H_OK) {
            ++bad;
            continue;
        }
        for (uint32_t k = 0; k <= copies[i].r_max; ++k) bad += s2[k] != expected(copies[i], k);
    }
    for (FakeCtx *c : core.stop()) delete c;
    if (bad) std::fprintf(stderr, "%d wrong values\n", bad);
    return bad ? 1 : 0;
}

}  // namespace

int main() {
    int rc = check_values();
    for (int round = 0; round < 6 && rc == 0; ++round) rc = round_trip(round, 3, 150, 3, round == 5);
    std::printf(rc == 0 ? "serve stress ok\n" : "serve stress FAILED\n");
    return rc;
}
