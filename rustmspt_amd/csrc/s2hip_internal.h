// Internal declarations shared by the translation units of libs2hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/s2hip.h"

namespace s2h {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
s2h_status fail(s2h_status code, const std::string &msg);

// Failure injection for error-path tests (S2HIP_FAIL_AT): S2H_ERR_HIP when
// this pass through the site is the one the variable names, else S2H_OK.
enum FailSite { kFailAlloc = 0, kFailLaunch = 1, kFailReadback = 2 };
s2h_status inject(FailSite site);

#define S2H_HIP(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ::s2h::fail(S2H_ERR_HIP, std::string(#expr " failed: ") +          \
                                                hipGetErrorString(e_));                 \
    } while (0)

#define S2H_TRY(expr)                   \
    do {                                \
        s2h_status st_ = (expr);        \
        if (st_ != S2H_OK) return st_;  \
    } while (0)

// popcount(x) + acc as one v_bcnt_u32_b32.  Written out because the compiler
// otherwise reassociates a chain of `acc += popcount(..)` into pairs of
// popcounts joined by v_add3_u32: one extra VALU instruction per two words in
// the counting loops.  Register operands only; no memory access.
__device__ __forceinline__ uint32_t popc_acc(uint32_t x, uint32_t acc) {
    uint32_t r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}

// Philox4x32-10 (Salmon et al., SC'11): counter-based random words.
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// ---------------------------------------------------------------- ordered fold
//
// The reference's curve is a sequential left fold of f64 terms, sum = sum + t_k
// in shell order (src/geometry/s2.rs:1187-1201): a chain of dependent adds that
// one lane runs at the add latency.  For non-negative terms the same bits come
// from a wave 64 terms at a time.  While the running sum S lies in one binade
// [2^m, 2^(m+1)), S = a*u with u = 2^(m-52) and an integer a in [2^52, 2^53],
// and for t >= 0 with S + t < 2^(m+1) round-to-nearest gives
//   RN(S + t) = (a + n + [f > 1/2]) * u,   t/u = n + f, n integer, 0 <= f < 1,
// independent of S except on exact ties (f = 1/2, where the even neighbour
// depends on a).  So the grid increments q = n + [f > 1/2] of a block of terms
// are summed with a wave prefix sum, exactly, in integers.  The first term whose
// exact sum reaches the next binade (a + n >= 2^53), and any tie, is added
// with the hardware add from the exact S before it, and the grid restarts at
// the new binade.  Every step therefore rounds exactly as the sequential
// fold.  A segment with a negative or non-finite term is folded sequentially.

// DPP move of a double (both halves move alike); disabled rows and lanes
// without a source read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xf, true);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Exclusive prefix sum over the 64 lanes: a wave shift right by one lane,
// then row shifts 1, 2, 4, 8 within rows of 16 and the row_bcast:15 and
// row_bcast:31 carries into rows 1, 3 and 2, 3.  Lane l sums only lanes below
// l, so it is exact for non-negative integer values while those sums stay
// below 2^53.
__device__ __forceinline__ double wave_exclusive_sum_f64(double x) {
    x = dpp_f64<0x138, 0xf>(x);
    x += dpp_f64<0x111, 0xf>(x);
    x += dpp_f64<0x112, 0xf>(x);
    x += dpp_f64<0x114, 0xf>(x);
    x += dpp_f64<0x118, 0xf>(x);
    x += dpp_f64<0x142, 0xa>(x);
    x += dpp_f64<0x143, 0xc>(x);
    return x;
}

// Inclusive prefix sum over the wave's lanes (lane l gets x_0 + ... + x_l),
// the same steps as above without the initial shift: exact for non-negative
// integer values while those sums stay below 2^53, and never below an exact
// sum that reaches it (every partial sum is of non-negative values, and
// rounding is monotone).
__device__ __forceinline__ double wave_inclusive_sum_f64(double x) {
    x += dpp_f64<0x111, 0xf>(x);
    x += dpp_f64<0x112, 0xf>(x);
    x += dpp_f64<0x114, 0xf>(x);
    x += dpp_f64<0x118, 0xf>(x);
    x += dpp_f64<0x142, 0xa>(x);
    x += dpp_f64<0x143, 0xc>(x);
    return x;
}

// A wave-uniform read of lane l's double.
__device__ __forceinline__ double read_lane_f64(double x, int l) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

struct FoldState {
    double S = 0.0;   // the exact running sum (wave-uniform)
    double a = 0.0;   // S / u while S > 0: an integer in [2^52, 2^53]
    double u = 0.0, inv_u = 0.0;
    __device__ void rebase() {
        if (S > 0.0) {
            const int m = ilogb(S);
            u = ldexp(1.0, m - 52);
            inv_u = ldexp(1.0, 52 - m);
            a = S * inv_u;
        }
    }
};

// Fold a block of 64*K terms into st: lane l holds terms l*K .. l*K+K-1 of the
// block in t[], and the first cnt terms of the block are active.  Each lane sums
// its run's grid increments; one wave prefix sum places the runs.  The
// increments and their sums are integers held in doubles, exact below 2^53
// (every sum used is below it; once a sum reaches it, rounding is monotone, so
// the comparison with 2^53 is still decided right).  All 64 lanes call this
// with the same cnt.
template <int K>
__device__ __forceinline__ void fold_block(FoldState &st, const double (&t)[K], uint32_t cnt, uint32_t lane) {
    constexpr double kTwo53 = 9007199254740992.0;
    uint32_t pos = 0;
    const uint32_t j0 = lane * K;  // block index of t[0]
    while (pos < cnt) {
        if (st.S == 0.0) {
            // 0 + t is t exactly; zero terms leave the sum at zero
            int first = -1;
#pragma unroll
            for (int i = K - 1; i >= 0; --i)
                if (j0 + i >= pos && j0 + i < cnt && t[i] != 0.0) first = i;
            const uint64_t m = __ballot(first >= 0);
            if (!m) return;
            const int l = __ffsll((long long)m) - 1;
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < K; ++i)
                if (i == first) v = t[i];
            const int fi = __builtin_amdgcn_readlane(first, l);
            st.S = read_lane_f64(v, l);
            st.rebase();
            pos = (uint32_t)l * K + (uint32_t)fi + 1;
            continue;
        }
        double q[K], nfl[K];
        bool flag[K], act[K];
        double T = 0.0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            act[i] = j0 + i >= pos && j0 + i < cnt;
            const double y = act[i] ? t[i] * st.inv_u : 0.0;  // exact: scaling by a power of two
            const bool big = !(y < kTwo53);                     // y >= 2^53: certainly the next binade
            nfl[i] = big ? 0.0 : floor(y);
            const double f = y - nfl[i];
            q[i] = nfl[i] + (f > 0.5 ? 1.0 : 0.0);
            flag[i] = act[i] && (big || f == 0.5);
            T += q[i];
        }
        double run = wave_exclusive_sum_f64(T);  // increments of the lanes before
        int stop = -1;
        double spex = 0.0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            if (stop < 0 && act[i] && (flag[i] || (st.a + run) + nfl[i] >= kTwo53)) {
                stop = i;
                spex = run;
            }
            run += q[i];
        }
        const uint64_t sm = __ballot(stop >= 0);
        if (!sm) {
            st.a += read_lane_f64(run, 63);  // run = all increments of the block at lane 63
            st.S = st.a * st.u;
            if (st.a >= kTwo53) st.rebase();  // rounded up to exactly 2^(m+1)
            return;
        }
        const int l = __ffsll((long long)sm) - 1;
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (i == stop) v = t[i];
        const int si = __builtin_amdgcn_readlane(stop, l);
        st.a += read_lane_f64(spex, l);  // the terms before the stopping one, on the grid
        st.S = st.a * st.u;
        st.S = st.S + read_lane_f64(v, l);  // the stopping term: one hardware add
        st.rebase();
        pos = (uint32_t)l * K + (uint32_t)si + 1;
    }
}

// The ordered f64 sum of term(0), ..., term(n - 1) on one wave (all 64 lanes
// call it; the result is wave-uniform): blocks of 64 * K terms, lane l holding
// terms l*K .. l*K+K-1 of a block, folded by fold_block.  Terms are produced
// AHEAD blocks before they are folded, so their loads overlap the fold.  term(j)
// is only called with j < n (past the end it is called with 0 and the value
// replaced by 0).  After a negative or non-finite term the rest of the sum
// folds one term at a time.
template <int K, int AHEAD, class Term>
__device__ __forceinline__ double fold_terms(uint64_t n, uint32_t lane, Term term) {
    constexpr uint64_t B = 64 * K, SB = B * AHEAD;
    FoldState st;
    if (n == 0) return 0.0;
    bool seq = false;
    double cur[AHEAD][K], nxt[AHEAD][K];
#pragma unroll
    for (int b = 0; b < AHEAD; ++b)
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const uint64_t j = b * B + (uint64_t)lane * K + i;
            const double v = term(j < n ? j : 0);
            cur[b][i] = j < n ? v : 0.0;
        }
    for (uint64_t s0 = 0; s0 < n; s0 += SB) {
#pragma unroll
        for (int b = 0; b < AHEAD; ++b)
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const uint64_t j = s0 + SB + b * B + (uint64_t)lane * K + i;
                const double v = term(j < n ? j : 0);
                nxt[b][i] = j < n ? v : 0.0;
            }
#pragma unroll
        for (int b = 0; b < AHEAD; ++b) {
            const uint64_t b0 = s0 + b * B;
            if (b0 >= n) break;
            const uint32_t cnt = (n - b0) < B ? (uint32_t)(n - b0) : (uint32_t)B;
            bool bad = false;
#pragma unroll
            for (int i = 0; i < K; ++i)
                bad = bad || (lane * K + i < cnt && !(cur[b][i] >= 0.0 && cur[b][i] <= 1.7976931348623157e308));
            if (!seq && __ballot(bad)) seq = true;
            if (seq) {
                // one term at a time, in order (lane l's terms are l*K .. l*K+K-1)
                for (uint32_t l = 0; l * K < cnt; ++l)
#pragma unroll
                    for (int i = 0; i < K; ++i)
                        if (l * K + i < cnt) st.S = st.S + read_lane_f64(cur[b][i], (int)l);
            } else {
                fold_block<K>(st, cur[b], cnt, lane);
            }
        }
#pragma unroll
        for (int b = 0; b < AHEAD; ++b)
#pragma unroll
            for (int i = 0; i < K; ++i) cur[b][i] = nxt[b][i];
    }
    return st.S;
}

// ---------------------------------------------------------------- shells (host)
struct Offset {
    int32_t dx, dy, dz;
};

// Index in the half cube [R_x + 1][2R_y + 1][2R_z + 1] of offset o: the
// correlation is even (C(d) = C(-d), it counts the same pairs), so the
// inverse passes keep dx >= 0 only and an offset with dx < 0 reads its
// mirror -o.
__device__ __forceinline__ uint64_t half_cube_index(Offset o, int Ry, int Rz) {
    if (o.dx < 0) o = Offset{-o.dx, -o.dy, -o.dz};
    const int nyo = 2 * Ry + 1, nzo = 2 * Rz + 1;
    const int iy = o.dy >= 0 ? o.dy : nyo + o.dy;
    const int iz = o.dz >= 0 ? o.dz : nzo + o.dz;
    return ((uint64_t)o.dx * nyo + iy) * nzo + iz;
}

// In-domain offsets of radii 1..r_max in the reference's shell order, with
// radius_start[r_max+2].
void shell_order_offsets(const uint64_t dims[3], uint32_t r_max, double pitch,
                         std::vector<Offset> &offs, std::vector<uint64_t> &radius_start);
// Every offset of shell r (not domain-filtered), in shell order.
void shell_offsets_full(double distance_vox, double half_width_vox, std::vector<Offset> &out);

void fill_missing(double *values, const uint8_t *has_support, uint64_t n, double vf);

// ---------------------------------------------------------------- device memory
struct DevBuf {
    void *ptr = nullptr;
    uint64_t bytes = 0;
    bool borrowed = false;  // a slice of the context's FFT arena: dev_free only forgets it
};

// Plan for exact counting on one (dims, r_max, pitch): shell-order offsets, the
// canonical half-ball column table and per-wave work items.
struct ColInfo {
    int32_t dx, dy;   // column (dx, dy) of the canonical half ball
    int32_t m;        // largest dz of the column (-1 = idle lane)
    uint32_t base;    // index of (dx, dy, 0) in the dense count table
};

struct WorkItem {
    uint32_t group;   // 64-column group
    uint16_t dz0;     // first dz (multiple of 32)
    uint16_t dzb;     // 8, 16, 24 or 32 dz values
    uint16_t x0, x1;  // b-row x range
    uint16_t y0, y1;  // b-row y range
    uint16_t nw;      // dwords of the shifted partner that can be non-zero
    uint16_t pad;
};

// One (column group, dz block) of a plan: the unit the batched delta counts
// iterate over.
struct DzBlock {
    uint32_t group;
    uint16_t dz0, dzb;
};

// Tunable shape of an exact plan.  Results do not depend on it (integer counts).
//   band  : columns are grouped in bands of `band` consecutive m values (ordered
//           by angle inside a band): 1 keeps every group's m equal (no dz waste,
//           long arcs -> idle boundary rows), larger bands shorten the arcs.
//   waves : target number of waves of the count launch (x-chunking of blocks).
struct ExactTune {
    int32_t band = 3;
    int32_t waves = 16384;
};

// Shell data of an exact plan that depends only on (r_max, pitch) and on the
// grid extent where it clips the shells: the in-domain offsets in shell order,
// the canonical half-ball columns (dense-table layout) and the gather map.
// Built once per process and shared (read-only) by every plan and context.
struct ShellBase {
    std::vector<Offset> offs;             // shell order
    std::vector<uint64_t> radius_start;   // r_max + 2
    std::vector<ColInfo> colset;          // every half-ball column, table order
    std::vector<int32_t> gather;          // shell order -> table index (-1 = zero offset)
    uint64_t table_size = 0;              // dense count-table entries
    int extent[3] = {0, 0, 0};            // largest |dx|, |dy|, |dz| of the offsets
    // columns ordered 64 to a wave for a band width (filled on first use)
    mutable std::mutex lanes_mu;
    mutable std::map<int, std::shared_ptr<const std::vector<ColInfo>>> lanes;
};

// A context's device copy of a ShellBase, shared by its plans.
struct ShellDev {
    std::shared_ptr<const ShellBase> host;
    DevBuf d_offs;     // Offset[K]
    DevBuf d_gather;   // int32[K], uploaded for the first direct-count run
    DevBuf d_rstart;   // uint64[r_max+2]
};

struct ExactPlan {
    explicit ExactPlan(std::shared_ptr<ShellDev> s)
        : shells(std::move(s)), offs(shells->host->offs), radius_start(shells->host->radius_start),
          colset(shells->host->colset), table_size(shells->host->table_size), d_offs(shells->d_offs),
          d_gather(shells->d_gather), d_rstart(shells->d_rstart) {}
    uint64_t dims[3];
    uint32_t r_max;
    double pitch;
    std::shared_ptr<ShellDev> shells;
    const std::vector<Offset> &offs;
    const std::vector<uint64_t> &radius_start;
    const std::vector<ColInfo> &colset;
    const uint64_t table_size;
    const DevBuf &d_offs, &d_gather, &d_rstart;
    // tuned part: groups, blocks and items for `tune`, built on the first
    // direct-count run (`shaped`); until then only the model's figures are set
    ExactTune tune;
    bool shaped = false;
    int tuned = 0;                        // 0 not yet, 1 timed on the device, -1 model pick
    int timed_candidates = 0;
    double model_ops = 0;                 // modelled executed VALU lane-ops per run
    double model_ms = 0;                  // modelled count-kernel time of the applied shape
    int method = 0;                       // kernel of the last evaluation: 1 direct, 2 FP64 FFT (rounded)
    uint32_t ngroups = 0;
    std::vector<DzBlock> blocks;
    std::vector<WorkItem> items;          // heaviest first
    // per-plan device buffers
    DevBuf d_cols;     // ColInfo[ngroups*64]
    DevBuf d_items;    // WorkItem[items]
    DevBuf d_table;    // uint64[table_size]
    DevBuf d_hits;     // uint64[K]
    DevBuf d_curve;    // double[r_max+1] values, then uint8 support
    DevBuf d_fold;     // the ordered fold's block tables and per-block scratch
    uint32_t fold_nblk = 0;
};

struct McPlan {
    uint32_t r_max;
    double pitch;
    std::vector<uint32_t> start;   // r_max + 2 (full shells, not domain-filtered)
    DevBuf d_shell;                // int16x4 per offset
    DevBuf d_start;                // uint32[r_max+2]
    DevBuf d_hv;                   // uint64[2*(r_max+1)]
};

struct BatchWS;  // defined in s2hip_batch.hip
struct FftPlan;  // defined in s2hip_fft.hip
void fft_plan_release(s2h_ctx *ctx, FftPlan *p);

}  // namespace s2h

struct s2h_vol {
    uint64_t dims[3];
    uint64_t row_words;   // u64 words per row
    uint64_t nwords;      // total u64 words
    uint64_t *d_words;
    uint64_t alloc_bytes; // nwords*8 plus the zeroed tail
    uint64_t occupied;    // cached popcount, UINT64_MAX = unknown
};

// Zeroed bytes after the last row of every volume buffer: kernels may read up to
// 64 dwords past a row start without a bounds check.
constexpr uint64_t kVolTailBytes = 512;

struct s2h_ctx {
    int device = 0;
    uint64_t budget = 0;
    uint64_t in_use = 0;
    hipStream_t stream = nullptr;
    std::string name;
    std::map<std::tuple<uint64_t, uint64_t, uint64_t, uint32_t, uint64_t>,
             std::unique_ptr<s2h::ExactPlan>>
        exact_plans;
    std::map<std::tuple<uint32_t, uint64_t>, std::unique_ptr<s2h::McPlan>> mc_plans;
    std::map<std::tuple<uint64_t, uint64_t, uint64_t>, s2h::FftPlan *> fft_plans;
    std::map<std::tuple<uint64_t, uint64_t, uint64_t>, s2h::FftPlan *> fftd_plans;  // region-correlation frames
    s2h::DevBuf d_scalar;   // small scratch (popcount result etc.)
    s2h::DevBuf terms;      // the exact path's per-offset terms, shared by the context's plans (grow-only)
    bool profiling = false;
    bool autotune = true;   // time launch-shape candidates on a new exact plan
    int exact_method = 0;   // exact kernel: 0 by the cost model, 1 direct counting, 2 FP64 FFT rounded
    uint64_t fft_budget = 0;  // largest FP64 FFT working set the model's choice may take (0: no limit)
    std::map<std::tuple<uint32_t, uint64_t, uint64_t, uint64_t, uint64_t>, std::shared_ptr<s2h::ShellDev>>
        shell_devs;   // device copies of shared shell data, by (r_max, pitch, clipped extent)
    s2h::BatchWS *batch_ws = nullptr;            // persistent batch workspace and graph
    std::multimap<uint64_t, void *> block_cache;  // released device blocks by size
    uint64_t cached_bytes = 0;
    uint64_t held_peak = 0;  // largest in_use + cached_bytes since it was last reset
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double prof_ms[4] = {0, 0, 0, 0};
    int launches = 0;
    void *pinned = nullptr;     // page-locked staging for small read-backs (hipHostMalloc)
    uint64_t pinned_bytes = 0;
    // Pre-reserved device block that an FFT plan's spectra and cube are carved
    // from when they fit (serving slots): no allocation on the request path.
    // At most one plan lives in it; a new shape releases the old plan first.
    s2h::DevBuf fft_arena;
};

// A welded, granule-split mesh resident on the device (s2hip_meshin.hip).
struct s2h_mesh {
    uint64_t nverts = 0, ntris = 0;
    uint32_t ngranules = 0;
    double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};  // mesh_bbox (f64::min / max)
    double amax = 0.0;                                 // largest |vertex coordinate|
    s2h::DevBuf xyz, tri, gid;  // f64 x3 per vertex, u32 x3 per face, u32 per face
};

namespace s2h {
s2h_status dev_alloc(s2h_ctx *ctx, DevBuf &b, uint64_t bytes);
void dev_free(s2h_ctx *ctx, DevBuf &b);
void dev_cache_release(s2h_ctx *ctx);  // hipFree every cached block
void dev_cache_trim(s2h_ctx *ctx, uint64_t keep);  // hipFree cached blocks, largest first, down to keep bytes
void batch_ws_release(s2h_ctx *ctx);
s2h_status upload(s2h_ctx *ctx, DevBuf &b, const void *src, uint64_t bytes);
s2h_status vol_alloc(s2h_ctx *ctx, const uint64_t dims[3], s2h_vol **out);
// Exclusive prefix sum of n (1..2^31-1) u32 on ctx->stream; waits and returns the total.
s2h_status exclusive_sum(s2h_ctx *ctx, const uint32_t *in, uint32_t *out, uint64_t n, uint64_t *total);
s2h_status vol_popcount(s2h_ctx *ctx, s2h_vol *vol, uint64_t *occupied);
// The context's page-locked staging buffer, grown to at least `bytes`.
s2h_status pinned_staging(s2h_ctx *ctx, uint64_t bytes, void **out);

s2h_status exact_plan_get(s2h_ctx *ctx, const uint64_t dims[3], uint32_t r_max, double pitch,
                          ExactPlan **out);
void exact_plan_release(s2h_ctx *ctx, ExactPlan *plan);
void shell_devs_release(s2h_ctx *ctx);  // device copies of shared shell data
// Launch the count kernels for vol into plan->d_table (zeroed first) on ctx->stream.
s2h_status exact_run_counts(s2h_ctx *ctx, s2h_vol *vol, ExactPlan *plan);
// Delta counts of batch candidates by FP64 correlation over a frame around
// each changed box (s2hip_fft.hip).  A candidate: its changed box [b0, b1)
// in voxels, its (A, D) rows in the batch's change buffer (rows of the box,
// x-major, nwd 32-bit word pairs each) and its delta table (canonical shell
// order).
struct FftdCand {
    int32_t b0[3], b1[3];
    uint64_t ad_off, delta_off;
};
// The seed side: packed volume, shell extent, shell-order offsets and their
// canonical indices (-1: the zero offset), occupied voxels.
struct FftdSeed {
    const uint64_t *words = nullptr;
    uint64_t row_words = 0;
    int dims[3] = {0, 0, 0};
    int R[3] = {0, 0, 0};
    const int32_t *sg = nullptr;
    const Offset *offs = nullptr;
    uint64_t noffs = 0;
    uint64_t occupied = 0;
};
// Frame lengths for changed boxes up to bmax (false: unsupported lengths).
bool fft_delta_frame(const int bmax[3], const int R[3], int F[3]);
// Worst-case |computed - exact| of a frame's delta values (exact counts below 1/2).
double fft_delta_error_bound(const int F[3], uint64_t ones, uint64_t changed);
// Modelled device time (ms) of ncand candidates on frames F.
double fft_delta_model_ms(const int F[3], const int R[3], uint32_t ncand);
// Fill the delta tables of ncand candidates (device array) on ctx->stream.
s2h_status fft_delta_run(s2h_ctx *ctx, const FftdSeed &seed, const FftdCand *cands, uint32_t ncand, const int F[3],
                         const uint2 *ad, int nwd, int32_t *delta);

// The pruned correlation cube an FFT evaluation leaves on the device: values
// of offsets |dx| <= R[0], |dy| <= R[1], |dz| <= R[2], x-major, each axis
// holding 0 .. R then the negative offsets -R .. -1.  FP32 values are used as
// they are; FP64 values round to the integer counts.
struct FftCube {
    const float *f32 = nullptr;
    const double *f64 = nullptr;
    int R[3] = {0, 0, 0};
};
// Gather per-offset f64 terms -- from the direct count table, or from an FFT
// cube (cube != nullptr and one of its pointers set) -- and fold per radius
// into d_curve, or into curve_dst (page-locked host memory of (r_max + 1) * 9
// bytes) when set; with want_hits the integer counts also go to d_hits.
s2h_status exact_gather_fold(s2h_ctx *ctx, ExactPlan *plan, uint64_t occupied, const FftCube *cube, bool want_hits,
                             void *curve_dst);
// FP64 FFT exact kernel (s2hip_fft.hip): worst-case |computed - exact| of the
// correlation values for a volume with `occupied` voxels (rounding is exact
// below 1/2), whether the shape's lengths are supported, the algorithmic bytes
// of one evaluation, and the run itself (device array in shell order).
double fft64_error_bound(const uint64_t dims[3], const int R[3], uint64_t occupied);
bool fft64_supported(const uint64_t dims[3], const int R[3]);
double fft64_model_bytes(const uint64_t dims[3], const int R[3]);
double fft64_plan_bytes(const uint64_t dims[3], const int R[3]);
// Padded transform length of an axis of n voxels whose offsets reach |d| <= R.
uint64_t fft_padded_length(uint64_t n, uint64_t R);
// Device bytes of an FFT plan's spectra and cube as allocated (real_bytes 4:
// FP32 path, 8: FP64), each buffer rounded up to 256 bytes.
uint64_t fft_plan_buffer_bytes(const uint64_t dims[3], const int R[3], int real_bytes);
// Largest |dx|, |dy|, |dz| of the in-domain shell offsets (host, cached).
void shell_extent(const uint64_t dims[3], uint32_t r_max, double pitch, int R[3]);
// Whether the context already holds the FP64 FFT plan of this shape (its
// buffers are then part of ctx->in_use).
bool fft64_plan_cached(s2h_ctx *ctx, const uint64_t dims[3], const int R[3]);
// Release every FFT plan of the context (their blocks go to the block cache).
void fft_plans_release(s2h_ctx *ctx);
s2h_status fft64_exact_device(s2h_ctx *ctx, s2h_vol *vol, ExactPlan *plan, FftCube *cube);
// Copy folded curve values and supported flags back (unless the fold wrote
// them to `staged`) and finish the curve.
s2h_status finish_curve_from_device(s2h_ctx *ctx, ExactPlan *plan, double vf, double *s2, void *staged);
// The context's page-locked buffer for a curve of the plan (values, then
// support bytes), which the fold can write directly.
s2h_status curve_staging(s2h_ctx *ctx, ExactPlan *plan, void **stage);
// Per-device address of the no-allocation entry point's scratch array
// (s2hip_mc.hip); the caller has made `device` current.
s2h_status noalloc_scratch_init(int device);
// Profiling marks: k = 0 (before counts), 1 (after counts), 2 (after fold).
void prof_mark(s2h_ctx *ctx, int k);
void prof_collect(s2h_ctx *ctx);
}  // namespace s2h
