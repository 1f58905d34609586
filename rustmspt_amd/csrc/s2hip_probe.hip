// Integer-VALU issue-rate probe: the roofline peak of the exact counting loop,
// measured on the device instead of derived from the clock.
//
// Every lane runs the counting loop's instruction pair -- v_and_b32 with a
// wave-uniform (SGPR) operand, then v_bcnt_u32_b32 accumulating -- for many
// iterations, with enough waves to fill every SIMD.  The pair is written as
// inline asm so each and/popcount pair is exactly two VALU instructions (left
// to itself the compiler adds VALU work around the shifts, which understates
// the rate); the shifts run on the scalar unit, as in the counting loop.
// Three register patterns: fresh operands per chain; one lane word against 32
// successive shifts of a wave-uniform pair (the counting loop's reuse); two
// lane words per shifted value.  The fastest is reported.  The only memory
// traffic is one store per lane at the end.
#include <algorithm>

#include "s2hip_internal.h"

namespace s2h {
namespace {

constexpr int kChains = 16;
constexpr int kProbeThreads = 256;

__device__ __forceinline__ void and_popc(uint32_t &acc, uint32_t lane_word, uint32_t uniform) {
    uint32_t t;
    asm volatile("v_and_b32 %0, %1, %2" : "=v"(t) : "s"(uniform), "v"(lane_word));
    asm volatile("v_bcnt_u32_b32 %0, %1, %0" : "+v"(acc) : "v"(t));
}

// Variant 0: kChains independent chains, a new uniform operand every iteration.
__global__ __launch_bounds__(kProbeThreads) void valu_probe_kernel(uint32_t seed, int iters,
                                                                   uint32_t *__restrict__ out) {
    uint32_t x[kChains], acc[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
        x[k] = seed * (threadIdx.x + 1u) + 0x9e3779b9u * (uint32_t)(k + 1);
        acc[k] = (uint32_t)k;
    }
    uint32_t m = __builtin_amdgcn_readfirstlane(seed ^ 0x5bd1e995u);
    for (int i = 0; i < iters; ++i) {
        m = m * 1664525u + 1013904223u;  // scalar: a new uniform mask every iteration
#pragma unroll
        for (int k = 0; k < kChains; ++k) and_popc(acc[k], x[k], m >> k);
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += acc[k];
    out[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = s;
}

// Variants 1 and 2: the counting loop's register pattern -- kWords lane words,
// each ANDed with kShifts successive shifts of a wave-uniform 64-bit pair
// (scalar s_lshr_b64) into kShifts accumulators; with TWO, a second set of
// lane words meets every shifted value too.
constexpr int kWords = 4, kShifts = 32;

template <bool TWO>
__global__ __launch_bounds__(kProbeThreads) void valu_probe_reuse_kernel(uint32_t seed, int iters,
                                                                         uint32_t *__restrict__ out) {
    uint32_t x[kWords], x2[kWords], acc[kShifts], acc2[kShifts];
#pragma unroll
    for (int w = 0; w < kWords; ++w) {
        x[w] = seed * (threadIdx.x + 1u) + 0x9e3779b9u * (uint32_t)(w + 1);
        x2[w] = x[w] * 747796405u;
    }
#pragma unroll
    for (int j = 0; j < kShifts; ++j) acc[j] = acc2[j] = (uint32_t)j;
    uint64_t m = __builtin_amdgcn_readfirstlane(seed ^ 0x5bd1e995u);
    for (int i = 0; i < iters; ++i) {
        m = m * 6364136223846793005ull + 1442695040888963407ull;  // scalar
#pragma unroll
        for (int w = 0; w < kWords; ++w) {
            const uint64_t pair = m ^ ((uint64_t)w << 32);
#pragma unroll
            for (int j = 0; j < kShifts; ++j) {
                const uint32_t sh = (uint32_t)(pair >> j);
                and_popc(acc[j], x[w], sh);
                if (TWO) and_popc(acc2[j], x2[w], sh);
            }
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kShifts; ++j) s += acc[j] + (TWO ? acc2[j] : 0u);
    out[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = s;
}

}  // namespace
}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_valu_probe(s2h_ctx *ctx, double *lane_ops_per_s) {
    if (!ctx || !lane_ops_per_s) return fail(S2H_ERR_ARG, "null argument");
    S2H_HIP(hipSetDevice(ctx->device));
    int cus = 0;
    S2H_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    // 8 waves per SIMD (4 SIMDs per CU, 4 waves per 256-thread block)
    const unsigned blocks = (unsigned)std::max(1, cus) * 8u;
    const int iters = 4096;
    DevBuf out;
    S2H_TRY(dev_alloc(ctx, out, (uint64_t)blocks * kProbeThreads * 4));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    s2h_status st = S2H_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = fail(S2H_ERR_HIP, "hipEventCreate");
    double best_rate = 0;
    for (int variant = 0; variant < 3 && st == S2H_OK; ++variant) {
        // about the same number of and/popcount pairs per lane in every variant
        const int pairs_per_iter = variant == 0 ? kChains : variant == 1 ? kWords * kShifts : 2 * kWords * kShifts;
        const int it = std::max(1, iters * kChains / pairs_per_iter);
        const double ops_v = (double)blocks * kProbeThreads * (double)it * pairs_per_iter * 2.0;
        float best = 1e30f;
        for (int rep = 0; rep < 4 && st == S2H_OK; ++rep) {
            hipError_t e = hipEventRecord(e0, ctx->stream);
            if (e == hipSuccess) {
                if (variant == 0)
                    valu_probe_kernel<<<blocks, kProbeThreads, 0, ctx->stream>>>(0x1234567u + rep, it,
                                                                                  (uint32_t *)out.ptr);
                else if (variant == 1)
                    valu_probe_reuse_kernel<false><<<blocks, kProbeThreads, 0, ctx->stream>>>(0x1234567u + rep, it,
                                                                                               (uint32_t *)out.ptr);
                else
                    valu_probe_reuse_kernel<true><<<blocks, kProbeThreads, 0, ctx->stream>>>(0x1234567u + rep, it,
                                                                                              (uint32_t *)out.ptr);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) st = fail(S2H_ERR_HIP, std::string("valu probe: ") + hipGetErrorString(e));
            if (rep > 0) best = std::min(best, ms);  // first run loads the code object
        }
        if (st == S2H_OK) best_rate = std::max(best_rate, ops_v / (best * 1e-3));
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(ctx, out);
    if (st != S2H_OK) return st;
    *lane_ops_per_s = best_rate;
    return S2H_OK;
}

}  // extern "C"
