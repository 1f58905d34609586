// Voxel Monte Carlo S2 on the packed volume with a counter-based RNG.
//
// Estimator of voxel_mc_radii (src/geometry/s2.rs:1412-1464): per radius, draw a
// uniform voxel y and a uniform offset d of the full shell (shell_offsets_for_
// distance, not domain-filtered); a partner outside the grid is rejected;
// otherwise valid += 1 and hits += occ(y) & occ(y + d).  The estimate is
// hits / valid per radius.
//
// Stream: sample s of radius r draws Philox4x32-10 with counter
// {s_lo, s_hi, r, 0x53324D43} and key {seed_lo, seed_hi}; word 0..2 pick x, y, z
// and word 3 the shell offset, each as (u32 * n) >> 32.  Every sample is an
// independent counter, so results do not depend on the launch shape.
#include <cmath>
#include <cstring>

#include "s2hip_internal.h"

namespace s2h {

struct Sh16 {
    int16_t dx, dy, dz, pad;
};

static constexpr int kMcThreads = 256;
static constexpr int kMcPerThread = 16;

__global__ __launch_bounds__(kMcThreads) void mc_kernel(const uint64_t *__restrict__ words, uint32_t nx, uint32_t ny,
                                                        uint32_t nz, uint64_t row_words,
                                                        const Sh16 *__restrict__ shell,
                                                        const uint32_t *__restrict__ start, uint64_t samples,
                                                        uint32_t seed_lo, uint32_t seed_hi, uint32_t blocks_per_r,
                                                        unsigned long long *__restrict__ hv) {
    const uint32_t r = blockIdx.x / blocks_per_r + 1;
    const uint32_t blk = blockIdx.x % blocks_per_r;
    const uint32_t s0 = start[r], n = start[r + 1] - s0;
    uint64_t hits = 0, valid = 0;
    if (n) {
        const uint64_t base = ((uint64_t)blk * kMcThreads * kMcPerThread) + threadIdx.x;
        for (int k = 0; k < kMcPerThread; ++k) {
            const uint64_t s = base + (uint64_t)k * kMcThreads;
            if (s >= samples) break;
            uint32_t o[4];
            philox4x32_10((uint32_t)s, (uint32_t)(s >> 32), r, 0x53324D43u, seed_lo, seed_hi, o);
            const uint32_t x = __umulhi(o[0], nx), y = __umulhi(o[1], ny), z = __umulhi(o[2], nz);
            const Sh16 d = shell[s0 + __umulhi(o[3], n)];
            const int x2 = (int)x + d.dx, y2 = (int)y + d.dy, z2 = (int)z + d.dz;
            if (x2 < 0 || y2 < 0 || z2 < 0 || x2 >= (int)nx || y2 >= (int)ny || z2 >= (int)nz) continue;
            ++valid;
            const uint64_t a = words[((uint64_t)x * ny + y) * row_words + (z >> 6)] >> (z & 63);
            const uint64_t b = words[((uint64_t)x2 * ny + (uint32_t)y2) * row_words + ((uint32_t)z2 >> 6)] >> (z2 & 63);
            hits += a & b & 1u;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        hits += __shfl_down(hits, off, 64);
        valid += __shfl_down(valid, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && valid) {
        atomicAdd(&hv[2 * r], (unsigned long long)hits);
        atomicAdd(&hv[2 * r + 1], (unsigned long long)valid);
    }
}

static s2h_status mc_plan_get(s2h_ctx *ctx, uint32_t r_max, double pitch, McPlan **out) {
    uint64_t pbits;
    std::memcpy(&pbits, &pitch, 8);
    auto key = std::make_tuple(r_max, pbits);
    auto it = ctx->mc_plans.find(key);
    if (it != ctx->mc_plans.end()) {
        *out = it->second.get();
        return S2H_OK;
    }
    auto p = std::make_unique<McPlan>();
    p->r_max = r_max;
    p->pitch = pitch;
    std::vector<Sh16> all;
    std::vector<Offset> shell;
    p->start.assign(r_max + 2, 0);
    for (uint32_t r = 1; r <= r_max; ++r) {
        shell_offsets_full((double)r / pitch, 0.5 / pitch, shell);
        for (const Offset &o : shell) {
            if (std::abs(o.dx) > 32767 || std::abs(o.dy) > 32767 || std::abs(o.dz) > 32767)
                return fail(S2H_ERR_UNSUPPORTED, "Monte Carlo shell radius exceeds 32767 voxels");
            all.push_back(Sh16{(int16_t)o.dx, (int16_t)o.dy, (int16_t)o.dz, 0});
        }
        if (all.size() > 0xFFFFFFF0ull) return fail(S2H_ERR_UNSUPPORTED, "Monte Carlo shell table too large");
        p->start[r + 1] = (uint32_t)all.size();
    }
    S2H_TRY(upload(ctx, p->d_shell, all.data(), std::max<size_t>(all.size(), 1) * sizeof(Sh16)));
    S2H_TRY(upload(ctx, p->d_start, p->start.data(), p->start.size() * 4));
    S2H_TRY(dev_alloc(ctx, p->d_hv, 16 * ((uint64_t)r_max + 1)));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    *out = p.get();
    ctx->mc_plans[key] = std::move(p);
    return S2H_OK;
}

static s2h_status mc_run(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, uint64_t samples, uint64_t seed,
                         std::vector<uint64_t> &hv) {
    McPlan *p = nullptr;
    S2H_TRY(mc_plan_get(ctx, r_max, pitch, &p));
    hv.assign(2 * ((size_t)r_max + 1), 0);
    S2H_HIP(hipMemsetAsync(p->d_hv.ptr, 0, 16 * ((uint64_t)r_max + 1), ctx->stream));
    if (r_max > 0 && samples > 0) {
        const uint64_t per_block = (uint64_t)kMcThreads * kMcPerThread;
        const uint64_t bpr = (samples + per_block - 1) / per_block;
        if (bpr * r_max > 0x7FFFFFFFull) return fail(S2H_ERR_UNSUPPORTED, "too many Monte Carlo samples");
        mc_kernel<<<(unsigned)(bpr * r_max), kMcThreads, 0, ctx->stream>>>(
            vol->d_words, (uint32_t)vol->dims[0], (uint32_t)vol->dims[1], (uint32_t)vol->dims[2], vol->row_words,
            (const Sh16 *)p->d_shell.ptr, (const uint32_t *)p->d_start.ptr, samples, (uint32_t)seed,
            (uint32_t)(seed >> 32), (uint32_t)bpr, (unsigned long long *)p->d_hv.ptr);
        S2H_HIP(hipGetLastError());
    }
    S2H_HIP(hipMemcpyAsync(hv.data(), p->d_hv.ptr, hv.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    return S2H_OK;
}

}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_mc_counts(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, uint64_t samples, uint64_t seed,
                         uint64_t *hits_valid, uint64_t n) {
    if (!ctx || !vol || !hits_valid) return fail(S2H_ERR_ARG, "null argument");
    if (n != 2 * ((uint64_t)r_max + 1)) return fail(S2H_ERR_SIZE, "hits_valid must hold 2*(r_max+1) entries");
    if (!(pitch > 0.0) || !std::isfinite(pitch)) return fail(S2H_ERR_ARG, "pitch must be finite and > 0");
    S2H_HIP(hipSetDevice(ctx->device));
    std::vector<uint64_t> hv;
    S2H_TRY(mc_run(ctx, vol, r_max, pitch, samples, seed, hv));
    std::memcpy(hits_valid, hv.data(), n * 8);
    return S2H_OK;
}

s2h_status s2h_mc_s2(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, uint64_t samples, uint64_t seed,
                     double *s2, uint64_t n) {
    if (!ctx || !vol || !s2) return fail(S2H_ERR_ARG, "null argument");
    if (n != (uint64_t)r_max + 1) return fail(S2H_ERR_SIZE, "s2 buffer must hold r_max+1 values");
    S2H_HIP(hipSetDevice(ctx->device));
    uint64_t occupied = 0;
    S2H_TRY(vol_popcount(ctx, vol, &occupied));
    if (occupied == 0) {
        std::fill(s2, s2 + n, 0.0);
        return S2H_OK;
    }
    const double pm = pitch < 1e-9 ? 1e-9 : pitch;
    const double vf = (double)occupied / (double)(vol->dims[0] * vol->dims[1] * vol->dims[2]);
    std::vector<uint64_t> hv;
    S2H_TRY(mc_run(ctx, vol, r_max, pm, samples < 200 ? 200 : samples, seed, hv));
    std::vector<uint8_t> sup(n, 0);
    s2[0] = vf;
    sup[0] = 1;
    for (uint32_t r = 1; r <= r_max; ++r) {
        const uint64_t h = hv[2 * r], v = hv[2 * r + 1];
        if (v == 0) {
            s2[r] = 0.0;
        } else {
            s2[r] = (double)h / (double)v;
            sup[r] = 1;
        }
    }
    fill_missing(s2, sup.data(), n, vf);
    s2[0] = vf;
    return S2H_OK;
}

}  // extern "C"
