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
#include <mutex>
#include <string>

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
    S2H_TRY(inject(kFailReadback));
    S2H_HIP(hipMemcpyAsync(hv.data(), p->d_hv.ptr, hv.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    return S2H_OK;
}

// ---------------------------------------------------------------- no allocation
//
// s2h_mc_counts_noalloc runs the same estimator with every device buffer it
// needs beyond the caller's volume inside one file-scope array: the shell table
// (built on the device, so no host buffer either), the radius starts and the
// counters.  Each device holds its own copy of the array until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns for the current device (the array's name is a host-side stand-in).
// Calls on one device share the array: two calls that overlap race.

static constexpr uint64_t kNoallocScratchBytes = 32ull << 20;
static constexpr uint32_t kNoallocMaxR = 4096;
static constexpr int kNoallocMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[kNoallocMaxDevices] = {};

// Per-device address of the scratch array (the caller has made `device`
// current).  Called from s2h_create and, if needed, on the entry point's first
// call on a device.
s2h_status noalloc_scratch_init(int device) {
    if (device < 0 || device >= kNoallocMaxDevices)
        return fail(S2H_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return S2H_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(S2H_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[device] = ptr;
    return S2H_OK;
}

// Scratch layout: [0, 256) header (u64 [0] table entries needed, [1] 1 when they
// exceed the capacity); then u32 start[r_max + 2]; then u64 hv[2 (r_max + 1)];
// then the shell table (Sh16 per offset) up to the end.
struct NoallocLayout {
    uint64_t start_off, hv_off, table_off, capacity;
};
static NoallocLayout noalloc_layout(uint32_t r_max) {
    NoallocLayout l;
    l.start_off = 256;
    l.hv_off = (l.start_off + 4ull * (r_max + 2) + 255) / 256 * 256;
    l.table_off = (l.hv_off + 16ull * (r_max + 1) + 255) / 256 * 256;
    l.capacity = (kNoallocScratchBytes - l.table_off) / sizeof(Sh16);
    return l;
}

// shell_offsets_for_distance (src/geometry/s2.rs:336-361) on the device, for
// radius r at (r / pitch, 0.5 / pitch) voxels: the cube [-lim, lim]^3 scanned
// dx, then dy, then dz ascending, keeping d != 0 with low2 <= |d|^2 < high2.
// Column (dx, dy) is index c = (dx + lim) (2 lim + 1) + (dy + lim); within it the
// admissible dz form at most two runs, bracketed from the float bounds with a
// slack of 2 and each candidate confirmed by the exact comparison, as the host
// enumeration does.
struct ShellDef {
    double low2, high2;
    int lim;
};
__device__ __forceinline__ ShellDef shell_def(uint32_t r, double pitch) {
    const double dist = (double)r / pitch, hw = 0.5 / pitch;
    double lo = dist - hw;
    if (lo < 0.0) lo = 0.0;
    ShellDef s;
    s.low2 = lo * lo;
    const double hi = dist + hw;
    s.high2 = hi * hi;
    s.lim = (int)ceil(dist + hw + 1.0);
    return s;
}

// Calls emit(dz) for the column's admissible dz in ascending order; returns the count.
template <class F>
__device__ __forceinline__ uint32_t shell_column(const ShellDef &s, int dx, int dy, F emit) {
    const int64_t r2 = (int64_t)dx * dx + (int64_t)dy * dy;
    if ((double)r2 >= s.high2) return 0;
    const double zmax_f = sqrt(fmax(0.0, s.high2 - (double)r2));
    const int zhi = min(s.lim, (int)zmax_f + 2);
    const double zmin_f = sqrt(fmax(0.0, s.low2 - (double)r2));
    const int zlo = max(0, (int)zmin_f - 2);
    uint32_t n = 0;
    auto ok = [&](int dz) {
        if (dx == 0 && dy == 0 && dz == 0) return false;
        const double d2 = (double)(r2 + (int64_t)dz * dz);
        return d2 >= s.low2 && d2 < s.high2;
    };
    for (int dz = -zhi; dz <= -zlo; ++dz)
        if (ok(dz)) {
            emit(dz, n);
            ++n;
        }
    for (int dz = max(zlo, 1); dz <= zhi; ++dz)
        if (ok(dz)) {
            emit(dz, n);
            ++n;
        }
    return n;
}

// Entries of each radius' shell: one workgroup per radius, columns strided.
__global__ __launch_bounds__(256) void noalloc_shell_count(double pitch, uint32_t *__restrict__ counts) {
    const uint32_t r = blockIdx.x + 1;
    const ShellDef s = shell_def(r, pitch);
    const int side = 2 * s.lim + 1;
    const int64_t cols = (int64_t)side * side;
    uint32_t n = 0;
    for (int64_t c = threadIdx.x; c < cols; c += blockDim.x)
        n += shell_column(s, (int)(c / side) - s.lim, (int)(c % side) - s.lim, [](int, uint32_t) {});
    for (int off = 32; off > 0; off >>= 1) n += __shfl_down(n, off, 64);
    __shared__ uint32_t part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) counts[r] = part[0] + part[1] + part[2] + part[3];
}

// Radius starts (exclusive scan of the counts, one lane), the capacity check
// and the zeroed counters.  When the table would not fit, every start is 0:
// the fill and the estimator then touch no table entry.
__global__ void noalloc_shell_starts(uint32_t r_max, uint64_t capacity, uint32_t *__restrict__ start,
                                     unsigned long long *__restrict__ hv, unsigned long long *__restrict__ hdr) {
    for (uint32_t i = threadIdx.x; i < 2 * (r_max + 1); i += blockDim.x) hv[i] = 0;
    if (threadIdx.x != 0) return;
    uint64_t total = 0;
    for (uint32_t r = 1; r <= r_max; ++r) total += start[r];  // start[r] holds radius r's count here
    hdr[0] = total;
    hdr[1] = total > capacity ? 1ull : 0ull;
    uint64_t run = 0;
    for (uint32_t r = 1; r <= r_max; ++r) {
        const uint64_t c = start[r];
        start[r] = total > capacity ? 0u : (uint32_t)run;
        run += c;
    }
    start[0] = 0;
    start[r_max + 1] = total > capacity ? 0u : (uint32_t)total;
}

// The shell table in order: per radius, columns in chunks of 256 with a
// workgroup exclusive scan placing each column's run.
__global__ __launch_bounds__(256) void noalloc_shell_fill(double pitch, const uint32_t *__restrict__ start,
                                                          const unsigned long long *__restrict__ hdr,
                                                          Sh16 *__restrict__ table) {
    if (hdr[1]) return;
    const uint32_t r = blockIdx.x + 1;
    const ShellDef s = shell_def(r, pitch);
    const int side = 2 * s.lim + 1;
    const int64_t cols = (int64_t)side * side;
    __shared__ uint32_t wsum[4];
    uint64_t base = start[r];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t c0 = 0; c0 < cols; c0 += blockDim.x) {
        const int64_t c = c0 + threadIdx.x;
        const int dx = c < cols ? (int)(c / side) - s.lim : 0, dy = c < cols ? (int)(c % side) - s.lim : 0;
        const uint32_t n = c < cols ? shell_column(s, dx, dy, [](int, uint32_t) {}) : 0u;
        uint32_t incl = n;  // inclusive scan in the wave
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint32_t before = 0, chunk = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) before += wsum[w];
            chunk += wsum[w];
        }
        if (n) {
            Sh16 *dst = table + base + before + (incl - n);
            shell_column(s, dx, dy, [&](int dz, uint32_t k) { dst[k] = Sh16{(int16_t)dx, (int16_t)dy, (int16_t)dz, 0}; });
        }
        base += chunk;
        __syncthreads();
    }
}

}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_mc_counts_noalloc(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, uint64_t samples,
                                 uint64_t seed, uint64_t *hits_valid, uint64_t n) {
    if (!ctx || !vol || !hits_valid) return fail(S2H_ERR_ARG, "null argument");
    if (n != 2 * ((uint64_t)r_max + 1)) return fail(S2H_ERR_SIZE, "hits_valid must hold 2*(r_max+1) entries");
    if (!(pitch > 0.0) || !std::isfinite(pitch)) return fail(S2H_ERR_ARG, "pitch must be finite and > 0");
    if (r_max > kNoallocMaxR)
        return fail(S2H_ERR_BUDGET, "r_max above " + std::to_string(kNoallocMaxR) + " for the no-allocation scratch");
    if ((double)r_max / pitch + 0.5 / pitch + 1.0 > 32767.0)
        return fail(S2H_ERR_UNSUPPORTED, "Monte Carlo shell radius exceeds 32767 voxels");
    const NoallocLayout l = noalloc_layout(r_max);
    // Consecutive shells tile the ball: every entry is a lattice point d with
    // |d| < r_max / pitch + 0.5 / pitch, and the unit cubes around those points
    // lie in the ball grown by half a cube diagonal, so its volume bounds the
    // table (1.18 million entries at r_max 64, 3.87 million at 96, pitch 1)
    const double reach = (double)r_max / pitch + 0.5 / pitch + 0.8661;
    const double bound = 4.18879020478639 * reach * reach * reach;
    if (bound > (double)l.capacity)
        return fail(S2H_ERR_BUDGET, "Monte Carlo shells of r_max " + std::to_string(r_max) + " at pitch " +
                                        std::to_string(pitch) + " may need up to " + std::to_string((uint64_t)bound) +
                                        " table entries; the no-allocation scratch holds " +
                                        std::to_string(l.capacity));
    S2H_HIP(hipSetDevice(ctx->device));
    S2H_TRY(inject(kFailLaunch));
    if (ctx->device < 0 || ctx->device >= kNoallocMaxDevices || !g_noalloc_ptr[ctx->device])
        S2H_TRY(noalloc_scratch_init(ctx->device));
    unsigned char *base = (unsigned char *)g_noalloc_ptr[ctx->device];
    auto *hdr = (unsigned long long *)base;
    auto *start = (uint32_t *)(base + l.start_off);
    auto *hv = (unsigned long long *)(base + l.hv_off);
    auto *table = (Sh16 *)(base + l.table_off);
    hipStream_t s = ctx->stream;
    if (r_max > 0) {
        noalloc_shell_count<<<r_max, 256, 0, s>>>(pitch, start);
        S2H_HIP(hipGetLastError());
    }
    noalloc_shell_starts<<<1, 256, 0, s>>>(r_max, l.capacity, start, hv, hdr);
    S2H_HIP(hipGetLastError());
    if (r_max > 0) {
        noalloc_shell_fill<<<r_max, 256, 0, s>>>(pitch, start, hdr, table);
        S2H_HIP(hipGetLastError());
    }
    if (r_max > 0 && samples > 0) {
        const uint64_t per_block = (uint64_t)kMcThreads * kMcPerThread;
        const uint64_t bpr = (samples + per_block - 1) / per_block;
        if (bpr * r_max > 0x7FFFFFFFull) return fail(S2H_ERR_UNSUPPORTED, "too many Monte Carlo samples");
        mc_kernel<<<(unsigned)(bpr * r_max), kMcThreads, 0, s>>>(
            vol->d_words, (uint32_t)vol->dims[0], (uint32_t)vol->dims[1], (uint32_t)vol->dims[2], vol->row_words,
            table, start, samples, (uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)bpr, hv);
        S2H_HIP(hipGetLastError());
    }
    unsigned long long flags[2] = {0, 0};
    S2H_TRY(inject(kFailReadback));
    S2H_HIP(hipMemcpyAsync(flags, hdr, sizeof(flags), hipMemcpyDeviceToHost, s));
    S2H_HIP(hipMemcpyAsync(hits_valid, hv, n * 8, hipMemcpyDeviceToHost, s));
    S2H_HIP(hipStreamSynchronize(s));
    if (flags[1])
        return fail(S2H_ERR_BUDGET, "Monte Carlo shell table of " + std::to_string(flags[0]) + " entries exceeds the " +
                                        std::to_string(l.capacity) + " that the no-allocation scratch holds");
    return S2H_OK;
}

s2h_status s2h_mc_counts(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, uint64_t samples, uint64_t seed,
                         uint64_t *hits_valid, uint64_t n) {
    if (!ctx || !vol || !hits_valid) return fail(S2H_ERR_ARG, "null argument");
    if (n != 2 * ((uint64_t)r_max + 1)) return fail(S2H_ERR_SIZE, "hits_valid must hold 2*(r_max+1) entries");
    if (!(pitch > 0.0) || !std::isfinite(pitch)) return fail(S2H_ERR_ARG, "pitch must be finite and > 0");
    S2H_HIP(hipSetDevice(ctx->device));
    S2H_TRY(inject(kFailLaunch));
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
    S2H_TRY(inject(kFailLaunch));
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
