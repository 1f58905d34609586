// Volume ingest on the device: z-major sample volumes (RAW slice folders, TIFF
// stacks or a buffer in memory) classified into occupancy and bit-packed into
// the S2 row layout.
//
// The samples go to HBM whole (a 768^3 uint8 volume is 453 MB): a folder is
// streamed slice by slice through two pinned staging buffers, so reading the
// next files overlaps the copy of the previous ones.  Then:
//   - FOREGROUND_AUTO: a boundary histogram (dense bins for 8/16-bit samples,
//     gathered values for 32-bit ones) gives the modal boundary value, smallest
//     on ties, over exactly the voxels detect_background_mode visits
//     (for_each_boundary_value, src/pipeline/crop.rs:312-330);
//   - pack_kernel: one wave per (64 x values, y, z word); lane x reads its 64
//     samples along z (each read is 64 consecutive x: coalesced), decodes the
//     file's width, sign and byte order (decode_raw_slice,
//     src/io/volume.rs:287-332), classifies, and writes the row word.  Bits at
//     z >= depth and the pad word are written as zero.
// File axes map one to one: S2 (x, y, z) = file (x, y, slice).
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "s2hip_internal.h"
#include "s2hip_volio.h"

namespace s2h {
namespace volin {

struct Rule {
    int32_t range;  // 1: lo <= v <= hi; 0: v != bg
    int64_t lo, hi, bg;
};

template <int B>
__device__ __forceinline__ int64_t sample(const uint8_t *__restrict__ s, uint64_t i, bool sgn, bool big) {
    if (B == 1) {
        const uint8_t v = s[i];
        return sgn ? (int64_t)(int8_t)v : (int64_t)v;
    } else if (B == 2) {
        uint16_t v = ((const uint16_t *)s)[i];
        if (big) v = (uint16_t)((v >> 8) | (v << 8));
        return sgn ? (int64_t)(int16_t)v : (int64_t)v;
    } else {
        uint32_t v = ((const uint32_t *)s)[i];
        if (big) v = __builtin_bswap32(v);
        return sgn ? (int64_t)(int32_t)v : (int64_t)v;
    }
}

__device__ __forceinline__ bool occupied(int64_t v, const Rule &r) {
    return r.range ? (r.lo <= v && v <= r.hi) : (v != r.bg);
}

// Generic shape: one sample per lane.
template <int B>
__global__ __launch_bounds__(64) void pack_kernel(const uint8_t *__restrict__ s, uint64_t W, uint64_t H, uint64_t D,
                                                  int sgn, int big, Rule rule, uint64_t row_words, uint64_t nblocks,
                                                  unsigned long long *__restrict__ words) {
    const uint64_t plane = W * H;
    for (uint64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const uint64_t w = b % row_words, t = b / row_words;
        const uint64_t y = t % H, x = (t / H) * 64 + threadIdx.x;
        if (x >= W) continue;
        const uint64_t z0 = w * 64;
        const uint32_t nz = z0 < D ? (uint32_t)std::min<uint64_t>(64, D - z0) : 0;
        uint64_t bits = 0;
        uint64_t i = z0 * plane + y * W + x;
        for (uint32_t k = 0; k < nz; ++k, i += plane)
            if (occupied(sample<B>(s, i, sgn, big), rule)) bits |= 1ull << k;
        words[(x * H + y) * row_words + w] = bits;
    }
}

// Rows of 8-byte multiples: each lane loads 8 bytes (8 / B samples at
// consecutive x) per slice, so a wave reads 512 contiguous bytes per load.
// One word per lane and x: the 8-byte row words land 104+ bytes apart, which
// the memory system writes as partial sectors (about 3.7x the packed bytes at
// 768^3); the kernel is still far below the host read and PCIe copy it follows.
template <int B>
__global__ __launch_bounds__(64) void pack_wide_kernel(const uint8_t *__restrict__ s, uint64_t W, uint64_t H,
                                                       uint64_t D, int sgn, int big, Rule rule, uint64_t row_words,
                                                       uint64_t nblocks, unsigned long long *__restrict__ words) {
    constexpr int S = 8 / B;
    const uint64_t plane = W * H;
    for (uint64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const uint64_t w = b % row_words, t = b / row_words;
        const uint64_t y = t % H, xs = ((t / H) * 64 + threadIdx.x) * S;
        if (xs >= W) continue;
        const uint64_t z0 = w * 64;
        const uint32_t nz = z0 < D ? (uint32_t)std::min<uint64_t>(64, D - z0) : 0;
        uint64_t bits[S];
#pragma unroll
        for (int j = 0; j < S; ++j) bits[j] = 0;
        const uint64_t base = z0 * plane + y * W + xs;  // in samples; W * B % 8 == 0 keeps this 8-byte aligned
        for (uint32_t k = 0; k < nz; ++k) {
            const uint2 v = *(const uint2 *)(s + (base + k * plane) * B);
            const uint8_t *vb = (const uint8_t *)&v;
#pragma unroll
            for (int j = 0; j < S; ++j)
                if (occupied(sample<B>(vb, j, sgn, big), rule)) bits[j] |= 1ull << k;
        }
#pragma unroll
        for (int j = 0; j < S; ++j)
            if (xs + j < W) words[((xs + j) * H + y) * row_words + w] = bits[j];
    }
}

// Boundary rows: (z, y) with z the first or last slice, or y the first or last
// row, contribute all W values; other rows their first and (W > 1) last value.
// Updates are aggregated per wave: lanes holding the same value add once (the
// boundary is mostly one value, so one atomic per wave is the common case).
template <int B>
__global__ __launch_bounds__(256) void boundary_kernel(const uint8_t *__restrict__ s, uint64_t W, uint64_t H,
                                                       uint64_t D, int sgn, int big, int64_t base, uint32_t nbins,
                                                       unsigned long long *__restrict__ hist,
                                                       int64_t *__restrict__ values, uint64_t cap,
                                                       unsigned long long *count) {
    const uint64_t rows = D * H;
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const uint64_t z = r / H, y = r % H;
        const bool full = z == 0 || z + 1 == D || y == 0 || y + 1 == H;
        const uint64_t n = full ? W : (W > 1 ? 2 : 1);
        // every wave of the block runs the same number of iterations (n is block-uniform)
        for (uint64_t k0 = 0; k0 < n; k0 += blockDim.x) {
            const uint64_t k = k0 + threadIdx.x;
            const bool have = k < n;
            int64_t v = 0;
            if (have) v = sample<B>(s, (z * H + y) * W + (full ? k : (k == 0 ? 0 : W - 1)), sgn, big);
            if (nbins) {
                uint64_t todo = __ballot(have);
                while (todo) {
                    const int leader = __ffsll((long long)todo) - 1;
                    const int64_t lv = __shfl(v, leader);
                    const uint64_t same = __ballot(have && v == lv) & todo;
                    if (lane == (uint32_t)leader) atomicAdd(&hist[(uint64_t)(lv - base)], (unsigned long long)__popcll(same));
                    todo &= ~same;
                }
            } else {
                const uint64_t act = __ballot(have);
                const int leader = __ffsll((long long)act) - 1;
                unsigned long long at0 = 0;
                if (act && lane == (uint32_t)leader) at0 = atomicAdd(count, (unsigned long long)__popcll(act));
                at0 = __shfl(at0, leader < 0 ? 0 : leader);
                if (have) {
                    const uint64_t at = at0 + __popcll(act & ((1ull << lane) - 1));
                    if (at < cap) values[at] = v;
                }
            }
        }
    }
}

uint64_t boundary_count(uint64_t W, uint64_t H, uint64_t D) {
    const uint64_t full_rows_per_mid = H == 1 ? 1 : 2;
    const uint64_t mid_rows = H > 2 ? H - 2 : 0;
    const uint64_t mid_slices = D > 2 ? D - 2 : 0;
    return (D == 1 ? 1 : 2) * W * H + mid_slices * (full_rows_per_mid * W + mid_rows * (W > 1 ? 2 : 1));
}

}  // namespace volin

namespace {

s2h_status check_layout(const s2h_volume_info *l) {
    if (l->bits != 8 && l->bits != 16 && l->bits != 32) return fail(S2H_ERR_ARG, "sample bits must be 8/16/32");
    if (l->width == 0 || l->height == 0 || l->depth == 0) return fail(S2H_ERR_ARG, "volume dimensions must be > 0");
    for (uint64_t v : {l->width, l->height, l->depth})
        if (v > 65535) return fail(S2H_ERR_UNSUPPORTED, "volume dimension " + std::to_string(v) + " outside 1..65535");
    return S2H_OK;
}

template <int B>
void launch_boundary(s2h_ctx *ctx, const uint8_t *s, const s2h_volume_info *l, int64_t base, uint32_t nbins,
                     unsigned long long *hist, int64_t *values, uint64_t cap, unsigned long long *count) {
    const uint64_t rows = l->depth * l->height;
    const unsigned grid = (unsigned)std::min<uint64_t>(rows, 1u << 20);
    volin::boundary_kernel<B><<<grid, 256, 0, ctx->stream>>>(s, l->width, l->height, l->depth, l->is_signed,
                                                             l->big_endian, base, nbins, hist, values, cap, count);
}

// Modal boundary value, smallest on ties (detect_background_mode,
// src/pipeline/crop.rs:338-373).
s2h_status background_mode(s2h_ctx *ctx, const uint8_t *d_samples, const s2h_volume_info *l, int64_t *bg) {
    const uint32_t bytes = l->bits / 8;
    const int64_t base = bytes == 4 ? 0 : (l->is_signed ? -(int64_t(1) << (l->bits - 1)) : 0);
    const uint32_t nbins = bytes == 4 ? 0 : (1u << l->bits);
    const uint64_t nb = volin::boundary_count(l->width, l->height, l->depth);
    DevBuf d_hist, d_vals, d_count;
    auto release = [&]() {
        dev_free(ctx, d_hist);
        dev_free(ctx, d_vals);
        dev_free(ctx, d_count);
    };
    s2h_status st = dev_alloc(ctx, d_count, 8);
    if (st == S2H_OK) st = nbins ? dev_alloc(ctx, d_hist, (uint64_t)nbins * 8) : dev_alloc(ctx, d_vals, nb * 8);
    if (st != S2H_OK) {
        release();
        return st;
    }
    hipError_t e = hipMemsetAsync(d_count.ptr, 0, 8, ctx->stream);
    if (e == hipSuccess && nbins) e = hipMemsetAsync(d_hist.ptr, 0, (uint64_t)nbins * 8, ctx->stream);
    if (e == hipSuccess) {
        auto *hist = (unsigned long long *)d_hist.ptr;
        auto *vals = (int64_t *)d_vals.ptr;
        auto *cnt = (unsigned long long *)d_count.ptr;
        const uint64_t cap = nbins ? 0 : nb;
        if (bytes == 1) launch_boundary<1>(ctx, d_samples, l, base, nbins, hist, vals, cap, cnt);
        else if (bytes == 2) launch_boundary<2>(ctx, d_samples, l, base, nbins, hist, vals, cap, cnt);
        else launch_boundary<4>(ctx, d_samples, l, base, nbins, hist, vals, cap, cnt);
        e = hipGetLastError();
    }
    std::vector<uint64_t> h;
    std::vector<int64_t> vals;
    uint64_t got = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&got, d_count.ptr, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && nbins) {
        h.resize(nbins);
        e = hipMemcpyAsync(h.data(), d_hist.ptr, (uint64_t)nbins * 8, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && !nbins) {
        if (got != nb) {
            release();
            return fail(S2H_ERR_HIP, "boundary enumeration mismatch");
        }
        vals.resize(got);
        e = hipMemcpy(vals.data(), d_vals.ptr, got * 8, hipMemcpyDeviceToHost);
    }
    release();
    if (e != hipSuccess) return fail(S2H_ERR_HIP, std::string("background detection: ") + hipGetErrorString(e));
    int64_t best = 0;
    uint64_t best_count = 0;
    if (nbins) {
        for (uint32_t i = 0; i < nbins; ++i)  // ascending value: the first maximum is the smallest value
            if (h[i] > best_count) {
                best_count = h[i];
                best = base + (int64_t)i;
            }
    } else {
        std::sort(vals.begin(), vals.end());
        for (size_t i = 0; i < vals.size();) {
            size_t j = i;
            while (j < vals.size() && vals[j] == vals[i]) ++j;
            if (j - i > best_count) {
                best_count = j - i;
                best = vals[i];
            }
            i = j;
        }
    }
    *bg = best;
    return S2H_OK;
}

s2h_status pack_samples(s2h_ctx *ctx, const uint8_t *d_samples, const s2h_volume_info *l, const s2h_phase *phase,
                        s2h_vol **out, int64_t *background_used, uint64_t *occupied) {
    volin::Rule rule{1, phase->lo, phase->hi, 0};
    if (phase->mode == S2H_PHASE_FOREGROUND || phase->mode == S2H_PHASE_FOREGROUND_AUTO) {
        rule.range = 0;
        rule.bg = phase->background;
        if (phase->mode == S2H_PHASE_FOREGROUND_AUTO) S2H_TRY(background_mode(ctx, d_samples, l, &rule.bg));
        if (background_used) *background_used = rule.bg;
    } else if (phase->mode != S2H_PHASE_RANGE) {
        return fail(S2H_ERR_ARG, "unknown phase mode " + std::to_string(phase->mode));
    }
    const uint64_t dims[3] = {l->width, l->height, l->depth};
    s2h_vol *vol = nullptr;
    S2H_TRY(vol_alloc(ctx, dims, &vol));
    auto *words = (unsigned long long *)vol->d_words;
    prof_mark(ctx, 0);
    const uint32_t bytes = l->bits / 8;
    if ((l->width * bytes) % 8 == 0 && ((uintptr_t)d_samples & 7) == 0) {
        const uint64_t per_block = 64 * (8 / bytes);
        const uint64_t nblocks = (l->width + per_block - 1) / per_block * l->height * vol->row_words;
        const unsigned grid = (unsigned)std::min<uint64_t>(nblocks, 1u << 22);
        if (bytes == 1)
            volin::pack_wide_kernel<1><<<grid, 64, 0, ctx->stream>>>(d_samples, l->width, l->height, l->depth,
                                                                     l->is_signed, l->big_endian, rule,
                                                                     vol->row_words, nblocks, words);
        else if (bytes == 2)
            volin::pack_wide_kernel<2><<<grid, 64, 0, ctx->stream>>>(d_samples, l->width, l->height, l->depth,
                                                                     l->is_signed, l->big_endian, rule,
                                                                     vol->row_words, nblocks, words);
        else
            volin::pack_wide_kernel<4><<<grid, 64, 0, ctx->stream>>>(d_samples, l->width, l->height, l->depth,
                                                                     l->is_signed, l->big_endian, rule,
                                                                     vol->row_words, nblocks, words);
    } else {
        const uint64_t nblocks = (l->width + 63) / 64 * l->height * vol->row_words;
        const unsigned grid = (unsigned)std::min<uint64_t>(nblocks, 1u << 22);
        if (l->bits == 8)
            volin::pack_kernel<1><<<grid, 64, 0, ctx->stream>>>(d_samples, l->width, l->height, l->depth, l->is_signed,
                                                                l->big_endian, rule, vol->row_words, nblocks, words);
        else if (l->bits == 16)
            volin::pack_kernel<2><<<grid, 64, 0, ctx->stream>>>(d_samples, l->width, l->height, l->depth, l->is_signed,
                                                                l->big_endian, rule, vol->row_words, nblocks, words);
        else
            volin::pack_kernel<4><<<grid, 64, 0, ctx->stream>>>(d_samples, l->width, l->height, l->depth, l->is_signed,
                                                                l->big_endian, rule, vol->row_words, nblocks, words);
    }
    ctx->launches = 1;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        s2h_vol_free(ctx, vol);
        return fail(S2H_ERR_HIP, std::string("pack kernel: ") + hipGetErrorString(e));
    }
    prof_mark(ctx, 1);
    prof_mark(ctx, 2);
    uint64_t occ = 0;
    s2h_status st = vol_popcount(ctx, vol, &occ);
    prof_collect(ctx);
    if (st != S2H_OK) {
        s2h_vol_free(ctx, vol);
        return st;
    }
    if (occupied) *occupied = occ;
    *out = vol;
    return S2H_OK;
}

// Two pinned staging buffers of whole slices; a full buffer is copied to its
// place in the device sample buffer while the other one fills.
struct Staging {
    s2h_ctx *ctx;
    uint8_t *dst;          // device sample buffer
    uint64_t slice_bytes, per_buf, cap_bytes;
    uint8_t *buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    int cur = 0;
    uint64_t fill = 0;     // slices in buf[cur]
    uint64_t written = 0;  // slices already queued for copy
    std::string error;

    s2h_status init(uint64_t sb, uint64_t depth) {
        slice_bytes = sb;
        per_buf = std::max<uint64_t>(1, std::min<uint64_t>(depth, (64ull << 20) / std::max<uint64_t>(sb, 1)));
        cap_bytes = per_buf * sb;
        for (int i = 0; i < 2; ++i) {
            if (hipHostMalloc((void **)&buf[i], cap_bytes, hipHostMallocDefault) != hipSuccess) {
                buf[i] = nullptr;
                return fail(S2H_ERR_HIP, "pinned staging allocation failed");
            }
            if (hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess)
                return fail(S2H_ERR_HIP, "event creation failed");
        }
        return S2H_OK;
    }
    uint8_t *slot() {
        if (fill == 0 && pending[cur]) {
            if (hipEventSynchronize(done[cur]) != hipSuccess) return nullptr;
            pending[cur] = false;
        }
        return buf[cur] + fill * slice_bytes;
    }
    bool flush() {
        if (fill == 0) return true;
        if (hipMemcpyAsync(dst + written * slice_bytes, buf[cur], fill * slice_bytes, hipMemcpyHostToDevice,
                           ctx->stream) != hipSuccess ||
            hipEventRecord(done[cur], ctx->stream) != hipSuccess)
            return false;
        pending[cur] = true;
        written += fill;
        fill = 0;
        cur ^= 1;
        return true;
    }
    bool commit() {
        ++fill;
        return fill < per_buf || flush();
    }
    ~Staging() {
        for (int i = 0; i < 2; ++i) {
            if (pending[i]) (void)hipEventSynchronize(done[i]);
            if (done[i]) (void)hipEventDestroy(done[i]);
            if (buf[i]) (void)hipHostFree(buf[i]);
        }
    }
};

}  // namespace
}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_vol_from_samples(s2h_ctx *ctx, const void *samples, uint64_t nbytes, const s2h_volume_info *layout,
                                const s2h_phase *phase, s2h_vol **out, int64_t *background_used,
                                uint64_t *occupied) {
    if (!ctx || !samples || !layout || !phase || !out) return fail(S2H_ERR_ARG, "null argument");
    S2H_TRY(check_layout(layout));
    const uint64_t need = layout->width * layout->height * layout->depth * (layout->bits / 8);
    if (nbytes != need)
        return fail(S2H_ERR_SIZE, "volume has " + std::to_string(need) + " sample bytes, buffer " + std::to_string(nbytes));
    S2H_HIP(hipSetDevice(ctx->device));
    DevBuf d;
    S2H_TRY(upload(ctx, d, samples, nbytes));
    s2h_status st = pack_samples(ctx, (const uint8_t *)d.ptr, layout, phase, out, background_used, occupied);
    dev_free(ctx, d);
    return st;
}

s2h_status s2h_vol_load(s2h_ctx *ctx, const s2h_volume_source *src, const s2h_phase *phase, s2h_vol **out,
                        s2h_volume_info *info, int64_t *background_used, uint64_t *occupied) {
    if (!ctx || !src || !phase || !out) return fail(S2H_ERR_ARG, "null argument");
    s2h_volume_info l;
    std::string err;
    if (!volio::probe(*src, l, err)) return fail(S2H_ERR_IO, err);
    S2H_TRY(check_layout(&l));
    if (info) *info = l;
    S2H_HIP(hipSetDevice(ctx->device));
    const uint64_t slice_bytes = l.width * l.height * (l.bits / 8);
    DevBuf d;
    S2H_TRY(dev_alloc(ctx, d, slice_bytes * l.depth));
    s2h_status st;
    {
        Staging stg;
        stg.ctx = ctx;
        stg.dst = (uint8_t *)d.ptr;
        st = stg.init(slice_bytes, l.depth);
        if (st == S2H_OK) {
            volio::SliceSink sink;
            sink.slot = [&](uint64_t bytes, std::string &e) -> uint8_t * {
                uint8_t *p = bytes == slice_bytes && stg.written + stg.fill < l.depth ? stg.slot() : nullptr;
                if (!p) e = "staging failed (slice size or count changed, or a copy failed)";
                return p;
            };
            sink.commit = [&](std::string &e) {
                if (!stg.commit()) {
                    e = "host-to-device copy failed";
                    return false;
                }
                return true;
            };
            if (!volio::stream(*src, l, sink, err)) st = fail(S2H_ERR_IO, err);
            else if (!stg.flush() || stg.written != l.depth) st = fail(S2H_ERR_IO, "volume copy incomplete");
        }
    }  // staging destructor waits for the copies
    if (st == S2H_OK) st = pack_samples(ctx, (const uint8_t *)d.ptr, &l, phase, out, background_used, occupied);
    dev_free(ctx, d);
    return st;
}

}  // extern "C"
