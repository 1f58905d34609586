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
//   - pack_tile_kernel (rows whose byte length is a multiple of 8): a
//     workgroup classifies a tile of x values by 4 rows for all z words,
//     lanes reading 8 bytes at consecutive x per slice (whole 128-byte lines),
//     decodes the file's width, sign and byte order (decode_raw_slice,
//     src/io/volume.rs:287-332), classifies, and writes the tile's row words
//     from LDS as contiguous runs; pack_kernel (any width): one wave per
//     (64 x values, y, z word), one sample per lane.  Bits at z >= depth and
//     the pad word are written as zero.
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

// Rule for samples of at most 16 bits, in "biased" unsigned form: the sample
// bits u (sign bit flipped for signed samples, so u orders like the value) are
// tested as u - lo <= span in uint32 (range rule: lo <= u <= lo + span) or
// u != bg (foreground).  Bounds outside the sample range are clamped to
// -1 / 2^bits, which keeps every comparison's outcome; an empty range becomes
// lo = 2^bits, span 0, which no sample matches.
struct Rule16 {
    uint32_t lo, span, bg;
    uint32_t flip;  // XOR mask applied to a whole 32-bit load (sign bits)
    uint32_t perm;  // v_perm selector: byte swap per 16-bit sample (big-endian) or identity
};

inline Rule16 rule16(const Rule &r, int bits, bool sgn, bool big) {
    const int64_t bias = sgn ? (int64_t(1) << (bits - 1)) : 0, top = int64_t(1) << bits;
    auto clamp = [&](int64_t v) {
        v = std::min<int64_t>(top, std::max<int64_t>(-top, v));  // no overflow in v + bias
        return std::min<int64_t>(top, std::max<int64_t>(-1, v + bias));
    };
    int64_t lo = clamp(r.lo), hi = clamp(r.hi);
    if (lo > hi) lo = hi = top;
    Rule16 q;
    q.lo = (uint32_t)lo;
    q.span = (uint32_t)(hi - lo);
    q.bg = (uint32_t)clamp(r.bg);
    q.flip = !sgn ? 0u : bits == 8 ? 0x80808080u : 0x80008000u;
    q.perm = bits == 16 && big ? 0x02030001u : 0x03020100u;
    return q;
}

// Occupancy of the S = 8 / B samples of one 8-byte load as S flags.
template <int B, bool RANGE>
__device__ __forceinline__ void classify8(uint2 v, bool *occ, int sgn, int big, const Rule &rule, const Rule16 &r16) {
    constexpr int S = 8 / B;
    if (B <= 2) {
        uint32_t w0 = v.x, w1 = v.y;
        if (B == 2) {
            w0 = __builtin_amdgcn_perm(w0, w0, r16.perm);
            w1 = __builtin_amdgcn_perm(w1, w1, r16.perm);
        }
        w0 ^= r16.flip;
        w1 ^= r16.flip;
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const uint32_t u = ((j < S / 2 ? w0 : w1) >> (8 * B * (j % (S / 2)))) & (B == 1 ? 0xffu : 0xffffu);
            occ[j] = RANGE ? (u - r16.lo <= r16.span) : (u != r16.bg);
        }
    } else {
        const uint8_t *vb = (const uint8_t *)&v;
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int64_t x = sample<B>(vb, j, sgn, big);
            occ[j] = RANGE ? ((rule.lo <= x) & (x <= rule.hi)) : (x != rule.bg);
        }
    }
}

// One row word per sample column: bit k of bits[j] is slice z0 + k of column
// j.  p points at slice z0 of the lane's 8 bytes; step is one slice in bytes.
template <int B, bool RANGE>
__device__ __forceinline__ void pack_word(const uint8_t *p, uint64_t step, uint32_t nz, uint64_t *bits, int sgn,
                                          int big, const Rule &rule, const Rule16 &r16) {
    constexpr int S = 8 / B;
    bool occ[S];
    if (nz == 64) {
        // 16 loads in flight per lane, then their 16 x S samples
#pragma unroll
        for (int kb = 0; kb < 64; kb += 16) {
            uint2 v[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                v[i] = *(const uint2 *)p;
                p += step;
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                classify8<B, RANGE>(v[i], occ, sgn, big, rule, r16);
#pragma unroll
                for (int j = 0; j < S; ++j) bits[j] |= (uint64_t)occ[j] << (kb + i);
            }
        }
    } else {
        for (uint32_t k = 0; k < nz; ++k, p += step) {
            classify8<B, RANGE>(*(const uint2 *)p, occ, sgn, big, rule, r16);
#pragma unroll
            for (int j = 0; j < S; ++j) bits[j] |= (uint64_t)occ[j] << k;
        }
    }
}

// Rows whose byte length is a multiple of 8.  A workgroup packs a tile of
// XT = 16 x S x values by YT = 4 rows y, for all z words of those rows.  Lane
// (g, r) of a wave loads 8 bytes (S = 8 / B samples at consecutive x) of row
// y0 + r per slice, so the 16 lanes of a row read one whole 128-byte line per
// slice; wave v packs the row words v, v + 4, ... of a chunk of wc words into
// LDS.  The chunk then leaves LDS in volume order: for one x the YT rows are
// adjacent rows of the volume, so each x's YT x wc words are one contiguous
// run when the chunk holds the whole row (wc == row_words), and the stores are
// whole runs instead of single 8-byte words of rows 100+ bytes apart.  The
// last word of every row is the zero pad and is written as zero.
template <int B>
__global__ __launch_bounds__(256, 3) void pack_tile_kernel(const uint8_t *__restrict__ s, uint64_t W, uint64_t H,
                                                        uint64_t D, int sgn, int big, Rule rule, Rule16 r16,
                                                        uint64_t row_words, uint32_t wc, uint64_t tiles_x,
                                                        uint64_t ntiles, unsigned long long *__restrict__ words) {
    constexpr int S = 8 / B, XT = 16 * S, YT = 4;
    extern __shared__ unsigned long long tile[];  // [XT][YT * wc + 1]
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t g = lane & 15, r = lane >> 4;
    const uint32_t pitch = YT * wc + 1;
    const uint64_t plane = W * H, step = plane * B, nw = row_words - 1;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t x0 = (t % tiles_x) * XT, y0 = (t / tiles_x) * YT;
        const uint64_t xs = x0 + g * S, y = y0 + r;
        const bool live = xs < W && y < H;  // W % S == 0: a live lane's S samples are all inside the row
        for (uint64_t c0 = 0; c0 < row_words; c0 += wc) {
            const uint32_t cw = (uint32_t)std::min<uint64_t>(wc, row_words - c0);
            for (uint32_t wi = wave; wi < cw; wi += 4) {
                const uint64_t w = c0 + wi;
                uint64_t bits[S];
#pragma unroll
                for (int j = 0; j < S; ++j) bits[j] = 0;
                if (live && w < nw) {
                    const uint64_t z0 = w * 64;
                    const uint32_t nz = (uint32_t)std::min<uint64_t>(64, D - z0);
                    const uint8_t *p = s + (z0 * plane + y * W + xs) * B;  // 8-byte aligned: W * B % 8 == 0
                    if (rule.range) pack_word<B, true>(p, step, nz, bits, sgn, big, rule, r16);
                    else pack_word<B, false>(p, step, nz, bits, sgn, big, rule, r16);
                }
#pragma unroll
                for (int j = 0; j < S; ++j) tile[(g * S + j) * pitch + r * wc + wi] = bits[j];
            }
            __syncthreads();
            const uint32_t run = YT * cw, total = XT * run;
            for (uint32_t l = threadIdx.x; l < total; l += 256) {
                const uint32_t xi = l / run, q = l - xi * run, yi = q / cw, wi = q - yi * cw;
                const uint64_t x = x0 + xi, yy = y0 + yi;
                if (x < W && yy < H) words[(x * H + yy) * row_words + c0 + wi] = tile[xi * pitch + yi * wc + wi];
            }
            __syncthreads();
        }
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
        // Tile of 16 x (8 / bytes) x values by 4 rows; chunks of wc row words keep
        // the tile's LDS within 64 KiB.
        const uint64_t xt = 16 * (8 / bytes), yt = 4;
        const uint32_t wc = (uint32_t)std::min<uint64_t>(vol->row_words, (65536 / (xt * 8) - 1) / yt);
        const uint64_t tiles_x = (l->width + xt - 1) / xt, ntiles = tiles_x * ((l->height + yt - 1) / yt);
        const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, 1u << 22);
        const size_t lds = (size_t)xt * (yt * wc + 1) * 8;
        const volin::Rule16 r16 = volin::rule16(rule, (int)l->bits, l->is_signed, l->big_endian);
        if (bytes == 1)
            volin::pack_tile_kernel<1><<<grid, 256, lds, ctx->stream>>>(d_samples, l->width, l->height, l->depth,
                                                                        l->is_signed, l->big_endian, rule, r16,
                                                                        vol->row_words, wc, tiles_x, ntiles, words);
        else if (bytes == 2)
            volin::pack_tile_kernel<2><<<grid, 256, lds, ctx->stream>>>(d_samples, l->width, l->height, l->depth,
                                                                        l->is_signed, l->big_endian, rule, r16,
                                                                        vol->row_words, wc, tiles_x, ntiles, words);
        else
            volin::pack_tile_kernel<4><<<grid, 256, lds, ctx->stream>>>(d_samples, l->width, l->height, l->depth,
                                                                        l->is_signed, l->big_endian, rule, r16,
                                                                        vol->row_words, wc, tiles_x, ntiles, words);
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
