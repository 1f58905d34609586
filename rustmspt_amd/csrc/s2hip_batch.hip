// Batched evaluation of simulated-annealing candidates by exact delta counts.
//
// The reference's optimize loop (src/pipeline/optimize.rs:359-833) moves one
// particle per iteration, updates the occupancy incrementally with per-voxel
// coverage counts (VoxelCoverage, src/geometry/s2.rs:266-328) and evaluates the
// full exact S2 of the candidate (src/pipeline/optimize_execution.rs:326-343).
// Here many candidates -- each one particle move from a resident seed packing,
// seeds of different grid sizes allowed -- are evaluated in one pass:
//
//  1. the moved particle's granules are re-voxelized at the old and the new
//     position with the bit-exact predicate, as -1 / +1 coverage deltas inside
//     the candidate's changed box B (triangle data prepared on the device from
//     the seed's resident geometry and the uploaded new vertices);
//  2. for every row crossing B, the added bits A = new & ~old and the removed
//     bits D = old & ~new (new bit = coverage + delta > 0), and each row's
//     z-span of changes;
//  3. exact delta counts.  With n = o + A - D and delta = A - D,
//        n(u) n(u+d) - o(u) o(u+d) = n(u) delta(u+d) + delta(u) o(u+d),
//     so C'(d) - C(d) = sum_u n(u) (A - D)(u + d) + sum_u (A - D)(u) o(u + d):
//     one pass over each changed row's partners on each side, with no
//     separate changed-changed term.  A wave takes one changed row q
//     (wave-uniform: its A and D words shifted by dz on the scalar unit) and 64
//     half-ball columns (dx, dy) as lanes, which read the rows q - (dx, dy) of
//     the new occupancy n (the seed row, with A and D applied inside the box)
//     and q + (dx, dy) of the seed -- so only the words inside q's own change
//     span are visited, and unchanged rows are skipped;
//  4. C' = C + dC, then the reference's ordered f64 fold per radius -> curves.
//     Delta counts and the seed's counts are kept in canonical shell order
//     (each radius' canonical offsets contiguous, in order of first
//     appearance in the reference's shell order), so a radius' fold reads one
//     contiguous slice of each instead of gathering across the whole table.
// All integer arithmetic, so every curve is bit-identical to a full evaluation
// of the candidate packing.
//
// Every kernel is grid-stride over work described in a device header, so the
// launch sequence does not depend on the batch: it is captured once into a HIP
// graph over a persistent workspace and replayed for every later batch (the
// inputs are copied in before each replay).  A committed move (s2h_seed_accept)
// applies one candidate's deltas to the seed's coverage, rows and counts.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "s2hip_internal.h"
#include "s2hip_voxel_dev.h"

namespace s2h {
namespace batch {

struct GB {            // (column group, dz block) of a seed's exact plan
    uint32_t group;
    uint16_t dz0, dzb;
};

struct SeedDev {
    const uint32_t *vol;        // packed seed rows (dwords)
    const uint32_t *cov;        // granule coverage count per voxel
    const unsigned long long *shtable;  // seed counts in canonical shell order
    const int32_t *sg;          // shell order -> canonical shell index (-1: zero offset)
    const int32_t *smap;        // count-table index -> canonical shell index (ncanon: no shell)
    const Offset *offs;         // shell order offsets
    const uint64_t *rstart;     // r_max + 2
    const ColInfo *cols;
    const GB *gbs;
    uint64_t occupied;
    uint32_t vol_bytes, ngb, r_max, pad;
    int32_t nx, ny, nz, row_dw;
};

struct CandDev {
    uint32_t seed;
    int32_t bx0, by0, bz0, bx1, by1, bz1;   // changed box, voxel index ranges
    uint32_t nchunks;                        // row chunks of the box
    uint64_t ad_off;                         // uint2 offset of the (A, D) words (box rows x nwd)
    uint64_t info_off;                       // uint2 offset of the row change spans (box rows)
    uint64_t box_off;                        // int32 offset of the coverage delta box
    uint64_t delta_off;                      // int32 offset of the delta table (canonical shell order)
    uint64_t item_start;                     // first delta item of this candidate
    uint64_t curve_off;                      // double offset of the curve values
};

struct BatchHeader {
    uint64_t nitems;       // delta-count items
    uint64_t ad_total;     // uint2 words of (A, D) rows
    uint64_t info_total;   // rows of all boxes
    uint64_t box_words;    // int32 entries of all coverage-delta boxes
    uint64_t delta_words;  // int32 entries of all delta tables
    uint32_t ncand;
    uint32_t nfold;        // (candidate, radius) fold tasks
    uint32_t ngran;        // prepared granules (both positions of every candidate)
    uint32_t ntiles;
    uint32_t row_chunk;    // box rows per delta item
    uint32_t pad;
};

constexpr uint32_t kOverflowCap = 1u << 16;

// ---------------------------------------------------------------- 0. clear

__global__ void clear_kernel(const BatchHeader *__restrict__ hdr, int32_t *__restrict__ delta,
                             int32_t *__restrict__ box, uint2 *__restrict__ info, int *__restrict__ occ_delta,
                             uint32_t *__restrict__ overflow) {
    const uint64_t nd = hdr->delta_words, nb = hdr->box_words, ni = hdr->info_total, nc = hdr->ncand;
    const uint64_t total = nd + nb + ni + nc + 1;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (i < nd) delta[i] = 0;
        else if (i < nd + nb) box[i - nd] = 0;
        else if (i < nd + nb + ni) info[i - nd - nb] = make_uint2(0u, 0u);
        else if (i < nd + nb + ni + nc) occ_delta[i - nd - nb - ni] = 0;
        else overflow[0] = 0;
    }
}

// ---------------------------------------------------------------- 1. triangle data and coverage deltas

// One block per granule: the predicate's precomputed triangle data from the
// granule's vertices (resident seed geometry or the uploaded new vertices).
__global__ __launch_bounds__(256) void prep_kernel(const vox::GranuleInfo *__restrict__ gran,
                                                   const BatchHeader *__restrict__ hdr,
                                                   vox::TriGeo *__restrict__ geo) {
    const uint32_t ng = hdr->ngran;
    for (uint32_t g = blockIdx.x; g < ng; g += gridDim.x) {
        const vox::GranuleInfo gi = gran[g];
        const uint32_t n = gi.tri_end - gi.tri_start;
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
            const uint32_t *t = gi.tri + 3 * (uint64_t)k;
            vox::tri_prepare(gi.xyz + 3 * (uint64_t)t[0], gi.xyz + 3 * (uint64_t)t[1], gi.xyz + 3 * (uint64_t)t[2],
                             gi.cmax, gi.smax, gi.amax, geo[gi.tri_start + k]);
        }
    }
}

struct AddDelta {
    const CandDev *cands;
    int32_t *box;
    __device__ void operator()(const vox::GranuleInfo &gi, uint32_t x, uint32_t y, uint32_t z) const {
        const CandDev &c = cands[gi.owner];
        const int sx = c.bx1 - c.bx0, sy = c.by1 - c.by0, sz = c.bz1 - c.bz0;
        const int lx = (int)x - c.bx0, ly = (int)y - c.by0, lz = (int)z - c.bz0;
        if (lx < 0 || ly < 0 || lz < 0 || lx >= sx || ly >= sy || lz >= sz) return;  // never: box covers ranges
        atomicAdd(&box[c.box_off + ((uint64_t)lx * sy + ly) * sz + lz], gi.sign);
    }
};

__global__ __launch_bounds__(256) void batch_voxel_kernel(const vox::TriGeo *__restrict__ tris,
                                                          const vox::GranuleInfo *__restrict__ gran,
                                                          const vox::Tile *__restrict__ tiles,
                                                          const BatchHeader *__restrict__ hdr,
                                                          const double *__restrict__ origin,  // per candidate: x,y,z,pitch
                                                          const CandDev *__restrict__ cands,
                                                          int32_t *__restrict__ box, uint32_t *__restrict__ overflow) {
    const uint32_t n = hdr->ntiles;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) {
        const vox::Tile tl = tiles[t];
        const uint32_t c = gran[tl.granule].owner;
        const double *o = origin + 4 * (uint64_t)c;
        vox::voxelize_tile(tris, gran, tl, t, o[0], o[1], o[2], o[3], overflow, kOverflowCap, AddDelta{cands, box});
        __syncthreads();
    }
}

__global__ void batch_voxel_overflow_kernel(const vox::TriGeo *__restrict__ tris,
                                            const vox::GranuleInfo *__restrict__ gran,
                                            const vox::Tile *__restrict__ tiles, const double *__restrict__ origin,
                                            const CandDev *__restrict__ cands, int32_t *__restrict__ box,
                                            const uint32_t *__restrict__ overflow) {
    const uint32_t count = min(overflow[0], kOverflowCap);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        const vox::Tile tl = tiles[overflow[1 + 2 * k]];
        const uint32_t c = gran[tl.granule].owner;
        const double *o = origin + 4 * (uint64_t)c;
        vox::voxel_overflow_ray(tris, gran, tiles, o[0], o[1], o[2], o[3], overflow, k, AddDelta{cands, box});
    }
}

// ---------------------------------------------------------------- 2. added / removed bits

// One thread per (candidate, box row, dword < nwd).  Inside [bz0, bz1) the new
// bit is coverage + delta > 0; outside, nothing changes.  Row change spans are
// kept as (nz - lowest changed z, highest changed z + 1) via atomic max, so a
// zeroed entry means "row unchanged".
__global__ void build_ad_kernel(const SeedDev *__restrict__ seeds, const CandDev *__restrict__ cands,
                                const BatchHeader *__restrict__ hdr, const uint64_t *__restrict__ ad_prefix,
                                const int32_t *__restrict__ box, uint2 *__restrict__ ad, uint2 *__restrict__ info,
                                int *__restrict__ occ_delta) {
    const uint32_t ncand = hdr->ncand;
    const uint64_t total = ad_prefix[ncand];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = ncand;  // candidate of word i
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (ad_prefix[mid] <= i) lo = mid;
            else hi = mid;
        }
        const CandDev &c = cands[lo];
        const SeedDev &s = seeds[c.seed];
        const int nwd = (s.nz + 31) / 32;
        const uint64_t local = i - ad_prefix[lo];
        const uint64_t r = local / nwd;
        const int w = (int)(local % nwd);
        const int sy = c.by1 - c.by0, sz = c.bz1 - c.bz0;
        const int x = c.bx0 + (int)(r / sy), y = c.by0 + (int)(r % sy);
        const uint32_t old = s.vol[((uint64_t)x * s.ny + y) * s.row_dw + w];
        uint32_t v = old;
        const int z0 = max(32 * w, c.bz0), z1 = min(32 * w + 32, c.bz1);
        for (int z = z0; z < z1; ++z) {
            const uint64_t vi = ((uint64_t)x * s.ny + y) * s.nz + z;
            const int lx = x - c.bx0, ly = y - c.by0, lz = z - c.bz0;
            const int cnt = (int)s.cov[vi] + box[c.box_off + ((uint64_t)lx * sy + ly) * sz + lz];
            const uint32_t bit = 1u << (z - 32 * w);
            v = cnt > 0 ? (v | bit) : (v & ~bit);
        }
        const uint32_t add = v & ~old, rem = old & ~v, chg = add | rem;
        ad[c.ad_off + local] = make_uint2(add, rem);
        if (chg) {
            const int zlo = 32 * w + __builtin_ctz(chg), zhi = 32 * w + 32 - __builtin_clz(chg);
            atomicMax(&info[c.info_off + r].x, (uint32_t)(s.nz - zlo));
            atomicMax(&info[c.info_off + r].y, (uint32_t)zhi);
            atomicAdd(&occ_delta[lo], __popc(add) - __popc(rem));
        }
    }
}

// ---------------------------------------------------------------- 3. delta counts

// (A, D) words k of a changed row (zero outside the row's words)
__device__ __forceinline__ uint2 adw(const uint2 *row, int k, int nwd) {
    const int kk = k < 0 ? 0 : (k < nwd ? k : nwd - 1);
    const uint2 v = row[kk];
    return (k >= 0 && k < nwd) ? v : make_uint2(0u, 0u);
}

template <int DZB>
__device__ void delta_item(const SeedDev &s, const CandDev &c, const GB gb, int chunk, int row_chunk,
                           const uint2 *__restrict__ ad, const __amdgpu_buffer_rsrc_t ad_rs,
                           const uint2 *__restrict__ info, int32_t *__restrict__ delta) {
    const int lane = threadIdx.x & 63;
    const ColInfo col = s.cols[gb.group * 64 + lane];
    const int nwd = (s.nz + 31) / 32;
    const int q0 = gb.dz0 >> 5, dz0 = gb.dz0;
    const int sx = c.bx1 - c.bx0, sy = c.by1 - c.by0;
    const int nrows = sx * sy;
    const int r0 = chunk * row_chunk, r1 = min(nrows, r0 + row_chunk);
    const __amdgpu_buffer_rsrc_t seed_rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)s.vol, 0, (int)s.vol_bytes, 0x00020000);
    const int oob = 0x7fffff00;  // + up to 255 bytes: past any buffer, no int overflow
    uint32_t cp[DZB], cm[DZB];  // added and subtracted pair counts per dz
#pragma unroll
    for (int j = 0; j < DZB; ++j) cp[j] = cm[j] = 0;
    const bool active = col.m >= 0;
    for (int r = r0; r < r1; ++r) {
        const uint2 span = info[c.info_off + r];
        if (span.y == 0) continue;  // row unchanged (uniform)
        const int zlo = s.nz - (int)span.x, zhi = (int)span.y;
        const int xq = c.bx0 + r / sy, yq = c.by0 + r % sy;
        const uint2 *m = ad + c.ad_off + (uint64_t)r * nwd;  // uniform (A, D) row of q
        // ---- X(+d) on the new occupancy: lanes u = q - (dx, dy) read n(u) = (o & ~D) | A
        // (n = o outside the box); M(q) >> dz meets u at z in [zlo - dz, zhi - dz)
        {
            const int xu = xq - col.dx, yu = yq - col.dy;
            const bool in = active && xu >= 0 && xu < s.nx && yu >= 0 && yu < s.ny;
            const int aoff = in ? (int)(((uint64_t)xu * s.ny + yu) * s.row_dw * 4) : oob;
            const bool inbox = in && xu >= c.bx0 && xu < c.bx1 && yu >= c.by0 && yu < c.by1;
            const int uoff = inbox ? (int)((c.ad_off + ((uint64_t)(xu - c.bx0) * sy + (yu - c.by0)) * nwd) * 8) : oob;
            const bool any_box = __any(inbox);
            const int w_lo = max(0, zlo - dz0 - DZB + 1) >> 5, w_hi = min(nwd, (zhi - dz0 + 31) >> 5);
            // word w + 1's loads are issued while word w counts; the changed
            // row's words m carry over (word w's upper word is w + 1's lower)
            if (w_lo < w_hi) {
                uint32_t an = __builtin_amdgcn_raw_buffer_load_b32(seed_rs, aoff + 4 * w_lo, 0, 0);
                uint32_t uan = 0, udn = 0;  // lanes outside the box read (0, 0): n = o
                if (any_box) {
                    uan = __builtin_amdgcn_raw_buffer_load_b32(ad_rs, uoff + 8 * w_lo, 0, 0);
                    udn = __builtin_amdgcn_raw_buffer_load_b32(ad_rs, uoff + 8 * w_lo + 4, 0, 0);
                }
                uint2 m0 = adw(m, w_lo + q0, nwd), m1n = adw(m, w_lo + q0 + 1, nwd);
                for (int w = w_lo; w < w_hi; ++w) {
                    uint32_t a = an;
                    const uint32_t ua = uan, ud = udn;
                    const uint2 m1 = m1n;
                    if (w + 1 < w_hi) {
                        an = __builtin_amdgcn_raw_buffer_load_b32(seed_rs, aoff + 4 * (w + 1), 0, 0);
                        if (any_box) {
                            uan = __builtin_amdgcn_raw_buffer_load_b32(ad_rs, uoff + 8 * (w + 1), 0, 0);
                            udn = __builtin_amdgcn_raw_buffer_load_b32(ad_rs, uoff + 8 * (w + 1) + 4, 0, 0);
                        }
                        m1n = adw(m, w + q0 + 2, nwd);
                    }
                    if (any_box) a = (a & ~ud) | ua;
                    const uint64_t pa = ((uint64_t)m1.x << 32) | m0.x, pd = ((uint64_t)m1.y << 32) | m0.y;
#pragma unroll
                    for (int j = 0; j < DZB; ++j) {
                        cp[j] = popc_acc(a & (uint32_t)(pa >> j), cp[j]);
                        cm[j] = popc_acc(a & (uint32_t)(pd >> j), cm[j]);
                    }
                    m0 = m1;
                }
            }
        }
        // ---- X(-d): lanes v = q + (dx, dy); M(q) << dz meets v at z in [zlo + dz, zhi + dz)
        {
            const int xv = xq + col.dx, yv = yq + col.dy;
            const bool in = active && xv >= 0 && xv < s.nx && yv >= 0 && yv < s.ny;
            const int voff = in ? (int)(((uint64_t)xv * s.ny + yv) * s.row_dw * 4) : oob;
            const int w_lo = min(nwd, (zlo + dz0) >> 5), w_hi = min(nwd, (zhi + dz0 + DZB - 1 + 31) >> 5);
            if (w_lo < w_hi) {
                uint32_t an = __builtin_amdgcn_raw_buffer_load_b32(seed_rs, voff + 4 * w_lo, 0, 0);
                uint2 m0 = adw(m, w_lo - q0 - 1, nwd), m1n = adw(m, w_lo - q0, nwd);
                for (int w = w_lo; w < w_hi; ++w) {
                    const uint32_t a = an;
                    const uint2 m1 = m1n;
                    if (w + 1 < w_hi) {
                        an = __builtin_amdgcn_raw_buffer_load_b32(seed_rs, voff + 4 * (w + 1), 0, 0);
                        m1n = adw(m, w + 1 - q0, nwd);
                    }
                    const uint64_t pa = ((uint64_t)m1.x << 32) | m0.x, pd = ((uint64_t)m1.y << 32) | m0.y;
#pragma unroll
                    for (int j = 0; j < DZB; ++j) {
                        cp[j] = popc_acc(a & (uint32_t)((pa << j) >> 32), cp[j]);
                        cm[j] = popc_acc(a & (uint32_t)((pd << j) >> 32), cm[j]);
                    }
                    m0 = m1;
                }
            }
        }
    }
    int32_t *dt = delta + c.delta_off;
#pragma unroll
    for (int j = 0; j < DZB; ++j) {
        const int dz = dz0 + j;
        const int d = (int)(cp[j] - cm[j]);
        if (active && dz <= col.m && d) atomicAdd(&dt[s.smap[col.base + dz]], d);
    }
}

static constexpr int kDeltaWaves = 4;

__global__ __launch_bounds__(64 * kDeltaWaves) void delta_count_kernel(const SeedDev *__restrict__ seeds,
                                                                       const CandDev *__restrict__ cands,
                                                                       const BatchHeader *__restrict__ hdr,
                                                                       const uint2 *__restrict__ ad,
                                                                       uint32_t ad_bytes,
                                                                       const uint2 *__restrict__ info,
                                                                       int32_t *__restrict__ delta) {
    const uint64_t nitems = hdr->nitems;
    const uint32_t ncand = hdr->ncand;
    const int row_chunk = (int)hdr->row_chunk;
    const __amdgpu_buffer_rsrc_t ad_rs = __builtin_amdgcn_make_buffer_rsrc((void *)ad, 0, (int)ad_bytes, 0x00020000);
    const uint64_t nwaves = (uint64_t)gridDim.x * kDeltaWaves;
    const uint64_t wave0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kDeltaWaves + threadIdx.x / 64));
    for (uint64_t item = wave0; item < nitems; item += nwaves) {
        uint32_t lo = 0, hi = ncand;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (cands[mid].item_start <= item) lo = mid;
            else hi = mid;
        }
        const CandDev c = cands[lo];
        const SeedDev &s = seeds[c.seed];
        const uint64_t local = item - c.item_start;
        const GB gb = s.gbs[local / c.nchunks];
        const int chunk = (int)(local % c.nchunks);
        switch (gb.dzb) {
            case 4: delta_item<4>(s, c, gb, chunk, row_chunk, ad, ad_rs, info, delta); break;
            case 8: delta_item<8>(s, c, gb, chunk, row_chunk, ad, ad_rs, info, delta); break;
            case 12: delta_item<12>(s, c, gb, chunk, row_chunk, ad, ad_rs, info, delta); break;
            case 16: delta_item<16>(s, c, gb, chunk, row_chunk, ad, ad_rs, info, delta); break;
            case 20: delta_item<20>(s, c, gb, chunk, row_chunk, ad, ad_rs, info, delta); break;
            case 24: delta_item<24>(s, c, gb, chunk, row_chunk, ad, ad_rs, info, delta); break;
            case 28: delta_item<28>(s, c, gb, chunk, row_chunk, ad, ad_rs, info, delta); break;
            default: delta_item<32>(s, c, gb, chunk, row_chunk, ad, ad_rs, info, delta); break;
        }
    }
}

// ---------------------------------------------------------------- 4. gather + ordered fold

static constexpr int kFoldWaves = 4;

struct FoldTask {
    uint32_t cand, r;
};

// One wave per (candidate, radius) task: term j of the radius is
// (seed count + delta) / valid of the radius' j-th offset in shell order, and
// the terms are folded in that order -- bit for bit the reference's sequential
// left fold (fold_terms).  The host orders the tasks longest shell first, and
// a shell's tasks (one per candidate of its seed) next to each other, so waves
// running together share the shell's data in L2.
__global__ __launch_bounds__(64 * kFoldWaves) void batch_fold_kernel(const SeedDev *__restrict__ seeds,
                                                                     const CandDev *__restrict__ cands,
                                                                     const BatchHeader *__restrict__ hdr,
                                                                     const FoldTask *__restrict__ tasks,
                                                                     const int32_t *__restrict__ delta,
                                                                     const int *__restrict__ occ_delta,
                                                                     double *__restrict__ curves,
                                                                     uint8_t *__restrict__ support) {
    const uint32_t nfold = hdr->nfold, lane = threadIdx.x & 63;
    const uint32_t nwaves = gridDim.x * kFoldWaves;
    for (uint32_t t = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kFoldWaves + threadIdx.x / 64)); t < nfold;
         t += nwaves) {
        const FoldTask tk = tasks[t];
        const CandDev &c = cands[tk.cand];
        const SeedDev &s = seeds[c.seed];
        const uint64_t b = s.rstart[tk.r], n = s.rstart[tk.r + 1] - b;
        const int32_t *__restrict__ dt = delta + c.delta_off;
        const unsigned long long *__restrict__ cnt = s.shtable;
        const int32_t *__restrict__ sg = s.sg + b;
        const Offset *__restrict__ offs = s.offs + b;
        const uint64_t occ = (uint64_t)((int64_t)s.occupied + occ_delta[tk.cand]);
        const uint64_t nx = (uint64_t)s.nx, ny = (uint64_t)s.ny, nz = (uint64_t)s.nz;
        const double sum = fold_terms<8, 1>(n, lane, [&](uint64_t j) {
            const int32_t k = sg[j];
            const Offset o = offs[j];
            const int32_t kk = k >= 0 ? k : 0;  // the zero offset counts every occupied voxel
            const uint64_t hk = (uint64_t)((long long)cnt[kk] + dt[kk]);
            const uint64_t h = k >= 0 ? hk : occ;
            const uint64_t valid = (nx - (uint64_t)abs(o.dx)) * (ny - (uint64_t)abs(o.dy)) * (nz - (uint64_t)abs(o.dz));
            return (double)h / (double)valid;
        });
        if (lane == 0) {
            curves[c.curve_off + tk.r] = n > 0 ? sum / (double)n : 0.0;
            support[c.curve_off + tk.r] = n > 0 ? 1 : 0;
        }
    }
}

// ---------------------------------------------------------------- 5. committing a move

// coverage += box delta of candidate 0
__global__ void apply_cov_kernel(const CandDev *__restrict__ cands, const int32_t *__restrict__ box,
                                 uint32_t *__restrict__ cov, int ny, int nz) {
    const CandDev c = cands[0];
    const int sx = c.bx1 - c.bx0, sy = c.by1 - c.by0, sz = c.bz1 - c.bz0;
    const uint64_t n = (uint64_t)sx * sy * sz;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int lz = (int)(i % sz), ly = (int)((i / sz) % sy), lx = (int)(i / ((uint64_t)sz * sy));
        const int d = box[c.box_off + i];
        if (d) {
            const uint64_t v = ((uint64_t)(c.bx0 + lx) * ny + (c.by0 + ly)) * nz + (c.bz0 + lz);
            cov[v] = (uint32_t)((int)cov[v] + d);
        }
    }
}

// seed rows of the box: new = (old & ~D) | A
__global__ void apply_rows_kernel(const CandDev *__restrict__ cands, const uint2 *__restrict__ ad,
                                  uint32_t *__restrict__ vol, int ny, int nwd, int row_dw) {
    const CandDev c = cands[0];
    const int sy = c.by1 - c.by0;
    const uint64_t n = (uint64_t)(c.bx1 - c.bx0) * sy * nwd;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = i / nwd;
        const int w = (int)(i % nwd);
        const uint2 m = ad[c.ad_off + i];
        const int x = c.bx0 + (int)(r / sy), y = c.by0 + (int)(r % sy);
        uint32_t &v = vol[((uint64_t)x * ny + y) * row_dw + w];
        v = (v & ~m.y) | m.x;
    }
}

// seed counts += delta counts of candidate 0 (both in canonical shell order)
__global__ void apply_table_kernel(const CandDev *__restrict__ cands, const int32_t *__restrict__ delta,
                                   unsigned long long *__restrict__ table, uint64_t n) {
    const uint64_t off = cands[0].delta_off;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        table[i] = (unsigned long long)((long long)table[i] + delta[off + i]);
}

struct CovAdd {
    uint32_t *cov;
    uint64_t ny, nz;
    __device__ void operator()(const vox::GranuleInfo &, uint32_t x, uint32_t y, uint32_t z) const {
        atomicAdd(&cov[((uint64_t)x * ny + y) * nz + z], 1u);
    }
};

__global__ __launch_bounds__(256) void cov_kernel(const vox::TriGeo *__restrict__ tris,
                                                  const vox::GranuleInfo *__restrict__ gran,
                                                  const vox::Tile *__restrict__ tiles, double bx, double by,
                                                  double bz, double pitch, uint32_t *__restrict__ cov, uint64_t ny,
                                                  uint64_t nz, uint32_t *__restrict__ overflow, uint32_t cap) {
    vox::voxelize_tile(tris, gran, tiles[blockIdx.x], blockIdx.x, bx, by, bz, pitch, overflow, cap,
                       CovAdd{cov, ny, nz});
}

__global__ void cov_overflow_kernel(const vox::TriGeo *__restrict__ tris, const vox::GranuleInfo *__restrict__ gran,
                                    const vox::Tile *__restrict__ tiles, double bx, double by, double bz, double pitch,
                                    uint32_t *__restrict__ cov, uint64_t ny, uint64_t nz,
                                    const uint32_t *__restrict__ overflow, uint32_t count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    vox::voxel_overflow_ray(tris, gran, tiles, bx, by, bz, pitch, overflow, k, CovAdd{cov, ny, nz});
}

__global__ void zero_u32_kernel(uint32_t *__restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Fixed launch shapes of the batch sequence (grid-stride kernels).
constexpr unsigned kClearBlocks = 2048, kPrepBlocks = 1024, kVoxelBlocks = 8192, kOverflowBlocks = 64,
                   kAdBlocks = 4096, kDeltaBlocks = 2048, kFoldBlocks = 2048;

}  // namespace batch

// Persistent batch workspace of a context: grow-only device buffers, pinned
// staging for the per-batch inputs and outputs, and the captured graph.
struct BatchWS {
    DevBuf seeds, cands, hdr, geo, gran, tiles, origin, adpre, tasks, delta, box, info, occ, ad, curves, over,
        newxyz, fcands;
    void *pin = nullptr;
    uint64_t pin_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> captured;  // buffer addresses the graph was captured with
    std::map<std::vector<uint64_t>, uint32_t> row_chunk;  // tuned rows per delta item, by seed shapes
};

void batch_ws_release(s2h_ctx *ctx) {
    BatchWS *w = ctx->batch_ws;
    if (!w) return;
    if (w->exec) (void)hipGraphExecDestroy(w->exec);
    if (w->graph) (void)hipGraphDestroy(w->graph);
    DevBuf *bufs[] = {&w->seeds, &w->cands, &w->hdr, &w->geo, &w->gran, &w->tiles, &w->origin, &w->adpre, &w->tasks,
                      &w->delta, &w->box, &w->info, &w->occ, &w->ad, &w->curves, &w->over, &w->newxyz, &w->fcands};
    for (DevBuf *b : bufs) dev_free(ctx, *b);
    if (w->pin) (void)hipHostFree(w->pin);
    delete w;
    ctx->batch_ws = nullptr;
}

}  // namespace s2h

// ================================================================ host side

struct s2h_seed {
    s2h_vol *vol = nullptr;
    s2h::DevBuf cov, shtable, sg, smap, gbs, d_xyz, d_tri;
    s2h::ExactPlan *plan = nullptr;
    uint64_t ncanon = 0;               // canonical offsets of the shells (delta table entries, + 1 unused)
    double bmin[3];
    double pitch;
    uint32_t r_max;
    uint64_t occupied;
    std::vector<s2h::batch::GB> gb_host;
    // particles: vertices [vstart[p], vstart[p+1]), triangles [tstart[p], tstart[p+1])
    uint32_t nparticles = 0;
    std::vector<uint32_t> vstart, tstart;
    std::vector<double> xyz;           // current vertex positions (host copy of d_xyz)
    std::vector<int32_t> vgran;        // local granule of each vertex (-1: not referenced)
    std::vector<uint32_t> pgran;       // per particle: index of its first granule start in gtri
    std::vector<uint32_t> gtri;        // per particle: its granules' first triangles in d_tri order, then the end
    std::vector<uint32_t> ngran;       // per particle: local granule count
};

namespace s2h {
namespace batch {

__global__ void shell_table_kernel(const unsigned long long *__restrict__ table, const int32_t *__restrict__ tab_of,
                                   uint64_t n, unsigned long long *__restrict__ out) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        out[k] = table[tab_of[k]];
}

// Canonical shell order of a seed's plan: each canonical offset numbered at its
// first appearance in the reference's shell order (a radius' canonical offsets
// are then contiguous, as d and -d share a radius).  Builds the shell-order ->
// canonical map (sg), the count-table -> canonical map (smap; ncanon for table
// entries no shell uses) and the seed's counts in canonical order.
static s2h_status seed_shell_tables(s2h_ctx *ctx, s2h_seed *sd) {
    const ExactPlan *p = sd->plan;
    const std::vector<int32_t> &gather = p->shells->host->gather;
    const uint64_t K = gather.size(), T = p->table_size;
    std::vector<int32_t> smap(std::max<uint64_t>(T, 1), -1), sg(std::max<uint64_t>(K, 1), -1), tab_of;
    for (uint64_t i = 0; i < K; ++i) {
        const int32_t g = gather[i];
        if (g < 0) continue;
        if (smap[g] < 0) {
            smap[g] = (int32_t)tab_of.size();
            tab_of.push_back(g);
        }
        sg[i] = smap[g];
    }
    const uint64_t nc = tab_of.size();
    if (nc >= (1ull << 31)) return fail(S2H_ERR_UNSUPPORTED, "too many offsets for a seed");
    for (int32_t &m : smap)
        if (m < 0) m = (int32_t)nc;
    sd->ncanon = nc;
    S2H_TRY(upload(ctx, sd->sg, sg.data(), sg.size() * 4));
    S2H_TRY(upload(ctx, sd->smap, smap.data(), smap.size() * 4));
    S2H_TRY(dev_alloc(ctx, sd->shtable, std::max<uint64_t>(nc, 1) * 8));
    DevBuf tab;
    s2h_status st = upload(ctx, tab, tab_of.data(), std::max<uint64_t>(nc, 1) * 4);
    if (st == S2H_OK && nc) {
        shell_table_kernel<<<1024, 256, 0, ctx->stream>>>((const unsigned long long *)p->d_table.ptr,
                                                          (const int32_t *)tab.ptr, nc,
                                                          (unsigned long long *)sd->shtable.ptr);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) st = fail(S2H_ERR_HIP, std::string("shell_table_kernel: ") + hipGetErrorString(e));
    }
    if (st == S2H_OK) {
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = fail(S2H_ERR_HIP, std::string("seed tables: ") + hipGetErrorString(e));
    }
    dev_free(ctx, tab);
    return st;
}

// Whole-mesh coverage counts of a seed (one count per granule covering a centre).
static s2h_status seed_coverage(s2h_ctx *ctx, s2h_seed *sd, const double *xyz, const uint32_t *tri, uint64_t ntris,
                                const uint32_t *granule_of_tri, uint32_t ngranules, const uint64_t dims[3]) {
    const uint64_t cells = dims[0] * dims[1] * dims[2];
    S2H_TRY(dev_alloc(ctx, sd->cov, cells * 4));
    zero_u32_kernel<<<4096, 256, 0, ctx->stream>>>((uint32_t *)sd->cov.ptr, cells);
    S2H_HIP(hipGetLastError());
    std::vector<uint32_t> start(ngranules + 1, 0);
    for (uint64_t f = 0; f < ntris; ++f) start[granule_of_tri[f] + 1]++;
    for (uint32_t g = 0; g < ngranules; ++g) start[g + 1] += start[g];
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    std::vector<double> txyz(9 * ntris);
    for (uint64_t f = 0; f < ntris; ++f) {
        const uint64_t slot = fill[granule_of_tri[f]]++;
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) txyz[9 * slot + 3 * c + a] = xyz[3 * (uint64_t)tri[3 * f + c] + a];
    }
    std::vector<vox::TriGeo> geo(ntris);
    std::vector<vox::GranuleInfo> gran(ngranules);
    std::vector<vox::Tile> tiles;
    for (uint32_t g = 0; g < ngranules; ++g) {
        vox::GranuleInfo &gi = gran[g];
        gi.tri_start = start[g];
        gi.tri_end = start[g + 1];
        gi.owner = 0;
        gi.sign = 1;
        int64_t lo[3], hi[3];
        if (vox::prepare_granule(txyz.data() + 9 * (uint64_t)start[g], start[g + 1] - start[g], sd->bmin, sd->pitch,
                                 dims, geo.data() + start[g], gi, lo, hi))
            vox::push_tiles(g, lo, hi, tiles);
    }
    if (tiles.empty()) return S2H_OK;
    DevBuf dg, dgr, dt, dov;
    auto release = [&]() {
        dev_free(ctx, dg);
        dev_free(ctx, dgr);
        dev_free(ctx, dt);
        dev_free(ctx, dov);
    };
    s2h_status st = upload(ctx, dg, geo.data(), geo.size() * sizeof(vox::TriGeo));
    if (st == S2H_OK) st = upload(ctx, dgr, gran.data(), gran.size() * sizeof(vox::GranuleInfo));
    if (st == S2H_OK) st = upload(ctx, dt, tiles.data(), tiles.size() * sizeof(vox::Tile));
    if (st == S2H_OK) st = dev_alloc(ctx, dov, (1 + 2 * (uint64_t)kOverflowCap) * 4);
    if (st != S2H_OK) {
        release();
        return st;
    }
    uint32_t nover = 0;
    hipError_t e = hipSuccess;
    zero_u32_kernel<<<1, 64, 0, ctx->stream>>>((uint32_t *)dov.ptr, 1);
    for (size_t t0 = 0; e == hipSuccess && t0 < tiles.size(); t0 += (1u << 30)) {
        const size_t nt = std::min<size_t>(tiles.size() - t0, 1u << 30);
        cov_kernel<<<(unsigned)nt, 256, 0, ctx->stream>>>(
            (const vox::TriGeo *)dg.ptr, (const vox::GranuleInfo *)dgr.ptr, (const vox::Tile *)dt.ptr + t0,
            sd->bmin[0], sd->bmin[1], sd->bmin[2], sd->pitch, (uint32_t *)sd->cov.ptr, dims[1], dims[2],
            (uint32_t *)dov.ptr, kOverflowCap);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&nover, dov.ptr, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && nover > kOverflowCap) st = fail(S2H_ERR_UNSUPPORTED, "too many overflow rays");
    if (e == hipSuccess && st == S2H_OK && nover) {
        cov_overflow_kernel<<<(nover + 63) / 64, 64, 0, ctx->stream>>>(
            (const vox::TriGeo *)dg.ptr, (const vox::GranuleInfo *)dgr.ptr, (const vox::Tile *)dt.ptr, sd->bmin[0],
            sd->bmin[1], sd->bmin[2], sd->pitch, (uint32_t *)sd->cov.ptr, dims[1], dims[2], (const uint32_t *)dov.ptr,
            nover);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) st = fail(S2H_ERR_HIP, std::string("coverage kernel: ") + hipGetErrorString(e));
    release();
    return st;
}

static s2h_status ws_reserve(s2h_ctx *ctx, DevBuf &b, uint64_t bytes) {
    if (b.ptr && b.bytes >= bytes) return S2H_OK;
    const uint64_t want = std::max<uint64_t>(std::max<uint64_t>(bytes, 256), b.bytes + b.bytes / 2);
    dev_free(ctx, b);
    return dev_alloc(ctx, b, want);
}

// The host side of one batch: descriptors of every candidate, its granules at
// both positions and their voxel tiles; device-side preparation does the rest.
struct HostBatch {
    std::vector<SeedDev> sdev;
    std::vector<CandDev> cd;
    std::vector<vox::GranuleInfo> gran;
    std::vector<vox::Tile> tiles;
    std::vector<double> origin;
    std::vector<uint64_t> ad_prefix;
    std::vector<FoldTask> tasks;
    std::vector<uint64_t> xyz_off;  // per candidate: first new vertex (3 doubles each)
    // delta counts by region correlation: per candidate whether it takes them,
    // the candidates' descriptors grouped by seed, and the groups
    std::vector<uint8_t> cfftd;
    std::vector<FftdCand> fcands;
    struct FftdGroup {
        uint32_t seed, start, count;
        int F[3];
    };
    std::vector<FftdGroup> groups;
    BatchHeader hdr{};
    uint64_t geo_total = 0, newxyz_total = 0, curve_total = 0;
};

// The item split of the direct delta counts for a row chunk (candidates whose
// deltas come from region correlation have none).
static void set_row_chunk(HostBatch &hb, uint32_t row_chunk) {
    uint64_t items = 0;
    for (size_t c = 0; c < hb.cd.size(); ++c) {
        CandDev &d = hb.cd[c];
        const uint64_t nrows = (uint64_t)(d.bx1 - d.bx0) * (uint64_t)(d.by1 - d.by0);
        d.nchunks = c < hb.cfftd.size() && hb.cfftd[c] ? 0u : (uint32_t)((nrows + row_chunk - 1) / row_chunk);
        d.item_start = items;
        items += (uint64_t)d.nchunks * hb.sdev[d.seed].ngb;
    }
    hb.hdr.nitems = items;
    hb.hdr.row_chunk = row_chunk;
}

// Modelled device time of direct delta counting per (changed row x canonical
// offset x span word), from the 256^3 r64 and 192^3 r48 seeds' timings.
static constexpr double kDirectMsPerUnit = 2.0e-10;

// Per seed, the delta counts of its candidates: direct counting, or region
// correlation where its frames are supported, provably exact (error bound
// below 1/4) and modelled faster.  Each candidate's own frame is the least
// supported length of box + 2R per axis; candidates are sorted by frame
// volume and grouped, each group on the largest frame of its members, a new
// group starting where that would exceed 1.2 x the volume of the group's
// first frame.  S2HIP_BATCH_DELTA=direct or =fft forces one (fft still only
// where supported and exact).
static void choose_delta_modes(s2h_seed *const *seeds, uint32_t nseeds, HostBatch &hb) {
    const size_t ncand = hb.cd.size();
    hb.cfftd.assign(ncand, 0);
    hb.fcands.clear();
    hb.groups.clear();
    const char *env = std::getenv("S2HIP_BATCH_DELTA");
    const int force = !env ? 0 : std::strcmp(env, "fft") == 0 ? 2 : std::strcmp(env, "direct") == 0 ? 1 : 0;
    if (force == 1) return;
    struct Own {
        uint32_t c;
        int F[3];
        double vol;
        uint64_t changed;
    };
    std::vector<Own> mine;
    for (uint32_t si = 0; si < nseeds; ++si) {
        const s2h_seed *sd = seeds[si];
        const int *ext = sd->plan->shells->host->extent;
        const int R[3] = {ext[0], ext[1], ext[2]};
        mine.clear();
        double direct_ms = 0.0;
        bool ok = true;
        for (uint32_t c = 0; c < ncand && ok; ++c) {
            const CandDev &d = hb.cd[c];
            if (d.seed != si) continue;
            const int b[3] = {d.bx1 - d.bx0, d.by1 - d.by0, d.bz1 - d.bz0};
            direct_ms += kDirectMsPerUnit * (double)b[0] * b[1] * (double)sd->ncanon * (b[2] / 32.0 + 2.0);
            Own o;
            o.c = c;
            o.changed = std::max<uint64_t>(1, (uint64_t)b[0] * b[1] * b[2]);
            ok = fft_delta_frame(b, R, o.F);
            o.vol = (double)o.F[0] * o.F[1] * o.F[2];
            mine.push_back(o);
        }
        if (!ok || mine.empty()) continue;
        std::sort(mine.begin(), mine.end(), [](const Own &a, const Own &b) { return a.vol < b.vol; });
        // groups over the sorted candidates
        std::vector<HostBatch::FftdGroup> gs;
        std::vector<uint64_t> gchanged;
        double fft_ms = 0.0;
        for (size_t k = 0; k < mine.size();) {
            HostBatch::FftdGroup g;
            g.seed = si;
            g.start = (uint32_t)k;  // position in `mine` for now
            for (int a = 0; a < 3; ++a) g.F[a] = mine[k].F[a];
            const double first = mine[k].vol;
            uint64_t changed = mine[k].changed;
            size_t e = k + 1;
            for (; e < mine.size(); ++e) {
                const int F2[3] = {std::max(g.F[0], mine[e].F[0]), std::max(g.F[1], mine[e].F[1]),
                                   std::max(g.F[2], mine[e].F[2])};
                if ((double)F2[0] * F2[1] * F2[2] > 1.2 * first) break;
                for (int a = 0; a < 3; ++a) g.F[a] = F2[a];
                changed = std::max(changed, mine[e].changed);
            }
            g.count = (uint32_t)(e - k);
            fft_ms += fft_delta_model_ms(g.F, R, g.count);
            gs.push_back(g);
            gchanged.push_back(changed);
            k = e;
        }
        bool exact = true;
        for (size_t i = 0; i < gs.size(); ++i)
            exact = exact && fft_delta_error_bound(gs[i].F, sd->occupied, gchanged[i]) <= 0.25;
        if (!exact) continue;
        if (force != 2 && !(fft_ms < direct_ms)) continue;
        for (HostBatch::FftdGroup &g : gs) {
            const uint32_t first = g.start;
            g.start = (uint32_t)hb.fcands.size();
            for (uint32_t k = first; k < first + g.count; ++k) {
                const CandDev &d = hb.cd[mine[k].c];
                hb.cfftd[mine[k].c] = 1;
                hb.fcands.push_back(FftdCand{{d.bx0, d.by0, d.bz0}, {d.bx1, d.by1, d.bz1}, d.ad_off, d.delta_off});
            }
            hb.groups.push_back(g);
        }
    }
}

static s2h_status build_host(s2h_ctx *ctx, BatchWS &w, s2h_seed *const *seeds, uint32_t nseeds,
                             const s2h_candidate *cands, uint32_t ncand, uint32_t row_chunk, HostBatch &hb) {
    hb.sdev.resize(nseeds);
    for (uint32_t i = 0; i < nseeds; ++i) {
        const s2h_seed *sd = seeds[i];
        if (!sd) return fail(S2H_ERR_ARG, "null seed");
        const ExactPlan *p = sd->plan;
        SeedDev &d = hb.sdev[i];
        d.vol = (const uint32_t *)sd->vol->d_words;
        d.cov = (const uint32_t *)sd->cov.ptr;
        d.shtable = (const unsigned long long *)sd->shtable.ptr;
        d.sg = (const int32_t *)sd->sg.ptr;
        d.smap = (const int32_t *)sd->smap.ptr;
        d.offs = (const Offset *)p->d_offs.ptr;
        d.rstart = (const uint64_t *)p->d_rstart.ptr;
        d.cols = (const ColInfo *)p->d_cols.ptr;
        d.gbs = (const GB *)sd->gbs.ptr;
        d.occupied = sd->occupied;
        d.vol_bytes = (uint32_t)(sd->vol->nwords * 8);
        d.ngb = (uint32_t)sd->gb_host.size();
        d.r_max = sd->r_max;
        d.pad = 0;
        d.nx = (int)sd->vol->dims[0];
        d.ny = (int)sd->vol->dims[1];
        d.nz = (int)sd->vol->dims[2];
        d.row_dw = (int)(sd->vol->row_words * 2);
    }
    // new-vertex layout first: granule descriptors point into the device copy
    hb.xyz_off.resize(ncand);
    for (uint32_t c = 0; c < ncand; ++c) {
        const s2h_candidate &k = cands[c];
        if (k.seed >= nseeds) return fail(S2H_ERR_ARG, "candidate seed index out of range");
        const s2h_seed *sd = seeds[k.seed];
        if (k.particle >= sd->nparticles) return fail(S2H_ERR_ARG, "candidate particle index out of range");
        if (!k.new_xyz) return fail(S2H_ERR_ARG, "candidate without vertices");
        hb.xyz_off[c] = hb.newxyz_total;
        hb.newxyz_total += (uint64_t)(sd->vstart[k.particle + 1] - sd->vstart[k.particle]);
    }
    S2H_TRY(ws_reserve(ctx, w.newxyz, std::max<uint64_t>(hb.newxyz_total, 1) * 24));
    hb.cd.resize(ncand);
    hb.origin.resize(4 * (size_t)ncand);
    hb.ad_prefix.assign(ncand + 1, 0);
    hb.tasks.clear();
    uint64_t ad_total = 0, info_total = 0, box_total = 0, delta_total = 0, items = 0;
    std::vector<double> mn, mx, amax;
    std::vector<uint8_t> have;
    for (uint32_t c = 0; c < ncand; ++c) {
        const s2h_candidate &k = cands[c];
        const s2h_seed *sd = seeds[k.seed];
        const uint64_t *dims = sd->vol->dims;
        const uint32_t p = k.particle, v0 = sd->vstart[p], nv = sd->vstart[p + 1] - v0, ng = sd->ngran[p];
        int64_t blo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, bhi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
        for (int side = 0; side < 2; ++side) {
            const double *hx = side == 0 ? sd->xyz.data() + 3 * (uint64_t)v0 : k.new_xyz;
            const double *dx = side == 0 ? (const double *)sd->d_xyz.ptr + 3 * (uint64_t)v0
                                         : (const double *)w.newxyz.ptr + 3 * hb.xyz_off[c];
            // vertex bbox of every local granule (over the vertices its triangles use)
            mn.assign(3 * (size_t)ng, 0.0);
            mx.assign(3 * (size_t)ng, 0.0);
            amax.assign(ng, 0.0);
            have.assign(ng, 0);
            for (uint32_t v = 0; v < nv; ++v) {
                const int32_t g = sd->vgran[v0 + v];
                if (g < 0) continue;
                const double *q = hx + 3 * (uint64_t)v;
                for (int a = 0; a < 3; ++a) {
                    mn[3 * g + a] = have[g] ? vox::rmin(mn[3 * g + a], q[a]) : q[a];
                    mx[3 * g + a] = have[g] ? vox::rmax(mx[3 * g + a], q[a]) : q[a];
                    amax[g] = std::max(amax[g], std::fabs(q[a]));
                }
                have[g] = 1;
            }
            for (uint32_t g = 0; g < ng; ++g) {
                const uint32_t t0 = sd->gtri[sd->pgran[p] + g], t1 = sd->gtri[sd->pgran[p] + g + 1];
                vox::GranuleInfo gi;
                for (int a = 0; a < 3; ++a) {
                    gi.mn[a] = mn[3 * g + a];
                    gi.mx[a] = mx[3 * g + a];
                }
                gi.tri_start = (uint32_t)hb.geo_total;
                gi.tri_end = (uint32_t)(hb.geo_total + (t1 - t0));
                gi.owner = c;
                gi.sign = side == 0 ? -1 : 1;
                gi.amax = amax[g];
                gi.cmax = gi.smax = 0.0;
                gi.xyz = dx;
                gi.tri = (const uint32_t *)sd->d_tri.ptr + 3 * (uint64_t)t0;
                int64_t lo[3], hi[3];
                if (!have[g] || !vox::granule_bounds(gi.mn, gi.mx, sd->bmin, sd->pitch, dims, lo, hi, &gi.cmax,
                                                     &gi.smax))
                    continue;
                hb.geo_total += t1 - t0;
                vox::push_tiles((uint32_t)hb.gran.size(), lo, hi, hb.tiles);
                hb.gran.push_back(gi);
                for (int a = 0; a < 3; ++a) {
                    blo[a] = std::min(blo[a], lo[a]);
                    bhi[a] = std::max(bhi[a], hi[a]);
                }
            }
        }
        CandDev &d = hb.cd[c];
        d.seed = k.seed;
        if (blo[0] > bhi[0])  // particle outside the grid at both positions: nothing changes
            for (int a = 0; a < 3; ++a) blo[a] = bhi[a] = 0;
        d.bx0 = (int32_t)blo[0];
        d.by0 = (int32_t)blo[1];
        d.bz0 = (int32_t)blo[2];
        d.bx1 = (int32_t)bhi[0];
        d.by1 = (int32_t)bhi[1];
        d.bz1 = (int32_t)bhi[2];
        const uint64_t nrows = (uint64_t)(bhi[0] - blo[0]) * (uint64_t)(bhi[1] - blo[1]);
        const uint64_t nwd = (dims[2] + 31) / 32;
        d.nchunks = (uint32_t)((nrows + row_chunk - 1) / row_chunk);
        d.ad_off = ad_total;
        ad_total += nrows * nwd;
        hb.ad_prefix[c + 1] = ad_total;
        d.info_off = info_total;
        info_total += nrows;
        d.box_off = box_total;
        box_total += (uint64_t)(bhi[0] - blo[0]) * (bhi[1] - blo[1]) * (bhi[2] - blo[2]);
        d.delta_off = delta_total;
        delta_total += sd->ncanon + 1;
        d.item_start = items;
        items += (uint64_t)d.nchunks * hb.sdev[k.seed].ngb;
        d.curve_off = hb.curve_total;
        hb.curve_total += sd->r_max + 1;
        for (uint32_t r = 1; r <= sd->r_max; ++r) hb.tasks.push_back(FoldTask{c, r});
        hb.origin[4 * c] = sd->bmin[0];
        hb.origin[4 * c + 1] = sd->bmin[1];
        hb.origin[4 * c + 2] = sd->bmin[2];
        hb.origin[4 * c + 3] = sd->pitch;
    }
    choose_delta_modes(seeds, nseeds, hb);
    set_row_chunk(hb, row_chunk);
    items = hb.hdr.nitems;
    if (hb.tiles.size() >= (1ull << 31) || hb.gran.size() >= (1ull << 31) || hb.geo_total >= (1ull << 32))
        return fail(S2H_ERR_UNSUPPORTED, "batch too large");
    if (ad_total * 8 + 64 >= 0x7ffffe00ull) return fail(S2H_ERR_UNSUPPORTED, "changed rows of the batch above 2 GiB");
    hb.hdr.nitems = items;
    hb.hdr.ad_total = ad_total;
    hb.hdr.info_total = info_total;
    hb.hdr.box_words = box_total;
    hb.hdr.delta_words = delta_total;
    hb.hdr.ncand = ncand;
    // fold tasks: longest shells first (the last waves to start are short), then
    // by seed, radius and candidate, so a shell's tasks run side by side
    auto shell_len = [&](const FoldTask &t) {
        const std::vector<uint64_t> &rs = seeds[hb.cd[t.cand].seed]->plan->radius_start;
        return rs[t.r + 1] - rs[t.r];
    };
    std::sort(hb.tasks.begin(), hb.tasks.end(), [&](const FoldTask &a, const FoldTask &b) {
        const uint64_t la = shell_len(a), lb = shell_len(b);
        if (la != lb) return la > lb;
        const uint32_t sa = hb.cd[a.cand].seed, sb = hb.cd[b.cand].seed;
        if (sa != sb) return sa < sb;
        if (a.r != b.r) return a.r < b.r;
        return a.cand < b.cand;
    });
    if (hb.tasks.size() >= (1ull << 31)) return fail(S2H_ERR_UNSUPPORTED, "batch too large");
    hb.hdr.nfold = (uint32_t)hb.tasks.size();
    hb.hdr.ngran = (uint32_t)hb.gran.size();
    hb.hdr.ntiles = (uint32_t)hb.tiles.size();
    hb.hdr.row_chunk = row_chunk;
    hb.hdr.pad = 0;
    return S2H_OK;
}

// Curve buffer: values, then one support byte per value, in one allocation.
static uint8_t *curve_support(const BatchWS &w) {
    return (uint8_t *)w.curves.ptr + w.curves.bytes / 9 * 8;
}

// Byte size of the (A, D) buffer for its buffer resource (below the
// out-of-range offset the kernels use for lanes without a row).
static uint32_t ad_records(const BatchWS &w) { return (uint32_t)std::min<uint64_t>(w.ad.bytes, 0x7ffffe00u); }

// The launch sequence: clear through the direct delta counts (front), then
// the fold.
static void launch_front(hipStream_t s, BatchWS &w) {
    const BatchHeader *hdr = (const BatchHeader *)w.hdr.ptr;
    const SeedDev *seeds = (const SeedDev *)w.seeds.ptr;
    const CandDev *cands = (const CandDev *)w.cands.ptr;
    int32_t *delta = (int32_t *)w.delta.ptr;
    int32_t *box = (int32_t *)w.box.ptr;
    uint2 *info = (uint2 *)w.info.ptr;
    int *occ = (int *)w.occ.ptr;
    uint32_t *over = (uint32_t *)w.over.ptr;
    uint2 *ad = (uint2 *)w.ad.ptr;
    clear_kernel<<<kClearBlocks, 256, 0, s>>>(hdr, delta, box, info, occ, over);
    prep_kernel<<<kPrepBlocks, 256, 0, s>>>((const vox::GranuleInfo *)w.gran.ptr, hdr, (vox::TriGeo *)w.geo.ptr);
    batch_voxel_kernel<<<kVoxelBlocks, 256, 0, s>>>((const vox::TriGeo *)w.geo.ptr, (const vox::GranuleInfo *)w.gran.ptr,
                                                    (const vox::Tile *)w.tiles.ptr, hdr, (const double *)w.origin.ptr,
                                                    cands, box, over);
    batch_voxel_overflow_kernel<<<kOverflowBlocks, 64, 0, s>>>(
        (const vox::TriGeo *)w.geo.ptr, (const vox::GranuleInfo *)w.gran.ptr, (const vox::Tile *)w.tiles.ptr,
        (const double *)w.origin.ptr, cands, box, over);
    build_ad_kernel<<<kAdBlocks, 256, 0, s>>>(seeds, cands, hdr, (const uint64_t *)w.adpre.ptr, box, ad, info, occ);
    delta_count_kernel<<<kDeltaBlocks, 64 * kDeltaWaves, 0, s>>>(seeds, cands, hdr, ad, ad_records(w), info,
                                                                 delta);
}

static void launch_fold(hipStream_t s, BatchWS &w) {
    batch_fold_kernel<<<kFoldBlocks, 64 * kFoldWaves, 0, s>>>(
        (const SeedDev *)w.seeds.ptr, (const CandDev *)w.cands.ptr, (const BatchHeader *)w.hdr.ptr,
        (const FoldTask *)w.tasks.ptr, (int32_t *)w.delta.ptr, (int *)w.occ.ptr, (double *)w.curves.ptr,
        curve_support(w));
}

static void launch_sequence(hipStream_t s, BatchWS &w) {
    launch_front(s, w);
    launch_fold(s, w);
}

// Region-correlation delta counts of each group's candidates (after the
// change rows exist, before the fold).
static s2h_status launch_fftd(s2h_ctx *ctx, BatchWS &w, const HostBatch &hb, s2h_seed *const *seeds) {
    for (const HostBatch::FftdGroup &g : hb.groups) {
        const s2h_seed *sd = seeds[g.seed];
        FftdSeed fs;
        fs.words = sd->vol->d_words;
        fs.row_words = sd->vol->row_words;
        for (int a = 0; a < 3; ++a) {
            fs.dims[a] = (int)sd->vol->dims[a];
            fs.R[a] = sd->plan->shells->host->extent[a];
        }
        fs.sg = (const int32_t *)sd->sg.ptr;
        fs.offs = (const Offset *)sd->plan->d_offs.ptr;
        fs.noffs = sd->plan->offs.size();
        fs.occupied = sd->occupied;
        S2H_TRY(fft_delta_run(ctx, fs, (const FftdCand *)w.fcands.ptr + g.start, g.count, g.F, (const uint2 *)w.ad.ptr,
                              (int)((sd->vol->dims[2] + 31) / 32), (int32_t *)w.delta.ptr));
    }
    return S2H_OK;
}

static std::vector<void *> ws_ptrs(const BatchWS &w) {
    return {w.seeds.ptr, w.cands.ptr, w.hdr.ptr, w.geo.ptr, w.gran.ptr, w.tiles.ptr, w.origin.ptr, w.adpre.ptr,
            w.tasks.ptr, w.delta.ptr, w.box.ptr, w.info.ptr, w.occ.ptr, w.ad.ptr, w.curves.ptr, w.over.ptr,
            (void *)(uintptr_t)w.ad.bytes, (void *)(uintptr_t)w.curves.bytes};
}

static bool use_graph() {
    const char *env = std::getenv("S2HIP_BATCH_GRAPH");
    return !(env && env[0] == '0');
}

// Launch the batch sequence: replay the captured graph, capturing it first when
// the workspace buffers moved since the last capture.  A batch with region-
// correlation groups launches directly (its frame passes depend on the batch).
static s2h_status run_sequence(s2h_ctx *ctx, BatchWS &w, const HostBatch &hb, s2h_seed *const *seeds) {
    if (!hb.groups.empty()) {
        launch_front(ctx->stream, w);
        S2H_HIP(hipGetLastError());
        S2H_TRY(launch_fftd(ctx, w, hb, seeds));
        launch_fold(ctx->stream, w);
        S2H_HIP(hipGetLastError());
        ctx->launches += 7;
        return S2H_OK;
    }
    if (!use_graph()) {
        launch_sequence(ctx->stream, w);
        S2H_HIP(hipGetLastError());
        ctx->launches += 7;
        return S2H_OK;
    }
    const std::vector<void *> now = ws_ptrs(w);
    if (!w.exec || now != w.captured) {
        if (w.exec) (void)hipGraphExecDestroy(w.exec);
        if (w.graph) (void)hipGraphDestroy(w.graph);
        w.exec = nullptr;
        w.graph = nullptr;
        S2H_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        launch_sequence(ctx->stream, w);
        hipGraph_t g = nullptr;
        const hipError_t le = hipGetLastError();
        const hipError_t ee = hipStreamEndCapture(ctx->stream, &g);
        if (le != hipSuccess || ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(S2H_ERR_HIP, std::string("batch graph capture failed: ") +
                                         hipGetErrorString(le != hipSuccess ? le : ee));
        }
        hipError_t ie = hipGraphInstantiate(&w.exec, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(g);
            w.exec = nullptr;
            return fail(S2H_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        }
        w.graph = g;
        w.captured = now;
    }
    S2H_HIP(hipGraphLaunch(w.exec, ctx->stream));
    ctx->launches += 1;
    return S2H_OK;
}

// Pinned staging layout: inputs from the start, outputs in the last part.
static uint64_t out_region(const HostBatch &hb, uint32_t ncand) {
    return ((hb.curve_total * 9 + 15) / 16 * 16) + ((uint64_t)ncand * 4 + 15) / 16 * 16 + 16;
}

// Reserve the workspace for a host batch and copy its inputs in (pinned
// staging, one async copy per buffer, on the context's stream).
static s2h_status upload_batch(s2h_ctx *ctx, BatchWS &w, const HostBatch &hb, const s2h_candidate *cands,
                               uint32_t ncand, s2h_seed *const *seeds) {
    const uint32_t nc = std::max<uint32_t>(ncand, 1);
    S2H_TRY(ws_reserve(ctx, w.seeds, std::max<size_t>(hb.sdev.size(), 1) * sizeof(SeedDev)));
    S2H_TRY(ws_reserve(ctx, w.cands, nc * sizeof(CandDev)));
    S2H_TRY(ws_reserve(ctx, w.hdr, sizeof(BatchHeader)));
    S2H_TRY(ws_reserve(ctx, w.geo, std::max<uint64_t>(hb.geo_total, 1) * sizeof(vox::TriGeo)));
    S2H_TRY(ws_reserve(ctx, w.gran, std::max<size_t>(hb.gran.size(), 1) * sizeof(vox::GranuleInfo)));
    S2H_TRY(ws_reserve(ctx, w.tiles, std::max<size_t>(hb.tiles.size(), 1) * sizeof(vox::Tile)));
    S2H_TRY(ws_reserve(ctx, w.origin, (size_t)nc * 32));
    S2H_TRY(ws_reserve(ctx, w.adpre, ((size_t)ncand + 1) * 8));
    S2H_TRY(ws_reserve(ctx, w.tasks, std::max<size_t>(hb.tasks.size(), 1) * sizeof(FoldTask)));
    S2H_TRY(ws_reserve(ctx, w.delta, std::max<uint64_t>(hb.hdr.delta_words, 1) * 4));
    S2H_TRY(ws_reserve(ctx, w.box, std::max<uint64_t>(hb.hdr.box_words, 1) * 4));
    S2H_TRY(ws_reserve(ctx, w.info, std::max<uint64_t>(hb.hdr.info_total, 1) * 8));
    S2H_TRY(ws_reserve(ctx, w.occ, (size_t)nc * 4));
    S2H_TRY(ws_reserve(ctx, w.ad, std::max<uint64_t>(hb.hdr.ad_total, 1) * 8 + 64));
    if (!w.curves.ptr || w.curves.bytes / 9 < hb.curve_total) {
        const uint64_t entries = std::max<uint64_t>(hb.curve_total, 64) * 3 / 2;
        dev_free(ctx, w.curves);
        S2H_TRY(dev_alloc(ctx, w.curves, entries * 9));
    }
    S2H_TRY(ws_reserve(ctx, w.over, (1 + 2 * (uint64_t)kOverflowCap) * 4));
    S2H_TRY(ws_reserve(ctx, w.fcands, std::max<size_t>(hb.fcands.size(), 1) * sizeof(FftdCand)));
    const uint64_t in_bytes = (hb.sdev.size() * sizeof(SeedDev) + 16) + ((uint64_t)ncand * sizeof(CandDev) + 16) +
                              (sizeof(BatchHeader) + 16) + (hb.gran.size() * sizeof(vox::GranuleInfo) + 16) +
                              (hb.tiles.size() * sizeof(vox::Tile) + 16) + ((uint64_t)ncand * 32 + 16) +
                              (((uint64_t)ncand + 1) * 8 + 16) + (hb.tasks.size() * sizeof(FoldTask) + 16) +
                              hb.newxyz_total * 24 + (hb.fcands.size() * sizeof(FftdCand) + 16) + 64;
    const uint64_t need = in_bytes + out_region(hb, ncand) + 64;
    if (w.pin_bytes < need) {
        if (w.pin) (void)hipHostFree(w.pin);
        w.pin = nullptr;
        w.pin_bytes = 0;
        const uint64_t want = need * 3 / 2;
        S2H_HIP(hipHostMalloc(&w.pin, want, hipHostMallocDefault));
        w.pin_bytes = want;
    }
    uint8_t *p = (uint8_t *)w.pin;
    auto put = [&](DevBuf &dst, const void *src, uint64_t bytes) -> s2h_status {
        if (!bytes) return S2H_OK;
        std::memcpy(p, src, bytes);
        S2H_HIP(hipMemcpyAsync(dst.ptr, p, bytes, hipMemcpyHostToDevice, ctx->stream));
        p += (bytes + 15) / 16 * 16;
        return S2H_OK;
    };
    S2H_TRY(put(w.seeds, hb.sdev.data(), hb.sdev.size() * sizeof(SeedDev)));
    S2H_TRY(put(w.cands, hb.cd.data(), (uint64_t)ncand * sizeof(CandDev)));
    S2H_TRY(put(w.hdr, &hb.hdr, sizeof(BatchHeader)));
    S2H_TRY(put(w.gran, hb.gran.data(), hb.gran.size() * sizeof(vox::GranuleInfo)));
    S2H_TRY(put(w.tiles, hb.tiles.data(), hb.tiles.size() * sizeof(vox::Tile)));
    S2H_TRY(put(w.origin, hb.origin.data(), (uint64_t)ncand * 32));
    S2H_TRY(put(w.adpre, hb.ad_prefix.data(), ((uint64_t)ncand + 1) * 8));
    S2H_TRY(put(w.tasks, hb.tasks.data(), hb.tasks.size() * sizeof(FoldTask)));
    S2H_TRY(put(w.fcands, hb.fcands.data(), hb.fcands.size() * sizeof(FftdCand)));
    // new vertices: one contiguous block in candidate order
    if (hb.newxyz_total) {
        for (uint32_t c = 0; c < ncand; ++c) {
            const s2h_seed *sd = seeds[cands[c].seed];
            const uint32_t nv = sd->vstart[cands[c].particle + 1] - sd->vstart[cands[c].particle];
            std::memcpy(p + hb.xyz_off[c] * 24, cands[c].new_xyz, (size_t)nv * 24);
        }
        S2H_HIP(hipMemcpyAsync(w.newxyz.ptr, p, hb.newxyz_total * 24, hipMemcpyHostToDevice, ctx->stream));
    }
    return S2H_OK;
}

struct BatchOut {
    std::vector<double> values;
    std::vector<uint8_t> support;
    std::vector<int> occ;
    uint32_t nover = 0;
};

static s2h_status download_batch(s2h_ctx *ctx, BatchWS &w, const HostBatch &hb, uint32_t ncand, BatchOut &out) {
    uint8_t *p = (uint8_t *)w.pin + (w.pin_bytes - out_region(hb, ncand)) / 16 * 16;
    double *vals = (double *)p;
    uint8_t *sup = p + hb.curve_total * 8;
    int *occ = (int *)(p + (hb.curve_total * 9 + 15) / 16 * 16);
    uint32_t *nover = (uint32_t *)(p + (hb.curve_total * 9 + 15) / 16 * 16 + ((uint64_t)ncand * 4 + 15) / 16 * 16);
    S2H_TRY(inject(kFailReadback));
    S2H_HIP(hipMemcpyAsync(vals, w.curves.ptr, hb.curve_total * 8, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipMemcpyAsync(sup, curve_support(w), hb.curve_total, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipMemcpyAsync(occ, w.occ.ptr, (uint64_t)ncand * 4, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipMemcpyAsync(nover, w.over.ptr, 4, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    out.values.assign(vals, vals + hb.curve_total);
    out.support.assign(sup, sup + hb.curve_total);
    out.occ.assign(occ, occ + ncand);
    out.nover = *nover;
    return S2H_OK;
}

static constexpr uint32_t kRowChunkDefault = 128;

// Rows per delta item: the first batch over a new set of seed shapes times
// each candidate value on the device (HIP events around the graph replay) and
// keeps the fastest; later batches over those shapes reuse it.  Contexts with
// autotuning off, and batches of fewer than 16 candidates, use the default.
static s2h_status pick_row_chunk(s2h_ctx *ctx, BatchWS &w, s2h_seed *const *seeds, uint32_t nseeds,
                                 const s2h_candidate *cands, uint32_t ncand, HostBatch &hb) {
    std::vector<uint64_t> key;
    for (uint32_t i = 0; i < nseeds; ++i) {
        const s2h_seed *sd = seeds[i];
        for (int a = 0; a < 3; ++a) key.push_back(sd->vol->dims[a]);
        key.push_back(sd->r_max);
    }
    auto it = w.row_chunk.find(key);
    if (it != w.row_chunk.end()) {
        set_row_chunk(hb, it->second);
        return S2H_OK;
    }
    if (hb.hdr.nitems == 0) return S2H_OK;  // no direct delta counts to tune
    const char *env = std::getenv("S2HIP_AUTOTUNE");
    if (!ctx->autotune || (env && env[0] == '0') || ncand < 16) return S2H_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    S2H_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(S2H_ERR_HIP, "hipEventCreate failed");
    }
    s2h_status st = S2H_OK;
    float best = 1e30f;
    uint32_t pick = kRowChunkDefault;
    for (uint32_t rc : {32u, 64u, 128u, 256u, 512u}) {
        set_row_chunk(hb, rc);
        st = upload_batch(ctx, w, hb, cands, ncand, seeds);
        if (st == S2H_OK) st = run_sequence(ctx, w, hb, seeds);  // warm-up (and capture)
        hipError_t e = hipSuccess;
        if (st == S2H_OK) e = hipEventRecord(e0, ctx->stream);
        if (st == S2H_OK && e == hipSuccess) st = run_sequence(ctx, w, hb, seeds);
        if (st == S2H_OK && e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
        if (st == S2H_OK && e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0;
        if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (st == S2H_OK && e != hipSuccess) st = fail(S2H_ERR_HIP, std::string("batch tuning: ") + hipGetErrorString(e));
        if (st != S2H_OK) break;
        if (ms < best) {
            best = ms;
            pick = rc;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != S2H_OK) return st;
    w.row_chunk[key] = pick;
    set_row_chunk(hb, pick);
    return S2H_OK;
}

// One batch end to end; the curves are finished on the host.
static s2h_status evaluate(s2h_ctx *ctx, s2h_seed *const *seeds, uint32_t nseeds, const s2h_candidate *cands,
                           uint32_t ncand, double *curves, uint64_t ncurve, HostBatch &hb, BatchOut &out) {
    if (!ctx->batch_ws) ctx->batch_ws = new BatchWS();
    BatchWS &w = *ctx->batch_ws;
    S2H_TRY(build_host(ctx, w, seeds, nseeds, cands, ncand, kRowChunkDefault, hb));
    if (ncurve != hb.curve_total) return fail(S2H_ERR_SIZE, "curves buffer must hold sum of r_max+1 over candidates");
    if (ncand == 0) return S2H_OK;
    S2H_TRY(pick_row_chunk(ctx, w, seeds, nseeds, cands, ncand, hb));
    S2H_TRY(upload_batch(ctx, w, hb, cands, ncand, seeds));
    ctx->launches = 0;
    prof_mark(ctx, 0);
    S2H_TRY(run_sequence(ctx, w, hb, seeds));
    prof_mark(ctx, 1);
    prof_mark(ctx, 2);
    S2H_TRY(download_batch(ctx, w, hb, ncand, out));
    prof_collect(ctx);
    if (out.nover > kOverflowCap) return fail(S2H_ERR_UNSUPPORTED, "more than 65536 candidate rays with over 16 hits");
    // finish each curve: vf, fill, S2(0) = vf
    for (uint32_t c = 0; c < ncand; ++c) {
        const s2h_seed *sd = seeds[hb.cd[c].seed];
        const uint64_t n = sd->r_max + 1, off = hb.cd[c].curve_off;
        const uint64_t cells = sd->vol->dims[0] * sd->vol->dims[1] * sd->vol->dims[2];
        const int64_t occ = (int64_t)sd->occupied + out.occ[c];
        double *o = curves + off;
        if (occ == 0) {
            std::fill(o, o + n, 0.0);
            continue;
        }
        const double vf = (double)occ / (double)cells;
        std::memcpy(o, out.values.data() + off, n * 8);
        std::vector<uint8_t> sup(out.support.begin() + (long)off, out.support.begin() + (long)(off + n));
        o[0] = vf;
        sup[0] = 1;
        fill_missing(o, sup.data(), n, vf);
        o[0] = vf;
    }
    return S2H_OK;
}

}  // namespace batch
}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_seed_create(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntris,
                           const uint32_t *granule_of_tri, uint32_t ngranules, const uint32_t *particle_vert_start,
                           const uint32_t *particle_tri_start, uint32_t nparticles, const double bbox_min[3],
                           double pitch, const uint64_t dims[3], uint32_t r_max, s2h_seed **out) {
    if (!ctx || !out || !dims || !bbox_min || !particle_vert_start || !particle_tri_start ||
        (ntris && (!xyz || !tri || !granule_of_tri)))
        return fail(S2H_ERR_ARG, "null argument");
    // particles: contiguous vertex and triangle ranges covering the mesh, each
    // triangle using only its particle's vertices
    if (particle_vert_start[0] != 0 || particle_tri_start[0] != 0 || particle_vert_start[nparticles] != nverts ||
        particle_tri_start[nparticles] != ntris)
        return fail(S2H_ERR_ARG, "particle ranges must cover the vertices and triangles");
    for (uint32_t p = 0; p < nparticles; ++p) {
        if (particle_vert_start[p + 1] < particle_vert_start[p] || particle_tri_start[p + 1] < particle_tri_start[p])
            return fail(S2H_ERR_ARG, "particle ranges must be non-decreasing");
        for (uint32_t f = particle_tri_start[p]; f < particle_tri_start[p + 1]; ++f)
            for (int k = 0; k < 3; ++k)
                if (tri[3 * (uint64_t)f + k] < particle_vert_start[p] ||
                    tri[3 * (uint64_t)f + k] >= particle_vert_start[p + 1])
                    return fail(S2H_ERR_ARG, "a triangle uses a vertex of another particle");
    }
    for (uint64_t f = 0; f < ntris; ++f)
        if (granule_of_tri[f] >= ngranules) return fail(S2H_ERR_ARG, "granule id out of range");
    S2H_HIP(hipSetDevice(ctx->device));
    auto *sd = new s2h_seed();
    auto cleanup = [&](s2h_status st) {
        s2h_seed_free(ctx, sd);
        return st;
    };
    s2h_status st = s2h_voxelize(ctx, xyz, nverts, tri, ntris, granule_of_tri, ngranules, bbox_min, pitch, dims,
                                 &sd->vol, &sd->occupied);
    if (st != S2H_OK) return cleanup(st);
    for (int a = 0; a < 3; ++a) sd->bmin[a] = bbox_min[a];
    sd->pitch = pitch;
    sd->r_max = r_max;
    // particle geometry: local granules (by first appearance), triangles sorted
    // by (particle, local granule) with particle-local vertex indices
    sd->nparticles = nparticles;
    sd->vstart.assign(particle_vert_start, particle_vert_start + nparticles + 1);
    sd->tstart.assign(particle_tri_start, particle_tri_start + nparticles + 1);
    sd->xyz.assign(xyz, xyz + 3 * nverts);
    sd->vgran.assign(nverts, -1);
    sd->pgran.assign(nparticles + 1, 0);
    sd->ngran.assign(nparticles, 0);
    std::vector<uint32_t> tri_sorted(3 * ntris);
    std::vector<int32_t> local(ngranules, -1);
    uint64_t tpos = 0;
    for (uint32_t p = 0; p < nparticles; ++p) {
        const uint32_t f0 = particle_tri_start[p], f1 = particle_tri_start[p + 1], v0 = particle_vert_start[p];
        std::vector<uint32_t> ids;
        for (uint32_t f = f0; f < f1; ++f) {
            const uint32_t g = granule_of_tri[f];
            if (local[g] < 0) {
                local[g] = (int32_t)ids.size();
                ids.push_back(g);
            }
        }
        std::vector<std::vector<uint32_t>> by(ids.size());
        for (uint32_t f = f0; f < f1; ++f) by[local[granule_of_tri[f]]].push_back(f);
        sd->pgran[p] = (uint32_t)sd->gtri.size();
        sd->ngran[p] = (uint32_t)ids.size();
        for (size_t g = 0; g < ids.size(); ++g) {
            sd->gtri.push_back((uint32_t)tpos);
            for (uint32_t f : by[g]) {
                for (int k = 0; k < 3; ++k) {
                    const uint32_t v = tri[3 * (uint64_t)f + k];
                    tri_sorted[3 * tpos + k] = v - v0;
                    sd->vgran[v] = (int32_t)g;
                }
                ++tpos;
            }
        }
        sd->gtri.push_back((uint32_t)tpos);  // end of the particle's last granule
        for (uint32_t g : ids) local[g] = -1;
    }
    if ((st = upload(ctx, sd->d_xyz, xyz, std::max<uint64_t>(nverts, 1) * 24)) != S2H_OK) return cleanup(st);
    if ((st = upload(ctx, sd->d_tri, tri_sorted.data(), std::max<uint64_t>(ntris, 1) * 12)) != S2H_OK)
        return cleanup(st);
    if ((st = batch::seed_coverage(ctx, sd, xyz, tri, ntris, granule_of_tri, ngranules, dims)) != S2H_OK)
        return cleanup(st);
    // seed exact counts (kept: the plan's table is reused by later evaluations)
    const double pm = pitch < 1e-9 ? 1e-9 : pitch;
    if ((st = exact_plan_get(ctx, dims, r_max, pm, &sd->plan)) != S2H_OK) return cleanup(st);
    if ((st = exact_run_counts(ctx, sd->vol, sd->plan)) != S2H_OK) return cleanup(st);
    if ((st = batch::seed_shell_tables(ctx, sd)) != S2H_OK) return cleanup(st);
    for (const DzBlock &b : sd->plan->blocks) sd->gb_host.push_back(batch::GB{b.group, b.dz0, b.dzb});
    if ((st = upload(ctx, sd->gbs, sd->gb_host.data(), std::max<size_t>(sd->gb_host.size(), 1) * sizeof(batch::GB))) !=
        S2H_OK)
        return cleanup(st);
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    *out = sd;
    return S2H_OK;
}

void s2h_seed_free(s2h_ctx *ctx, s2h_seed *sd) {
    if (!ctx || !sd) return;
    (void)hipStreamSynchronize(ctx->stream);
    if (sd->vol) s2h_vol_free(ctx, sd->vol);
    dev_free(ctx, sd->cov);
    dev_free(ctx, sd->shtable);
    dev_free(ctx, sd->sg);
    dev_free(ctx, sd->smap);
    dev_free(ctx, sd->gbs);
    dev_free(ctx, sd->d_xyz);
    dev_free(ctx, sd->d_tri);
    delete sd;
}

s2h_status s2h_seed_volume(s2h_seed *sd, s2h_vol **vol) {
    if (!sd || !vol) return fail(S2H_ERR_ARG, "null argument");
    *vol = sd->vol;
    return S2H_OK;
}

s2h_status s2h_batch_eval(s2h_ctx *ctx, s2h_seed *const *seeds, uint32_t nseeds, const s2h_candidate *cands,
                          uint32_t ncand, double *curves, uint64_t ncurve) {
    if (!ctx || (!seeds && nseeds) || (!cands && ncand) || (!curves && ncurve)) return fail(S2H_ERR_ARG, "null argument");
    S2H_HIP(hipSetDevice(ctx->device));
    S2H_TRY(inject(kFailLaunch));
    batch::HostBatch hb;
    batch::BatchOut out;
    return batch::evaluate(ctx, seeds, nseeds, cands, ncand, curves, ncurve, hb, out);
}

s2h_status s2h_seed_accept(s2h_ctx *ctx, s2h_seed *sd, uint32_t particle, const double *new_xyz, double *curve,
                           uint64_t n) {
    using namespace s2h::batch;
    if (!ctx || !sd || !new_xyz) return fail(S2H_ERR_ARG, "null argument");
    if (particle >= sd->nparticles) return fail(S2H_ERR_ARG, "particle index out of range");
    if (curve && n != sd->r_max + 1) return fail(S2H_ERR_SIZE, "curve buffer must hold r_max+1 values");
    S2H_HIP(hipSetDevice(ctx->device));
    S2H_TRY(inject(kFailLaunch));
    std::vector<double> tmp(sd->r_max + 1);
    s2h_seed *const seeds[1] = {sd};
    const s2h_candidate cand{0, particle, new_xyz};
    HostBatch hb;
    BatchOut out;
    S2H_TRY(evaluate(ctx, seeds, 1, &cand, 1, tmp.data(), tmp.size(), hb, out));
    BatchWS &w = *ctx->batch_ws;
    const uint64_t *dims = sd->vol->dims;
    const int nwd = (int)((dims[2] + 31) / 32);
    apply_cov_kernel<<<1024, 256, 0, ctx->stream>>>((const CandDev *)w.cands.ptr, (const int32_t *)w.box.ptr,
                                                    (uint32_t *)sd->cov.ptr, (int)dims[1], (int)dims[2]);
    apply_rows_kernel<<<1024, 256, 0, ctx->stream>>>((const CandDev *)w.cands.ptr, (const uint2 *)w.ad.ptr,
                                                     (uint32_t *)sd->vol->d_words, (int)dims[1], nwd,
                                                     (int)(sd->vol->row_words * 2));
    apply_table_kernel<<<1024, 256, 0, ctx->stream>>>((const CandDev *)w.cands.ptr, (const int32_t *)w.delta.ptr,
                                                      (unsigned long long *)sd->shtable.ptr, sd->ncanon);
    S2H_HIP(hipGetLastError());
    const uint32_t v0 = sd->vstart[particle], nv = sd->vstart[particle + 1] - v0;
    std::memcpy(sd->xyz.data() + 3 * (uint64_t)v0, new_xyz, (size_t)nv * 24);
    if (nv)
        S2H_HIP(hipMemcpyAsync((double *)sd->d_xyz.ptr + 3 * (uint64_t)v0, sd->xyz.data() + 3 * (uint64_t)v0,
                               (size_t)nv * 24, hipMemcpyHostToDevice, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    sd->occupied = (uint64_t)((int64_t)sd->occupied + out.occ[0]);
    sd->vol->occupied = sd->occupied;
    if (curve) std::memcpy(curve, tmp.data(), tmp.size() * 8);
    return S2H_OK;
}

}  // extern "C"
