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
//     the candidate's changed box B;
//  2. the rows crossing B are rebuilt: new bit = (coverage + delta) > 0;
//  3. exact delta counts.  C'(d) - C(d) is a sum over row pairs (u, u+(dx,dy))
//     with at least one changed row.  T1 handles pairs whose partner row
//     q = u+(dx,dy) changed: q is the wave-uniform row (shifted right by dz),
//     lanes hold u = q-(dx,dy), new and old versions.  T2 handles pairs whose
//     first row q changed and whose partner v = q+(dx,dy) lies outside B: q is
//     uniform and shifted LEFT by dz against the unchanged lane row v.  Every
//     pair with a changed row is counted exactly once; only the words whose bits
//     can differ are visited;
//  4. C' = C + dC, then the reference's ordered f64 fold per radius -> curves.
// All integer arithmetic, so every curve is bit-identical to a full evaluation
// of the candidate packing.  The kernels are grid-stride over device-resident
// work descriptors, so one launch shape serves any batch and the sequence can
// be captured in a HIP graph and replayed.
#include <algorithm>
#include <cmath>
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
    const unsigned long long *table;  // seed counts, dense canonical table
    const int32_t *gather;      // shell order -> table index (-1: zero offset)
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
    uint32_t nchunks;                        // row chunks of the footprint
    uint64_t rows_off;                       // u32 offset of the rebuilt rows
    uint64_t box_off;                        // int32 offset of the coverage delta box
    uint64_t delta_off;                      // int32 offset of the delta table
    uint64_t item_start;                     // first delta item of this candidate
    uint64_t curve_off;                      // double offset of the curve values
};

static constexpr int kRowChunk = 64;

struct BatchHeader {
    uint64_t nitems;      // delta-count items
    uint32_t ncand;
    uint32_t nfold;       // (candidate, radius) fold tasks
};

// ---------------------------------------------------------------- 1. coverage deltas

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
                                                          const uint32_t *__restrict__ ntiles,
                                                          const double *__restrict__ origin,  // per candidate: x,y,z,pitch
                                                          const CandDev *__restrict__ cands,
                                                          int32_t *__restrict__ box, uint32_t *__restrict__ overflow,
                                                          uint32_t overflow_cap) {
    const uint32_t n = *ntiles;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) {
        const vox::Tile tl = tiles[t];
        const uint32_t c = gran[tl.granule].owner;
        const double *o = origin + 4 * (uint64_t)c;
        vox::voxelize_tile(tris, gran, tl, t, o[0], o[1], o[2], o[3], overflow, overflow_cap, AddDelta{cands, box});
        __syncthreads();
    }
}

__global__ void batch_voxel_overflow_kernel(const vox::TriGeo *__restrict__ tris,
                                            const vox::GranuleInfo *__restrict__ gran,
                                            const vox::Tile *__restrict__ tiles, const double *__restrict__ origin,
                                            const CandDev *__restrict__ cands, int32_t *__restrict__ box,
                                            const uint32_t *__restrict__ overflow, uint32_t cap) {
    const uint32_t count = min(overflow[0], cap);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        const vox::Tile tl = tiles[overflow[1 + 2 * k]];
        const uint32_t c = gran[tl.granule].owner;
        const double *o = origin + 4 * (uint64_t)c;
        vox::voxel_overflow_ray(tris, gran, tiles, o[0], o[1], o[2], o[3], overflow, k, AddDelta{cands, box});
    }
}

// ---------------------------------------------------------------- 2. rebuilt rows

// One thread per (candidate, footprint row, dword).  Bits outside [bz0, bz1)
// are copied from the seed; inside, bit = coverage + delta > 0.
__global__ void build_rows_kernel(const SeedDev *__restrict__ seeds, const CandDev *__restrict__ cands,
                                  uint32_t ncand, const uint64_t *__restrict__ row_prefix,
                                  const int32_t *__restrict__ box, uint32_t *__restrict__ rows,
                                  int *__restrict__ occ_delta) {
    const uint64_t total = row_prefix[ncand];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = ncand;  // candidate of dword i
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (row_prefix[mid] <= i) lo = mid;
            else hi = mid;
        }
        const CandDev &c = cands[lo];
        const SeedDev &s = seeds[c.seed];
        const uint64_t local = i - row_prefix[lo];
        const uint64_t r = local / s.row_dw;
        const int w = (int)(local % s.row_dw);
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
        rows[c.rows_off + local] = v;
        const int d = __popc(v) - __popc(old);
        if (d) atomicAdd(&occ_delta[lo], d);
    }
}

// ---------------------------------------------------------------- 3. delta counts

// uniform dword k of a row (zero beyond the row's voxels)
__device__ __forceinline__ uint32_t urow(const uint32_t *row, int k, int nwd) {
    const int kk = k < 0 ? 0 : (k < nwd ? k : nwd - 1);
    const uint32_t v = row[kk];
    return (k >= 0 && k < nwd) ? v : 0u;
}

template <int DZB>
__device__ void delta_item(const SeedDev &s, const CandDev &c, uint32_t cand, const GB gb, int chunk,
                           const uint32_t *__restrict__ rows, int32_t *__restrict__ delta) {
    const int lane = threadIdx.x & 63;
    const ColInfo col = s.cols[gb.group * 64 + lane];
    const int nwd = (s.nz + 31) / 32;
    const int q0 = gb.dz0 >> 5;
    const int sy = c.by1 - c.by0;
    const int nrows = (c.bx1 - c.bx0) * sy;
    const int r0 = chunk * kRowChunk, r1 = min(nrows, r0 + kRowChunk);
    const __amdgpu_buffer_rsrc_t seed_rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)s.vol, 0, (int)s.vol_bytes, 0x00020000);
    const int oob = 0x7ffffff0;
    // T1 words: a changed in [bz0, bz1); (b >> dz) changed at z in [bz0 - dz, bz1 - dz)
    const int t1_lo = max(0, c.bz0 - gb.dz0 - DZB + 1) >> 5, t1_hi = min(nwd, (c.bz1 + 31) >> 5);
    // T2 words: (b << dz) changed at z in [bz0 + dz, bz1 + dz)
    const int t2_lo = min(nwd, (c.bz0 + gb.dz0) >> 5), t2_hi = min(nwd, (c.bz1 + gb.dz0 + DZB - 1 + 31) >> 5);
    uint32_t cn[DZB], co[DZB];
#pragma unroll
    for (int j = 0; j < DZB; ++j) cn[j] = co[j] = 0;
    const bool active = col.m >= 0;
    for (int r = r0; r < r1; ++r) {
        const int xq = c.bx0 + r / sy, yq = c.by0 + r % sy;
        const uint32_t *qn = rows + c.rows_off + (uint64_t)r * s.row_dw;               // new q row
        const uint32_t *qo = s.vol + ((uint64_t)xq * s.ny + yq) * s.row_dw;         // old q row
        // ---- T1: lanes u = q - (dx, dy)
        {
            const int xu = xq - col.dx, yu = yq - col.dy;
            const bool in = active && xu >= 0 && xu < s.nx && yu >= 0 && yu < s.ny;
            const bool fb = in && xu >= c.bx0 && xu < c.bx1 && yu >= c.by0 && yu < c.by1;
            const int aoff = in ? (int)(((uint64_t)xu * s.ny + yu) * s.row_dw * 4) : oob;
            const uint32_t *anew = fb ? rows + c.rows_off + ((uint64_t)(xu - c.bx0) * sy + (yu - c.by0)) * s.row_dw
                                      : nullptr;
            for (int w = t1_lo; w < t1_hi; ++w) {
                const uint32_t ao = __builtin_amdgcn_raw_buffer_load_b32(seed_rs, aoff + 4 * w, 0, 0);
                const uint32_t an = fb ? anew[w] : ao;
                const uint32_t bn0 = urow(qn, w + q0, nwd), bn1 = urow(qn, w + q0 + 1, nwd);
                const uint32_t bo0 = urow(qo, w + q0, nwd), bo1 = urow(qo, w + q0 + 1, nwd);
                const uint64_t pn = ((uint64_t)bn1 << 32) | bn0, po = ((uint64_t)bo1 << 32) | bo0;
#pragma unroll
                for (int j = 0; j < DZB; ++j) {
                    cn[j] = popc_acc(an & (uint32_t)(pn >> j), cn[j]);
                    co[j] = popc_acc(ao & (uint32_t)(po >> j), co[j]);
                }
            }
        }
        // ---- T2: lanes v = q + (dx, dy) outside the changed box
        {
            const int xv = xq + col.dx, yv = yq + col.dy;
            const bool in = active && xv >= 0 && xv < s.nx && yv >= 0 && yv < s.ny &&
                            !(xv >= c.bx0 && xv < c.bx1 && yv >= c.by0 && yv < c.by1);
            const int aoff = in ? (int)(((uint64_t)xv * s.ny + yv) * s.row_dw * 4) : oob;
            for (int w = t2_lo; w < t2_hi; ++w) {
                const uint32_t a = __builtin_amdgcn_raw_buffer_load_b32(seed_rs, aoff + 4 * w, 0, 0);
                const uint32_t bn1 = urow(qn, w - q0, nwd), bn0 = urow(qn, w - q0 - 1, nwd);
                const uint32_t bo1 = urow(qo, w - q0, nwd), bo0 = urow(qo, w - q0 - 1, nwd);
                const uint64_t pn = ((uint64_t)bn1 << 32) | bn0, po = ((uint64_t)bo1 << 32) | bo0;
#pragma unroll
                for (int j = 0; j < DZB; ++j) {
                    cn[j] = popc_acc(a & (uint32_t)((pn << j) >> 32), cn[j]);
                    co[j] = popc_acc(a & (uint32_t)((po << j) >> 32), co[j]);
                }
            }
        }
    }
    int32_t *dt = delta + c.delta_off;
#pragma unroll
    for (int j = 0; j < DZB; ++j) {
        const int dz = gb.dz0 + j;
        const int d = (int)(cn[j] - co[j]);
        if (active && dz <= col.m && d) atomicAdd(&dt[col.base + dz], d);
    }
    (void)cand;
}

static constexpr int kDeltaWaves = 4;

__global__ __launch_bounds__(64 * kDeltaWaves) void delta_count_kernel(const SeedDev *__restrict__ seeds,
                                                                       const CandDev *__restrict__ cands,
                                                                       const BatchHeader *__restrict__ hdr,
                                                                       const uint32_t *__restrict__ rows,
                                                                       int32_t *__restrict__ delta) {
    const uint64_t nitems = hdr->nitems;
    const uint32_t ncand = hdr->ncand;
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
            case 8: delta_item<8>(s, c, lo, gb, chunk, rows, delta); break;
            case 16: delta_item<16>(s, c, lo, gb, chunk, rows, delta); break;
            case 24: delta_item<24>(s, c, lo, gb, chunk, rows, delta); break;
            default: delta_item<32>(s, c, lo, gb, chunk, rows, delta); break;
        }
    }
}

// ---------------------------------------------------------------- 4. gather + ordered fold

static constexpr int kFoldChunkB = 1024;

// One workgroup per (candidate, radius): terms are produced in parallel into LDS
// (hits = seed count + delta, over analytic valid) and folded in shell order by
// lane 0, the reference's sequential left fold.
__global__ __launch_bounds__(256) void batch_fold_kernel(const SeedDev *__restrict__ seeds,
                                                         const CandDev *__restrict__ cands,
                                                         const BatchHeader *__restrict__ hdr,
                                                         const uint32_t *__restrict__ fold_prefix,
                                                         const int32_t *__restrict__ delta,
                                                         const int *__restrict__ occ_delta,
                                                         double *__restrict__ curves, uint8_t *__restrict__ support) {
    __shared__ double buf[kFoldChunkB];
    const uint32_t ncand = hdr->ncand, nfold = hdr->nfold;
    for (uint32_t task = blockIdx.x; task < nfold; task += gridDim.x) {
        uint32_t lo = 0, hi = ncand;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (fold_prefix[mid] <= task) lo = mid;
            else hi = mid;
        }
        const CandDev &c = cands[lo];
        const SeedDev &s = seeds[c.seed];
        const uint32_t r = task - fold_prefix[lo] + 1;
        const uint64_t b = s.rstart[r], e = s.rstart[r + 1];
        const int32_t *dt = delta + c.delta_off;
        const uint64_t occ = (uint64_t)((int64_t)s.occupied + occ_delta[lo]);
        double sum = 0.0;
        for (uint64_t c0 = b; c0 < e; c0 += kFoldChunkB) {
            const uint64_t len = min((uint64_t)kFoldChunkB, e - c0);
            __syncthreads();
            for (uint64_t k = threadIdx.x; k < len; k += blockDim.x) {
                const uint64_t i = c0 + k;
                const Offset o = s.offs[i];
                const int32_t g = s.gather[i];
                const uint64_t h = g >= 0 ? (uint64_t)((int64_t)s.table[g] + dt[g]) : occ;
                const uint64_t valid = (uint64_t)(s.nx - abs(o.dx)) * (uint64_t)(s.ny - abs(o.dy)) *
                                       (uint64_t)(s.nz - abs(o.dz));
                buf[k] = (double)h / (double)valid;
            }
            __syncthreads();
            if (threadIdx.x == 0)
                for (uint64_t k = 0; k < len; ++k) sum = sum + buf[k];
        }
        if (threadIdx.x == 0) {
            curves[c.curve_off + r] = e > b ? sum / (double)(e - b) : 0.0;
            support[c.curve_off + r] = e > b ? 1 : 0;
        }
        __syncthreads();
    }
}

__global__ void zero_kernel(uint32_t *__restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

}  // namespace batch
}  // namespace s2h

// ================================================================ host side

struct s2h_seed {
    s2h_vol *vol = nullptr;
    s2h::DevBuf cov, table, gbs;
    s2h::ExactPlan *plan = nullptr;
    double bmin[3];
    double pitch;
    uint32_t r_max;
    uint64_t occupied;
    std::vector<s2h::batch::GB> gb_host;
};

namespace s2h {
namespace batch {

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

// Group the triangles of one mesh by granule and prepare tiles (owner/sign set).
static void prepare_mesh(const double *tri_xyz, uint32_t ntri, const uint32_t *granule, uint32_t ngran,
                         const double bmin[3], double pitch, const uint64_t dims[3], uint32_t owner, int sign,
                         std::vector<vox::TriGeo> &geo, std::vector<vox::GranuleInfo> &gran,
                         std::vector<vox::Tile> &tiles, int64_t box_lo[3], int64_t box_hi[3]) {
    std::vector<uint32_t> start(ngran + 1, 0);
    for (uint32_t f = 0; f < ntri; ++f) start[granule[f] + 1]++;
    for (uint32_t g = 0; g < ngran; ++g) start[g + 1] += start[g];
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    std::vector<double> txyz(9 * (size_t)ntri);
    for (uint32_t f = 0; f < ntri; ++f) std::memcpy(&txyz[9 * (size_t)fill[granule[f]]++], tri_xyz + 9 * (size_t)f, 72);
    const uint32_t base = (uint32_t)geo.size();
    geo.resize(geo.size() + ntri);
    for (uint32_t g = 0; g < ngran; ++g) {
        vox::GranuleInfo gi;
        gi.tri_start = base + start[g];
        gi.tri_end = base + start[g + 1];
        gi.owner = owner;
        gi.sign = sign;
        int64_t lo[3], hi[3];
        const bool ok = vox::prepare_granule(txyz.data() + 9 * (size_t)start[g], start[g + 1] - start[g], bmin, pitch,
                                             dims, geo.data() + base + start[g], gi, lo, hi);
        const uint32_t gidx = (uint32_t)gran.size();
        gran.push_back(gi);
        if (!ok) continue;
        vox::push_tiles(gidx, lo, hi, tiles);
        for (int a = 0; a < 3; ++a) {
            box_lo[a] = std::min(box_lo[a], lo[a]);
            box_hi[a] = std::max(box_hi[a], hi[a]);
        }
    }
}

}  // namespace batch
}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_seed_create(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntris,
                           const uint32_t *granule_of_tri, uint32_t ngranules, const double bbox_min[3], double pitch,
                           const uint64_t dims[3], uint32_t r_max, s2h_seed **out) {
    if (!ctx || !out || !dims || !bbox_min) return fail(S2H_ERR_ARG, "null argument");
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
    // coverage counts: the same predicate, counting granules per centre
    const uint64_t cells = dims[0] * dims[1] * dims[2];
    if ((st = dev_alloc(ctx, sd->cov, cells * 4)) != S2H_OK) return cleanup(st);
    batch::zero_kernel<<<4096, 256, 0, ctx->stream>>>((uint32_t *)sd->cov.ptr, cells);
    {
        std::vector<double> txyz(9 * ntris);
        for (uint64_t f = 0; f < ntris; ++f)
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 3; ++a) txyz[9 * f + 3 * c + a] = xyz[3 * (uint64_t)tri[3 * f + c] + a];
        std::vector<vox::TriGeo> geo;
        std::vector<vox::GranuleInfo> gran;
        std::vector<vox::Tile> tiles;
        int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
        batch::prepare_mesh(txyz.data(), (uint32_t)ntris, granule_of_tri, ngranules, bbox_min, pitch, dims, 0, 1, geo,
                            gran, tiles, lo, hi);
        if (!tiles.empty()) {
            DevBuf dg, dgr, dt, dov;
            const uint32_t cap = 1u << 16;
            st = upload(ctx, dg, geo.data(), geo.size() * sizeof(vox::TriGeo));
            if (st == S2H_OK) st = upload(ctx, dgr, gran.data(), gran.size() * sizeof(vox::GranuleInfo));
            if (st == S2H_OK) st = upload(ctx, dt, tiles.data(), tiles.size() * sizeof(vox::Tile));
            if (st == S2H_OK) st = dev_alloc(ctx, dov, (1 + 2 * (uint64_t)cap) * 4);
            uint32_t nover = 0;
            hipError_t e = hipSuccess;
            if (st == S2H_OK) {
                e = hipMemsetAsync(dov.ptr, 0, 4, ctx->stream);
                if (e == hipSuccess) {
                    batch::cov_kernel<<<(unsigned)tiles.size(), 256, 0, ctx->stream>>>(
                        (const vox::TriGeo *)dg.ptr, (const vox::GranuleInfo *)dgr.ptr, (const vox::Tile *)dt.ptr,
                        bbox_min[0], bbox_min[1], bbox_min[2], pitch, (uint32_t *)sd->cov.ptr, dims[1], dims[2],
                        (uint32_t *)dov.ptr, cap);
                    e = hipGetLastError();
                }
                if (e == hipSuccess) e = hipMemcpyAsync(&nover, dov.ptr, 4, hipMemcpyDeviceToHost, ctx->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
                if (e == hipSuccess && nover > cap) st = fail(S2H_ERR_UNSUPPORTED, "too many overflow rays");
                if (e == hipSuccess && st == S2H_OK && nover) {
                    batch::cov_overflow_kernel<<<(nover + 63) / 64, 64, 0, ctx->stream>>>(
                        (const vox::TriGeo *)dg.ptr, (const vox::GranuleInfo *)dgr.ptr, (const vox::Tile *)dt.ptr,
                        bbox_min[0], bbox_min[1], bbox_min[2], pitch, (uint32_t *)sd->cov.ptr, dims[1], dims[2],
                        (const uint32_t *)dov.ptr, nover);
                    e = hipGetLastError();
                }
                if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
                if (e != hipSuccess) st = fail(S2H_ERR_HIP, std::string("coverage kernel: ") + hipGetErrorString(e));
            }
            dev_free(ctx, dg);
            dev_free(ctx, dgr);
            dev_free(ctx, dt);
            dev_free(ctx, dov);
            if (st != S2H_OK) return cleanup(st);
        }
    }
    // seed exact counts (kept: the plan's table is reused by later evaluations)
    const double pm = pitch < 1e-9 ? 1e-9 : pitch;
    if ((st = exact_plan_get(ctx, dims, r_max, pm, &sd->plan)) != S2H_OK) return cleanup(st);
    if ((st = exact_run_counts(ctx, sd->vol, sd->plan)) != S2H_OK) return cleanup(st);
    if ((st = dev_alloc(ctx, sd->table, std::max<uint64_t>(sd->plan->table_size, 1) * 8)) != S2H_OK)
        return cleanup(st);
    S2H_HIP(hipMemcpyAsync(sd->table.ptr, sd->plan->d_table.ptr, sd->plan->table_size * 8,
                           hipMemcpyDeviceToDevice, ctx->stream));
    for (const DzBlock &b : sd->plan->blocks) sd->gb_host.push_back(batch::GB{b.group, b.dz0, b.dzb});
    if ((st = upload(ctx, sd->gbs, sd->gb_host.data(), sd->gb_host.size() * sizeof(batch::GB))) != S2H_OK)
        return cleanup(st);
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    *out = sd;
    return S2H_OK;
}

void s2h_seed_free(s2h_ctx *ctx, s2h_seed *sd) {
    if (!ctx || !sd) return;
    if (sd->vol) s2h_vol_free(ctx, sd->vol);
    dev_free(ctx, sd->cov);
    dev_free(ctx, sd->table);
    dev_free(ctx, sd->gbs);
    delete sd;
}

s2h_status s2h_seed_volume(s2h_seed *sd, s2h_vol **vol) {
    if (!sd || !vol) return fail(S2H_ERR_ARG, "null argument");
    *vol = sd->vol;
    return S2H_OK;
}

s2h_status s2h_batch_eval(s2h_ctx *ctx, s2h_seed *const *seeds, uint32_t nseeds, const s2h_candidate *cands,
                          uint32_t ncand, double *curves, uint64_t ncurve) {
    using namespace s2h::batch;
    if (!ctx || (!seeds && nseeds) || (!cands && ncand) || (!curves && ncurve)) return fail(S2H_ERR_ARG, "null argument");
    S2H_HIP(hipSetDevice(ctx->device));
    // ---- host: descriptors
    std::vector<SeedDev> sdev(nseeds);
    for (uint32_t i = 0; i < nseeds; ++i) {
        const s2h_seed *sd = seeds[i];
        const ExactPlan *p = sd->plan;
        SeedDev &d = sdev[i];
        d.vol = (const uint32_t *)sd->vol->d_words;
        d.cov = (const uint32_t *)sd->cov.ptr;
        d.table = (const unsigned long long *)sd->table.ptr;
        d.gather = (const int32_t *)p->d_gather.ptr;
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
    std::vector<CandDev> cd(ncand);
    std::vector<vox::TriGeo> geo;
    std::vector<vox::GranuleInfo> gran;
    std::vector<vox::Tile> tiles;
    std::vector<double> origin(4 * (size_t)ncand);
    std::vector<uint64_t> row_prefix(ncand + 1, 0);
    std::vector<uint32_t> fold_prefix(ncand + 1, 0);
    uint64_t rows_total = 0, box_total = 0, delta_total = 0, items = 0, curve_total = 0;
    for (uint32_t c = 0; c < ncand; ++c) {
        const s2h_candidate &k = cands[c];
        if (k.seed >= nseeds) return fail(S2H_ERR_ARG, "candidate seed index out of range");
        const s2h_seed *sd = seeds[k.seed];
        const uint64_t *dims = sd->vol->dims;
        for (uint32_t f = 0; f < k.ntris; ++f)
            if (k.granule[f] >= k.ngranules) return fail(S2H_ERR_ARG, "candidate granule id out of range");
        int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
        prepare_mesh(k.old_tri, k.ntris, k.granule, k.ngranules, sd->bmin, sd->pitch, dims, c, -1, geo, gran, tiles, lo,
                     hi);
        prepare_mesh(k.new_tri, k.ntris, k.granule, k.ngranules, sd->bmin, sd->pitch, dims, c, +1, geo, gran, tiles, lo,
                     hi);
        CandDev &d = cd[c];
        d.seed = k.seed;
        if (lo[0] > hi[0]) {  // particle outside the grid at both positions: nothing changes
            for (int a = 0; a < 3; ++a) lo[a] = hi[a] = 0;
        }
        d.bx0 = (int32_t)lo[0];
        d.by0 = (int32_t)lo[1];
        d.bz0 = (int32_t)lo[2];
        d.bx1 = (int32_t)hi[0];
        d.by1 = (int32_t)hi[1];
        d.bz1 = (int32_t)hi[2];
        const uint64_t nrows = (uint64_t)(hi[0] - lo[0]) * (uint64_t)(hi[1] - lo[1]);
        d.nchunks = (uint32_t)((nrows + kRowChunk - 1) / kRowChunk);
        d.rows_off = rows_total;
        rows_total += nrows * (uint64_t)sdev[k.seed].row_dw;
        row_prefix[c + 1] = rows_total;
        d.box_off = box_total;
        box_total += (uint64_t)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
        d.delta_off = delta_total;
        delta_total += sd->plan->table_size;
        d.item_start = items;
        items += (uint64_t)d.nchunks * sdev[k.seed].ngb;
        d.curve_off = curve_total;
        curve_total += sd->r_max + 1;
        fold_prefix[c + 1] = fold_prefix[c] + sd->r_max;
        origin[4 * c] = sd->bmin[0];
        origin[4 * c + 1] = sd->bmin[1];
        origin[4 * c + 2] = sd->bmin[2];
        origin[4 * c + 3] = sd->pitch;
    }
    if (ncurve != curve_total) return fail(S2H_ERR_SIZE, "curves buffer must hold sum of r_max+1 over candidates");
    if (ncand == 0) return S2H_OK;
    BatchHeader hdr{items, ncand, fold_prefix[ncand]};
    uint32_t ntiles = (uint32_t)tiles.size();
    // ---- device buffers (grow-only scratch owned by this call)
    DevBuf b_seeds, b_cands, b_hdr, b_geo, b_gran, b_tiles, b_ntiles, b_origin, b_rowpre, b_foldpre, b_zero, b_rows,
        b_curve, b_over;
    auto release = [&]() {
        DevBuf *all[] = {&b_seeds, &b_cands, &b_hdr, &b_geo, &b_gran, &b_tiles, &b_ntiles, &b_origin, &b_rowpre,
                         &b_foldpre, &b_zero, &b_rows, &b_curve, &b_over};
        for (DevBuf *b : all) dev_free(ctx, *b);
    };
    // zeroed region: [delta tables][boxes][occ_delta][overflow counter]
    const uint64_t zero_words = delta_total + box_total + ncand + 1;
    const uint32_t over_cap = 1u << 16;
    s2h_status st = S2H_OK;
    auto up = [&](DevBuf &b, const void *src, uint64_t bytes) {
        if (st == S2H_OK) st = upload(ctx, b, src, std::max<uint64_t>(bytes, 8));
    };
    up(b_seeds, sdev.data(), sdev.size() * sizeof(SeedDev));
    up(b_cands, cd.data(), cd.size() * sizeof(CandDev));
    up(b_hdr, &hdr, sizeof hdr);
    up(b_geo, geo.data(), geo.size() * sizeof(vox::TriGeo));
    up(b_gran, gran.data(), gran.size() * sizeof(vox::GranuleInfo));
    up(b_tiles, tiles.data(), tiles.size() * sizeof(vox::Tile));
    up(b_ntiles, &ntiles, 4);
    up(b_origin, origin.data(), origin.size() * 8);
    up(b_rowpre, row_prefix.data(), row_prefix.size() * 8);
    up(b_foldpre, fold_prefix.data(), fold_prefix.size() * 4);
    if (st == S2H_OK) st = dev_alloc(ctx, b_zero, zero_words * 4);
    if (st == S2H_OK) st = dev_alloc(ctx, b_rows, std::max<uint64_t>(rows_total, 1) * 4);
    if (st == S2H_OK) st = dev_alloc(ctx, b_curve, curve_total * 9 + 16);
    if (st == S2H_OK) st = dev_alloc(ctx, b_over, (1 + 2 * (uint64_t)over_cap) * 4);
    if (st != S2H_OK) {
        release();
        return st;
    }
    int32_t *delta = (int32_t *)b_zero.ptr;
    int32_t *box = delta + delta_total;
    int *occ_delta = box + box_total;
    uint32_t *over = (uint32_t *)b_over.ptr;
    hipStream_t s = ctx->stream;
    prof_mark(ctx, 0);
    zero_kernel<<<2048, 256, 0, s>>>((uint32_t *)b_zero.ptr, zero_words);
    zero_kernel<<<1, 64, 0, s>>>(over, 1);
    batch_voxel_kernel<<<8192, 256, 0, s>>>((const vox::TriGeo *)b_geo.ptr, (const vox::GranuleInfo *)b_gran.ptr,
                                            (const vox::Tile *)b_tiles.ptr, (const uint32_t *)b_ntiles.ptr,
                                            (const double *)b_origin.ptr, (const CandDev *)b_cands.ptr, box, over,
                                            over_cap);
    batch_voxel_overflow_kernel<<<64, 64, 0, s>>>((const vox::TriGeo *)b_geo.ptr, (const vox::GranuleInfo *)b_gran.ptr,
                                                  (const vox::Tile *)b_tiles.ptr, (const double *)b_origin.ptr,
                                                  (const CandDev *)b_cands.ptr, box, over, over_cap);
    build_rows_kernel<<<4096, 256, 0, s>>>((const SeedDev *)b_seeds.ptr, (const CandDev *)b_cands.ptr, ncand,
                                           (const uint64_t *)b_rowpre.ptr, box, (uint32_t *)b_rows.ptr, occ_delta);
    delta_count_kernel<<<2048, 64 * kDeltaWaves, 0, s>>>((const SeedDev *)b_seeds.ptr, (const CandDev *)b_cands.ptr,
                                                         (const BatchHeader *)b_hdr.ptr, (const uint32_t *)b_rows.ptr,
                                                         delta);
    prof_mark(ctx, 1);
    double *d_curves = (double *)b_curve.ptr;
    uint8_t *d_support = (uint8_t *)(d_curves + curve_total);
    batch_fold_kernel<<<8192, 256, 0, s>>>((const SeedDev *)b_seeds.ptr, (const CandDev *)b_cands.ptr,
                                           (const BatchHeader *)b_hdr.ptr, (const uint32_t *)b_foldpre.ptr, delta,
                                           occ_delta, d_curves, d_support);
    prof_mark(ctx, 2);
    ctx->launches = 7;
    hipError_t e = hipGetLastError();
    std::vector<double> hc(curve_total);
    std::vector<uint8_t> hs(curve_total);
    std::vector<int> hocc(ncand);
    uint32_t nover = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(hc.data(), d_curves, curve_total * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(hs.data(), d_support, curve_total, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(hocc.data(), occ_delta, ncand * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&nover, over, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    prof_collect(ctx);
    release();
    if (e != hipSuccess) return fail(S2H_ERR_HIP, std::string("batch kernels failed: ") + hipGetErrorString(e));
    if (nover > over_cap) return fail(S2H_ERR_UNSUPPORTED, "more than 65536 candidate rays with over 16 hits");
    // ---- host: finish each curve (vf, fill, S2(0) = vf)
    for (uint32_t c = 0; c < ncand; ++c) {
        const s2h_seed *sd = seeds[cd[c].seed];
        const uint64_t n = sd->r_max + 1, off = cd[c].curve_off;
        const uint64_t cells = sd->vol->dims[0] * sd->vol->dims[1] * sd->vol->dims[2];
        const int64_t occ = (int64_t)sd->occupied + hocc[c];
        double *out = curves + off;
        if (occ == 0) {
            std::fill(out, out + n, 0.0);
            continue;
        }
        const double vf = (double)occ / (double)cells;
        std::memcpy(out, hc.data() + off, n * 8);
        std::vector<uint8_t> sup(hs.begin() + (long)off, hs.begin() + (long)(off + n));
        out[0] = vf;
        sup[0] = 1;
        fill_missing(out, sup.data(), n, vf);
        out[0] = vf;
    }
    return S2H_OK;
}

}  // extern "C"
