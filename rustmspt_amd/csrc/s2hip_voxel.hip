// Bit-exact voxelization of a granulated triangle mesh into the packed volume.
//
// Restates build_bbox_occupancy (src/geometry/s2.rs:160-201): for each granule
// (connected component), every voxel centre min + (i + 0.5)*pitch in the
// granule's clamped index range is tested with the granule's parity predicate
// (src/geometry/mesh_query.rs:73-116): outside the granule bbox +-1e-9 -> out;
// otherwise every forward Moller-Trumbore hit t (src/geometry/s2.rs:27-56) is
// collected, sorted, counted with the anchored 1e-8 rule, and odd parity means
// inside.  Granules are OR-ed.  Every f64 operation keeps the reference's
// order and this file is compiled with -ffp-contract=off, so results are
// bit-identical.
//
// Work shape: one workgroup per (granule, 4x8x8 voxel tile).  The workgroup
// first keeps the granule's triangles whose projection on the plane normal to
// the ray direction can reach the tile (a conservative 2-D bound whose margin
// dominates every rounding error of the predicate -- culled triangles provably
// miss, so the hit multiset is unchanged), then each thread runs the predicate
// for its voxel centre over that short list, reading triangle data through
// uniform (scalar) loads.  Rays with more than kMaxHits hits go to an overflow
// list that a second kernel resolves exactly with a repeated-minimum scan.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "s2hip_internal.h"

namespace s2h {

namespace {

constexpr double kRayX = 0.9428090415820634, kRayY = 0.2705980500730985, kRayZ = 0.19611613513818402;
constexpr int kTileX = 4, kTileY = 8, kTileZ = 8;  // 256 voxel centres per workgroup
constexpr int kMaxHits = 16;
constexpr int kListCap = 2048;  // culled triangle indices per pass

struct TriGeo {
    double ax, ay, az;     // vertex a
    double e1x, e1y, e1z;  // b - a
    double e2x, e2y, e2z;  // c - a
    double hx, hy, hz;     // dir x e2
    double inv;            // 1 / (e1 . h)
    double umin, umax, vmin, vmax;  // projected bound incl. margin (infinite: never culled)
};

struct GranuleInfo {
    double mn[3], mx[3];   // granule bbox (over referenced vertices)
    uint32_t tri_start, tri_end;
};

struct Tile {
    uint32_t granule;
    uint16_t x0, y0, z0;   // tile origin (voxel index)
    uint16_t x1, y1, z1;   // exclusive end, clipped to the granule range
};

// Projection basis normal to the ray direction: u = normalize(dir x e_z),
// v = dir x u (|dir . u|, |dir . v| < 1e-16; the cull margin covers it).
constexpr double kUx = 0.27587465517742293, kUy = -0.9611936197409646, kUz = 0.0;
constexpr double kVx = 0.18850557782307928, kVy = 0.05410347115597539, kVz = -0.9808731791583722;

// Moller-Trumbore in the reference's operation order on precomputed
// e1, e2, h = dir x e2 and inv = 1/det (identical values to computing them per
// test).  Returns t > 0 on a hit, -1 on a miss.
__device__ __forceinline__ double mt_hit(const TriGeo &g, double px, double py, double pz) {
    const double eps = 1e-10;
    const double sx = px - g.ax, sy = py - g.ay, sz = pz - g.az;
    const double u = g.inv * (sx * g.hx + sy * g.hy + sz * g.hz);
    if (!(u >= 0.0 - eps && u <= 1.0 + eps)) return -1.0;
    const double qx = sy * g.e1z - sz * g.e1y;
    const double qy = sz * g.e1x - sx * g.e1z;
    const double qz = sx * g.e1y - sy * g.e1x;
    const double v = g.inv * (kRayX * qx + kRayY * qy + kRayZ * qz);
    if (v < -eps || (u + v) > 1.0 + eps) return -1.0;
    const double t = g.inv * (g.e2x * qx + g.e2y * qy + g.e2z * qz);
    return t > eps ? t : -1.0;
}

__device__ __forceinline__ bool outside_bbox(const GranuleInfo &gi, double px, double py, double pz) {
    const double eps = 1e-9;
    return px < gi.mn[0] - eps || px > gi.mx[0] + eps || py < gi.mn[1] - eps || py > gi.mx[1] + eps ||
           pz < gi.mn[2] - eps || pz > gi.mx[2] + eps;
}

__device__ __forceinline__ void set_voxel(unsigned long long *words, uint64_t ny, uint64_t row_words, uint32_t x,
                                          uint32_t y, uint32_t z) {
    atomicOr(&words[((uint64_t)x * ny + y) * row_words + (z >> 6)], 1ull << (z & 63));
}

__global__ __launch_bounds__(256) void voxelize_kernel(const TriGeo *__restrict__ tris,
                                                       const GranuleInfo *__restrict__ gran,
                                                       const Tile *__restrict__ tiles, double bx, double by,
                                                       double bz, double pitch, uint64_t ny, uint64_t row_words,
                                                       unsigned long long *__restrict__ words,
                                                       uint32_t *__restrict__ overflow, uint32_t overflow_cap) {
    __shared__ uint32_t list[kListCap];
    __shared__ uint32_t list_n;
    const Tile tl = tiles[blockIdx.x];
    const GranuleInfo gi = gran[tl.granule];
    const int tx = threadIdx.x / (kTileY * kTileZ), ty = (threadIdx.x / kTileZ) % kTileY, tz = threadIdx.x % kTileZ;
    const uint32_t x = tl.x0 + tx, y = tl.y0 + ty, z = tl.z0 + tz;
    const bool active = x < tl.x1 && y < tl.y1 && z < tl.z1;
    // voxel centre, src/geometry/s2.rs:150-152
    const double px = bx + ((double)x + 0.5) * pitch;
    const double py = by + ((double)y + 0.5) * pitch;
    const double pz = bz + ((double)z + 0.5) * pitch;
    const bool live = active && !outside_bbox(gi, px, py, pz);
    const double pu = kUx * px + kUy * py + kUz * pz;
    const double pv = kVx * px + kVy * py + kVz * pz;
    // projected bound of the tile's centres (exact up to rounding; the per-
    // triangle margin covers it)
    double tu0 = live ? pu : INFINITY, tu1 = live ? pu : -INFINITY;
    double tv0 = live ? pv : INFINITY, tv1 = live ? pv : -INFINITY;
    for (int o = 32; o > 0; o >>= 1) {
        tu0 = fmin(tu0, __shfl_xor(tu0, o, 64));
        tu1 = fmax(tu1, __shfl_xor(tu1, o, 64));
        tv0 = fmin(tv0, __shfl_xor(tv0, o, 64));
        tv1 = fmax(tv1, __shfl_xor(tv1, o, 64));
    }
    __shared__ double wb[4][4];
    const int wid = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        wb[wid][0] = tu0;
        wb[wid][1] = tu1;
        wb[wid][2] = tv0;
        wb[wid][3] = tv1;
    }
    __syncthreads();
    tu0 = fmin(fmin(wb[0][0], wb[1][0]), fmin(wb[2][0], wb[3][0]));
    tu1 = fmax(fmax(wb[0][1], wb[1][1]), fmax(wb[2][1], wb[3][1]));
    tv0 = fmin(fmin(wb[0][2], wb[1][2]), fmin(wb[2][2], wb[3][2]));
    tv1 = fmax(fmax(wb[0][3], wb[1][3]), fmax(wb[2][3], wb[3][3]));
    if (!(tu0 <= tu1)) return;  // no live centre in this tile (uniform)

    double hits[kMaxHits];
    int nh = 0;
    bool over = false;
    for (uint32_t c0 = gi.tri_start; c0 < gi.tri_end; c0 += kListCap) {
        const uint32_t c1 = min(gi.tri_end, c0 + (uint32_t)kListCap);
        if (threadIdx.x == 0) list_n = 0;
        __syncthreads();
        for (uint32_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
            const TriGeo &g = tris[i];
            if (!(g.umax < tu0 || g.umin > tu1 || g.vmax < tv0 || g.vmin > tv1)) {
                const uint32_t k = atomicAdd(&list_n, 1u);
                list[k] = i;
            }
        }
        __syncthreads();
        const uint32_t n = list_n;
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t i = __builtin_amdgcn_readfirstlane(list[k]);
            const TriGeo g = tris[i];
            if (!live || pu < g.umin || pu > g.umax || pv < g.vmin || pv > g.vmax) continue;
            const double t = mt_hit(g, px, py, pz);
            if (t < 0.0) continue;
            if (nh >= kMaxHits) {
                over = true;
                continue;
            }
            // sorted insert with static register indices
            double cur = t;
#pragma unroll
            for (int s = 0; s < kMaxHits; ++s) {
                if (s < nh) {
                    const double lo = hits[s] < cur ? hits[s] : cur;
                    const double hi = hits[s] < cur ? cur : hits[s];
                    hits[s] = lo;
                    cur = hi;
                } else if (s == nh) {
                    hits[s] = cur;
                }
            }
            ++nh;
        }
        __syncthreads();
    }
    if (!live) return;
    if (over) {
        const uint32_t k = atomicAdd(&overflow[0], 1u);
        if (k < overflow_cap) {
            overflow[1 + 2 * k] = blockIdx.x;
            overflow[2 + 2 * k] = threadIdx.x;
        }
        return;
    }
    // anchored dedup, src/geometry/mesh_query.rs:105-114
    int unique = 0;
    double last = -INFINITY;
#pragma unroll
    for (int s = 0; s < kMaxHits; ++s) {
        if (s < nh && fabs(hits[s] - last) > 1e-8) {
            ++unique;
            last = hits[s];
        }
    }
    if (unique & 1) set_voxel(words, ny, row_words, x, y, z);
}

// Exact parity for rays with more than kMaxHits hits.  In the sorted hit list
// the anchored rule keeps t_0 = min, then repeatedly the smallest t with
// t - last > 1e-8; that is a repeated minimum over all triangles, needing no
// storage.  One thread per overflowing ray.
__global__ void voxelize_overflow_kernel(const TriGeo *__restrict__ tris, const GranuleInfo *__restrict__ gran,
                                         const Tile *__restrict__ tiles, double bx, double by, double bz,
                                         double pitch, uint64_t ny, uint64_t row_words,
                                         unsigned long long *__restrict__ words,
                                         const uint32_t *__restrict__ overflow, uint32_t count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const Tile tl = tiles[overflow[1 + 2 * k]];
    const uint32_t lt = overflow[2 + 2 * k];
    const GranuleInfo gi = gran[tl.granule];
    const int tx = lt / (kTileY * kTileZ), ty = (lt / kTileZ) % kTileY, tz = lt % kTileZ;
    const uint32_t x = tl.x0 + tx, y = tl.y0 + ty, z = tl.z0 + tz;
    const double px = bx + ((double)x + 0.5) * pitch;
    const double py = by + ((double)y + 0.5) * pitch;
    const double pz = bz + ((double)z + 0.5) * pitch;
    double last = -INFINITY;
    int unique = 0;
    while (true) {
        double best = INFINITY;
        bool found = false;
        for (uint32_t i = gi.tri_start; i < gi.tri_end; ++i) {
            const double t = mt_hit(tris[i], px, py, pz);
            if (t < 0.0) continue;
            if (t - last > 1e-8 && (!found || t < best)) {
                best = t;
                found = true;
            }
        }
        if (!found) break;
        ++unique;
        last = best;
    }
    if (unique & 1) set_voxel(words, ny, row_words, x, y, z);
}

// f64::min / f64::max as Rust defines them: a NaN operand is ignored.
inline double rmin(double a, double b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a < b ? a : b)); }
inline double rmax(double a, double b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a > b ? a : b)); }

// Rust `f as isize` (saturating, NaN -> 0).
inline int64_t f2i(double f) {
    if (std::isnan(f)) return 0;
    if (f >= 9.2233720368547758e18) return std::numeric_limits<int64_t>::max();
    if (f <= -9.2233720368547758e18) return std::numeric_limits<int64_t>::min();
    return (int64_t)f;
}

}  // namespace
}  // namespace s2h

using namespace s2h;

extern "C" s2h_status s2h_voxelize(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri,
                                   uint64_t ntris, const uint32_t *granule_of_tri, uint32_t ngranules,
                                   const double bbox_min[3], double pitch, const uint64_t dims[3], s2h_vol **out,
                                   uint64_t *occupied) {
    if (!ctx || !dims || !bbox_min || !out || (ntris && (!xyz || !tri || !granule_of_tri)))
        return fail(S2H_ERR_ARG, "null argument");
    if (!(pitch > 0.0) || !std::isfinite(pitch)) return fail(S2H_ERR_ARG, "pitch must be finite and > 0");
    S2H_HIP(hipSetDevice(ctx->device));
    for (uint64_t f = 0; f < ntris; ++f) {
        if (granule_of_tri[f] >= ngranules) return fail(S2H_ERR_ARG, "granule id out of range");
        for (int k = 0; k < 3; ++k)
            if (tri[3 * f + k] >= nverts) return fail(S2H_ERR_ARG, "vertex index out of range");
    }
    s2h_vol *vol = nullptr;
    S2H_TRY(vol_alloc(ctx, dims, &vol));
    auto cleanup = [&](s2h_status st) {
        s2h_vol_free(ctx, vol);
        return st;
    };
    if (hipMemsetAsync(vol->d_words, 0, vol->nwords * 8, ctx->stream) != hipSuccess)
        return cleanup(fail(S2H_ERR_HIP, "memset failed"));

    // ---- host preparation: triangles grouped by granule, per-granule bbox and range
    std::vector<uint32_t> start(ngranules + 1, 0);
    for (uint64_t f = 0; f < ntris; ++f) start[granule_of_tri[f] + 1]++;
    for (uint32_t g = 0; g < ngranules; ++g) start[g + 1] += start[g];
    std::vector<uint32_t> fill(start.begin(), start.end() - 1), order(ntris);
    for (uint64_t f = 0; f < ntris; ++f) order[fill[granule_of_tri[f]]++] = (uint32_t)f;

    std::vector<GranuleInfo> gran(ngranules);
    std::vector<Tile> tiles;
    std::vector<TriGeo> geo(ntris);
    const double dirx = kRayX, diry = kRayY, dirz = kRayZ;
    for (uint32_t g = 0; g < ngranules; ++g) {
        GranuleInfo &gi = gran[g];
        gi.tri_start = start[g];
        gi.tri_end = start[g + 1];
        bool have = false;
        double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0}, amax = 0.0;
        for (uint32_t k = start[g]; k < start[g + 1]; ++k)
            for (int c = 0; c < 3; ++c) {
                const double *v = xyz + 3 * (uint64_t)tri[3 * (uint64_t)order[k] + c];
                for (int a = 0; a < 3; ++a) {
                    mn[a] = have ? rmin(mn[a], v[a]) : v[a];
                    mx[a] = have ? rmax(mx[a], v[a]) : v[a];
                    amax = std::max(amax, std::fabs(v[a]));
                }
                have = true;
            }
        for (int a = 0; a < 3; ++a) {
            gi.mn[a] = mn[a];
            gi.mx[a] = mx[a];
        }
        // src/geometry/s2.rs:126-147 part_voxel_ranges
        int64_t lo[3], hi[3];
        bool empty = !have;
        for (int a = 0; a < 3 && !empty; ++a) {
            lo[a] = std::max<int64_t>(0, f2i(std::floor((mn[a] - bbox_min[a]) / pitch)));
            hi[a] = std::min<int64_t>((int64_t)dims[a], std::max<int64_t>(0, f2i(std::ceil((mx[a] - bbox_min[a]) / pitch))));
            if (lo[a] >= hi[a]) empty = true;
        }
        if (empty) continue;
        // largest |coordinate| of a centre this granule tests, and the largest
        // distance from a centre to a vertex (bounds for the cull margin)
        double cmax = 0.0;
        for (int a = 0; a < 3; ++a) {
            cmax = std::max(cmax, std::fabs(bbox_min[a] + ((double)lo[a] + 0.5) * pitch));
            cmax = std::max(cmax, std::fabs(bbox_min[a] + ((double)hi[a] + 0.5) * pitch));
        }
        double smax = 0.0;
        for (int a = 0; a < 3; ++a)
            smax += std::pow(std::max({std::fabs(mx[a] - (bbox_min[a] + (double)lo[a] * pitch)),
                                       std::fabs((bbox_min[a] + (double)(hi[a] + 1) * pitch) - mn[a]), pitch}),
                             2.0);
        smax = std::sqrt(smax);
        for (uint32_t k = start[g]; k < start[g + 1]; ++k) {
            const uint32_t f = order[k];
            const double *A = xyz + 3 * (uint64_t)tri[3 * (uint64_t)f];
            const double *B = xyz + 3 * (uint64_t)tri[3 * (uint64_t)f + 1];
            const double *C = xyz + 3 * (uint64_t)tri[3 * (uint64_t)f + 2];
            TriGeo &t = geo[k];
            t.ax = A[0];
            t.ay = A[1];
            t.az = A[2];
            t.e1x = B[0] - A[0];
            t.e1y = B[1] - A[1];
            t.e1z = B[2] - A[2];
            t.e2x = C[0] - A[0];
            t.e2y = C[1] - A[1];
            t.e2z = C[2] - A[2];
            t.hx = diry * t.e2z - dirz * t.e2y;
            t.hy = dirz * t.e2x - dirx * t.e2z;
            t.hz = dirx * t.e2y - diry * t.e2x;
            const double det = t.e1x * t.hx + t.e1y * t.hy + t.e1z * t.hz;
            t.inv = 1.0 / det;
            if (std::fabs(det) <= 1e-10 || std::isnan(det)) {
                // the reference rejects |det| <= eps before anything else; NaN det
                // reaches the comparisons, which then all fail -- never a hit
                t.umin = t.vmin = INFINITY;
                t.umax = t.vmax = -INFINITY;
                continue;
            }
            const double pa[3][3] = {{A[0], A[1], A[2]}, {B[0], B[1], B[2]}, {C[0], C[1], C[2]}};
            bool finite = std::isfinite(t.inv);
            double u0 = INFINITY, u1 = -INFINITY, v0 = INFINITY, v1 = -INFINITY, ext = 0.0;
            for (auto &p : pa) {
                for (double c : p) finite = finite && std::isfinite(c);
                const double u = kUx * p[0] + kUy * p[1] + kUz * p[2];
                const double v = kVx * p[0] + kVy * p[1] + kVz * p[2];
                u0 = std::min(u0, u);
                u1 = std::max(u1, u);
                v0 = std::min(v0, v);
                v1 = std::max(v1, v);
            }
            if (!finite || !std::isfinite(smax)) {
                t.umin = t.vmin = -INFINITY;
                t.umax = t.vmax = INFINITY;
                continue;
            }
            ext = std::max(u1 - u0, v1 - v0);
            const double len1 = std::sqrt(t.e1x * t.e1x + t.e1y * t.e1y + t.e1z * t.e1z);
            const double len2 = std::sqrt(t.e2x * t.e2x + t.e2y * t.e2y + t.e2z * t.e2z);
            const double L = std::max({len1, len2, ext, 1e-300});
            // barycentric slack the predicate can admit: eps plus a bound on the
            // rounding error of u and v (|s| |edge| / |det| * 16 ulps), times 4
            const double tau = 1e-10 + 16.0 * 1.1102230246251565e-16 * smax * std::max(len1, len2) * std::fabs(t.inv);
            const double margin = 4.0 * 3.0 * L * tau +
                                  64.0 * 1.1102230246251565e-16 * (cmax + amax + L + smax) + 1e-12;
            t.umin = u0 - margin;
            t.umax = u1 + margin;
            t.vmin = v0 - margin;
            t.vmax = v1 + margin;
        }
        for (int64_t x = lo[0]; x < hi[0]; x += kTileX)
            for (int64_t y = lo[1]; y < hi[1]; y += kTileY)
                for (int64_t z = lo[2]; z < hi[2]; z += kTileZ) {
                    Tile t;
                    t.granule = g;
                    t.x0 = (uint16_t)x;
                    t.y0 = (uint16_t)y;
                    t.z0 = (uint16_t)z;
                    t.x1 = (uint16_t)std::min<int64_t>(hi[0], x + kTileX);
                    t.y1 = (uint16_t)std::min<int64_t>(hi[1], y + kTileY);
                    t.z1 = (uint16_t)std::min<int64_t>(hi[2], z + kTileZ);
                    tiles.push_back(t);
                }
    }
    DevBuf d_geo, d_gran, d_tiles, d_over;
    auto release = [&]() {
        dev_free(ctx, d_geo);
        dev_free(ctx, d_gran);
        dev_free(ctx, d_tiles);
        dev_free(ctx, d_over);
    };
    s2h_status st = S2H_OK;
    const uint32_t over_cap = 1u << 16;
    if (!tiles.empty()) {
        if ((st = upload(ctx, d_geo, geo.data(), geo.size() * sizeof(TriGeo))) != S2H_OK ||
            (st = upload(ctx, d_gran, gran.data(), gran.size() * sizeof(GranuleInfo))) != S2H_OK ||
            (st = upload(ctx, d_tiles, tiles.data(), tiles.size() * sizeof(Tile))) != S2H_OK ||
            (st = dev_alloc(ctx, d_over, (1 + 2 * (uint64_t)over_cap) * 4)) != S2H_OK) {
            release();
            return cleanup(st);
        }
        hipError_t e = hipMemsetAsync(d_over.ptr, 0, 4, ctx->stream);
        for (size_t t0 = 0; e == hipSuccess && t0 < tiles.size(); t0 += (1u << 30)) {
            const size_t nt = std::min<size_t>(tiles.size() - t0, 1u << 30);
            voxelize_kernel<<<(unsigned)nt, 256, 0, ctx->stream>>>(
                (const TriGeo *)d_geo.ptr, (const GranuleInfo *)d_gran.ptr, (const Tile *)d_tiles.ptr + t0,
                bbox_min[0], bbox_min[1], bbox_min[2], pitch, dims[1], vol->row_words,
                (unsigned long long *)vol->d_words, (uint32_t *)d_over.ptr, over_cap);
            e = hipGetLastError();
        }
        uint32_t nover = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&nover, d_over.ptr, 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess && nover > over_cap) {
            release();
            return cleanup(fail(S2H_ERR_UNSUPPORTED, "more than 65536 voxel rays with over 16 hits"));
        }
        if (e == hipSuccess && nover) {
            voxelize_overflow_kernel<<<(nover + 63) / 64, 64, 0, ctx->stream>>>(
                (const TriGeo *)d_geo.ptr, (const GranuleInfo *)d_gran.ptr, (const Tile *)d_tiles.ptr,
                bbox_min[0], bbox_min[1], bbox_min[2], pitch, dims[1], vol->row_words,
                (unsigned long long *)vol->d_words, (const uint32_t *)d_over.ptr, nover);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        }
        if (e != hipSuccess) {
            release();
            return cleanup(fail(S2H_ERR_HIP, std::string("voxelize kernel failed: ") + hipGetErrorString(e)));
        }
    }
    release();
    uint64_t occ = 0;
    st = vol_popcount(ctx, vol, &occ);
    if (st != S2H_OK) return cleanup(st);
    if (occupied) *occupied = occ;
    *out = vol;
    return S2H_OK;
}
