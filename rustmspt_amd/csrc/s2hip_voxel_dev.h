// Device-side pieces of the bit-exact voxel predicate shared by the full
// voxelizer and the batched per-particle re-voxelization (see s2hip_voxel.hip
// for the algorithm and its reference citations).
#pragma once
#include "s2hip_internal.h"

#include <cmath>
#include <vector>

namespace s2h {
namespace vox {

constexpr double kRayX = 0.9428090415820634, kRayY = 0.2705980500730985, kRayZ = 0.19611613513818402;
constexpr int kTileX = 4, kTileY = 8, kTileZ = 8;  // 256 voxel centres per workgroup
constexpr int kMaxHits = 16;
constexpr int kListCap = 1024;  // culled triangle indices per pass

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
    uint32_t owner;        // batch path: candidate index (unused by the full voxelizer)
    int32_t sign;          // batch path: +1 new position, -1 old position
    // batch path, device-side triangle preparation: cull-margin bounds and the
    // granule's triangles (3 vertex indices each, into xyz)
    double cmax, smax, amax;
    const double *xyz;
    const uint32_t *tri;
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

__host__ __device__ inline bool f64_isnan(double x) { return x != x; }
__host__ __device__ inline bool f64_isfinite(double x) { return fabs(x) <= 1.7976931348623157e308; }

// Precomputed triangle data for the predicate: edges, h = dir x e2 and
// inv = 1/det in the reference's operation order (bit-identical to computing
// them per test), plus the triangle's projected bound on the plane normal to
// the ray, widened by a margin that dominates every rounding error of the
// predicate for centres within the granule's range.  cmax is the largest
// |coordinate| of such a centre, smax the largest centre-vertex distance and
// amax the largest |vertex coordinate| of the granule.  Shared by the host
// voxelizer and the batched device preparation.
__host__ __device__ inline void tri_prepare(const double *A, const double *B, const double *C, double cmax,
                                            double smax, double amax, TriGeo &t) {
    t.ax = A[0];
    t.ay = A[1];
    t.az = A[2];
    t.e1x = B[0] - A[0];
    t.e1y = B[1] - A[1];
    t.e1z = B[2] - A[2];
    t.e2x = C[0] - A[0];
    t.e2y = C[1] - A[1];
    t.e2z = C[2] - A[2];
    t.hx = kRayY * t.e2z - kRayZ * t.e2y;
    t.hy = kRayZ * t.e2x - kRayX * t.e2z;
    t.hz = kRayX * t.e2y - kRayY * t.e2x;
    const double det = t.e1x * t.hx + t.e1y * t.hy + t.e1z * t.hz;
    t.inv = 1.0 / det;
    if (fabs(det) <= 1e-10 || f64_isnan(det)) {
        // the reference rejects |det| <= eps before anything else; a NaN det
        // reaches comparisons that all fail -- never a hit
        t.umin = t.vmin = INFINITY;
        t.umax = t.vmax = -INFINITY;
        return;
    }
    const double *pa[3] = {A, B, C};
    bool finite = f64_isfinite(t.inv);
    double u0 = INFINITY, u1 = -INFINITY, v0 = INFINITY, v1 = -INFINITY;
    for (int k = 0; k < 3; ++k) {
        const double *p = pa[k];
        for (int c = 0; c < 3; ++c) finite = finite && f64_isfinite(p[c]);
        const double u = kUx * p[0] + kUy * p[1] + kUz * p[2];
        const double v = kVx * p[0] + kVy * p[1] + kVz * p[2];
        u0 = fmin(u0, u);
        u1 = fmax(u1, u);
        v0 = fmin(v0, v);
        v1 = fmax(v1, v);
    }
    if (!finite || !f64_isfinite(smax)) {
        t.umin = t.vmin = -INFINITY;
        t.umax = t.vmax = INFINITY;
        return;
    }
    const double ext = fmax(u1 - u0, v1 - v0);
    const double len1 = sqrt(t.e1x * t.e1x + t.e1y * t.e1y + t.e1z * t.e1z);
    const double len2 = sqrt(t.e2x * t.e2x + t.e2y * t.e2y + t.e2z * t.e2z);
    const double L = fmax(fmax(len1, len2), fmax(ext, 1e-300));
    // barycentric slack the predicate can admit: eps plus a bound on the
    // rounding error of u and v (|s| |edge| / |det| * 16 ulps), times 4
    const double tau = 1e-10 + 16.0 * 1.1102230246251565e-16 * smax * fmax(len1, len2) * fabs(t.inv);
    const double margin = 4.0 * 3.0 * L * tau + 64.0 * 1.1102230246251565e-16 * (cmax + amax + L + smax) + 1e-12;
    t.umin = u0 - margin;
    t.umax = u1 + margin;
    t.vmin = v0 - margin;
    t.vmax = v1 + margin;
}

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


// Local index window [lo, hi] (clamped to [0, n - 1]) of the lattice points
// q(i) = o + (base + i + 0.5) * pitch whose coordinate can lie in [a, b]: the
// real-valued indices of a and b widened by `slack` (far above the rounding
// error of every coordinate and projection here).  Infinite or NaN bounds give
// the whole range.
__device__ __forceinline__ void index_window(double a, double b, double o, double ipitch, uint32_t base, int n,
                                             double slack, int &lo, int &hi) {
    const double ia = (a - o) * ipitch - 0.5 - (double)base, ib = (b - o) * ipitch - 0.5 - (double)base;
    const double l = floor(fmin(ia, ib) - slack), h = ceil(fmax(ia, ib) + slack);
    lo = (int)fmax(0.0, fmin((double)(n - 1), l));   // fmin / fmax drop a NaN operand
    hi = (int)fmax(0.0, fmin((double)(n - 1), h));
    if (!(l <= h)) lo = 0, hi = n - 1;               // NaN: everything
}

// Sorted hits of one centre (N >= nh slots, unused ones +inf) with the
// anchored 1e-8 dedup (src/geometry/mesh_query.rs:105-114); returns the count.
template <int N>
__device__ __forceinline__ int unique_hits(const double *slot, uint32_t stride, uint32_t nh) {
    double h[N];
#pragma unroll
    for (int s = 0; s < N; ++s) h[s] = s < (int)nh ? slot[s * stride] : INFINITY;
#pragma unroll
    for (int i = 1; i < N; ++i)
#pragma unroll
        for (int j = i; j > 0; --j) {
            const double lo = h[j - 1] < h[j] ? h[j - 1] : h[j];
            const double hi = h[j - 1] < h[j] ? h[j] : h[j - 1];
            h[j - 1] = lo;
            h[j] = hi;
        }
    int unique = 0;
    double last = -INFINITY;
#pragma unroll
    for (int s = 0; s < N; ++s) {
        if (s < (int)nh && fabs(h[s] - last) > 1e-8) {
            ++unique;
            last = h[s];
        }
    }
    return unique;
}

// One workgroup (256 threads) evaluates the parity predicate at the centres of
// one tile; out(gi, x, y, z) receives every centre found inside.  Rays with more
// than kMaxHits hits are appended (tile_id, thread) to `overflow`.
//
// The tile's triangles are first culled to those whose projected bound (on the
// plane normal to the ray) meets the tile's centres.  Work items are then
// (triangle, x slice of the tile): a lane enumerates only the slice's centres
// whose projection can fall inside the triangle's bound (a y window from the u
// bound, then a z window from the v bound), applies the per-centre bound test
// and the Moller-Trumbore predicate to those, and appends each hit t to the
// centre's list in LDS.  A centre outside the windows fails the bound test,
// and the bound test is conservative (every hit passes it), so each centre
// collects the same hit multiset as testing every triangle; a triangle
// typically touches a few centres of a slice, so no lane runs the predicate
// for a centre it cannot hit.  Finally each thread sorts its centre's hits and
// counts them with the anchored rule.
template <class Out>
__device__ void voxelize_tile(const TriGeo *__restrict__ tris, const GranuleInfo *__restrict__ gran, const Tile tl,
                              uint32_t tile_id, double bx, double by, double bz, double pitch,
                              uint32_t *__restrict__ overflow, uint32_t overflow_cap, Out out) {
    constexpr int kCentres = kTileX * kTileY * kTileZ;
    __shared__ uint32_t list[kListCap];
    __shared__ uint32_t list_n;
    __shared__ double hit[kMaxHits * kCentres];  // [slot][centre]
    __shared__ uint32_t nhit[kCentres];
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

    nhit[threadIdx.x] = 0;
    const double ipitch = 1.0 / pitch;
    // index slack: 1e-9 of the largest coordinate magnitude in voxels, against
    // rounding errors of a few ulps of it
    const double slack = 1e-6 + 1e-9 * (fabs(bx) + fabs(by) + fabs(bz) + 3.0 * 65536.0 * pitch) * ipitch;
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
        const uint32_t items = list_n * kTileX;
        for (uint32_t it = threadIdx.x; it < items; it += blockDim.x) {
            const int sx = (int)(it % kTileX);
            const uint32_t qxi = tl.x0 + sx;
            if (qxi >= tl.x1) continue;
            const TriGeo g = tris[list[it / kTileX]];
            const double qx = bx + ((double)qxi + 0.5) * pitch;
            // u = kUx x + kUy y (kUz = 0): the y window of u in [umin, umax]
            int ya, yb;
            index_window((g.umin - kUx * qx) / kUy, (g.umax - kUx * qx) / kUy, by, ipitch, tl.y0, kTileY, slack, ya,
                         yb);
            for (int ly = ya; ly <= yb; ++ly) {
                const uint32_t qyi = tl.y0 + ly;
                if (qyi >= tl.y1) break;
                const double qy = by + ((double)qyi + 0.5) * pitch;
                int za, zb;
                const double rest = kVx * qx + kVy * qy;
                index_window((g.vmin - rest) / kVz, (g.vmax - rest) / kVz, bz, ipitch, tl.z0, kTileZ, slack, za, zb);
                for (int lz = za; lz <= zb; ++lz) {
                    const uint32_t qzi = tl.z0 + lz;
                    if (qzi >= tl.z1) break;
                    const double qz = bz + ((double)qzi + 0.5) * pitch;
                    if (outside_bbox(gi, qx, qy, qz)) continue;
                    const double qu = kUx * qx + kUy * qy + kUz * qz;
                    const double qv = kVx * qx + kVy * qy + kVz * qz;
                    if (qu < g.umin || qu > g.umax || qv < g.vmin || qv > g.vmax) continue;
                    const double t = mt_hit(g, qx, qy, qz);
                    if (t < 0.0) continue;
                    const int c = (sx * kTileY + ly) * kTileZ + lz;
                    const uint32_t slot = atomicAdd(&nhit[c], 1u);
                    if (slot < (uint32_t)kMaxHits) hit[slot * kCentres + c] = t;
                }
            }
        }
        __syncthreads();
    }
    if (!live) return;
    const uint32_t nh = nhit[threadIdx.x];
    if (nh > (uint32_t)kMaxHits) {
        const uint32_t k = atomicAdd(&overflow[0], 1u);
        if (k < overflow_cap) {
            overflow[1 + 2 * k] = tile_id;
            overflow[2 + 2 * k] = threadIdx.x;
        }
        return;
    }
    const double *slot = hit + threadIdx.x;
    const int unique = nh <= 4 ? unique_hits<4>(slot, kCentres, nh)
                               : (nh <= 8 ? unique_hits<8>(slot, kCentres, nh) : unique_hits<kMaxHits>(slot, kCentres, nh));
    if (unique & 1) out(gi, x, y, z);
}

// Exact parity for rays with more than kMaxHits hits.  In the sorted hit list
// the anchored rule keeps t_0 = min, then repeatedly the smallest t with
// t - last > 1e-8; that is a repeated minimum over all triangles, needing no
// storage.  One thread per overflowing ray.
template <class Out>
__device__ void voxel_overflow_ray(const TriGeo *__restrict__ tris, const GranuleInfo *__restrict__ gran,
                                   const Tile *__restrict__ tiles, double bx, double by, double bz, double pitch,
                                   const uint32_t *__restrict__ overflow, uint32_t k, Out out) {
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
    if (unique & 1) out(gi, x, y, z);
}


struct SetBit {
    unsigned long long *words;
    uint64_t ny, row_words;
    __device__ void operator()(const GranuleInfo &, uint32_t x, uint32_t y, uint32_t z) const {
        set_voxel(words, ny, row_words, x, y, z);
    }
};

// Rust `f as isize` (saturating, NaN -> 0).
__host__ __device__ inline int64_t f2i_sat(double f) {
    if (f64_isnan(f)) return 0;
    if (f >= 9.2233720368547758e18) return INT64_MAX;
    if (f <= -9.2233720368547758e18) return INT64_MIN;
    return (int64_t)f;
}
// Rust's f64::min / f64::max (a NaN operand is ignored).
__host__ __device__ inline double rmin(double a, double b) { return f64_isnan(a) ? b : (f64_isnan(b) ? a : (a < b ? a : b)); }
__host__ __device__ inline double rmax(double a, double b) { return f64_isnan(a) ? b : (f64_isnan(b) ? a : (a > b ? a : b)); }
// std::max on doubles: (a < b) ? b : a.
__host__ __device__ inline double smax2(double a, double b) { return a < b ? b : a; }

// Voxel index range [lo, hi) of a granule with vertex bbox [mn, mx]
// (part_voxel_ranges, src/geometry/s2.rs:126-147) and the cull-margin bounds
// cmax (largest |coordinate| of a centre the granule tests) and smax (largest
// centre-vertex distance) of tri_prepare.  Returns false when the range is
// empty.  Shared by the host preparation and the device mesh ingest.
__host__ __device__ inline bool granule_bounds(const double mn[3], const double mx[3], const double bbox_min[3],
                                               double pitch, const uint64_t dims[3], int64_t lo[3], int64_t hi[3],
                                               double *cmax, double *smax) {
    for (int a = 0; a < 3; ++a) {
        const int64_t l = f2i_sat(floor((mn[a] - bbox_min[a]) / pitch));
        const int64_t h = f2i_sat(ceil((mx[a] - bbox_min[a]) / pitch));
        lo[a] = l > 0 ? l : 0;
        const int64_t hp = h > 0 ? h : 0;
        hi[a] = hp < (int64_t)dims[a] ? hp : (int64_t)dims[a];
        if (lo[a] >= hi[a]) return false;
    }
    double c = 0.0;
    for (int a = 0; a < 3; ++a) {
        c = smax2(c, fabs(bbox_min[a] + ((double)lo[a] + 0.5) * pitch));
        c = smax2(c, fabs(bbox_min[a] + ((double)hi[a] + 0.5) * pitch));
    }
    double s = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double d = smax2(smax2(fabs(mx[a] - (bbox_min[a] + (double)lo[a] * pitch)),
                                     fabs((bbox_min[a] + (double)(hi[a] + 1) * pitch) - mn[a])),
                               pitch);
        s += d * d;
    }
    *cmax = c;
    *smax = sqrt(s);
    return true;
}

// Launch the voxelize kernels (and the exact overflow pass) for prepared
// triangles, granules and tiles on ctx->stream, OR-ing into vol; waits for
// completion.  Shared by s2h_voxelize and the resident-mesh voxelizer.
s2h_status run_voxelize(s2h_ctx *ctx, const TriGeo *d_geo, const GranuleInfo *d_gran, const Tile *d_tiles,
                        uint64_t ntiles, const double bbox_min[3], double pitch, s2h_vol *vol);

// Host-side preparation of one granule's triangles (precomputed edges, 1/det and
// the conservative projected bound) and its voxel range.  Returns false when the
// granule's range is empty.
bool prepare_granule(const double *tri_xyz /* 9 per triangle */, uint32_t ntri, const double bbox_min[3],
                     double pitch, const uint64_t dims[3], TriGeo *geo, GranuleInfo &gi, int64_t lo[3],
                     int64_t hi[3]);
// Tiles of kTileX x kTileY x kTileZ centres covering [lo, hi) for granule g.
void push_tiles(uint32_t g, const int64_t lo[3], const int64_t hi[3], std::vector<Tile> &tiles);

}  // namespace vox
}  // namespace s2h
