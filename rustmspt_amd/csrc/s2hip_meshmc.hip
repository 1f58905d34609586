// Continuous mesh Monte Carlo S2 (no voxelization).
//
// Estimator of calculate_s2_mesh_mc_seeded_with_vf (src/geometry/s2.rs:1274-1320):
// for each radius r = 1..r_max and each of max(samples, 200) samples, a point p
// uniform in the measurement box and a direction uniform on the sphere (by
// rejection in the cube: accept 1e-12 < |d|^2 <= 1, then normalize); q = p + r d.
// A q outside the box is rejected; otherwise valid += 1, and hits += 1 when both
// p and q are inside the mesh.  Inside means the whole-mesh parity query of
// PreparedMeshQuery::contains_point (src/geometry/mesh_query.rs:73-116): outside
// the mesh's vertex bbox +-1e-9 -> out; otherwise every forward hit t of the ray
// along RAY_DIR_GPU (the predicate of src/geometry/s2.rs:27-56), sorted, counted
// with the anchored 1e-8 rule, odd -> in.
//
// Stream (the reference's ChaCha12 streams cannot be reproduced; the estimand
// is the same): sample i of radius r draws Philox4x32-10 with key
// {seed_lo, seed_hi} and counter {i_lo, i_hi, r, 0x4D455348 + k} for call k.
// Calls 0-1 give p (53-bit uniforms from word pairs (0,1), (2,3) of call 0 and
// (0,1) of call 1, each u = ((a >> 5) * 2^26 + (b >> 6)) / 2^53, coordinate
// min + (max - min) u).  Direction attempt a uses calls 2+2a and 3+2a the same
// way for x, y, z in [-1, 1); after 64 rejected attempts the direction is +x.
//
// Acceleration: every query ray has the same direction, so a triangle can only
// be hit when the query point's projection on the plane normal to the ray lies
// in the triangle's projected bound (widened by a margin that dominates every
// rounding error of the predicate -- the voxelizer's conservative cull).  The
// triangles are prepared and binned by that bound on a 2-D grid over the plane
// on the device (count per cell, prefix sum, scatter); a query tests only its
// cell's list, so the hit multiset equals a full scan.
//
// Schedule: the queries are answered in cell order.  A first kernel draws
// every sample's p and q (and counts the valid ones; the later kernels draw a
// sample again from its index rather than storing it), a radix sort orders the
// p queries by grid cell, a second kernel answers them, a second sort orders
// the q queries of the samples whose p is inside, and a third kernel answers
// those and counts the hits.  Neighbouring lanes then read the same cell lists
// and triangles, which stay in the caches; answered in sample order, every
// query was a fresh gather from memory.  hits and valid are sums over samples,
// so the order changes no count.
// Compiled with -ffp-contract=off: every f64 operation keeps its order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "s2hip_internal.h"
#include "s2hip_voxel_dev.h"

namespace s2h {
namespace meshmc {

constexpr uint32_t kTag = 0x4D455348u;
constexpr int kDirAttempts = 64;
constexpr int kThreads = 256;
constexpr int kPerThread = 8;
constexpr uint64_t kPassSamples = 16ull << 20;  // samples per pass of the sorted schedule

struct Grid {
    const vox::TriGeo *tris;
    const uint32_t *start;   // gu * gv + 1
    const uint32_t *ids;     // triangle index per cell entry
    double u0, v0, inv_du, inv_dv;
    int gu, gv;
    int have_bbox;
    double mn[3], mx[3];     // whole-mesh vertex bbox
};

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Exact parity for a ray with more than kMaxHits hits: repeated minimum over
// the cell's list (the anchored rule keeps t_0 = min, then the smallest t with
// t - last > 1e-8).
__device__ int overflow_count(const Grid &g, uint32_t c0, uint32_t c1, double px, double py, double pz) {
    double last = -INFINITY;
    int unique = 0;
    while (true) {
        double best = INFINITY;
        bool found = false;
        for (uint32_t k = c0; k < c1; ++k) {
            const double t = vox::mt_hit(g.tris[g.ids[k]], px, py, pz);
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
    return unique;
}

// The key past every cell (sorts after all of them): no query.
__device__ __forceinline__ uint32_t no_cell(const Grid &g) { return (uint32_t)g.gu * (uint32_t)g.gv; }

// The grid cell a query at p reads, or no_cell when p lies outside the mesh's
// vertex bbox +-1e-9 (then it is outside the mesh).
__device__ uint32_t query_cell(const Grid &g, double px, double py, double pz) {
    if (!g.have_bbox) return no_cell(g);
    const double eps = 1e-9;
    if (px < g.mn[0] - eps || px > g.mx[0] + eps || py < g.mn[1] - eps || py > g.mx[1] + eps ||
        pz < g.mn[2] - eps || pz > g.mx[2] + eps)
        return no_cell(g);
    const double pu = vox::kUx * px + vox::kUy * py + vox::kUz * pz;
    const double pv = vox::kVx * px + vox::kVy * py + vox::kVz * pz;
    int iu = (int)floor((pu - g.u0) * g.inv_du), iv = (int)floor((pv - g.v0) * g.inv_dv);
    iu = min(max(iu, 0), g.gu - 1);
    iv = min(max(iv, 0), g.gv - 1);
    return (uint32_t)iu * g.gv + iv;
}

// The parity query of a point whose cell is `cell` (from query_cell).
__device__ bool contains_in_cell(const Grid &g, uint32_t cell, double px, double py, double pz) {
    const double pu = vox::kUx * px + vox::kUy * py + vox::kUz * pz;
    const double pv = vox::kVx * px + vox::kVy * py + vox::kVz * pz;
    const uint32_t c0 = g.start[cell], c1 = g.start[cell + 1];
    double hits[vox::kMaxHits];
    int nh = 0;
    bool over = false;
    // the list walk is a chain of dependent loads (index, then bound): the
    // index two entries ahead and the bound one entry ahead load while this
    // entry is tested
    uint32_t id1 = c0 < c1 ? g.ids[c0] : 0u, id2 = c0 + 1 < c1 ? g.ids[c0 + 1] : 0u;
    double b1[4] = {0.0, 0.0, 0.0, 0.0};
    if (c0 < c1) {
        const vox::TriGeo &t1 = g.tris[id1];
        b1[0] = t1.umin;
        b1[1] = t1.umax;
        b1[2] = t1.vmin;
        b1[3] = t1.vmax;
    }
    for (uint32_t k = c0; k < c1; ++k) {
        const uint32_t id = id1;
        const double umin = b1[0], umax = b1[1], vmin = b1[2], vmax = b1[3];
        id1 = id2;
        if (k + 2 < c1) id2 = g.ids[k + 2];
        if (k + 1 < c1) {
            const vox::TriGeo &t1 = g.tris[id1];
            b1[0] = t1.umin;
            b1[1] = t1.umax;
            b1[2] = t1.vmin;
            b1[3] = t1.vmax;
        }
        if (pu < umin || pu > umax || pv < vmin || pv > vmax) continue;
        const vox::TriGeo &t = g.tris[id];
        const double th = vox::mt_hit(t, px, py, pz);
        if (th < 0.0) continue;
        if (nh >= vox::kMaxHits) {
            over = true;
            break;
        }
        double cur = th;
#pragma unroll
        for (int s = 0; s < vox::kMaxHits; ++s) {
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
    int unique = 0;
    if (over) {
        unique = overflow_count(g, c0, c1, px, py, pz);
    } else {
        double last = -INFINITY;
#pragma unroll
        for (int s = 0; s < vox::kMaxHits; ++s) {
            if (s < nh && fabs(hits[s] - last) > 1e-8) {
                ++unique;
                last = hits[s];
            }
        }
    }
    return unique & 1;
}

// The sample stream of one pass: sample sidx (0 .. radii x samples) is sample
// i = sidx % samples of radius r0 + sidx / samples.  p and q are functions of
// (seed, r, i) alone, so every kernel of the schedule draws them again from the
// sample's index instead of storing them.
struct McStream {
    double b0[3], b1[3];  // measurement box
    uint32_t samples;     // per radius
    uint32_t r0;
    uint32_t seed_lo, seed_hi;
};

__device__ __forceinline__ void draw_p(const McStream &s, uint32_t r, uint32_t i, double &px, double &py,
                                       double &pz) {
    uint32_t a[4], b[4];
    philox4x32_10(i, 0u, r, kTag + 0u, s.seed_lo, s.seed_hi, a);
    philox4x32_10(i, 0u, r, kTag + 1u, s.seed_lo, s.seed_hi, b);
    px = s.b0[0] + (s.b1[0] - s.b0[0]) * u53(a[0], a[1]);
    py = s.b0[1] + (s.b1[1] - s.b0[1]) * u53(a[2], a[3]);
    pz = s.b0[2] + (s.b1[2] - s.b0[2]) * u53(b[0], b[1]);
}

// q = p + r d for the sample's direction d (rejection in the cube; +x after
// kDirAttempts rejected attempts).
__device__ __forceinline__ void draw_q(const McStream &s, uint32_t r, uint32_t i, double px, double py, double pz,
                                       double &qx, double &qy, double &qz) {
    double dx = 1.0, dy = 0.0, dz = 0.0;
    for (int t = 0; t < kDirAttempts; ++t) {
        uint32_t a[4], b[4];
        philox4x32_10(i, 0u, r, kTag + 2u + 2u * (uint32_t)t, s.seed_lo, s.seed_hi, a);
        philox4x32_10(i, 0u, r, kTag + 3u + 2u * (uint32_t)t, s.seed_lo, s.seed_hi, b);
        const double x = -1.0 + 2.0 * u53(a[0], a[1]);
        const double y = -1.0 + 2.0 * u53(a[2], a[3]);
        const double z = -1.0 + 2.0 * u53(b[0], b[1]);
        const double n2 = x * x + y * y + z * z;
        if (n2 > 1e-12 && n2 <= 1.0) {
            const double inv = 1.0 / sqrt(n2);
            dx = x * inv;
            dy = y * inv;
            dz = z * inv;
            break;
        }
    }
    const double rr = (double)r;
    qx = px + dx * rr;
    qy = py + dy * rr;
    qz = pz + dz * rr;
}

__device__ __forceinline__ bool in_box(const McStream &s, double x, double y, double z) {
    return !(x < s.b0[0] || x > s.b1[0] || y < s.b0[1] || y > s.b1[1] || z < s.b0[2] || z > s.b1[2]);
}

// 1. Every sample of radii r0 .. of the pass: a valid sample (q inside the
// box) gets the cell of its p query as key (no_cell: no query, or p outside
// the mesh bbox), an invalid one no_cell; every sample carries its index.
// Counts the valid samples per radius.
__global__ __launch_bounds__(kThreads) void mc_points_kernel(Grid g, McStream s, uint32_t blocks_per_r,
                                                             uint32_t *__restrict__ key, uint32_t *__restrict__ idx,
                                                             unsigned long long *__restrict__ hv) {
    const uint32_t rl = blockIdx.x / blocks_per_r;  // radius within this pass
    const uint32_t r = s.r0 + rl;
    const uint32_t blk = blockIdx.x % blocks_per_r;
    uint64_t valid = 0;
    const uint32_t base = blk * kThreads * kPerThread + threadIdx.x;
    for (int k = 0; k < kPerThread; ++k) {
        const uint32_t i = base + (uint32_t)k * kThreads;
        if (i >= s.samples) break;
        const uint32_t sidx = rl * s.samples + i;
        double px, py, pz, qx, qy, qz;
        draw_p(s, r, i, px, py, pz);
        draw_q(s, r, i, px, py, pz, qx, qy, qz);
        const bool ok = in_box(s, qx, qy, qz);
        if (ok) ++valid;
        key[sidx] = ok ? query_cell(g, px, py, pz) : no_cell(g);
        idx[sidx] = sidx;
    }
    for (int o = 32; o > 0; o >>= 1) valid += __shfl_xor(valid, o, 64);
    if ((threadIdx.x & 63) == 0 && valid) atomicAdd(&hv[2 * r + 1], (unsigned long long)valid);
}

// 2. The p queries in cell order; a sample whose p is inside gets the cell of
// its q query as key for the next sort (no_cell otherwise).
__global__ __launch_bounds__(kThreads) void mc_query_p_kernel(Grid g, McStream s, uint32_t n,
                                                              const uint32_t *__restrict__ key,
                                                              const uint32_t *__restrict__ idx,
                                                              uint32_t *__restrict__ qkey,
                                                              uint32_t *__restrict__ qidx) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t c = key[j], sidx = idx[j];
        uint32_t k2 = no_cell(g);
        if (c != no_cell(g)) {
            const uint32_t rl = sidx / s.samples, i = sidx - rl * s.samples, r = s.r0 + rl;
            double px, py, pz;
            draw_p(s, r, i, px, py, pz);
            if (contains_in_cell(g, c, px, py, pz)) {
                double qx, qy, qz;
                draw_q(s, r, i, px, py, pz, qx, qy, qz);
                k2 = query_cell(g, qx, qy, qz);
            }
        }
        qkey[j] = k2;
        qidx[j] = sidx;
    }
}

// 3. The q queries in cell order: a hit when q is inside too.
__global__ __launch_bounds__(kThreads) void mc_query_q_kernel(Grid g, McStream s, uint32_t n,
                                                              const uint32_t *__restrict__ key,
                                                              const uint32_t *__restrict__ idx,
                                                              unsigned long long *__restrict__ hv) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t c = key[j];
        if (c == no_cell(g)) continue;
        const uint32_t sidx = idx[j];
        const uint32_t rl = sidx / s.samples, i = sidx - rl * s.samples, r = s.r0 + rl;
        double px, py, pz, qx, qy, qz;
        draw_p(s, r, i, px, py, pz);
        draw_q(s, r, i, px, py, pz, qx, qy, qz);
        if (contains_in_cell(g, c, qx, qy, qz)) atomicAdd(&hv[2 * r], 1ull);
    }
}

__global__ void clear_u64_kernel(unsigned long long *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void fill_u32_kernel(uint32_t *__restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Cells of a G-cell axis whose span meets [lo, hi]; empty when the bound misses
// the grid (c0 > c1).
__device__ __forceinline__ void cell_range(double lo, double hi, double o, double inv, int G, int &c0, int &c1) {
    if (!(hi >= lo)) {
        c0 = 1;
        c1 = 0;
        return;
    }
    const double a = (lo - o) * inv, b = (hi - o) * inv;
    c0 = a <= 0.0 ? 0 : (a >= G ? G : (int)floor(a));
    c1 = b < 0.0 ? -1 : (b >= G - 1 ? G - 1 : (int)floor(b));
}

// Per triangle: the predicate's precomputed data and conservative projected
// bound (tri_prepare), its cell range, and one count per covered cell.
__global__ void bin_prepare_kernel(const double *__restrict__ xyz, const uint32_t *__restrict__ tri, uint64_t ntris,
                                   double cmax, double smax, double amax, double u0, double v0, double inv_du,
                                   double inv_dv, int G, vox::TriGeo *__restrict__ geo, int4 *__restrict__ range,
                                   uint32_t *count) {
    for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < ntris; f += (uint64_t)gridDim.x * blockDim.x) {
        vox::TriGeo t;
        vox::tri_prepare(xyz + 3 * (uint64_t)tri[3 * f], xyz + 3 * (uint64_t)tri[3 * f + 1],
                         xyz + 3 * (uint64_t)tri[3 * f + 2], cmax, smax, amax, t);
        geo[f] = t;
        int a0, a1, b0, b1;
        cell_range(t.umin, t.umax, u0, inv_du, G, a0, a1);
        cell_range(t.vmin, t.vmax, v0, inv_dv, G, b0, b1);
        range[f] = make_int4(a0, a1, b0, b1);
        for (int i = a0; i <= a1; ++i)
            for (int j = b0; j <= b1; ++j) atomicAdd(&count[(uint64_t)i * G + j], 1u);
    }
}

// Scatter triangle indices into their cells' lists.  The order inside a list
// does not matter: a query sorts its hits, so its result is the same for every
// order.
__global__ void bin_fill_kernel(const int4 *__restrict__ range, uint64_t ntris, int G,
                                const uint32_t *__restrict__ start, uint32_t *cursor, uint32_t *__restrict__ ids) {
    for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < ntris; f += (uint64_t)gridDim.x * blockDim.x) {
        const int4 rg = range[f];
        for (int i = rg.x; i <= rg.y; ++i)
            for (int j = rg.z; j <= rg.w; ++j) {
                const uint64_t cell = (uint64_t)i * G + j;
                ids[start[cell] + atomicAdd(&cursor[cell], 1u)] = (uint32_t)f;
            }
    }
}

// Grid geometry and cull-margin bounds for a mesh with vertex bbox [mn, mx]
// (have = false: no vertices) and largest |coordinate| amax, queried inside
// [bmin, bmax].
void setup_grid(const double bmin[3], const double bmax[3], bool have, const double mn[3], const double mx[3],
                uint64_t ntris, Grid &g, double &cmax, double &smax) {
    g.have_bbox = have;
    for (int a = 0; a < 3; ++a) {
        g.mn[a] = have ? mn[a] : 0.0;
        g.mx[a] = have ? mx[a] : 0.0;
    }
    // any query point lies in the measurement box
    cmax = 0.0;
    double span2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        cmax = std::max({cmax, std::fabs(bmin[a]), std::fabs(bmax[a])});
        const double lo = std::min(bmin[a], have ? mn[a] : bmin[a]);
        const double hi = std::max(bmax[a], have ? mx[a] : bmax[a]);
        span2 += (hi - lo) * (hi - lo);
    }
    smax = std::sqrt(span2) + 1.0;
    // grid over the projection of the measurement box
    double u0 = INFINITY, u1 = -INFINITY, v0 = INFINITY, v1 = -INFINITY;
    for (int c = 0; c < 8; ++c) {
        const double x = (c & 1) ? bmax[0] : bmin[0], y = (c & 2) ? bmax[1] : bmin[1], z = (c & 4) ? bmax[2] : bmin[2];
        const double u = vox::kUx * x + vox::kUy * y + vox::kUz * z, v = vox::kVx * x + vox::kVy * y + vox::kVz * z;
        u0 = std::min(u0, u);
        u1 = std::max(u1, u);
        v0 = std::min(v0, v);
        v1 = std::max(v1, v);
    }
    const double pad = 1e-6 * std::max(1.0, std::max(u1 - u0, v1 - v0));
    u0 -= pad;
    u1 += pad;
    v0 -= pad;
    v1 += pad;
    const int G = (int)std::min<double>(1024.0, std::max(8.0, std::sqrt((double)ntris / 2.0)));
    g.gu = g.gv = G;
    g.u0 = u0;
    g.v0 = v0;
    g.inv_du = G / std::max(u1 - u0, 1e-300);
    g.inv_dv = G / std::max(v1 - v0, 1e-300);
}

unsigned grid_for(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 1u << 20)); }

// Bin the resident triangles on the device and run the estimator.
s2h_status bin_and_run(s2h_ctx *ctx, Grid g, double cmax, double smax, double amax, const double *d_xyz,
                       const uint32_t *d_tri, uint64_t ntris, const double bbox_min[3], const double bbox_max[3],
                       uint32_t r_max, uint64_t samples, uint64_t seed, uint64_t *hits_valid, uint64_t n) {
    const uint64_t spr = std::max<uint64_t>(samples, 200);
    const uint64_t per_block = (uint64_t)kThreads * kPerThread;
    const uint64_t bpr = (spr + per_block - 1) / per_block;
    if (bpr * r_max >= (1ull << 31)) return fail(S2H_ERR_UNSUPPORTED, "too many samples");
    const uint64_t cells = (uint64_t)g.gu * g.gv;
    DevBuf d_geo, d_range, d_count, d_start, d_cursor, d_ids, d_hv;
    auto release = [&]() {
        for (DevBuf *b : {&d_geo, &d_range, &d_count, &d_start, &d_cursor, &d_ids, &d_hv}) dev_free(ctx, *b);
    };
    s2h_status st = dev_alloc(ctx, d_geo, std::max<uint64_t>(ntris, 1) * sizeof(vox::TriGeo));
    if (st == S2H_OK) st = dev_alloc(ctx, d_range, std::max<uint64_t>(ntris, 1) * sizeof(int4));
    if (st == S2H_OK) st = dev_alloc(ctx, d_count, cells * 4);
    if (st == S2H_OK) st = dev_alloc(ctx, d_start, (cells + 1) * 4);
    if (st == S2H_OK) st = dev_alloc(ctx, d_cursor, cells * 4);
    if (st == S2H_OK) st = dev_alloc(ctx, d_hv, n * 8);
    if (st != S2H_OK) {
        release();
        return st;
    }
    prof_mark(ctx, 0);
    auto *count = (uint32_t *)d_count.ptr, *start = (uint32_t *)d_start.ptr, *cursor = (uint32_t *)d_cursor.ptr;
    fill_u32_kernel<<<grid_for(cells), 256, 0, ctx->stream>>>(count, cells);
    fill_u32_kernel<<<grid_for(cells), 256, 0, ctx->stream>>>(cursor, cells);
    if (ntris)
        bin_prepare_kernel<<<grid_for(ntris), 256, 0, ctx->stream>>>(
            d_xyz, d_tri, ntris, cmax, smax, amax, g.u0, g.v0, g.inv_du, g.inv_dv, g.gu, (vox::TriGeo *)d_geo.ptr,
            (int4 *)d_range.ptr, count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        release();
        return fail(S2H_ERR_HIP, std::string("mesh MC binning: ") + hipGetErrorString(e));
    }
    uint64_t total = 0;
    st = exclusive_sum(ctx, count, start, cells, &total);
    if (st == S2H_OK && total >= (1ull << 31)) st = fail(S2H_ERR_UNSUPPORTED, "mesh too large for the query grid");
    if (st == S2H_OK) st = dev_alloc(ctx, d_ids, std::max<uint64_t>(total, 1) * 4);
    if (st != S2H_OK) {
        release();
        return st;
    }
    uint32_t tot32 = (uint32_t)total;
    e = hipMemcpyAsync(start + cells, &tot32, 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && ntris)
        bin_fill_kernel<<<grid_for(ntris), 256, 0, ctx->stream>>>((const int4 *)d_range.ptr, ntris, g.gu, start,
                                                                   cursor, (uint32_t *)d_ids.ptr);
    if (e == hipSuccess) e = hipGetLastError();
    g.tris = (const vox::TriGeo *)d_geo.ptr;
    g.start = start;
    g.ids = (const uint32_t *)d_ids.ptr;
    prof_mark(ctx, 1);
    if (e == hipSuccess) {
        clear_u64_kernel<<<1, 256, 0, ctx->stream>>>((unsigned long long *)d_hv.ptr, n);
        e = hipGetLastError();
    }
    // radii in passes of at most kPassSamples samples (keys and indices: 16 bytes each)
    const uint64_t rpp = std::max<uint64_t>(1, std::min<uint64_t>(r_max, kPassSamples / spr));
    const uint64_t cap = rpp * spr;
    if (cap >= (1ull << 31)) e = hipErrorInvalidValue;
    int bits = 1;
    while (bits < 32 && (1ull << bits) <= cells) ++bits;  // keys 0 .. cells (cells: no query)
    size_t sort_bytes = 0;
    DevBuf d_k0, d_i0, d_k1, d_i1, d_sort;
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                               (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)cap, 0, bits,
                                               ctx->stream);
    if (e == hipSuccess) {
        st = dev_alloc(ctx, d_k0, cap * 4);
        if (st == S2H_OK) st = dev_alloc(ctx, d_i0, cap * 4);
        if (st == S2H_OK) st = dev_alloc(ctx, d_k1, cap * 4);
        if (st == S2H_OK) st = dev_alloc(ctx, d_i1, cap * 4);
        if (st == S2H_OK) st = dev_alloc(ctx, d_sort, std::max<size_t>(sort_bytes, 16));
    }
    auto release_pass = [&]() {
        for (DevBuf *b : {&d_k0, &d_i0, &d_k1, &d_i1, &d_sort}) dev_free(ctx, *b);
    };
    if (st != S2H_OK) {
        release_pass();
        release();
        return st;
    }
    auto *k0 = (uint32_t *)d_k0.ptr, *i0 = (uint32_t *)d_i0.ptr, *k1 = (uint32_t *)d_k1.ptr, *i1 = (uint32_t *)d_i1.ptr;
    auto *hv = (unsigned long long *)d_hv.ptr;
    int passes = 0;
    for (uint64_t r0 = 1; e == hipSuccess && r0 <= r_max; r0 += rpp) {
        const uint64_t nr = std::min<uint64_t>(rpp, r_max - r0 + 1), T = nr * spr;
        McStream ms;
        for (int a = 0; a < 3; ++a) {
            ms.b0[a] = bbox_min[a];
            ms.b1[a] = bbox_max[a];
        }
        ms.samples = (uint32_t)spr;
        ms.r0 = (uint32_t)r0;
        ms.seed_lo = (uint32_t)seed;
        ms.seed_hi = (uint32_t)(seed >> 32);
        mc_points_kernel<<<(unsigned)(bpr * nr), kThreads, 0, ctx->stream>>>(g, ms, (uint32_t)bpr, k0, i0, hv);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortPairs(d_sort.ptr, sort_bytes, k0, k1, i0, i1, (int)T, 0, bits,
                                                   ctx->stream);
        if (e == hipSuccess) {
            mc_query_p_kernel<<<grid_for(T), kThreads, 0, ctx->stream>>>(g, ms, (uint32_t)T, k1, i1, k0, i0);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortPairs(d_sort.ptr, sort_bytes, k0, k1, i0, i1, (int)T, 0, bits,
                                                   ctx->stream);
        if (e == hipSuccess) {
            mc_query_q_kernel<<<grid_for(T), kThreads, 0, ctx->stream>>>(g, ms, (uint32_t)T, k1, i1, hv);
            e = hipGetLastError();
        }
        ++passes;
    }
    ctx->launches = 7 + 5 * passes;
    prof_mark(ctx, 2);
    const s2h_status injected = e == hipSuccess ? inject(kFailReadback) : S2H_OK;
    if (e == hipSuccess && injected == S2H_OK)
        e = hipMemcpyAsync(hits_valid, d_hv.ptr, n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    prof_collect(ctx);
    release_pass();
    release();
    if (injected != S2H_OK) return injected;
    if (e != hipSuccess) return fail(S2H_ERR_HIP, std::string("mesh Monte Carlo: ") + hipGetErrorString(e));
    return S2H_OK;
}

s2h_status check_mc_args(const double bbox_min[3], const double bbox_max[3], uint32_t r_max, uint64_t *hits_valid,
                         uint64_t n) {
    if (!bbox_min || !bbox_max || !hits_valid) return fail(S2H_ERR_ARG, "null argument");
    if (n != 2 * ((uint64_t)r_max + 1)) return fail(S2H_ERR_SIZE, "hits_valid must hold 2*(r_max+1) values");
    for (int a = 0; a < 3; ++a)
        if (!(bbox_max[a] >= bbox_min[a]) || !std::isfinite(bbox_min[a]) || !std::isfinite(bbox_max[a]))
            return fail(S2H_ERR_ARG, "bounding box must be finite with max >= min");
    return S2H_OK;
}

}  // namespace meshmc
}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_mesh_mc_counts(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntris,
                              const double bbox_min[3], const double bbox_max[3], uint32_t r_max, uint64_t samples,
                              uint64_t seed, uint64_t *hits_valid, uint64_t n) {
    using namespace s2h::meshmc;
    if (!ctx || (ntris && (!xyz || !tri)) || (nverts && !xyz)) return fail(S2H_ERR_ARG, "null argument");
    S2H_TRY(check_mc_args(bbox_min, bbox_max, r_max, hits_valid, n));
    for (uint64_t f = 0; f < 3 * ntris; ++f)
        if (tri[f] >= nverts) return fail(S2H_ERR_ARG, "vertex index out of range");
    if (ntris > 0x7FFFFFFFull) return fail(S2H_ERR_UNSUPPORTED, "mesh too large");
    S2H_HIP(hipSetDevice(ctx->device));
    S2H_TRY(inject(kFailLaunch));
    std::fill(hits_valid, hits_valid + n, 0);
    if (r_max == 0) return S2H_OK;
    // whole-mesh vertex bbox (mesh_bbox, src/geometry/bbox.rs:8-25) and amax
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0}, amax = 0.0;
    for (uint64_t v = 0; v < nverts; ++v)
        for (int a = 0; a < 3; ++a) {
            const double c = xyz[3 * v + a];
            mn[a] = v ? vox::rmin(mn[a], c) : c;
            mx[a] = v ? vox::rmax(mx[a], c) : c;
            amax = std::max(amax, std::fabs(c));
        }
    Grid g{};
    double cmax, smax;
    setup_grid(bbox_min, bbox_max, nverts > 0, mn, mx, ntris, g, cmax, smax);
    DevBuf d_xyz, d_tri;
    s2h_status st = upload(ctx, d_xyz, xyz, std::max<uint64_t>(nverts, 1) * 24);
    if (st == S2H_OK) st = upload(ctx, d_tri, tri, std::max<uint64_t>(ntris, 1) * 12);
    if (st == S2H_OK)
        st = bin_and_run(ctx, g, cmax, smax, amax, (const double *)d_xyz.ptr, (const uint32_t *)d_tri.ptr, ntris,
                         bbox_min, bbox_max, r_max, samples, seed, hits_valid, n);
    dev_free(ctx, d_xyz);
    dev_free(ctx, d_tri);
    return st;
}

s2h_status s2h_mesh_mc_counts_resident(s2h_ctx *ctx, const s2h_mesh *m, const double bbox_min[3],
                                       const double bbox_max[3], uint32_t r_max, uint64_t samples, uint64_t seed,
                                       uint64_t *hits_valid, uint64_t n) {
    using namespace s2h::meshmc;
    if (!ctx || !m) return fail(S2H_ERR_ARG, "null argument");
    S2H_TRY(check_mc_args(bbox_min, bbox_max, r_max, hits_valid, n));
    S2H_HIP(hipSetDevice(ctx->device));
    S2H_TRY(inject(kFailLaunch));
    std::fill(hits_valid, hits_valid + n, 0);
    if (r_max == 0) return S2H_OK;
    Grid g{};
    double cmax, smax;
    setup_grid(bbox_min, bbox_max, m->nverts > 0, m->bmin, m->bmax, m->ntris, g, cmax, smax);
    return bin_and_run(ctx, g, cmax, smax, m->amax, (const double *)m->xyz.ptr, (const uint32_t *)m->tri.ptr, m->ntris,
                       bbox_min, bbox_max, r_max, samples, seed, hits_valid, n);
}

s2h_status s2h_mesh_mc_s2(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntris,
                          const double bbox_min[3], const double bbox_max[3], uint32_t r_max, uint64_t samples,
                          uint64_t seed, double vf, double *s2, uint64_t n) {
    if (!s2) return fail(S2H_ERR_ARG, "null argument");
    if (n != (uint64_t)r_max + 1) return fail(S2H_ERR_SIZE, "s2 buffer must hold r_max+1 values");
    std::vector<uint64_t> hv(2 * ((uint64_t)r_max + 1));
    S2H_TRY(s2h_mesh_mc_counts(ctx, xyz, nverts, tri, ntris, bbox_min, bbox_max, r_max, samples, seed, hv.data(),
                               hv.size()));
    s2[0] = vf;
    for (uint32_t r = 1; r <= r_max; ++r)
        s2[r] = hv[2 * r + 1] == 0 ? 0.0 : (double)hv[2 * r] / (double)hv[2 * r + 1];
    return S2H_OK;
}

}  // extern "C"
