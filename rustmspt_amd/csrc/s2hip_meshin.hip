// Device-resident meshes: STL weld, granule split and voxelization without a
// host round trip.
//
// Weld (load_stl, src/io/stl.rs:55-80, 157-189, 214-239): binary files are
// decoded on the device (f32 little-endian vertices widened to f64); ASCII files
// are parsed on the host into corners.  Every corner gets the key
// (round(x*1e6), round(y*1e6), round(z*1e6)) as saturating i64, plus its file:
// welds never cross files, as load_stl_or_merge_folder (src/io/stl.rs:282-302)
// welds each file and then concatenates.  Corners are inserted into an
// open-addressing table (one CAS per new key); each key keeps its smallest
// corner index (atomicMin); a prefix sum over "first occurrence" flags numbers
// the vertices in order of first occurrence -- exactly the sequential rule
// "the first occurrence keeps its index".
//
// Granules (split_mesh_into_granules, src/geometry/mesh_ops.rs:48-~120): a
// lock-free union-find over shared vertices (hook the larger root under the
// smaller with CAS), then each component is numbered by the rank of its
// smallest face index: the order in which the reference's BFS starts.
//
// Voxelization (build_bbox_occupancy, src/geometry/s2.rs:160-201): faces are
// stably sorted by granule (the host path's order), per-granule bounds and
// triangle preparation run on the device, and the voxelize kernels are the
// ones s2h_voxelize uses.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "s2hip_internal.h"
#include "s2hip_stlio.h"
#include "s2hip_voxel_dev.h"

namespace s2h {
namespace meshin {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;

struct Part {            // one input file
    uint64_t face_start;  // first face (global)
    uint64_t nfaces;
    uint64_t byte_off;    // binary: offset of the file in the uploaded blob; ASCII: ~0
};

__device__ __forceinline__ uint32_t part_of(const Part *__restrict__ parts, uint32_t nparts, uint64_t f) {
    uint32_t lo = 0, hi = nparts - 1;  // last part with face_start <= f
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (parts[mid].face_start <= f) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Binary faces: 50-byte records after the 84-byte header; vertex k of a face is
// 3 f32 at record + 12 + 12k (2-byte aligned: read as u16 halves).
__global__ void decode_kernel(const uint8_t *__restrict__ blob, const Part *__restrict__ parts, uint32_t nparts,
                              uint64_t nfaces, double *__restrict__ corners, uint32_t *__restrict__ fpart) {
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nfaces;
         f += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = part_of(parts, nparts, f);
        fpart[f] = p;
        const Part pt = parts[p];
        if (pt.byte_off == ~0ull) continue;  // ASCII: corners already in place
        const uint16_t *q = (const uint16_t *)(blob + pt.byte_off + 84 + 50 * (f - pt.face_start) + 12);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const uint32_t bits = (uint32_t)q[2 * k] | ((uint32_t)q[2 * k + 1] << 16);
            corners[9 * f + k] = (double)__uint_as_float(bits);
        }
    }
}

__device__ __forceinline__ int64_t key_of(double v) { return vox::f2i_sat(round(v * 1000000.0)); }

__device__ __forceinline__ uint64_t mix(uint64_t h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

__global__ void key_kernel(const double *__restrict__ corners, const uint32_t *__restrict__ fpart, uint64_t n,
                           int64_t *__restrict__ keys /* 3 per corner */, uint32_t *__restrict__ kpart) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < n; c += (uint64_t)gridDim.x * blockDim.x) {
        keys[3 * c] = key_of(corners[3 * c]);
        keys[3 * c + 1] = key_of(corners[3 * c + 1]);
        keys[3 * c + 2] = key_of(corners[3 * c + 2]);
        kpart[c] = fpart[c / 3];
    }
}

__global__ void fill_u32_kernel(uint32_t *__restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void insert_kernel(const int64_t *__restrict__ keys, const uint32_t *__restrict__ kpart, uint64_t n,
                              uint32_t *slots, uint64_t mask, uint32_t *__restrict__ slot_of) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < n; c += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t x = keys[3 * c], y = keys[3 * c + 1], z = keys[3 * c + 2];
        const uint32_t p = kpart[c];
        uint64_t h = mix((uint64_t)x * 0x9E3779B97F4A7C15ull ^ mix((uint64_t)y + 0x632BE59BD9B4E019ull) ^
                         mix((uint64_t)z * 0x85EBCA77C2B2AE63ull + p)) & mask;
        while (true) {
            uint32_t o = __hip_atomic_load(&slots[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (o == kEmpty) {
                const uint32_t prev = atomicCAS(&slots[h], kEmpty, (uint32_t)c);
                if (prev == kEmpty) {
                    slot_of[c] = (uint32_t)h;
                    break;
                }
                o = prev;
            }
            if (keys[3 * (uint64_t)o] == x && keys[3 * (uint64_t)o + 1] == y && keys[3 * (uint64_t)o + 2] == z &&
                kpart[o] == p) {
                slot_of[c] = (uint32_t)h;
                break;
            }
            h = (h + 1) & mask;
        }
    }
}

__global__ void first_kernel(const uint32_t *__restrict__ slot_of, uint64_t n, uint32_t *first) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < n; c += (uint64_t)gridDim.x * blockDim.x)
        atomicMin(&first[slot_of[c]], (uint32_t)c);
}

__global__ void flag_kernel(const uint32_t *__restrict__ slot_of, const uint32_t *__restrict__ first, uint64_t n,
                            uint32_t *__restrict__ flag) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < n; c += (uint64_t)gridDim.x * blockDim.x)
        flag[c] = first[slot_of[c]] == (uint32_t)c;
}

__global__ void emit_kernel(const double *__restrict__ corners, const uint32_t *__restrict__ slot_of,
                            const uint32_t *__restrict__ first, const uint32_t *__restrict__ rank, uint64_t n,
                            double *__restrict__ xyz, uint32_t *__restrict__ tri) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < n; c += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t f = first[slot_of[c]];
        if (f == (uint32_t)c) {
            const uint64_t v = rank[c];
            xyz[3 * v] = corners[3 * c];
            xyz[3 * v + 1] = corners[3 * c + 1];
            xyz[3 * v + 2] = corners[3 * c + 2];
        }
        tri[c] = rank[f];
    }
}

// ---- union-find over vertices

__device__ __forceinline__ uint32_t uf_find(uint32_t *parent, uint32_t v) {
    while (true) {
        const uint32_t p = __hip_atomic_load(&parent[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p == v) return v;
        v = p;
    }
}

__global__ void uf_init_kernel(uint32_t *__restrict__ parent, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        parent[i] = (uint32_t)i;
}

// Join the components of a and b: the larger root is hooked under the smaller,
// so every parent index is smaller than its child and the forest has no cycles.
__device__ void uf_union(uint32_t *parent, uint32_t a, uint32_t b) {
    while (true) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&parent[b], b, a) == b) return;
    }
}

__global__ void uf_hook_kernel(const uint32_t *__restrict__ tri, uint64_t nfaces, uint32_t *parent) {
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nfaces;
         f += (uint64_t)gridDim.x * blockDim.x) {
        uf_union(parent, tri[3 * f], tri[3 * f + 1]);
        uf_union(parent, tri[3 * f], tri[3 * f + 2]);
    }
}

__global__ void uf_root_kernel(uint32_t *parent, uint64_t n, uint32_t *__restrict__ root) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        root[i] = uf_find(parent, (uint32_t)i);
}

__global__ void minface_kernel(const uint32_t *__restrict__ tri, const uint32_t *__restrict__ root, uint64_t nfaces,
                               uint32_t *minf) {
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nfaces;
         f += (uint64_t)gridDim.x * blockDim.x)
        atomicMin(&minf[root[tri[3 * f]]], (uint32_t)f);
}

__global__ void gflag_kernel(const uint32_t *__restrict__ tri, const uint32_t *__restrict__ root,
                             const uint32_t *__restrict__ minf, uint64_t nfaces, uint32_t *__restrict__ flag) {
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nfaces;
         f += (uint64_t)gridDim.x * blockDim.x)
        flag[f] = minf[root[tri[3 * f]]] == (uint32_t)f;
}

__global__ void gid_kernel(const uint32_t *__restrict__ tri, const uint32_t *__restrict__ root,
                           const uint32_t *__restrict__ minf, const uint32_t *__restrict__ grank, uint64_t nfaces,
                           uint32_t *__restrict__ gid) {
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nfaces;
         f += (uint64_t)gridDim.x * blockDim.x)
        gid[f] = grank[minf[root[tri[3 * f]]]];
}

// ---- mesh bbox: f64::min / f64::max over all vertices (mesh_bbox, src/geometry/bbox.rs:8-25),
// and the largest |coordinate| (std::max from 0, NaN ignored)

__global__ __launch_bounds__(256) void bbox_kernel(const double *__restrict__ xyz, uint64_t nverts,
                                                   double *__restrict__ partial /* 7 per block */) {
    __shared__ double red[7][256];
    double v[7] = {NAN, NAN, NAN, NAN, NAN, NAN, 0.0};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nverts;
         i += (uint64_t)gridDim.x * blockDim.x)
        for (int a = 0; a < 3; ++a) {
            const double c = xyz[3 * i + a];
            v[a] = vox::rmin(v[a], c);
            v[3 + a] = vox::rmax(v[3 + a], c);
            v[6] = vox::smax2(v[6], fabs(c));
        }
    for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int a = 0; a < 3; ++a) {
                red[a][threadIdx.x] = vox::rmin(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = vox::rmax(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
            red[6][threadIdx.x] = vox::smax2(red[6][threadIdx.x], red[6][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) partial[7 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// ---- voxelization from the resident mesh

__global__ void granule_start_kernel(const uint32_t *__restrict__ sorted_gid, uint64_t nfaces, uint32_t ngran,
                                     uint32_t *__restrict__ start) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= nfaces;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (i == nfaces) start[ngran] = (uint32_t)nfaces;
        else if (i == 0 || sorted_gid[i] != sorted_gid[i - 1]) start[sorted_gid[i]] = (uint32_t)i;
    }
}

struct Bounds {
    int64_t lo[3], hi[3];
};

// One block per granule: its vertex bbox and largest |coordinate| over its faces
// in sorted order (prepare_granule's reductions), then its voxel range, margins
// and tile count.
__global__ __launch_bounds__(256) void granule_kernel(const double *__restrict__ xyz, const uint32_t *__restrict__ tri,
                                                      const uint32_t *__restrict__ order,
                                                      const uint32_t *__restrict__ start, double bx, double by,
                                                      double bz, double pitch, uint64_t d0, uint64_t d1, uint64_t d2,
                                                      vox::GranuleInfo *__restrict__ gran, Bounds *__restrict__ bounds,
                                                      uint32_t *__restrict__ ntiles) {
    __shared__ double red[7][256];
    const uint32_t g = blockIdx.x;
    const uint32_t s0 = start[g], s1 = start[g + 1];
    double v[7] = {NAN, NAN, NAN, NAN, NAN, NAN, 0.0};
    for (uint32_t k = s0 + threadIdx.x; k < s1; k += blockDim.x) {
        const uint32_t f = order[k];
        for (int c = 0; c < 3; ++c) {
            const double *p = xyz + 3 * (uint64_t)tri[3 * (uint64_t)f + c];
            for (int a = 0; a < 3; ++a) {
                v[a] = vox::rmin(v[a], p[a]);
                v[3 + a] = vox::rmax(v[3 + a], p[a]);
                v[6] = vox::smax2(v[6], fabs(p[a]));
            }
        }
    }
    for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int a = 0; a < 3; ++a) {
                red[a][threadIdx.x] = vox::rmin(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = vox::rmax(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
            red[6][threadIdx.x] = vox::smax2(red[6][threadIdx.x], red[6][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    vox::GranuleInfo gi;
    for (int a = 0; a < 3; ++a) {
        gi.mn[a] = red[a][0];
        gi.mx[a] = red[3 + a][0];
    }
    gi.tri_start = s0;
    gi.tri_end = s1;
    gi.owner = 0;
    gi.sign = 0;
    gi.amax = red[6][0];
    gi.cmax = gi.smax = 0.0;
    gi.xyz = nullptr;
    gi.tri = nullptr;
    const double bmin[3] = {bx, by, bz};
    const uint64_t dims[3] = {d0, d1, d2};
    Bounds b;
    const bool ok = vox::granule_bounds(gi.mn, gi.mx, bmin, pitch, dims, b.lo, b.hi, &gi.cmax, &gi.smax);
    gran[g] = gi;
    bounds[g] = b;
    ntiles[g] = ok ? (uint32_t)(((b.hi[0] - b.lo[0] + vox::kTileX - 1) / vox::kTileX) *
                                ((b.hi[1] - b.lo[1] + vox::kTileY - 1) / vox::kTileY) *
                                ((b.hi[2] - b.lo[2] + vox::kTileZ - 1) / vox::kTileZ))
                   : 0u;
}

// Tiles of granule g in push_tiles order (x, then y, then z).
__global__ void tiles_kernel(const Bounds *__restrict__ bounds, const uint32_t *__restrict__ ntiles,
                             const uint32_t *__restrict__ tile_start, vox::Tile *__restrict__ tiles) {
    const uint32_t g = blockIdx.x;
    const uint32_t n = ntiles[g];
    if (n == 0) return;
    const Bounds b = bounds[g];
    const uint32_t ny = (uint32_t)((b.hi[1] - b.lo[1] + vox::kTileY - 1) / vox::kTileY);
    const uint32_t nz = (uint32_t)((b.hi[2] - b.lo[2] + vox::kTileZ - 1) / vox::kTileZ);
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
        const uint32_t iz = t % nz, iy = (t / nz) % ny, ix = t / (nz * ny);
        vox::Tile tl;
        tl.granule = g;
        const int64_t x = b.lo[0] + (int64_t)ix * vox::kTileX, y = b.lo[1] + (int64_t)iy * vox::kTileY,
                      z = b.lo[2] + (int64_t)iz * vox::kTileZ;
        tl.x0 = (uint16_t)x;
        tl.y0 = (uint16_t)y;
        tl.z0 = (uint16_t)z;
        tl.x1 = (uint16_t)(x + vox::kTileX < b.hi[0] ? x + vox::kTileX : b.hi[0]);
        tl.y1 = (uint16_t)(y + vox::kTileY < b.hi[1] ? y + vox::kTileY : b.hi[1]);
        tl.z1 = (uint16_t)(z + vox::kTileZ < b.hi[2] ? z + vox::kTileZ : b.hi[2]);
        tiles[tile_start[g] + t] = tl;
    }
}

__global__ void prepare_kernel(const double *__restrict__ xyz, const uint32_t *__restrict__ tri,
                               const uint32_t *__restrict__ order, const uint32_t *__restrict__ sorted_gid,
                               const vox::GranuleInfo *__restrict__ gran, const uint32_t *__restrict__ ntiles,
                               uint64_t nfaces, vox::TriGeo *__restrict__ geo) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nfaces;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t g = sorted_gid[k];
        if (ntiles[g] == 0) continue;  // never tested
        const uint32_t f = order[k];
        const vox::GranuleInfo &gi = gran[g];
        const double *A = xyz + 3 * (uint64_t)tri[3 * (uint64_t)f];
        const double *B = xyz + 3 * (uint64_t)tri[3 * (uint64_t)f + 1];
        const double *C = xyz + 3 * (uint64_t)tri[3 * (uint64_t)f + 2];
        vox::TriGeo t;
        vox::tri_prepare(A, B, C, gi.cmax, gi.smax, gi.amax, t);
        geo[k] = t;
    }
}

}  // namespace meshin

namespace {

unsigned grid_for(uint64_t n, unsigned block = 256) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + block - 1) / block, 1u << 20));
}

struct Bufs {
    s2h_ctx *ctx;
    std::vector<DevBuf *> held;
    explicit Bufs(s2h_ctx *c) : ctx(c) {}
    s2h_status get(DevBuf &b, uint64_t bytes) {
        S2H_TRY(dev_alloc(ctx, b, std::max<uint64_t>(bytes, 8)));
        held.push_back(&b);
        return S2H_OK;
    }
    ~Bufs() {
        for (DevBuf *b : held) dev_free(ctx, *b);
    }
};

#define LAUNCH_OK()                                                                                     \
    do {                                                                                                \
        hipError_t e_ = hipGetLastError();                                                              \
        if (e_ != hipSuccess) return fail(S2H_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

// Exclusive prefix sum of n u32 flags/counts into out; returns the total.
s2h_status exclusive_sum(s2h_ctx *ctx, const uint32_t *in, uint32_t *out, uint64_t n, uint64_t *total) {
    if (n == 0 || n > 0x7FFFFFFFull) return fail(S2H_ERR_UNSUPPORTED, "scan length out of range");
    size_t bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, ctx->stream) != hipSuccess)
        return fail(S2H_ERR_HIP, "scan sizing failed");
    DevBuf tmp;
    S2H_TRY(dev_alloc(ctx, tmp, std::max<size_t>(bytes, 8)));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp.ptr, bytes, in, out, (int)n, ctx->stream);
    uint32_t last[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(&last[0], out + n - 1, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&last[1], in + n - 1, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, tmp);
    if (e != hipSuccess) return fail(S2H_ERR_HIP, std::string("scan failed: ") + hipGetErrorString(e));
    *total = (uint64_t)last[0] + last[1];
    return S2H_OK;
}

// Vertex bbox and largest |coordinate| of a resident mesh, reduced on the device.
s2h_status mesh_bounds(s2h_ctx *ctx, s2h_mesh *m) {
    using namespace meshin;
    const uint64_t nv = m->nverts;
    if (nv == 0) return S2H_OK;
    const unsigned blocks = (unsigned)std::min<uint64_t>((nv + 255) / 256, 1024);
    DevBuf part;
    S2H_TRY(dev_alloc(ctx, part, (uint64_t)blocks * 7 * 8));
    bbox_kernel<<<blocks, 256, 0, ctx->stream>>>((const double *)m->xyz.ptr, nv, (double *)part.ptr);
    std::vector<double> h((size_t)blocks * 7);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part.ptr, h.size() * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, part);
    if (e != hipSuccess) return fail(S2H_ERR_HIP, std::string("mesh bounds: ") + hipGetErrorString(e));
    double amax = 0.0;
    for (int a = 0; a < 3; ++a) {
        double lo = NAN, hi = NAN;
        for (unsigned b = 0; b < blocks; ++b) {
            lo = vox::rmin(lo, h[7 * b + a]);
            hi = vox::rmax(hi, h[7 * b + 3 + a]);
        }
        m->bmin[a] = lo;
        m->bmax[a] = hi;
    }
    for (unsigned b = 0; b < blocks; ++b) amax = vox::smax2(amax, h[7 * b + 6]);
    m->amax = amax;
    return S2H_OK;
}

namespace {

// Granules and bbox of a mesh whose xyz/tri are resident.
s2h_status finish_mesh(s2h_ctx *ctx, s2h_mesh *m) {
    using namespace meshin;
    const uint64_t nv = m->nverts, nf = m->ntris;
    if (nf > 0x7FFFFFFEull || nv > 0xFFFFFFFEull) return fail(S2H_ERR_UNSUPPORTED, "mesh too large");
    DevBuf parent, root, minf, flag, grank;  // declared before tmp, which frees them
    Bufs tmp(ctx);
    S2H_TRY(dev_alloc(ctx, m->gid, std::max<uint64_t>(nf, 1) * 4));
    if (nf > 0) {
        S2H_TRY(tmp.get(parent, nv * 4));
        S2H_TRY(tmp.get(root, nv * 4));
        S2H_TRY(tmp.get(minf, nv * 4));
        S2H_TRY(tmp.get(flag, nf * 4));
        S2H_TRY(tmp.get(grank, nf * 4));
        auto *P = (uint32_t *)parent.ptr, *R = (uint32_t *)root.ptr, *M = (uint32_t *)minf.ptr;
        auto *T = (const uint32_t *)m->tri.ptr;
        uf_init_kernel<<<grid_for(nv), 256, 0, ctx->stream>>>(P, nv);
        uf_hook_kernel<<<grid_for(nf), 256, 0, ctx->stream>>>(T, nf, P);
        uf_root_kernel<<<grid_for(nv), 256, 0, ctx->stream>>>(P, nv, R);
        fill_u32_kernel<<<grid_for(nv), 256, 0, ctx->stream>>>(M, nv, kEmpty);
        minface_kernel<<<grid_for(nf), 256, 0, ctx->stream>>>(T, R, nf, M);
        gflag_kernel<<<grid_for(nf), 256, 0, ctx->stream>>>(T, R, M, nf, (uint32_t *)flag.ptr);
        LAUNCH_OK();
        uint64_t ng = 0;
        S2H_TRY(exclusive_sum(ctx, (const uint32_t *)flag.ptr, (uint32_t *)grank.ptr, nf, &ng));
        gid_kernel<<<grid_for(nf), 256, 0, ctx->stream>>>(T, R, M, (const uint32_t *)grank.ptr, nf,
                                                          (uint32_t *)m->gid.ptr);
        LAUNCH_OK();
        m->ngranules = (uint32_t)ng;
    }
    if (nv > 0) S2H_TRY(mesh_bounds(ctx, m));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    return S2H_OK;
}

// Weld corners (3 f64 each, three per face) whose file index per face is fpart.
s2h_status weld(s2h_ctx *ctx, const double *d_corners, const uint32_t *d_fpart, uint64_t nfaces, s2h_mesh *m) {
    using namespace meshin;
    const uint64_t n = 3 * nfaces;
    if (n > 0x7FFFFFFEull) return fail(S2H_ERR_UNSUPPORTED, "mesh too large");
    uint64_t cap = 1;
    while (cap < 2 * n) cap <<= 1;
    DevBuf keys, kpart, slots, slot_of, first, flag, rank;
    Bufs tmp(ctx);
    S2H_TRY(tmp.get(keys, n * 24));
    S2H_TRY(tmp.get(kpart, n * 4));
    S2H_TRY(tmp.get(slots, cap * 4));
    S2H_TRY(tmp.get(slot_of, n * 4));
    S2H_TRY(tmp.get(first, cap * 4));
    S2H_TRY(tmp.get(flag, n * 4));
    S2H_TRY(tmp.get(rank, n * 4));
    key_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(d_corners, d_fpart, n, (int64_t *)keys.ptr, (uint32_t *)kpart.ptr);
    fill_u32_kernel<<<grid_for(cap), 256, 0, ctx->stream>>>((uint32_t *)slots.ptr, cap, kEmpty);
    fill_u32_kernel<<<grid_for(cap), 256, 0, ctx->stream>>>((uint32_t *)first.ptr, cap, kEmpty);
    insert_kernel<<<grid_for(n), 256, 0, ctx->stream>>>((const int64_t *)keys.ptr, (const uint32_t *)kpart.ptr, n,
                                                        (uint32_t *)slots.ptr, cap - 1, (uint32_t *)slot_of.ptr);
    first_kernel<<<grid_for(n), 256, 0, ctx->stream>>>((const uint32_t *)slot_of.ptr, n, (uint32_t *)first.ptr);
    flag_kernel<<<grid_for(n), 256, 0, ctx->stream>>>((const uint32_t *)slot_of.ptr, (const uint32_t *)first.ptr, n,
                                                      (uint32_t *)flag.ptr);
    LAUNCH_OK();
    uint64_t nv = 0;
    S2H_TRY(exclusive_sum(ctx, (const uint32_t *)flag.ptr, (uint32_t *)rank.ptr, n, &nv));
    m->nverts = nv;
    m->ntris = nfaces;
    S2H_TRY(dev_alloc(ctx, m->xyz, std::max<uint64_t>(nv, 1) * 24));
    S2H_TRY(dev_alloc(ctx, m->tri, std::max<uint64_t>(n, 1) * 4));
    emit_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(d_corners, (const uint32_t *)slot_of.ptr,
                                                      (const uint32_t *)first.ptr, (const uint32_t *)rank.ptr, n,
                                                      (double *)m->xyz.ptr, (uint32_t *)m->tri.ptr);
    LAUNCH_OK();
    return finish_mesh(ctx, m);
}

void mesh_release(s2h_ctx *ctx, s2h_mesh *m) {
    if (!m) return;
    dev_free(ctx, m->xyz);
    dev_free(ctx, m->tri);
    dev_free(ctx, m->gid);
    delete m;
}

s2h_status from_stl_bytes(s2h_ctx *ctx, const std::vector<const std::vector<unsigned char> *> &files,
                          s2h_mesh **out) {
    using namespace meshin;
    if (files.empty()) return fail(S2H_ERR_ARG, "no STL input");
    std::vector<Part> parts(files.size());
    std::vector<std::vector<double>> ascii(files.size());
    uint64_t faces = 0, blob_bytes = 0;
    for (size_t i = 0; i < files.size(); ++i) {
        const auto &b = *files[i];
        parts[i].face_start = faces;
        parts[i].byte_off = ~0ull;
        if (stlio::looks_ascii(b) && stlio::ascii_corners(b, ascii[i])) {
            parts[i].nfaces = ascii[i].size() / 9;
        } else {
            ascii[i].clear();
            uint32_t count;
            std::string err;
            if (!stlio::binary_count(b.data(), b.size(), count, err))
                return fail(S2H_ERR_IO, err + " in file " + std::to_string(i));
            parts[i].nfaces = count;
            parts[i].byte_off = blob_bytes;
            blob_bytes += (84 + (uint64_t)count * 50 + 7) & ~7ull;
        }
        faces += parts[i].nfaces;
    }
    DevBuf blob, corners, fpart, d_parts;
    Bufs tmp(ctx);
    S2H_TRY(tmp.get(blob, blob_bytes));
    S2H_TRY(tmp.get(corners, std::max<uint64_t>(faces, 1) * 72));
    S2H_TRY(tmp.get(fpart, std::max<uint64_t>(faces, 1) * 4));
    S2H_TRY(tmp.get(d_parts, parts.size() * sizeof(Part)));
    S2H_HIP(hipMemcpyAsync(d_parts.ptr, parts.data(), parts.size() * sizeof(Part), hipMemcpyHostToDevice, ctx->stream));
    for (size_t i = 0; i < files.size(); ++i) {
        if (parts[i].byte_off != ~0ull)
            S2H_HIP(hipMemcpyAsync((uint8_t *)blob.ptr + parts[i].byte_off, files[i]->data(),
                                   84 + parts[i].nfaces * 50, hipMemcpyHostToDevice, ctx->stream));
        else if (parts[i].nfaces)
            S2H_HIP(hipMemcpyAsync((double *)corners.ptr + 9 * parts[i].face_start, ascii[i].data(),
                                   ascii[i].size() * 8, hipMemcpyHostToDevice, ctx->stream));
    }
    if (faces)
        decode_kernel<<<grid_for(faces), 256, 0, ctx->stream>>>((const uint8_t *)blob.ptr, (const Part *)d_parts.ptr,
                                                                (uint32_t)parts.size(), faces, (double *)corners.ptr,
                                                                (uint32_t *)fpart.ptr);
    LAUNCH_OK();
    // the host buffers (files, ascii) must outlive the copies
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    auto *m = new s2h_mesh();
    s2h_status st = faces ? weld(ctx, (const double *)corners.ptr, (const uint32_t *)fpart.ptr, faces, m)
                          : finish_mesh(ctx, m);
    if (st != S2H_OK) {
        mesh_release(ctx, m);
        return st;
    }
    *out = m;
    return S2H_OK;
}

}  // namespace
}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_mesh_from_stl(s2h_ctx *ctx, const void *const *files, const uint64_t *sizes, uint32_t nfiles,
                             s2h_mesh **out) {
    if (!ctx || !files || !sizes || !out || nfiles == 0) return fail(S2H_ERR_ARG, "null argument");
    S2H_HIP(hipSetDevice(ctx->device));
    std::vector<std::vector<unsigned char>> copies(nfiles);
    std::vector<const std::vector<unsigned char> *> views(nfiles);
    for (uint32_t i = 0; i < nfiles; ++i) {
        if (!files[i] && sizes[i]) return fail(S2H_ERR_ARG, "null file buffer");
        copies[i].assign((const unsigned char *)files[i], (const unsigned char *)files[i] + sizes[i]);
        views[i] = &copies[i];
    }
    return from_stl_bytes(ctx, views, out);
}

s2h_status s2h_mesh_load_stl(s2h_ctx *ctx, const char *const *paths, uint32_t npaths, s2h_mesh **out) {
    if (!ctx || !paths || !out || npaths == 0) return fail(S2H_ERR_ARG, "null argument");
    S2H_HIP(hipSetDevice(ctx->device));
    std::vector<std::vector<unsigned char>> bytes(npaths);
    std::vector<const std::vector<unsigned char> *> views(npaths);
    for (uint32_t i = 0; i < npaths; ++i) {
        if (!paths[i] || !stlio::read_file(paths[i], bytes[i]))
            return fail(S2H_ERR_IO, std::string("cannot read ") + (paths[i] ? paths[i] : "(null)"));
        views[i] = &bytes[i];
    }
    s2h_status st = from_stl_bytes(ctx, views, out);
    if (st != S2H_OK && npaths == 1) {
        std::string msg = s2h_last_error();
        return fail(st, msg + " (" + paths[0] + ")");
    }
    return st;
}

s2h_status s2h_mesh_from_arrays(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntris,
                                s2h_mesh **out) {
    if (!ctx || !out || (nverts && !xyz) || (ntris && !tri)) return fail(S2H_ERR_ARG, "null argument");
    for (uint64_t i = 0; i < 3 * ntris; ++i)
        if (tri[i] >= nverts) return fail(S2H_ERR_ARG, "vertex index out of range");
    S2H_HIP(hipSetDevice(ctx->device));
    auto *m = new s2h_mesh();
    m->nverts = nverts;
    m->ntris = ntris;
    s2h_status st = upload(ctx, m->xyz, xyz, std::max<uint64_t>(nverts, 1) * 24);
    if (st == S2H_OK) st = upload(ctx, m->tri, tri, std::max<uint64_t>(ntris, 1) * 12);
    if (st == S2H_OK) st = finish_mesh(ctx, m);
    if (st != S2H_OK) {
        mesh_release(ctx, m);
        return st;
    }
    *out = m;
    return S2H_OK;
}

s2h_status s2h_mesh_info(const s2h_mesh *m, uint64_t *nverts, uint64_t *ntris, uint32_t *ngranules,
                         double bbox_min[3], double bbox_max[3]) {
    if (!m) return fail(S2H_ERR_ARG, "null argument");
    if (nverts) *nverts = m->nverts;
    if (ntris) *ntris = m->ntris;
    if (ngranules) *ngranules = m->ngranules;
    for (int a = 0; a < 3; ++a) {
        if (bbox_min) bbox_min[a] = m->bmin[a];
        if (bbox_max) bbox_max[a] = m->bmax[a];
    }
    return S2H_OK;
}

s2h_status s2h_mesh_read(s2h_ctx *ctx, const s2h_mesh *m, double *xyz, uint64_t nxyz, uint32_t *tri, uint64_t ntri,
                         uint32_t *gid, uint64_t ngid) {
    if (!ctx || !m) return fail(S2H_ERR_ARG, "null argument");
    if ((xyz && nxyz != 3 * m->nverts) || (tri && ntri != 3 * m->ntris) || (gid && ngid != m->ntris))
        return fail(S2H_ERR_SIZE, "buffer lengths must be 3*nverts, 3*ntris and ntris");
    S2H_HIP(hipSetDevice(ctx->device));
    if (xyz && nxyz) S2H_HIP(hipMemcpyAsync(xyz, m->xyz.ptr, nxyz * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (tri && ntri) S2H_HIP(hipMemcpyAsync(tri, m->tri.ptr, ntri * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (gid && ngid) S2H_HIP(hipMemcpyAsync(gid, m->gid.ptr, ngid * 4, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    return S2H_OK;
}

s2h_status s2h_mesh_voxelize(s2h_ctx *ctx, const s2h_mesh *m, const double bbox_min[3], double pitch,
                             const uint64_t dims[3], s2h_vol **out, uint64_t *occupied) {
    using namespace meshin;
    if (!ctx || !m || !bbox_min || !dims || !out) return fail(S2H_ERR_ARG, "null argument");
    if (!(pitch > 0.0) || !(pitch <= 1.7976931348623157e308)) return fail(S2H_ERR_ARG, "pitch must be finite and > 0");
    S2H_HIP(hipSetDevice(ctx->device));
    S2H_TRY(inject(kFailLaunch));
    if ((double)dims[0] * (double)dims[1] * (double)dims[2] / 256.0 >= 2147483647.0)
        return fail(S2H_ERR_UNSUPPORTED, "grid too large for the resident-mesh voxelizer's tile count");
    s2h_vol *vol = nullptr;
    S2H_TRY(vol_alloc(ctx, dims, &vol));
    auto cleanup = [&](s2h_status st) {
        s2h_vol_free(ctx, vol);
        return st;
    };
    if (hipMemsetAsync(vol->d_words, 0, vol->nwords * 8, ctx->stream) != hipSuccess)
        return cleanup(fail(S2H_ERR_HIP, "memset failed"));
    const uint64_t nf = m->ntris;
    const uint32_t ng = m->ngranules;
    if (nf > 0 && ng > 0) {
        DevBuf order_in, order, sgid, start, gran, bounds, ntiles, tstart, tiles, geo, sort_tmp;
        Bufs tmp(ctx);
        s2h_status st = S2H_OK;
        if ((st = tmp.get(order_in, nf * 4)) != S2H_OK || (st = tmp.get(order, nf * 4)) != S2H_OK ||
            (st = tmp.get(sgid, nf * 4)) != S2H_OK || (st = tmp.get(start, ((uint64_t)ng + 1) * 4)) != S2H_OK ||
            (st = tmp.get(gran, (uint64_t)ng * sizeof(vox::GranuleInfo))) != S2H_OK ||
            (st = tmp.get(bounds, (uint64_t)ng * sizeof(Bounds))) != S2H_OK ||
            (st = tmp.get(ntiles, (uint64_t)ng * 4)) != S2H_OK || (st = tmp.get(tstart, (uint64_t)ng * 4)) != S2H_OK ||
            (st = tmp.get(geo, nf * sizeof(vox::TriGeo))) != S2H_OK)
            return cleanup(st);
        // identity order, then a stable sort of (granule, face)
        uf_init_kernel<<<grid_for(nf), 256, 0, ctx->stream>>>((uint32_t *)order_in.ptr, nf);
        size_t bytes = 0;
        int bits = 1;
        while (bits < 32 && (1ull << bits) < ng) ++bits;
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)m->gid.ptr, (uint32_t *)sgid.ptr,
                                               (const uint32_t *)order_in.ptr, (uint32_t *)order.ptr, (int)nf, 0, bits,
                                               ctx->stream) != hipSuccess)
            return cleanup(fail(S2H_ERR_HIP, "sort sizing failed"));
        if ((st = tmp.get(sort_tmp, bytes)) != S2H_OK) return cleanup(st);
        if (hipcub::DeviceRadixSort::SortPairs(sort_tmp.ptr, bytes, (const uint32_t *)m->gid.ptr, (uint32_t *)sgid.ptr,
                                               (const uint32_t *)order_in.ptr, (uint32_t *)order.ptr, (int)nf, 0, bits,
                                               ctx->stream) != hipSuccess)
            return cleanup(fail(S2H_ERR_HIP, "sort failed"));
        granule_start_kernel<<<grid_for(nf + 1), 256, 0, ctx->stream>>>((const uint32_t *)sgid.ptr, nf, ng,
                                                                        (uint32_t *)start.ptr);
        granule_kernel<<<ng, 256, 0, ctx->stream>>>((const double *)m->xyz.ptr, (const uint32_t *)m->tri.ptr,
                                                    (const uint32_t *)order.ptr, (const uint32_t *)start.ptr,
                                                    bbox_min[0], bbox_min[1], bbox_min[2], pitch, dims[0], dims[1],
                                                    dims[2], (vox::GranuleInfo *)gran.ptr, (Bounds *)bounds.ptr,
                                                    (uint32_t *)ntiles.ptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return cleanup(fail(S2H_ERR_HIP, std::string("granule kernel: ") + hipGetErrorString(e)));
        uint64_t total = 0;
        if ((st = exclusive_sum(ctx, (const uint32_t *)ntiles.ptr, (uint32_t *)tstart.ptr, ng, &total)) != S2H_OK)
            return cleanup(st);
        if (total > 0) {
            if ((st = tmp.get(tiles, total * sizeof(vox::Tile))) != S2H_OK) return cleanup(st);
            tiles_kernel<<<ng, 256, 0, ctx->stream>>>((const Bounds *)bounds.ptr, (const uint32_t *)ntiles.ptr,
                                                      (const uint32_t *)tstart.ptr, (vox::Tile *)tiles.ptr);
            prepare_kernel<<<grid_for(nf), 256, 0, ctx->stream>>>(
                (const double *)m->xyz.ptr, (const uint32_t *)m->tri.ptr, (const uint32_t *)order.ptr,
                (const uint32_t *)sgid.ptr, (const vox::GranuleInfo *)gran.ptr, (const uint32_t *)ntiles.ptr, nf,
                (vox::TriGeo *)geo.ptr);
            e = hipGetLastError();
            if (e != hipSuccess) return cleanup(fail(S2H_ERR_HIP, std::string("prepare: ") + hipGetErrorString(e)));
            if ((st = vox::run_voxelize(ctx, (const vox::TriGeo *)geo.ptr, (const vox::GranuleInfo *)gran.ptr,
                                        (const vox::Tile *)tiles.ptr, total, bbox_min, pitch, vol)) != S2H_OK)
                return cleanup(st);
        }
    }
    uint64_t occ = 0;
    s2h_status st = vol_popcount(ctx, vol, &occ);
    if (st != S2H_OK) return cleanup(st);
    if (occupied) *occupied = occ;
    *out = vol;
    return S2H_OK;
}

s2h_status s2h_mesh_set_vertices(s2h_ctx *ctx, s2h_mesh *m, uint64_t vstart, const double *xyz, uint64_t n) {
    if (!ctx || !m || (n && !xyz)) return fail(S2H_ERR_ARG, "null argument");
    if (vstart > m->nverts || n > m->nverts - vstart) return fail(S2H_ERR_SIZE, "vertex range outside the mesh");
    S2H_HIP(hipSetDevice(ctx->device));
    if (n) S2H_HIP(hipMemcpyAsync((double *)m->xyz.ptr + 3 * vstart, xyz, n * 24, hipMemcpyHostToDevice, ctx->stream));
    return mesh_bounds(ctx, m);
}

void s2h_mesh_free(s2h_ctx *ctx, s2h_mesh *m) {
    if (!ctx || !m) return;
    (void)hipSetDevice(ctx->device);
    mesh_release(ctx, m);
}

}  // extern "C"
