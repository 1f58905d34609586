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
// miss, so the hit multiset is unchanged).  Lanes then take (triangle, x slice)
// items and run the predicate only at the slice's centres whose projection can
// fall in the triangle's bound, appending hits to per-centre lists in LDS; each
// thread finally sorts and counts its centre's hits.  Rays with more than
// kMaxHits hits go to an overflow list that a second kernel resolves exactly
// with a repeated-minimum scan.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "s2hip_internal.h"
#include "s2hip_voxel_dev.h"

namespace s2h {
namespace vox {

__global__ __launch_bounds__(256) void voxelize_kernel(const TriGeo *__restrict__ tris,
                                                       const GranuleInfo *__restrict__ gran,
                                                       const Tile *__restrict__ tiles, double bx, double by,
                                                       double bz, double pitch, uint64_t ny, uint64_t row_words,
                                                       unsigned long long *__restrict__ words,
                                                       uint32_t *__restrict__ overflow, uint32_t overflow_cap) {
    voxelize_tile(tris, gran, tiles[blockIdx.x], blockIdx.x, bx, by, bz, pitch, overflow, overflow_cap,
                  SetBit{words, ny, row_words});
}

__global__ void voxelize_overflow_kernel(const TriGeo *__restrict__ tris, const GranuleInfo *__restrict__ gran,
                                         const Tile *__restrict__ tiles, double bx, double by, double bz,
                                         double pitch, uint64_t ny, uint64_t row_words,
                                         unsigned long long *__restrict__ words,
                                         const uint32_t *__restrict__ overflow, uint32_t count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    voxel_overflow_ray(tris, gran, tiles, bx, by, bz, pitch, overflow, k, SetBit{words, ny, row_words});
}


s2h_status run_voxelize(s2h_ctx *ctx, const TriGeo *d_geo, const GranuleInfo *d_gran, const Tile *d_tiles,
                        uint64_t ntiles, const double bbox_min[3], double pitch, s2h_vol *vol) {
    if (ntiles == 0) return S2H_OK;
    const uint32_t over_cap = 1u << 16;
    DevBuf d_over;
    S2H_TRY(dev_alloc(ctx, d_over, (1 + 2 * (uint64_t)over_cap) * 4));
    hipError_t e = hipMemsetAsync(d_over.ptr, 0, 4, ctx->stream);
    for (uint64_t t0 = 0; e == hipSuccess && t0 < ntiles; t0 += (1u << 30)) {
        const uint64_t nt = std::min<uint64_t>(ntiles - t0, 1u << 30);
        voxelize_kernel<<<(unsigned)nt, 256, 0, ctx->stream>>>(d_geo, d_gran, d_tiles + t0, bbox_min[0], bbox_min[1],
                                                              bbox_min[2], pitch, vol->dims[1], vol->row_words,
                                                              (unsigned long long *)vol->d_words,
                                                              (uint32_t *)d_over.ptr, over_cap);
        e = hipGetLastError();
    }
    uint32_t nover = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nover, d_over.ptr, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && nover > over_cap) {
        dev_free(ctx, d_over);
        return fail(S2H_ERR_UNSUPPORTED, "more than 65536 voxel rays with over 16 hits");
    }
    if (e == hipSuccess && nover) {
        voxelize_overflow_kernel<<<(nover + 63) / 64, 64, 0, ctx->stream>>>(
            d_geo, d_gran, d_tiles, bbox_min[0], bbox_min[1], bbox_min[2], pitch, vol->dims[1], vol->row_words,
            (unsigned long long *)vol->d_words, (const uint32_t *)d_over.ptr, nover);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    dev_free(ctx, d_over);
    if (e != hipSuccess) return fail(S2H_ERR_HIP, std::string("voxelize kernel failed: ") + hipGetErrorString(e));
    return S2H_OK;
}

bool prepare_granule(const double *tri_xyz, uint32_t ntri, const double bbox_min[3], double pitch,
                     const uint64_t dims[3], TriGeo *geo, GranuleInfo &gi, int64_t lo[3], int64_t hi[3]) {
    bool have = false;
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0}, amax = 0.0;
    for (uint32_t k = 0; k < ntri; ++k)
        for (int c = 0; c < 3; ++c) {
            const double *v = tri_xyz + 9 * (uint64_t)k + 3 * c;
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
    gi.amax = amax;
    gi.cmax = gi.smax = 0.0;
    gi.xyz = nullptr;
    gi.tri = nullptr;
    if (!have) return false;
    if (!granule_bounds(mn, mx, bbox_min, pitch, dims, lo, hi, &gi.cmax, &gi.smax)) return false;
    for (uint32_t k = 0; k < ntri; ++k) {
        const double *A = tri_xyz + 9 * (uint64_t)k;
        tri_prepare(A, A + 3, A + 6, gi.cmax, gi.smax, amax, geo[k]);
    }
    return true;
}

void push_tiles(uint32_t g, const int64_t lo[3], const int64_t hi[3], std::vector<Tile> &tiles) {
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

}  // namespace vox
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
    S2H_TRY(inject(kFailLaunch));
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
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    std::vector<double> txyz(9 * ntris);
    for (uint64_t f = 0; f < ntris; ++f) {
        double *dst = txyz.data() + 9 * (uint64_t)fill[granule_of_tri[f]]++;
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) dst[3 * c + a] = xyz[3 * (uint64_t)tri[3 * f + c] + a];
    }
    std::vector<vox::GranuleInfo> gran(ngranules);
    std::vector<vox::Tile> tiles;
    std::vector<vox::TriGeo> geo(ntris);
    for (uint32_t g = 0; g < ngranules; ++g) {
        vox::GranuleInfo &gi = gran[g];
        gi.tri_start = start[g];
        gi.tri_end = start[g + 1];
        int64_t lo[3], hi[3];
        if (vox::prepare_granule(txyz.data() + 9 * (uint64_t)start[g], start[g + 1] - start[g], bbox_min, pitch,
                                 dims, geo.data() + start[g], gi, lo, hi))
            vox::push_tiles(g, lo, hi, tiles);
    }
    DevBuf d_geo, d_gran, d_tiles;
    auto release = [&]() {
        dev_free(ctx, d_geo);
        dev_free(ctx, d_gran);
        dev_free(ctx, d_tiles);
    };
    s2h_status st = S2H_OK;
    if (!tiles.empty()) {
        if ((st = upload(ctx, d_geo, geo.data(), geo.size() * sizeof(vox::TriGeo))) != S2H_OK ||
            (st = upload(ctx, d_gran, gran.data(), gran.size() * sizeof(vox::GranuleInfo))) != S2H_OK ||
            (st = upload(ctx, d_tiles, tiles.data(), tiles.size() * sizeof(vox::Tile))) != S2H_OK ||
            (st = vox::run_voxelize(ctx, (const vox::TriGeo *)d_geo.ptr, (const vox::GranuleInfo *)d_gran.ptr,
                                    (const vox::Tile *)d_tiles.ptr, tiles.size(), bbox_min, pitch, vol)) != S2H_OK) {
            release();
            return cleanup(st);
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
