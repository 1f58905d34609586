// Exact S2 pair counting on a bit-packed volume (gfx950).
//
// Counts C(d) = #{y : occ(y) and occ(y + d)} for every in-domain offset d of the
// shells 1..r_max.  C(d) = C(-d), so only the canonical half ball is counted:
// dz > 0, or dz == 0 with (dx, dy) lexicographically positive.  The half ball is
// a set of columns (dx, dy), each with dz = 0..m(dx, dy).
//
// Kernel shape (one wave per work item):
//   * lanes    = 64 columns (dx, dy) with nearly equal m (sorted by m);
//   * uniform  = the partner ("b") row p = (xb, yb) and a block of 8/16/32 dz
//                values starting at a multiple of 32, so the b row shifted right
//                by dz is wave-uniform: it lives in SGPRs and is produced by the
//                scalar unit (64-bit shift of a dword pair), free of VALU cost;
//   * per lane = the "a" row at p - (dx, dy), held in VGPRs for the whole dz
//                block, so the inner loop is v_and_b32 + v_bcnt_u32_b32 per
//                32 voxel pairs, with no memory traffic.
//   Per-lane u32 counters (one per dz of the block) accumulate over the item's
//   b rows and are added once into a u64 dense table.
//
// The dense table is then gathered into shell order, divided by the analytic
// valid count in f64, and folded per radius in shell order -- the reference's
// ordered f64 sum (src/geometry/s2.rs:1202-1207) -- so the curve is bit-identical.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "s2hip_internal.h"

namespace s2h {

static constexpr int kWavesPerBlock = 4;

// ---------------------------------------------------------------- count kernel

// Volume buffers carry kVolTailBytes of zero padding after the last row, so an
// unguarded read of up to 64 dwords past any row start stays inside the
// allocation; bits that do not belong to the row are masked to zero.
template <int NW, int DZB>
__global__ __launch_bounds__(64 * kWavesPerBlock) void exact_count_kernel(
    const uint32_t *__restrict__ vol, uint32_t vol_bytes, int nx, int ny, int row_dw,
    const WorkItem *__restrict__ items, int nitems, const ColInfo *__restrict__ cols,
    unsigned long long *__restrict__ table) {
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWavesPerBlock + threadIdx.x / 64));
    if (wave >= nitems) return;
    const int lane = threadIdx.x & 63;
    const WorkItem it = items[wave];
    const ColInfo c = cols[it.group * 64 + lane];
    const int q0 = it.dz0 >> 5;
    const int nw = it.nw;  // partner dwords q0 .. q0+nw-1 hold voxels; later ones read as 0
    // per-lane rows through a buffer resource: an out-of-range offset reads 0
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)vol, 0, (int)vol_bytes, 0x00020000);
    const int oob = 0x7ffffff0;

    uint32_t cnt[DZB];
#pragma unroll
    for (int j = 0; j < DZB; ++j) cnt[j] = 0;

    const int y0 = it.y0, y1 = it.y1;
    for (int xb = it.x0; xb < it.x1; ++xb) {
        const int xa = xb - c.dx;
        const bool xok = c.m >= 0 && xa >= 0 && xa < nx;
        const uint32_t *brow = vol + ((size_t)xb * ny + y0) * row_dw + q0;
        const int abase = xa * ny * row_dw * 4;
        auto a_off = [&](int yb) {
            const int ya = yb - c.dy;
            return (xok && ya >= 0 && ya < ny) ? abase + ya * row_dw * 4 : oob;
        };
        uint32_t an[NW];
        {
            const int off = a_off(y0);
#pragma unroll
            for (int w = 0; w < NW; ++w) an[w] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, off + 4 * w, 0, 0);
        }
        for (int yb = y0; yb < y1; ++yb, brow += row_dw) {
            uint32_t a[NW];
#pragma unroll
            for (int w = 0; w < NW; ++w) a[w] = an[w];
            if (yb + 1 < y1) {
                const int off = a_off(yb + 1);
#pragma unroll
                for (int w = 0; w < NW; ++w) an[w] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, off + 4 * w, 0, 0);
            }
            uint32_t b[NW + 1];
#pragma unroll
            for (int i = 0; i <= NW; ++i) {
                const uint32_t v = brow[i];
                b[i] = i < nw ? v : 0u;
            }
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint64_t pair = ((uint64_t)b[w + 1] << 32) | b[w];
#pragma unroll
                for (int j = 0; j < DZB; ++j) cnt[j] += __popc(a[w] & (uint32_t)(pair >> j));
            }
        }
    }
#pragma unroll
    for (int j = 0; j < DZB; ++j) {
        const int dz = it.dz0 + j;
        if (c.m >= 0 && dz <= c.m && cnt[j]) atomicAdd(&table[c.base + dz], (unsigned long long)cnt[j]);
    }
}

// ---------------------------------------------------------------- gather / fold

// term[i] = hits / valid in f64 (division correctly rounded, as Rust's `/`).
__global__ void gather_terms_kernel(const Offset *__restrict__ offs, const int32_t *__restrict__ gather,
                                    uint64_t n, const unsigned long long *__restrict__ table,
                                    const float *__restrict__ fcounts, uint64_t occupied, uint64_t nx,
                                    uint64_t ny, uint64_t nz, unsigned long long *__restrict__ hits,
                                    double *__restrict__ terms) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Offset o = offs[i];
    const uint64_t valid = (nx - (uint64_t)abs(o.dx)) * (ny - (uint64_t)abs(o.dy)) * (nz - (uint64_t)abs(o.dz));
    double h;
    if (fcounts) {
        h = (double)fcounts[i];
    } else {
        const int32_t g = gather[i];
        const uint64_t hv = g >= 0 ? table[g] : occupied;
        hits[i] = hv;
        h = (double)hv;
    }
    terms[i] = h / (double)valid;
}

// One workgroup per radius: stage the radius' terms through LDS and fold them
// in shell order on lane 0 (the reference's sequential left fold).
static constexpr int kFoldChunk = 2048;

__global__ __launch_bounds__(256) void fold_kernel(const double *__restrict__ terms,
                                                   const uint64_t *__restrict__ rstart, uint32_t r_max,
                                                   double *__restrict__ values, uint8_t *__restrict__ support) {
    __shared__ double buf[kFoldChunk];
    const uint32_t r = blockIdx.x + 1;
    if (r > r_max) return;
    const uint64_t s = rstart[r], e = rstart[r + 1];
    double sum = 0.0;
    for (uint64_t c0 = s; c0 < e; c0 += kFoldChunk) {
        const uint64_t len = (e - c0) < (uint64_t)kFoldChunk ? (e - c0) : (uint64_t)kFoldChunk;
        __syncthreads();
        for (uint64_t k = threadIdx.x; k < len; k += blockDim.x) buf[k] = terms[c0 + k];
        __syncthreads();
        if (threadIdx.x == 0) {
            for (uint64_t k = 0; k < len; ++k) sum = sum + buf[k];
        }
    }
    if (threadIdx.x == 0) {
        if (e > s) {
            values[r] = sum / (double)(e - s);
            support[r] = 1;
        } else {
            values[r] = 0.0;
            support[r] = 0;
        }
    }
}

// ---------------------------------------------------------------- plan (host)

static int supported_nw(int need) {
    static const int sizes[] = {1, 2, 3, 4, 6, 8, 10, 12, 16, 20, 24, 32};
    for (int s : sizes)
        if (s >= need) return s;
    return -1;
}

static s2h_status build_plan(s2h_ctx *ctx, ExactPlan *p) {
    shell_order_offsets(p->dims, p->r_max, p->pitch, p->offs, p->radius_start);
    const uint64_t K = p->offs.size();
    const int nx = (int)p->dims[0], ny = (int)p->dims[1], nz = (int)p->dims[2];
    // canonical half ball: column (dx, dy) -> largest dz
    int dmax = 0;
    for (const Offset &o : p->offs) dmax = std::max({dmax, std::abs(o.dx), std::abs(o.dy)});
    const int span = 2 * dmax + 1;
    std::vector<int32_t> colm((size_t)span * span, -1);
    auto canon = [](Offset o) {
        if (o.dz < 0 || (o.dz == 0 && (o.dx < 0 || (o.dx == 0 && o.dy < 0)))) return Offset{-o.dx, -o.dy, -o.dz};
        return o;
    };
    for (const Offset &o0 : p->offs) {
        if (o0.dx == 0 && o0.dy == 0 && o0.dz == 0) continue;
        Offset o = canon(o0);
        int32_t &m = colm[(size_t)(o.dx + dmax) * span + (o.dy + dmax)];
        m = std::max(m, o.dz);
    }
    struct Col {
        int dx, dy, m;
        double ang;
    };
    std::vector<Col> cols;
    for (int dx = -dmax; dx <= dmax; ++dx)
        for (int dy = -dmax; dy <= dmax; ++dy) {
            int m = colm[(size_t)(dx + dmax) * span + (dy + dmax)];
            if (m >= 0) cols.push_back({dx, dy, m, std::atan2((double)dy, (double)dx)});
        }
    // dense table: column c holds dz = 0..m
    std::vector<uint32_t> base(cols.size());
    std::vector<int32_t> colidx((size_t)span * span, -1);
    uint64_t tsize = 0;
    for (size_t i = 0; i < cols.size(); ++i) {
        base[i] = (uint32_t)tsize;
        colidx[(size_t)(cols[i].dx + dmax) * span + (cols[i].dy + dmax)] = (int32_t)i;
        tsize += (uint64_t)cols[i].m + 1;
    }
    if (tsize >= (1ull << 31)) return fail(S2H_ERR_UNSUPPORTED, "offset table too large");
    p->table_size = tsize;
    std::vector<int32_t> gather(K);
    for (uint64_t i = 0; i < K; ++i) {
        const Offset &o0 = p->offs[i];
        if (o0.dx == 0 && o0.dy == 0 && o0.dz == 0) {
            gather[i] = -1;
            continue;
        }
        Offset o = canon(o0);
        int32_t ci = colidx[(size_t)(o.dx + dmax) * span + (o.dy + dmax)];
        gather[i] = (int32_t)(base[ci] + (uint32_t)o.dz);
    }
    // groups of 64 columns with nearly equal m (then by angle, keeping each
    // group's (dx, dy) spread, and so its idle boundary rows, small)
    std::vector<size_t> order(cols.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        if (cols[a].m != cols[b].m) return cols[a].m > cols[b].m;
        return cols[a].ang < cols[b].ang;
    });
    const size_t ngroups = (cols.size() + 63) / 64;
    p->ngroups = (uint32_t)ngroups;
    std::vector<ColInfo> ci(ngroups * 64, ColInfo{0, 0, -1, 0});
    for (size_t k = 0; k < order.size(); ++k) {
        const Col &c = cols[order[k]];
        ci[k] = ColInfo{c.dx, c.dy, c.m, base[order[k]]};
    }
    // work items
    const int need_dw = (nz + 31) / 32;
    for (auto &v : p->items) v.clear();
    // x chunking: aim for enough waves to fill 256 CUs several times over
    double total_rows = 0;
    for (size_t g = 0; g < ngroups; ++g) total_rows += (double)nx * ny;
    const int chunk = std::max(1, std::min(nx, (int)std::ceil((double)nx * ngroups * 3 / 16384.0)));
    for (size_t g = 0; g < ngroups; ++g) {
        int M = -1, dxmin = 1 << 30, dxmax = -(1 << 30), dymin = 1 << 30, dymax = -(1 << 30);
        for (int l = 0; l < 64; ++l) {
            const ColInfo &c = ci[g * 64 + l];
            if (c.m < 0) continue;
            M = std::max(M, c.m);
            dxmin = std::min(dxmin, c.dx);
            dxmax = std::max(dxmax, c.dx);
            dymin = std::min(dymin, c.dy);
            dymax = std::max(dymax, c.dy);
        }
        if (M < 0) continue;
        // b row p valid for some lane: xb - dx in [0, nx) for some dx in the group
        const int x0 = std::max(0, dxmin), x1 = std::min(nx, nx + dxmax);
        const int y0 = std::max(0, dymin), y1 = std::min(ny, ny + dymax);
        if (x0 >= x1 || y0 >= y1) continue;
        for (int dz0 = 0; dz0 <= M; dz0 += 32) {
            const int rem = M + 1 - dz0;
            const int dzb = rem > 16 ? 32 : (rem > 8 ? 16 : 8);
            const int slot = dzb == 8 ? 0 : (dzb == 16 ? 1 : 2);
            const int nwi = std::max(0, need_dw - dz0 / 32);
            if (nwi == 0) continue;
            for (int xs = x0; xs < x1; xs += chunk) {
                WorkItem w;
                w.group = (uint32_t)g;
                w.dz0 = (uint16_t)dz0;
                w.dzb = (uint16_t)dzb;
                w.x0 = (uint16_t)xs;
                w.x1 = (uint16_t)std::min(x1, xs + chunk);
                w.y0 = (uint16_t)y0;
                w.y1 = (uint16_t)y1;
                w.nw = (uint16_t)nwi;
                w.pad = 0;
                p->items[slot].push_back(w);
            }
        }
    }
    (void)total_rows;
    // device copies
    std::vector<uint64_t> rs(p->radius_start.begin(), p->radius_start.end());
    S2H_TRY(upload(ctx, p->d_offs, p->offs.data(), K * sizeof(Offset)));
    S2H_TRY(upload(ctx, p->d_gather, gather.data(), K * sizeof(int32_t)));
    S2H_TRY(upload(ctx, p->d_rstart, rs.data(), rs.size() * 8));
    S2H_TRY(upload(ctx, p->d_cols, ci.data(), ci.size() * sizeof(ColInfo)));
    for (int s = 0; s < 3; ++s)
        S2H_TRY(upload(ctx, p->d_items[s], p->items[s].data(), p->items[s].size() * sizeof(WorkItem)));
    S2H_TRY(dev_alloc(ctx, p->d_table, std::max<uint64_t>(tsize, 1) * 8));
    S2H_TRY(dev_alloc(ctx, p->d_terms, std::max<uint64_t>(K, 1) * 8));
    S2H_TRY(dev_alloc(ctx, p->d_hits, std::max<uint64_t>(K, 1) * 8));
    S2H_TRY(dev_alloc(ctx, p->d_curve, (p->r_max + 1) * 9 + 16));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    return S2H_OK;
}

s2h_status exact_plan_get(s2h_ctx *ctx, const uint64_t dims[3], uint32_t r_max, double pitch,
                          ExactPlan **out) {
    if (!(pitch > 0.0) || !std::isfinite(pitch)) return fail(S2H_ERR_ARG, "pitch must be finite and > 0");
    uint64_t pbits;
    std::memcpy(&pbits, &pitch, 8);
    auto key = std::make_tuple(dims[0], dims[1], dims[2], r_max, pbits);
    auto itr = ctx->exact_plans.find(key);
    if (itr != ctx->exact_plans.end()) {
        *out = itr->second.get();
        return S2H_OK;
    }
    auto p = std::make_unique<ExactPlan>();
    for (int a = 0; a < 3; ++a) p->dims[a] = dims[a];
    p->r_max = r_max;
    p->pitch = pitch;
    s2h_status st = build_plan(ctx, p.get());
    if (st != S2H_OK) {
        DevBuf *bufs[] = {&p->d_offs, &p->d_gather, &p->d_rstart, &p->d_cols, &p->d_items[0],
                          &p->d_items[1], &p->d_items[2], &p->d_table, &p->d_terms, &p->d_hits,
                          &p->d_curve};
        for (DevBuf *b : bufs) dev_free(ctx, *b);
        return st;
    }
    *out = p.get();
    ctx->exact_plans[key] = std::move(p);
    return S2H_OK;
}

template <int NW>
static s2h_status launch_counts_nw(s2h_ctx *ctx, s2h_vol *vol, ExactPlan *p) {
    const int row_dw = (int)(vol->row_words * 2);
    const uint32_t *v = (const uint32_t *)vol->d_words;
    auto *table = (unsigned long long *)p->d_table.ptr;
    const ColInfo *cols = (const ColInfo *)p->d_cols.ptr;
    const int nx = (int)vol->dims[0], ny = (int)vol->dims[1];
    const uint32_t vbytes = (uint32_t)(vol->nwords * 8);
    for (int s = 0; s < 3; ++s) {
        const int n = (int)p->items[s].size();
        if (n == 0) continue;
        const WorkItem *items = (const WorkItem *)p->d_items[s].ptr;
        const unsigned blocks = (unsigned)((n + kWavesPerBlock - 1) / kWavesPerBlock);
        const dim3 block(64 * kWavesPerBlock);
        if (s == 0)
            exact_count_kernel<NW, 8><<<blocks, block, 0, ctx->stream>>>(v, vbytes, nx, ny, row_dw, items, n, cols, table);
        else if (s == 1)
            exact_count_kernel<NW, 16><<<blocks, block, 0, ctx->stream>>>(v, vbytes, nx, ny, row_dw, items, n, cols, table);
        else
            exact_count_kernel<NW, 32><<<blocks, block, 0, ctx->stream>>>(v, vbytes, nx, ny, row_dw, items, n, cols, table);
        S2H_HIP(hipGetLastError());
        ++ctx->launches;
    }
    return S2H_OK;
}

s2h_status exact_run_counts(s2h_ctx *ctx, s2h_vol *vol, ExactPlan *p) {
    if (vol->nwords * 8 + kVolTailBytes >= (1ull << 31))
        return fail(S2H_ERR_UNSUPPORTED, "packed volume above 2 GiB");
    S2H_HIP(hipMemsetAsync(p->d_table.ptr, 0, std::max<uint64_t>(p->table_size, 1) * 8, ctx->stream));
    prof_mark(ctx, 0);
    const int need = (int)((vol->dims[2] + 31) / 32);
    switch (supported_nw(need)) {
        case 1: return launch_counts_nw<1>(ctx, vol, p);
        case 2: return launch_counts_nw<2>(ctx, vol, p);
        case 3: return launch_counts_nw<3>(ctx, vol, p);
        case 4: return launch_counts_nw<4>(ctx, vol, p);
        case 6: return launch_counts_nw<6>(ctx, vol, p);
        case 8: return launch_counts_nw<8>(ctx, vol, p);
        case 10: return launch_counts_nw<10>(ctx, vol, p);
        case 12: return launch_counts_nw<12>(ctx, vol, p);
        case 16: return launch_counts_nw<16>(ctx, vol, p);
        case 20: return launch_counts_nw<20>(ctx, vol, p);
        case 24: return launch_counts_nw<24>(ctx, vol, p);
        case 32: return launch_counts_nw<32>(ctx, vol, p);
        default:
            return fail(S2H_ERR_UNSUPPORTED, "nz=" + std::to_string(vol->dims[2]) + " exceeds 1024");
    }
}

s2h_status exact_gather_fold(s2h_ctx *ctx, ExactPlan *p, uint64_t occupied, const float *d_fcounts) {
    const uint64_t K = p->offs.size();
    prof_mark(ctx, 1);
    if (K) {
        gather_terms_kernel<<<(unsigned)((K + 255) / 256), 256, 0, ctx->stream>>>(
            (const Offset *)p->d_offs.ptr, (const int32_t *)p->d_gather.ptr, K,
            (const unsigned long long *)p->d_table.ptr, d_fcounts, occupied, p->dims[0], p->dims[1],
            p->dims[2], (unsigned long long *)p->d_hits.ptr, (double *)p->d_terms.ptr);
        S2H_HIP(hipGetLastError());
    }
    double *values = (double *)p->d_curve.ptr;
    uint8_t *support = (uint8_t *)(values + p->r_max + 1);
    if (p->r_max > 0) {
        fold_kernel<<<p->r_max, 256, 0, ctx->stream>>>((const double *)p->d_terms.ptr,
                                                       (const uint64_t *)p->d_rstart.ptr, p->r_max,
                                                       values, support);
        S2H_HIP(hipGetLastError());
    }
    prof_mark(ctx, 2);
    return S2H_OK;
}

s2h_status finish_curve_from_device(s2h_ctx *ctx, ExactPlan *p, double vf, double *s2) {
    const uint32_t n = p->r_max + 1;
    std::vector<uint8_t> host((size_t)n * 9);
    S2H_HIP(hipMemcpyAsync(host.data(), p->d_curve.ptr, (size_t)n * 9, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    std::vector<uint8_t> sup(n, 0);
    std::memcpy(s2, host.data(), (size_t)n * 8);
    std::memcpy(sup.data(), host.data() + (size_t)n * 8, n);
    s2[0] = vf;
    sup[0] = 1;
    fill_missing(s2, sup.data(), n, vf);
    s2[0] = vf;
    return S2H_OK;
}

}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_exact_counts(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, uint64_t *hits,
                            uint64_t n_offsets) {
    if (!ctx || !vol || (!hits && n_offsets)) return fail(S2H_ERR_ARG, "null argument");
    S2H_HIP(hipSetDevice(ctx->device));
    ExactPlan *p = nullptr;
    S2H_TRY(exact_plan_get(ctx, vol->dims, r_max, pitch, &p));
    if (p->offs.size() != n_offsets)
        return fail(S2H_ERR_SIZE, "hits buffer holds " + std::to_string(n_offsets) + " entries, plan has " +
                                      std::to_string(p->offs.size()));
    uint64_t occupied = 0;
    S2H_TRY(vol_popcount(ctx, vol, &occupied));
    S2H_TRY(exact_run_counts(ctx, vol, p));
    S2H_TRY(exact_gather_fold(ctx, p, occupied, nullptr));
    if (n_offsets)
        S2H_HIP(hipMemcpyAsync(hits, p->d_hits.ptr, n_offsets * 8, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    return S2H_OK;
}

s2h_status s2h_exact_s2(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, double *s2, uint64_t n) {
    if (!ctx || !vol || !s2) return fail(S2H_ERR_ARG, "null argument");
    if (n != (uint64_t)r_max + 1) return fail(S2H_ERR_SIZE, "s2 buffer must hold r_max+1 values");
    S2H_HIP(hipSetDevice(ctx->device));
    uint64_t occupied = 0;
    S2H_TRY(vol_popcount(ctx, vol, &occupied));
    if (occupied == 0) {
        std::fill(s2, s2 + n, 0.0);
        return S2H_OK;
    }
    const double pm = pitch < 1e-9 ? 1e-9 : pitch;
    ExactPlan *p = nullptr;
    S2H_TRY(exact_plan_get(ctx, vol->dims, r_max, pm, &p));
    const double vf = (double)occupied / (double)(vol->dims[0] * vol->dims[1] * vol->dims[2]);
    S2H_TRY(exact_run_counts(ctx, vol, p));
    S2H_TRY(exact_gather_fold(ctx, p, occupied, nullptr));
    return finish_curve_from_device(ctx, p, vf, s2);
}

}  // extern "C"
