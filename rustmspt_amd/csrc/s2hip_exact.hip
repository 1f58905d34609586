// Exact S2 pair counting on a bit-packed volume (gfx950).
//
// Counts C(d) = #{y : occ(y) and occ(y + d)} for every in-domain offset d of the
// shells 1..r_max.  C(d) = C(-d), so only the canonical half ball is counted:
// dz > 0, or dz == 0 with (dx, dy) lexicographically positive.  The half ball is
// a set of columns (dx, dy), each with dz = 0..m(dx, dy).
//
// Kernel shape (one wave per work item):
//   * lanes    = 64 columns (dx, dy) with nearly equal m (sorted by m);
//   * uniform  = the partner ("b") row p = (xb, yb) and a block of 4..32 dz
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
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>

#include "s2hip_internal.h"

namespace s2h {

static constexpr int kWavesPerBlock = 4;

// ---------------------------------------------------------------- count kernel

// Volume buffers carry kVolTailBytes of zero padding after the last row, so an
// unguarded read of up to 64 dwords past any row start stays inside the
// allocation; bits that do not belong to the row are masked to zero.
// The workgroup's flush: items are ordered by cost, so the waves of one
// workgroup usually hold x-chunks of the same (group, dz block).  Then wave 0
// adds the four waves' counters (through LDS) and issues one atomic per
// (lane, dz) instead of four; otherwise every wave adds its own.  Every wave
// of the workgroup passes both barriers exactly once.
struct FlushLds {
    uint32_t (*red)[32][64];  // counters of waves 1 .. kWavesPerBlock-1: [wave - 1][dz][lane]
    const uint32_t *key;      // per wave: (group, dz0) of its item, or a value no item has
    int w;                    // this wave's index in the workgroup
};

template <int DZB>
__device__ __forceinline__ void flush_counts(const uint32_t (&cnt)[DZB], const WorkItem &it, const ColInfo &c,
                                             unsigned long long *__restrict__ table, const FlushLds &f) {
    const int lane = threadIdx.x & 63;
    if (f.w > 0) {
#pragma unroll
        for (int j = 0; j < DZB; ++j) f.red[f.w - 1][j][lane] = cnt[j];
    }
    __syncthreads();
    bool same = true;
#pragma unroll
    for (int q = 1; q < kWavesPerBlock; ++q) same = same && f.key[q] == f.key[0];
    if (same) {
        if (f.w == 0) {
#pragma unroll
            for (int j = 0; j < DZB; ++j) {
                unsigned long long sum = cnt[j];
#pragma unroll
                for (int q = 0; q < kWavesPerBlock - 1; ++q) sum += f.red[q][j][lane];
                const int dz = it.dz0 + j;
                if (c.m >= 0 && dz <= c.m && sum) atomicAdd(&table[c.base + dz], sum);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < DZB; ++j) {
            const int dz = it.dz0 + j;
            if (c.m >= 0 && dz <= c.m && cnt[j]) atomicAdd(&table[c.base + dz], (unsigned long long)cnt[j]);
        }
    }
    __syncthreads();
}

template <int NW, int DZB>
__device__ __forceinline__ void count_item(const uint32_t *__restrict__ vol, const __amdgpu_buffer_rsrc_t rsrc, int nx,
                                           int ny, int row_dw, const WorkItem &it, const ColInfo &c,
                                           unsigned long long *__restrict__ table, const FlushLds &f) {
    const int q0 = it.dz0 >> 5;
    const int nw = it.nw;  // partner dwords q0 .. q0+nw-1 hold voxels; later ones read as 0
    const int oob = 0x7fffff00;  // + up to 255 bytes: past any buffer, no int overflow

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
                for (int j = 0; j < DZB; ++j) cnt[j] = popc_acc(a[w] & (uint32_t)(pair >> j), cnt[j]);
            }
        }
    }
    flush_counts<DZB>(cnt, it, c, table, f);
}

// dz blocks are 4..32 wide in steps of 4 (the wave-uniform width of the item).
template <int NW>
__device__ __forceinline__ void count_dzb(const uint32_t *__restrict__ vol, const __amdgpu_buffer_rsrc_t rsrc, int nx,
                                          int ny, int row_dw, const WorkItem &it, const ColInfo &c,
                                          unsigned long long *__restrict__ table, const FlushLds &f) {
    switch (it.dzb) {
        case 4: count_item<NW, 4>(vol, rsrc, nx, ny, row_dw, it, c, table, f); break;
        case 8: count_item<NW, 8>(vol, rsrc, nx, ny, row_dw, it, c, table, f); break;
        case 12: count_item<NW, 12>(vol, rsrc, nx, ny, row_dw, it, c, table, f); break;
        case 16: count_item<NW, 16>(vol, rsrc, nx, ny, row_dw, it, c, table, f); break;
        case 20: count_item<NW, 20>(vol, rsrc, nx, ny, row_dw, it, c, table, f); break;
        case 24: count_item<NW, 24>(vol, rsrc, nx, ny, row_dw, it, c, table, f); break;
        case 28: count_item<NW, 28>(vol, rsrc, nx, ny, row_dw, it, c, table, f); break;
        default: count_item<NW, 32>(vol, rsrc, nx, ny, row_dw, it, c, table, f); break;
    }
}

// One wave per work item; every dz-block width runs in the same launch (the
// branch is wave-uniform), items are ordered heaviest first.
template <int NW>
__global__ __launch_bounds__(64 * kWavesPerBlock) void exact_count_kernel(
    const uint32_t *__restrict__ vol, uint32_t vol_bytes, int nx, int ny, int row_dw,
    const WorkItem *__restrict__ items, int nitems, const ColInfo *__restrict__ cols,
    unsigned long long *__restrict__ table) {
    __shared__ uint32_t s_red[kWavesPerBlock - 1][32][64];
    __shared__ uint32_t s_key[kWavesPerBlock];
    const int w = (int)(threadIdx.x / 64);
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWavesPerBlock + w));
    const int lane = threadIdx.x & 63;
    if (wave >= nitems) {  // no item: a key no item has, then the flush's two barriers
        if (lane == 0) s_key[w] = 0xffffffffu - (uint32_t)w;
        __syncthreads();
        __syncthreads();
        return;
    }
    const WorkItem it = items[wave];
    if (lane == 0) s_key[w] = (it.group << 10) | it.dz0;  // dz0 < 1024
    const ColInfo c = cols[it.group * 64 + lane];
    // per-lane rows through a buffer resource: an out-of-range offset reads 0
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)vol, 0, (int)vol_bytes, 0x00020000);
    // Blocks whose partner words end one dword early (dz0 >= 32, or a row of
    // fewer dwords than NW) run the NW - 1 word body: the last word would
    // only meet zeros.
    const FlushLds f{s_red, s_key, w};
    if (NW > 1 && it.nw < NW)
        count_dzb<(NW > 1 ? NW - 1 : 1)>(vol, rsrc, nx, ny, row_dw, it, c, table, f);
    else
        count_dzb<NW>(vol, rsrc, nx, ny, row_dw, it, c, table, f);
}

// ---------------------------------------------------------------- gather / fold

// The ordered fold of segments of terms (the curve's radii, or the segments
// of s2h_ordered_sum): each segment's sequential left fold, bit for bit.  The
// terms are cut into blocks of 128 (64 lanes x kFoldK), numbered across all
// segments.  Inside one binade of the running sum S (grid
// u = 2^(ilogb(S) - 52)), a block's effect on S / u is the sum G of its terms'
// rounded grid increments -- which depends on u and the terms only -- as long
// as no term is a tie on the grid or too large and S stays in the binade
// (fold_block's non-stopping case).  Three launches:
//   fold_sums_kernel   one wave per block: the block's plain f64 sum (any
//                      order) and whether it holds a negative or non-finite
//                      term;
//   fold_incr_kernel   one wave per block: the plain sums of the segment's
//                      earlier blocks (any order) guess the binade at the
//                      block's start; G on that grid (ties counted at their
//                      floor), whether a large term (y >= 2^43, so every sum
//                      of 128 is exact) occurs, and up to kFoldTies ties with
//                      the increments before each;
//   fold_walk_kernel   one wave per segment, in order: a block whose guess
//                      matches the exact S's binade, with no flag and
//                      S / u + G < 2^53, adds G (the same integers fold_block
//                      would add); a block with ties adds G plus, tie by tie,
//                      1 where S / u before the tie plus its floor is odd
//                      (round half to even); any other block runs fold_block itself, or
//                      one term at a time after a negative or non-finite term
//                      (as a one-wave fold would, block for block).
// The guesses only pick which blocks take the short path; every value comes
// from exact arithmetic.  The first two launches spread over the whole GPU; the
// walk's time is its blocks that are not short (binade crossings, the first
// blocks of a segment, ties), which cost least at 128 terms per block.
static constexpr int kFoldK = 2;                         // terms per lane per block
static constexpr uint32_t kFoldB = 64 * kFoldK;          // terms per block
static constexpr int kFoldWaves = 4;                     // waves per workgroup of the per-block launches
static constexpr int kNoBinade = -100000;                // guessed start S == 0
static constexpr int kFoldTies = 4;                      // ties a block may hold and still skip fold_block
static constexpr uint32_t kFoldTieFlag = 4;              // flag 4 + t: t ties, recorded in FoldView::ties

__device__ __forceinline__ double wave_sum_any(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ void load_block(const double *__restrict__ t, uint64_t n, uint64_t start, uint32_t lane,
                                           double (&v)[kFoldK]) {
#pragma unroll
    for (int i = 0; i < kFoldK; ++i) {
        const uint64_t j = start + (uint64_t)lane * kFoldK + i;
        const double x = t[j < n ? j : 0];
        v[i] = j < n ? x : 0.0;
    }
}

// A block's terms in any order (lane-interleaved, so each load instruction
// reads 64 consecutive terms): for the launches whose results do not depend on
// the order -- the plain block sum that only guesses binades, the exact
// integer increments and the flags.
__device__ __forceinline__ void load_block_any(const double *__restrict__ t, uint64_t n, uint64_t start, uint32_t lane,
                                               double (&v)[kFoldK]) {
#pragma unroll
    for (int i = 0; i < kFoldK; ++i) {
        const uint64_t j = start + (uint64_t)i * 64 + lane;
        const double x = t[j < n ? j : 0];
        v[i] = j < n ? x : 0.0;
    }
}

// Device tables and per-block scratch of one segment list: segment g owns
// blocks bbase[g] .. bbase[g+1] - 1, block b belongs to segment bseg[b].
struct FoldView {
    const uint32_t *bbase;
    const uint32_t *bseg;
    double *sum;     // plain block sums
    double *inc;     // grid increments G on the guessed grid
    int32_t *m;      // guessed binade at the block's start
    uint32_t *flag;  // 2: negative or non-finite term; 1: large term or too many ties; 4 + t: t ties
    double *ties;    // per block, kFoldTies pairs (increments before the tie, the tie's floor)
    uint32_t nblk;
};

// The wave's block: its segment's first term s0, the segment length n and the
// block's first term within the segment.
__device__ __forceinline__ uint32_t fold_block_of(const uint64_t *__restrict__ seg_start, const FoldView &v,
                                                  uint32_t w, uint64_t &s0, uint64_t &n, uint32_t &i) {
    const uint32_t seg = v.bseg[w];
    s0 = seg_start[seg];
    n = seg_start[seg + 1] - s0;
    i = w - v.bbase[seg];
    return seg;
}

__global__ __launch_bounds__(64 * kFoldWaves) void fold_sums_kernel(const double *__restrict__ terms,
                                                                     const uint64_t *__restrict__ seg_start,
                                                                     FoldView v) {
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * kFoldWaves + threadIdx.x / 64));
    if (w >= v.nblk) return;
    const uint32_t lane = threadIdx.x & 63;
    uint64_t s0, n;
    uint32_t i;
    fold_block_of(seg_start, v, w, s0, n, i);
    double x[kFoldK];
    load_block_any(terms + s0, n, (uint64_t)i * kFoldB, lane, x);
    double sum = 0.0;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < kFoldK; ++k) {
        sum += x[k];
        bad = bad || !(x[k] >= 0.0 && x[k] <= 1.7976931348623157e308);
    }
    sum = wave_sum_any(sum);
    const bool anybad = __ballot(bad) != 0;
    if (lane == 0) {
        v.sum[w] = sum;
        v.flag[w] = anybad ? 2u : 0u;
    }
}

// term = hits / valid in f64 (division correctly rounded, as Rust's `/`) of
// every offset, one wave per fold block, with the block's plain sum and flag
// (fold_sums_kernel's outputs).  With an FFT cube (dx >= 0 only; dx < 0 reads
// the mirror) the hits are read at the offset's cube position: FP32 values as they are, FP64 values rounded to the
// nearest integer, at least 0 (`.re.round().max(0.0)` of the reference's FFT
// kernel, src/geometry/s2.rs:1245-1248); else from the count table through
// the offset's gather index (the zero offset: the occupied count).  hits, when
// set, receives the integer counts (not for an FP32 cube).
__global__ __launch_bounds__(64 * kFoldWaves) void terms_kernel(
    const Offset *__restrict__ offs, const int32_t *__restrict__ gather, const unsigned long long *__restrict__ table,
    const float *__restrict__ fcube, const double *__restrict__ dcube, int Rx, int Ry, int Rz, uint64_t occupied,
    uint64_t nx, uint64_t ny, uint64_t nz, unsigned long long *__restrict__ hits, double *__restrict__ terms,
    const uint64_t *__restrict__ seg_start, FoldView v) {
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * kFoldWaves + threadIdx.x / 64));
    if (w >= v.nblk) return;
    const uint32_t lane = threadIdx.x & 63;
    uint64_t s0, n;
    uint32_t i;
    fold_block_of(seg_start, v, w, s0, n, i);
    double sum = 0.0;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < kFoldK; ++k) {
        const uint64_t j = (uint64_t)i * kFoldB + (uint64_t)k * 64 + lane;
        if (j >= n) continue;
        const uint64_t gi = s0 + j;
        const Offset o = offs[gi];
        const uint64_t valid =
            (nx - (uint64_t)abs(o.dx)) * (ny - (uint64_t)abs(o.dy)) * (nz - (uint64_t)abs(o.dz));
        const uint64_t ci = (fcube || dcube) ? half_cube_index(o, Ry, Rz) : 0;
        double h;
        if (fcube) {
            h = (double)fcube[ci];
        } else if (dcube) {
            h = fmax(round(dcube[ci]), 0.0);
            if (hits) hits[gi] = (unsigned long long)h;
        } else {
            const int32_t g = gather[gi];
            const uint64_t hv = g >= 0 ? table[g] : occupied;
            if (hits) hits[gi] = hv;
            h = (double)hv;
        }
        const double t = h / (double)valid;
        terms[gi] = t;
        sum += t;
        bad = bad || !(t >= 0.0 && t <= 1.7976931348623157e308);
    }
    sum = wave_sum_any(sum);
    const bool anybad = __ballot(bad) != 0;
    if (lane == 0) {
        v.sum[w] = sum;
        v.flag[w] = anybad ? 2u : 0u;
    }
}

__global__ __launch_bounds__(64 * kFoldWaves) void fold_incr_kernel(const double *__restrict__ terms,
                                                                     const uint64_t *__restrict__ seg_start,
                                                                     FoldView v) {
    constexpr double kTwo43 = 8796093022208.0;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * kFoldWaves + threadIdx.x / 64));
    if (w >= v.nblk) return;
    const uint32_t lane = threadIdx.x & 63;
    uint64_t s0, n;
    uint32_t i;
    fold_block_of(seg_start, v, w, s0, n, i);
    const uint32_t b0 = w - i;
    double pre = 0.0;  // the earlier blocks' plain sums, any order: a guess of S at the block's start
    for (uint32_t j = lane; j < i; j += 64) pre += v.sum[b0 + j];
    pre = wave_sum_any(pre);
    const int m = (pre > 0.0 && pre <= 1.7976931348623157e308) ? ilogb(pre) : kNoBinade;
    uint32_t flag = v.flag[w];
    double T = 0.0;
    if (flag == 0 && m != kNoBinade) {
        // shell order: lane l holds terms l K .. l K + K - 1
        double x[kFoldK];
        load_block(terms + s0, n, (uint64_t)i * kFoldB, lane, x);
        const double inv_u = ldexp(1.0, 52 - m);
        bool big = false;
        bool tie[kFoldK];
        double q[kFoldK], nfl[kFoldK];
#pragma unroll
        for (int k = 0; k < kFoldK; ++k) {
            const double y = x[k] * inv_u;  // exact: scaling by a power of two (0 past the end)
            const bool b = !(y < kTwo43);
            nfl[k] = b ? 0.0 : floor(y);
            const double f = y - nfl[k];
            q[k] = nfl[k] + (f > 0.5 ? 1.0 : 0.0);
            tie[k] = !b && f == 0.5;
            big = big || b;
            T += q[k];
        }
        // increments of the lanes before (integers below 2^53: exact)
        const double before = wave_exclusive_sum_f64(T);
        T = wave_sum_any(T);
        uint32_t nt = 0;
        uint64_t tm[kFoldK];
#pragma unroll
        for (int k = 0; k < kFoldK; ++k) {
            tm[k] = __ballot(tie[k]);
            nt += (uint32_t)__popcll(tm[k]);
        }
        if (__ballot(big) != 0 || nt > (uint32_t)kFoldTies) {
            flag = 1u;
        } else if (nt > 0) {
            // the ties in shell order: (increments of the block's terms before
            // the tie, the tie's floor) -- the walk rounds each to even
            flag = kFoldTieFlag + nt;
            double pre_k = before;
            double pk[kFoldK];
#pragma unroll
            for (int k = 0; k < kFoldK; ++k) {
                pk[k] = pre_k;
                pre_k += q[k];
            }
            double *tz = v.ties + (uint64_t)w * (2 * kFoldTies);
            for (uint32_t e = 0; e < nt; ++e) {
                // the earliest tie left: lowest lane, then lowest k within it
                int bl = 64, bk = 0;
#pragma unroll
                for (int k = 0; k < kFoldK; ++k) {
                    if (tm[k]) {
                        const int l = __ffsll((long long)tm[k]) - 1;
                        if (l < bl) {
                            bl = l;
                            bk = k;
                        }
                    }
                }
                double pv = 0.0, nv = 0.0;
#pragma unroll
                for (int k = 0; k < kFoldK; ++k)
                    if (k == bk) {
                        pv = read_lane_f64(pk[k], bl);
                        nv = read_lane_f64(nfl[k], bl);
                        tm[k] &= tm[k] - 1;
                    }
                if (lane == 0) {
                    tz[2 * e] = pv;
                    tz[2 * e + 1] = nv;
                }
            }
        }
    }
    if (lane == 0) {
        v.inc[w] = T;
        v.m[w] = m;
        v.flag[w] = flag;
    }
}

// The walk of one segment on one wave (all 64 lanes; the result is
// wave-uniform).
__device__ __forceinline__ double fold_walk(const double *__restrict__ t, uint64_t n, const FoldView &v, uint32_t b0,
                                            uint32_t nb, uint32_t lane) {
    constexpr double kTwo53 = 9007199254740992.0;
    FoldState st;
    bool seq = false;
    for (uint32_t g0 = 0; g0 < nb; g0 += 64) {
        const uint32_t bl = g0 + lane;
        double gl = 0.0;
        int ml = kNoBinade;
        uint32_t fl = 1;
        if (bl < nb) {
            gl = v.inc[b0 + bl];
            ml = v.m[b0 + bl];
            fl = v.flag[b0 + bl];
        }
        const uint32_t cntb = min(64u, nb - g0);
        for (uint32_t k = 0; k < cntb; ++k) {
            // The blocks from k on that take the short path, all at once:
            // lane j's block is short when its guess names the exact binade
            // (unchanged while the blocks before it are short), it has no
            // flag, and S / u plus the increments of blocks k .. j stays below
            // 2^53.  Their increments are added in one step, up to the first
            // block that is not.
            const bool live = lane >= k && lane < cntb;
            const double inc = wave_inclusive_sum_f64(live ? gl : 0.0);
            const bool shortp = live && !seq && fl == 0 && ml != kNoBinade && st.S > 0.0 &&
                                ldexp(1.0, ml - 52) == st.u && st.a + inc < kTwo53;
            const uint64_t stops = __ballot(!shortp) & (~0ull << k);
            const uint32_t j = stops ? (uint32_t)(__ffsll((long long)stops) - 1) : 64u;
            if (j > k) {
                st.a += read_lane_f64(inc, (int)(j - 1));
                st.S = st.a * st.u;
            }
            if (j >= cntb) break;
            k = j;  // block j runs the exact path
            // a block with a few ties, on the exact binade: its increments,
            // and each tie rounded to the even neighbour of S / u before it
            const uint32_t fj = (uint32_t)__builtin_amdgcn_readlane((int)fl, (int)j);
            if (fj > kFoldTieFlag && !seq && st.S > 0.0 &&
                ldexp(1.0, __builtin_amdgcn_readlane(ml, (int)j) - 52) == st.u) {
                const uint32_t nt = fj - kFoldTieFlag;
                const double gj = read_lane_f64(gl, (int)j);
                if (st.a + (gj + (double)nt) < kTwo53) {
                    const double *tz = v.ties + (uint64_t)(b0 + g0 + j) * (2 * kFoldTies);
                    double extra = 0.0;
                    for (uint32_t e = 0; e < nt; ++e) {
                        const double x = st.a + tz[2 * e] + extra + tz[2 * e + 1];  // an integer below 2^53
                        extra += (double)((uint64_t)x & 1ull);
                    }
                    st.a += gj + extra;
                    st.S = st.a * st.u;
                    continue;
                }
            }
            const uint64_t start = (uint64_t)(g0 + k) * kFoldB;
            const uint32_t cnt = (uint32_t)min<uint64_t>(kFoldB, n - start);
            double x[kFoldK];
            load_block(t, n, start, lane, x);
            bool bad = false;
#pragma unroll
            for (int i = 0; i < kFoldK; ++i)
                bad = bad || (lane * kFoldK + i < cnt && !(x[i] >= 0.0 && x[i] <= 1.7976931348623157e308));
            if (!seq && __ballot(bad)) seq = true;
            if (seq) {
                for (uint32_t l = 0; l * kFoldK < cnt; ++l)
#pragma unroll
                    for (int i = 0; i < kFoldK; ++i)
                        if (l * kFoldK + i < cnt) st.S = st.S + read_lane_f64(x[i], (int)l);
            } else {
                fold_block<kFoldK>(st, x, cnt, lane);
            }
        }
    }
    return st.S;
}

// One wave per segment.  With values set, segment g is radius g + 1 of the
// curve: values[g + 1] = sum / n and support[g + 1] = (n > 0); else sums[g].
__global__ __launch_bounds__(64) void fold_walk_kernel(const double *__restrict__ terms,
                                                       const uint64_t *__restrict__ seg_start, uint32_t nseg,
                                                       FoldView v, double *__restrict__ values,
                                                       uint8_t *__restrict__ support, double *__restrict__ sums) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t g = blockIdx.x; g < nseg; g += gridDim.x) {
        const uint64_t s0 = seg_start[g], n = seg_start[g + 1] - s0;
        const uint32_t b0 = v.bbase[g], nb = v.bbase[g + 1] - b0;
        const double sum = fold_walk(terms + s0, n, v, b0, nb, lane);
        if (lane == 0) {
            if (values) {
                values[g + 1] = n > 0 ? sum / (double)n : 0.0;
                support[g + 1] = n > 0 ? 1 : 0;
            } else {
                sums[g] = sum;
            }
        }
    }
}

// Host side of the fold: block tables of a segment list (segment g holds
// terms starts[g] .. starts[g + 1] - 1), laid out with the scratch in one
// device buffer, and the three launches.
struct FoldLayout {
    uint64_t off_bseg, off_sum, off_inc, off_ties, off_m, off_flag, bytes;
};

static FoldLayout fold_layout(uint32_t nseg, uint32_t nblk) {
    auto up8 = [](uint64_t x) { return (x + 7) / 8 * 8; };
    FoldLayout L;
    L.off_bseg = up8(((uint64_t)nseg + 1) * 4);
    L.off_sum = L.off_bseg + up8((uint64_t)nblk * 4);
    L.off_inc = L.off_sum + (uint64_t)nblk * 8;
    L.off_ties = L.off_inc + (uint64_t)nblk * 8;
    L.off_m = L.off_ties + (uint64_t)nblk * 16 * kFoldTies;
    L.off_flag = L.off_m + (uint64_t)nblk * 4;
    L.bytes = std::max<uint64_t>(L.off_flag + (uint64_t)nblk * 4, 8);
    return L;
}

// Blocks of a segment list, or an error past 2^31 blocks.
static s2h_status fold_block_count(const uint64_t *starts, uint32_t nseg, uint32_t *nblk) {
    uint64_t nb = 0;
    for (uint32_t g = 0; g < nseg; ++g) nb += (starts[g + 1] - starts[g] + kFoldB - 1) / kFoldB;
    if (nb >= (1ull << 31)) return fail(S2H_ERR_UNSUPPORTED, "ordered fold: too many terms");
    *nblk = (uint32_t)nb;
    return S2H_OK;
}

// Allocates buf and uploads the block tables of the segment list.
static s2h_status fold_prepare(s2h_ctx *ctx, DevBuf &buf, const uint64_t *starts, uint32_t nseg, uint32_t *nblk) {
    S2H_TRY(fold_block_count(starts, nseg, nblk));
    const FoldLayout L = fold_layout(nseg, *nblk);
    std::vector<uint8_t> host(L.off_sum, 0);
    uint32_t *bbase = (uint32_t *)host.data();
    uint32_t *bseg = (uint32_t *)(host.data() + L.off_bseg);
    uint32_t b = 0;
    for (uint32_t g = 0; g < nseg; ++g) {
        bbase[g] = b;
        const uint32_t nb = (uint32_t)((starts[g + 1] - starts[g] + kFoldB - 1) / kFoldB);
        for (uint32_t k = 0; k < nb; ++k) bseg[b + k] = g;
        b += nb;
    }
    bbase[nseg] = b;
    S2H_TRY(dev_alloc(ctx, buf, L.bytes));
    S2H_HIP(hipMemcpyAsync(buf.ptr, host.data(), L.off_sum, hipMemcpyHostToDevice, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));  // the host tables go out of scope here
    return S2H_OK;
}

static FoldView fold_view(const DevBuf &buf, uint32_t nseg, uint32_t nblk) {
    const FoldLayout L = fold_layout(nseg, nblk);
    uint8_t *base = (uint8_t *)buf.ptr;
    return FoldView{(const uint32_t *)base,         (const uint32_t *)(base + L.off_bseg),
                    (double *)(base + L.off_sum),   (double *)(base + L.off_inc),
                    (int32_t *)(base + L.off_m),    (uint32_t *)(base + L.off_flag),
                    (double *)(base + L.off_ties),  nblk};
}

// sums_ready: the block sums and flags are already in v (terms_kernel).
static s2h_status fold_launch(s2h_ctx *ctx, const double *terms, const uint64_t *seg_start, uint32_t nseg,
                              const FoldView &v, double *values, uint8_t *support, double *sums, bool sums_ready) {
    if (nseg == 0) return S2H_OK;
    if (v.nblk) {
        const unsigned grid = (v.nblk + kFoldWaves - 1) / kFoldWaves;
        if (!sums_ready) {
            fold_sums_kernel<<<grid, 64 * kFoldWaves, 0, ctx->stream>>>(terms, seg_start, v);
            S2H_HIP(hipGetLastError());
            ++ctx->launches;
        }
        fold_incr_kernel<<<grid, 64 * kFoldWaves, 0, ctx->stream>>>(terms, seg_start, v);
        S2H_HIP(hipGetLastError());
        ++ctx->launches;
    }
    fold_walk_kernel<<<std::min<uint32_t>(nseg, 65536), 64, 0, ctx->stream>>>(terms, seg_start, nseg, v, values,
                                                                             support, sums);
    S2H_HIP(hipGetLastError());
    ++ctx->launches;
    return S2H_OK;
}

// ---------------------------------------------------------------- plan (host)

static int supported_nw(int need) {
    static const int sizes[] = {1, 2, 3, 4, 6, 8, 10, 12, 16, 20, 24, 32};
    for (int s : sizes)
        if (s >= need) return s;
    return -1;
}

static Offset canonical(Offset o) {
    if (o.dz < 0 || (o.dz == 0 && (o.dx < 0 || (o.dx == 0 && o.dy < 0)))) return Offset{-o.dx, -o.dy, -o.dz};
    return o;
}

// ---- shared shell data

// Grid extent that matters for the domain filter |d_i| < n_i: beyond the
// largest |d_i| of the shells, every offset is in domain.
static uint64_t clip_extent(uint64_t n, uint32_t r_max, double pitch) {
    const uint64_t lim = (uint64_t)std::ceil((double)r_max / pitch + 0.5 / pitch) + 2;
    return std::min(n, lim);
}

using ShellKey = std::tuple<uint32_t, uint64_t, uint64_t, uint64_t, uint64_t>;

static ShellKey shell_key(const uint64_t dims[3], uint32_t r_max, double pitch) {
    uint64_t pbits;
    std::memcpy(&pbits, &pitch, 8);
    return ShellKey{r_max, pbits, clip_extent(dims[0], r_max, pitch), clip_extent(dims[1], r_max, pitch),
                    clip_extent(dims[2], r_max, pitch)};
}

static std::shared_ptr<ShellBase> build_shell_base(const uint64_t dims[3], uint32_t r_max, double pitch) {
    auto b = std::make_shared<ShellBase>();
    shell_order_offsets(dims, r_max, pitch, b->offs, b->radius_start);
    const uint64_t K = b->offs.size();
    int dmax = 0;
    for (const Offset &o : b->offs) dmax = std::max({dmax, std::abs(o.dx), std::abs(o.dy)});
    const int span = 2 * dmax + 1;
    std::vector<int32_t> colm((size_t)span * span, -1);
    for (const Offset &o0 : b->offs) {
        if (o0.dx == 0 && o0.dy == 0 && o0.dz == 0) continue;
        const Offset o = canonical(o0);
        int32_t &m = colm[(size_t)(o.dx + dmax) * span + (o.dy + dmax)];
        m = std::max(m, o.dz);
    }
    std::vector<int32_t> colidx((size_t)span * span, -1);
    uint64_t tsize = 0;
    for (int dx = -dmax; dx <= dmax; ++dx)
        for (int dy = -dmax; dy <= dmax; ++dy) {
            const int m = colm[(size_t)(dx + dmax) * span + (dy + dmax)];
            if (m < 0) continue;
            colidx[(size_t)(dx + dmax) * span + (dy + dmax)] = (int32_t)b->colset.size();
            b->colset.push_back(ColInfo{dx, dy, m, (uint32_t)tsize});
            tsize += (uint64_t)m + 1;
        }
    b->table_size = tsize;
    for (const Offset &o : b->offs) {
        b->extent[0] = std::max(b->extent[0], std::abs(o.dx));
        b->extent[1] = std::max(b->extent[1], std::abs(o.dy));
        b->extent[2] = std::max(b->extent[2], std::abs(o.dz));
    }
    b->gather.resize(K);
    for (uint64_t i = 0; i < K; ++i) {
        const Offset &o0 = b->offs[i];
        if (o0.dx == 0 && o0.dy == 0 && o0.dz == 0) {
            b->gather[i] = -1;
            continue;
        }
        const Offset o = canonical(o0);
        const int32_t ci = colidx[(size_t)(o.dx + dmax) * span + (o.dy + dmax)];
        b->gather[i] = (int32_t)(b->colset[ci].base + (uint32_t)o.dz);
    }
    return b;
}

// Process-wide shell bases, kept while their total offset count stays under
// kShellCacheOffsets (least recently used dropped first; plans keep their own
// reference).
static constexpr uint64_t kShellCacheOffsets = 160ull << 20;
static std::mutex g_shell_mu;
static std::map<ShellKey, std::pair<std::shared_ptr<const ShellBase>, uint64_t>> g_shells;  // base, last use
static uint64_t g_shell_clock = 0, g_shell_offsets = 0;

static std::shared_ptr<const ShellBase> shell_base(const uint64_t dims[3], uint32_t r_max, double pitch) {
    const ShellKey key = shell_key(dims, r_max, pitch);
    {
        std::lock_guard<std::mutex> lk(g_shell_mu);
        auto it = g_shells.find(key);
        if (it != g_shells.end()) {
            it->second.second = ++g_shell_clock;
            return it->second.first;
        }
    }
    // the clipped extent gives the same in-domain offsets as the real one
    const uint64_t cd[3] = {std::get<2>(key), std::get<3>(key), std::get<4>(key)};
    std::shared_ptr<const ShellBase> b = build_shell_base(cd, r_max, pitch);
    std::lock_guard<std::mutex> lk(g_shell_mu);
    auto it = g_shells.find(key);
    if (it != g_shells.end()) return it->second.first;  // built meanwhile by another thread
    g_shells[key] = {b, ++g_shell_clock};
    g_shell_offsets += b->offs.size();
    while (g_shell_offsets > kShellCacheOffsets && g_shells.size() > 1) {
        auto old = g_shells.begin();
        for (auto j = g_shells.begin(); j != g_shells.end(); ++j)
            if (j->second.second < old->second.second) old = j;
        g_shell_offsets -= old->second.first->offs.size();
        g_shells.erase(old);
    }
    return b;
}

static s2h_status shell_dev(s2h_ctx *ctx, const uint64_t dims[3], uint32_t r_max, double pitch,
                            std::shared_ptr<ShellDev> *out) {
    const ShellKey key = shell_key(dims, r_max, pitch);
    auto it = ctx->shell_devs.find(key);
    if (it != ctx->shell_devs.end()) {
        *out = it->second;
        return S2H_OK;
    }
    auto d = std::make_shared<ShellDev>();
    d->host = shell_base(dims, r_max, pitch);
    const ShellBase &b = *d->host;
    if (b.table_size >= (1ull << 31)) return fail(S2H_ERR_UNSUPPORTED, "offset table too large");
    const uint64_t K = b.offs.size();
    s2h_status st = upload(ctx, d->d_offs, b.offs.data(), std::max<uint64_t>(K, 1) * sizeof(Offset));
    if (st == S2H_OK) st = upload(ctx, d->d_rstart, b.radius_start.data(), b.radius_start.size() * 8);
    if (st != S2H_OK) {
        dev_free(ctx, d->d_offs);
        dev_free(ctx, d->d_gather);
        dev_free(ctx, d->d_rstart);
        return st;
    }
    ctx->shell_devs[key] = d;
    *out = d;
    return S2H_OK;
}

void shell_devs_release(s2h_ctx *ctx) {
    for (auto &kv : ctx->shell_devs) {
        dev_free(ctx, kv.second->d_offs);
        dev_free(ctx, kv.second->d_gather);
        dev_free(ctx, kv.second->d_rstart);
    }
    ctx->shell_devs.clear();
}

// Per-plan device buffers.  The direct kernel's count table, the shells'
// gather indices and the per-offset hits are allocated on first use
// (exact_run_counts, exact_gather_fold with hits): a plan the FFT kernel
// serves never holds them.  The terms live in the context's scratch.
static s2h_status plan_buffers(s2h_ctx *ctx, ExactPlan *p) {
    S2H_TRY(dev_alloc(ctx, p->d_curve, (p->r_max + 1) * 9 + 16));
    S2H_TRY(fold_prepare(ctx, p->d_fold, p->radius_start.data() + 1, p->r_max, &p->fold_nblk));
    return S2H_OK;
}

// Per-row and per-item VALU overheads of the count kernel beyond the
// and/popcount pairs (address math, prefetch, counter flush), in
// wave-instructions; from the instruction counts of the kernel.
static constexpr double kRowOverhead = 12.0, kItemOverhead = 400.0;

// Columns grouped 64 to a wave for one band width, with the (group, dz block)
// list and each block's modelled cost in VALU wave-instructions.
struct Grouping {
    int band = 0;
    uint32_t ngroups = 0;
    std::vector<ColInfo> lanes;   // ngroups * 64
    struct Blk {
        DzBlock b;
        int x0, x1, y0, y1, nw;
        double cost;
    };
    std::vector<Blk> blks;
    double total = 0;
};

// Columns in bands of `band` consecutive m values (counted from the largest m
// down), by angle inside a band, padded to whole groups of 64 with idle lanes.
static std::shared_ptr<const std::vector<ColInfo>> band_lanes(const ShellBase &b, int band) {
    {
        std::lock_guard<std::mutex> lk(b.lanes_mu);
        auto it = b.lanes.find(band);
        if (it != b.lanes.end()) return it->second;
    }
    const std::vector<ColInfo> &cs = b.colset;
    std::vector<int32_t> ms;
    for (const ColInfo &c : cs) ms.push_back(c.m);
    std::sort(ms.begin(), ms.end(), std::greater<int32_t>());
    ms.erase(std::unique(ms.begin(), ms.end()), ms.end());
    std::vector<int> bnd(cs.size());
    std::vector<double> ang(cs.size());
    for (size_t i = 0; i < cs.size(); ++i) {
        const size_t rank =
            (size_t)(std::lower_bound(ms.begin(), ms.end(), cs[i].m, std::greater<int32_t>()) - ms.begin());
        bnd[i] = (int)(rank / (size_t)std::max(1, band));
        ang[i] = std::atan2((double)cs[i].dy, (double)cs[i].dx);
    }
    std::vector<size_t> order(cs.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](size_t a, size_t c) {
        if (bnd[a] != bnd[c]) return bnd[a] < bnd[c];
        if (ang[a] != ang[c]) return ang[a] < ang[c];
        return cs[a].m > cs[c].m;
    });
    auto lanes = std::make_shared<std::vector<ColInfo>>((cs.size() + 63) / 64 * 64, ColInfo{0, 0, -1, 0});
    for (size_t k = 0; k < order.size(); ++k) (*lanes)[k] = cs[order[k]];
    std::lock_guard<std::mutex> lk(b.lanes_mu);
    auto it = b.lanes.find(band);
    if (it != b.lanes.end()) return it->second;
    b.lanes[band] = lanes;
    return lanes;
}

static void group_columns(const ExactPlan *p, int band, Grouping &g) {
    const int nx = (int)p->dims[0], ny = (int)p->dims[1], nz = (int)p->dims[2];
    g.band = band;
    g.lanes = *band_lanes(*p->shells->host, band);
    g.ngroups = (uint32_t)(g.lanes.size() / 64);
    const int need_dw = (nz + 31) / 32;
    const int NW = std::max(1, supported_nw(need_dw));
    g.blks.clear();
    g.total = 0;
    for (uint32_t gi = 0; gi < g.ngroups; ++gi) {
        int M = -1, dxmin = 1 << 30, dxmax = -(1 << 30), dymin = 1 << 30, dymax = -(1 << 30);
        for (int l = 0; l < 64; ++l) {
            const ColInfo &c = g.lanes[(size_t)gi * 64 + l];
            if (c.m < 0) continue;
            M = std::max(M, c.m);
            dxmin = std::min(dxmin, c.dx);
            dxmax = std::max(dxmax, c.dx);
            dymin = std::min(dymin, c.dy);
            dymax = std::max(dymax, c.dy);
        }
        if (M < 0) continue;
        // b row p is useful to some lane: p - (dx, dy) inside the grid
        const int x0 = std::max(0, dxmin), x1 = std::min(nx, nx + dxmax);
        const int y0 = std::max(0, dymin), y1 = std::min(ny, ny + dymax);
        if (x0 >= x1 || y0 >= y1) continue;
        for (int dz0 = 0; dz0 <= M; dz0 += 32) {
            const int rem = std::min(32, M + 1 - dz0);
            const int dzb = (rem + 3) / 4 * 4;
            const int nwi = need_dw - dz0 / 32;
            if (nwi <= 0) continue;
            const int words = (NW > 1 && nwi < NW) ? NW - 1 : NW;  // the kernel's body for this block
            const double rows = (double)(x1 - x0) * (y1 - y0);
            const double cost = rows * (words * dzb * 2.0 + kRowOverhead);
            g.blks.push_back(Grouping::Blk{DzBlock{gi, (uint16_t)dz0, (uint16_t)dzb}, x0, x1, y0, y1, nwi, cost});
            g.total += cost;
        }
    }
}

// x-chunks of a block for a wave target: proportional to its cost.
// At least enough chunks that no lane's u32 counter can wrap: an item's
// counter adds at most 32 per partner dword of each of its rows.
static int block_chunks(const Grouping &g, const Grouping::Blk &b, int waves) {
    const int span = b.x1 - b.x0;
    const uint64_t per_x = (uint64_t)(b.y1 - b.y0) * (uint64_t)b.nw * 32u;
    const int min_n = (int)(((uint64_t)span * per_x + 0xFFFFFFFEull) / 0xFFFFFFFFull);
    const int n = std::max(min_n, (int)std::lround(b.cost / std::max(g.total, 1.0) * waves));
    return std::max(1, std::min(span, n));
}

// Workgroups (kWavesPerBlock waves) resident on the whole chip at once: 256 CUs
// x 5 (the count kernel's 5 waves per SIMD).
static constexpr double kBlockSlots = 256.0 * 5.0;

// Each item ends with one u64 atomic per (lane, dz), which reaches memory at
// 32-byte granularity; measured at 256^3 r64 (one flush per wave), doubling the
// items from 32768 to 65536 added 0.46 ms for 1.6 GB more of these writes.
// The workgroup flush usually merges its waves' items into one.
static constexpr double kFlushBytesPerS = 3.5e12;

// Cost model: executed VALU wave-instructions at the issue rate of 1024 SIMDs
// (4 cycles each), scaled down when there are too few waves to hide latency
// and for the partly filled last round of workgroups (r >= 1 rounds of
// resident workgroups lose about half a round: r / (r + 1/2); less than one
// round is covered by the fill term), bounded below by the longest item, plus
// the items' counter flushes.
static double model_ms(const Grouping &g, int waves, double *ops_out) {
    double items = 0, maxitem = 0, flush_bytes = 0;
    for (const Grouping::Blk &b : g.blks) {
        const int n = block_chunks(g, b, waves);
        items += n;
        flush_bytes += (double)n * 64.0 * b.b.dzb * 32.0 / kWavesPerBlock;
        maxitem = std::max(maxitem, b.cost * std::ceil((double)(b.x1 - b.x0) / n) / (b.x1 - b.x0) + kItemOverhead);
    }
    const double ops = g.total + kItemOverhead * items;
    if (ops_out) *ops_out = ops;
    const double simds = 1024.0, clk = 2.4e9;
    const double fill = std::min(1.0, items / (simds * 6.0));
    const double rounds = items / kWavesPerBlock / kBlockSlots;
    const double tail = rounds >= 1.0 ? rounds / (rounds + 0.5) : 1.0;  // a single partial round loses nothing
    return 1e3 * (std::max(ops * 4.0 / (simds * clk * std::max(fill, 0.05) * tail), maxitem * 4.0 * 2.0 / clk) +
                  flush_bytes / kFlushBytesPerS);
}

static s2h_status apply_shape(s2h_ctx *ctx, ExactPlan *p, const Grouping &g, int waves) {
    p->tune = ExactTune{g.band, waves};
    p->ngroups = g.ngroups;
    p->blocks.clear();
    p->items.clear();
    std::vector<double> icost;
    for (const Grouping::Blk &b : g.blks) {
        p->blocks.push_back(b.b);
        const int span = b.x1 - b.x0, n = block_chunks(g, b, waves);
        for (int k = 0; k < n; ++k) {
            const int xa = b.x0 + (int)((int64_t)span * k / n), xb = b.x0 + (int)((int64_t)span * (k + 1) / n);
            if (xa >= xb) continue;
            WorkItem w;
            w.group = b.b.group;
            w.dz0 = b.b.dz0;
            w.dzb = b.b.dzb;
            w.x0 = (uint16_t)xa;
            w.x1 = (uint16_t)xb;
            w.y0 = (uint16_t)b.y0;
            w.y1 = (uint16_t)b.y1;
            w.nw = (uint16_t)b.nw;
            w.pad = 0;
            p->items.push_back(w);
            icost.push_back(b.cost * (double)(xb - xa) / span);
        }
    }
    // heaviest first: the last waves to start are the short ones
    std::vector<size_t> io(p->items.size());
    std::iota(io.begin(), io.end(), 0);
    std::stable_sort(io.begin(), io.end(), [&](size_t a, size_t b) { return icost[a] > icost[b]; });
    std::vector<WorkItem> sorted(p->items.size());
    for (size_t k = 0; k < io.size(); ++k) sorted[k] = p->items[io[k]];
    p->items.swap(sorted);
    p->model_ms = model_ms(g, waves, &p->model_ops);
    p->shaped = true;
    dev_free(ctx, p->d_cols);
    dev_free(ctx, p->d_items);
    S2H_TRY(upload(ctx, p->d_cols, g.lanes.data(), g.lanes.size() * sizeof(ColInfo)));
    S2H_TRY(upload(ctx, p->d_items, p->items.data(), std::max<size_t>(p->items.size(), 1) * sizeof(WorkItem)));
    return S2H_OK;
}

static s2h_status apply_tune(s2h_ctx *ctx, ExactPlan *p, ExactTune t) {
    Grouping g;
    group_columns(p, t.band, g);
    return apply_shape(ctx, p, g, t.waves);
}

// A plan's shape without its work items: the model's figures for the kernel
// choice.  The items (and their upload) wait for the first direct-count run,
// so a plan served by the FFT kernel never builds them.
static void model_tune(ExactPlan *p, ExactTune t) {
    Grouping g;
    group_columns(p, t.band, g);
    p->tune = t;
    p->ngroups = g.ngroups;
    p->model_ms = model_ms(g, t.waves, &p->model_ops);
    p->shaped = false;
}

static s2h_status plan_shape(s2h_ctx *ctx, ExactPlan *p) {
    return p->shaped ? S2H_OK : apply_tune(ctx, p, p->tune);
}

struct Ranked {
    ExactTune tune;
    double ms;
};

// Every (band, waves) candidate, best modelled first.
static std::vector<Ranked> rank_tunes(const ExactPlan *p) {
    std::vector<Ranked> out;
    for (int band : {1, 2, 3, 4, 6}) {
        Grouping g;
        group_columns(p, band, g);
        for (int waves : {8192, 16384, 32768, 65536})
            out.push_back(Ranked{ExactTune{band, waves}, model_ms(g, waves, nullptr)});
    }
    std::stable_sort(out.begin(), out.end(), [](const Ranked &a, const Ranked &b) { return a.ms < b.ms; });
    return out;
}

// Process-wide table of chosen shapes per (dims, r_max, pitch): a shape timed
// once on the device serves every later plan of that key in any context.
struct TuneEntry {
    ExactTune tune;
    int tuned;  // 1 timed on the device, -1 model pick
    int timed;
};
using TuneKey = std::tuple<uint64_t, uint64_t, uint64_t, uint32_t, uint64_t>;
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneEntry> g_tunes;

static TuneKey tune_key(const ExactPlan *p) {
    uint64_t pbits;
    std::memcpy(&pbits, &p->pitch, 8);
    return TuneKey{p->dims[0], p->dims[1], p->dims[2], p->r_max, pbits};
}

static void tune_store(const ExactPlan *p, const TuneEntry &e) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    TuneEntry &slot = g_tunes[tune_key(p)];
    if (slot.tuned != 1 || e.tuned == 1) slot = e;  // never replace a timed pick by a model pick
}

static bool tune_lookup(const ExactPlan *p, TuneEntry &e) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tunes.find(tune_key(p));
    if (it == g_tunes.end()) return false;
    e = it->second;
    return true;
}

static bool autotune_enabled(const s2h_ctx *ctx) {
    const char *env = std::getenv("S2HIP_AUTOTUNE");
    return ctx->autotune && !(env && env[0] == '0');
}

void exact_plan_release(s2h_ctx *ctx, ExactPlan *p) {
    DevBuf *bufs[] = {&p->d_cols, &p->d_items, &p->d_table, &p->d_hits, &p->d_curve, &p->d_fold};
    for (DevBuf *b : bufs) dev_free(ctx, *b);
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
    std::shared_ptr<ShellDev> sdv;
    S2H_TRY(shell_dev(ctx, dims, r_max, pitch, &sdv));
    auto p = std::make_unique<ExactPlan>(sdv);
    for (int a = 0; a < 3; ++a) p->dims[a] = dims[a];
    p->r_max = r_max;
    p->pitch = pitch;
    s2h_status st = plan_buffers(ctx, p.get());
    if (st == S2H_OK) {
        TuneEntry e;
        if (tune_lookup(p.get(), e)) {
            model_tune(p.get(), e.tune);
            // a model pick may still be timed by a context that tunes
            p->tuned = (e.tuned == 1 || !autotune_enabled(ctx)) ? e.tuned : 0;
            p->timed_candidates = e.timed;
        } else {
            const std::vector<Ranked> ranked = rank_tunes(p.get());
            model_tune(p.get(), ranked.front().tune);
            p->tuned = autotune_enabled(ctx) ? 0 : -1;
            tune_store(p.get(), TuneEntry{ranked.front().tune, -1, 0});
        }
    }
    if (st == S2H_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) st = fail(S2H_ERR_HIP, "plan upload failed");
    if (st != S2H_OK) {
        exact_plan_release(ctx, p.get());
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
    const int n = (int)p->items.size();
    if (n == 0) return S2H_OK;
    const WorkItem *items = (const WorkItem *)p->d_items.ptr;
    const unsigned blocks = (unsigned)((n + kWavesPerBlock - 1) / kWavesPerBlock);
    exact_count_kernel<NW><<<blocks, dim3(64 * kWavesPerBlock), 0, ctx->stream>>>(v, vbytes, nx, ny, row_dw, items,
                                                                                   n, cols, table);
    S2H_HIP(hipGetLastError());
    ++ctx->launches;
    return S2H_OK;
}

static s2h_status launch_counts(s2h_ctx *ctx, s2h_vol *vol, ExactPlan *p) {
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

// Autotuner: the cost model ranks every (band, waves) candidate; when one run
// is short enough (modelled below kTuneMaxMs), the best kTuneCandidates are
// timed on the device with HIP events on the context's stream and the fastest
// is kept for the plan and recorded process-wide.  Contexts with autotuning
// off (s2h_set_autotune, or S2HIP_AUTOTUNE=0) keep the model's choice.  Counts
// are integers, so the choice never changes a result.
static constexpr double kTuneMaxMs = 80.0;
static constexpr int kTuneCandidates = 4;
static constexpr int kTuneRounds = 3;

static s2h_status autotune(s2h_ctx *ctx, s2h_vol *vol, ExactPlan *p) {
    const std::vector<Ranked> ranked = rank_tunes(p);
    if (ranked.empty() || ranked.front().ms > kTuneMaxMs) return S2H_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    S2H_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(S2H_ERR_HIP, "hipEventCreate failed");
    }
    s2h_status st = S2H_OK;
    int best = -1, timed = 0;
    const int ncand = std::min<int>(kTuneCandidates, (int)ranked.size());
    // kTuneRounds interleaved rounds, the fastest run of each candidate counts:
    // one run each is too noisy to separate shapes a few percent apart
    std::vector<float> best_ms(ncand, 1e30f);
    for (int round = 0; round < kTuneRounds && st == S2H_OK; ++round) {
        for (int k = 0; k < ncand && st == S2H_OK; ++k) {
            st = apply_tune(ctx, p, ranked[k].tune);
            if (st == S2H_OK && round == 0) st = launch_counts(ctx, vol, p);  // warm-up
            if (st != S2H_OK) break;
            hipError_t e = hipEventRecord(e0, ctx->stream);
            if (e == hipSuccess) st = launch_counts(ctx, vol, p);
            if (e == hipSuccess && st == S2H_OK) e = hipEventRecord(e1, ctx->stream);
            if (e == hipSuccess && st == S2H_OK) e = hipEventSynchronize(e1);
            float ms = 0;
            if (e == hipSuccess && st == S2H_OK) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) st = fail(S2H_ERR_HIP, std::string("autotune timing: ") + hipGetErrorString(e));
            if (st == S2H_OK) best_ms[k] = std::min(best_ms[k], ms);
            if (round == 0) ++timed;
        }
    }
    for (int k = 0; k < ncand; ++k)
        if (best < 0 || best_ms[k] < best_ms[best]) best = k;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != S2H_OK) return st;
    if (best >= 0) S2H_TRY(apply_tune(ctx, p, ranked[best].tune));
    p->tuned = 1;
    p->timed_candidates = timed;
    tune_store(p, TuneEntry{p->tune, 1, timed});
    return S2H_OK;
}

s2h_status exact_run_counts(s2h_ctx *ctx, s2h_vol *vol, ExactPlan *p) {
    if (vol->nwords * 8 + kVolTailBytes >= (1ull << 31))
        return fail(S2H_ERR_UNSUPPORTED, "packed volume above 2 GiB");
    if (supported_nw((int)((vol->dims[2] + 31) / 32)) < 0)
        return fail(S2H_ERR_UNSUPPORTED, "nz=" + std::to_string(vol->dims[2]) + " exceeds 1024");
    S2H_TRY(plan_shape(ctx, p));
    if (!p->d_table.ptr) S2H_TRY(dev_alloc(ctx, p->d_table, std::max<uint64_t>(p->table_size, 1) * 8));
    if (!p->shells->d_gather.ptr) {
        const std::vector<int32_t> &g = p->shells->host->gather;
        S2H_TRY(upload(ctx, p->shells->d_gather, g.data(), std::max<uint64_t>(g.size(), 1) * sizeof(int32_t)));
    }
    if (p->tuned == 0) {
        if (autotune_enabled(ctx)) S2H_TRY(autotune(ctx, vol, p));
        if (p->tuned == 0) p->tuned = -1;  // the model's choice is kept; not retried
    }
    S2H_HIP(hipMemsetAsync(p->d_table.ptr, 0, std::max<uint64_t>(p->table_size, 1) * 8, ctx->stream));
    prof_mark(ctx, 0);
    return launch_counts(ctx, vol, p);
}

s2h_status exact_gather_fold(s2h_ctx *ctx, ExactPlan *p, uint64_t occupied, const FftCube *cube, bool want_hits,
                             void *curve_dst) {
    const FftCube none;
    const FftCube &c = cube ? *cube : none;
    prof_mark(ctx, 1);
    const FoldView v = fold_view(p->d_fold, p->r_max, p->fold_nblk);
    // segments: radii 1 .. r_max (radius_start[r] .. radius_start[r + 1])
    const uint64_t *seg_start = (const uint64_t *)p->d_rstart.ptr + 1;
    if (want_hits && !p->d_hits.ptr) S2H_TRY(dev_alloc(ctx, p->d_hits, std::max<uint64_t>(p->offs.size(), 1) * 8));
    if (!c.f32 && !c.f64 && (!p->d_table.ptr || !p->d_gather.ptr))
        return fail(S2H_ERR_ARG, "no counts to gather: neither cube nor table");
    S2H_TRY(dev_alloc(ctx, ctx->terms, std::max<uint64_t>(p->offs.size(), 1) * 8));
    if (v.nblk) {
        terms_kernel<<<(v.nblk + kFoldWaves - 1) / kFoldWaves, 64 * kFoldWaves, 0, ctx->stream>>>(
            (const Offset *)p->d_offs.ptr, (const int32_t *)p->d_gather.ptr, (const unsigned long long *)p->d_table.ptr,
            c.f32, c.f64, c.R[0], c.R[1], c.R[2], occupied, p->dims[0], p->dims[1], p->dims[2],
            want_hits ? (unsigned long long *)p->d_hits.ptr : nullptr, (double *)ctx->terms.ptr, seg_start, v);
        S2H_HIP(hipGetLastError());
        ++ctx->launches;
    }
    double *values = (double *)(curve_dst ? curve_dst : p->d_curve.ptr);
    uint8_t *support = (uint8_t *)(values + p->r_max + 1);
    S2H_TRY(fold_launch(ctx, (const double *)ctx->terms.ptr, seg_start, p->r_max, v, values, support, nullptr, true));
    prof_mark(ctx, 2);
    return S2H_OK;
}

void shell_extent(const uint64_t dims[3], uint32_t r_max, double pitch, int R[3]) {
    const std::shared_ptr<const ShellBase> b = shell_base(dims, r_max, pitch < 1e-9 ? 1e-9 : pitch);
    for (int a = 0; a < 3; ++a) R[a] = b->extent[a];
}

// FP64 passes modelled at this share of the HBM peak over their algorithmic
// bytes, plus launch and gather time.
static constexpr double kFft64BytesPerS = 5.0e12;
static constexpr double kFft64FixedMs = 0.05;
static constexpr double kFft64MaxBound = 0.25;  // required worst-case error (rounding is exact below 0.5)

// The exact kernel of one evaluation: direct counting into the table, or the
// FP64 FFT whose values round to the same integer counts (cube->f64 set).  The
// reference's planner picks between its two exact kernels by modelled time
// under a memory budget (src/geometry/s2.rs:733-766) because both return the
// same counts; here the FFT is eligible when its worst-case round-off bound
// for this volume is below 1/4 and (for the model's choice) its working set
// fits the context's FFT budget, and the faster model wins.
static s2h_status exact_counts_pick(s2h_ctx *ctx, s2h_vol *vol, ExactPlan *p, uint64_t occupied, FftCube *cube) {
    *cube = FftCube();
    const int *ext = p->shells->host->extent;
    const int R[3] = {ext[0], ext[1], ext[2]};
    const bool fft_ok = fft64_supported(p->dims, R) && fft64_error_bound(p->dims, R, occupied) <= kFft64MaxBound;
    int m = ctx->exact_method;
    if (m == 2 && !fft_ok)
        return fail(S2H_ERR_UNSUPPORTED, "FP64 FFT counts are not provably exact for this volume (error bound " +
                                             std::to_string(fft64_error_bound(p->dims, R, occupied)) + ")");
    if (m == 0) {
        const double plan_bytes = fft64_plan_bytes(p->dims, R);
        bool fits = !ctx->fft_budget || plan_bytes <= (double)ctx->fft_budget;
        const bool in_arena = ctx->fft_arena.ptr && plan_bytes <= (double)ctx->fft_arena.bytes;
        if (ctx->budget && !in_arena) {
            // the spectra and the cube (unless this context already holds
            // them) must fit what the context's budget leaves
            const double need = fft64_plan_cached(ctx, p->dims, R) ? 0.0 : plan_bytes;
            fits = fits && (double)ctx->in_use + need <= (double)ctx->budget;
        }
        const double fft_ms = fft64_model_bytes(p->dims, R) / kFft64BytesPerS * 1e3 + kFft64FixedMs;
        m = fft_ok && fits && fft_ms < p->model_ms ? 2 : 1;
        if (m == 2) {
            p->method = 2;
            const s2h_status st = fft64_exact_device(ctx, vol, p, cube);
            if (st != S2H_ERR_BUDGET) return st;
            // the model's choice only: give the spectra back and count directly,
            // as the reference's planner keeps to its budget (src/geometry/s2.rs:733-766)
            (void)hipStreamSynchronize(ctx->stream);
            fft_plans_release(ctx);
            *cube = FftCube();
            m = 1;
        }
    }
    p->method = m;
    if (m == 2) return fft64_exact_device(ctx, vol, p, cube);
    return exact_run_counts(ctx, vol, p);
}

s2h_status curve_staging(s2h_ctx *ctx, ExactPlan *p, void **stage) {
    return pinned_staging(ctx, ((uint64_t)p->r_max + 1) * 9, stage);
}

s2h_status finish_curve_from_device(s2h_ctx *ctx, ExactPlan *p, double vf, double *s2, void *staged) {
    const uint32_t n = p->r_max + 1;
    void *stage = staged;  // page-locked
    if (!stage) S2H_TRY(pinned_staging(ctx, (uint64_t)n * 9, &stage));
    S2H_TRY(inject(kFailReadback));
    // without a staged curve: one DMA of the device curve the stream sync waits for
    if (!staged) S2H_HIP(hipMemcpyAsync(stage, p->d_curve.ptr, (size_t)n * 9, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    std::vector<uint8_t> sup(n, 0);
    std::memcpy(s2, stage, (size_t)n * 8);
    std::memcpy(sup.data(), (const uint8_t *)stage + (size_t)n * 8, n);
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
    S2H_TRY(inject(kFailLaunch));
    ExactPlan *p = nullptr;
    S2H_TRY(exact_plan_get(ctx, vol->dims, r_max, pitch, &p));
    if (p->offs.size() != n_offsets)
        return fail(S2H_ERR_SIZE, "hits buffer holds " + std::to_string(n_offsets) + " entries, plan has " +
                                      std::to_string(p->offs.size()));
    uint64_t occupied = 0;
    S2H_TRY(vol_popcount(ctx, vol, &occupied));
    FftCube cube;
    S2H_TRY(exact_counts_pick(ctx, vol, p, occupied, &cube));
    S2H_TRY(exact_gather_fold(ctx, p, occupied, &cube, true, nullptr));
    S2H_TRY(inject(kFailReadback));
    if (n_offsets)
        S2H_HIP(hipMemcpyAsync(hits, p->d_hits.ptr, n_offsets * 8, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    return S2H_OK;
}

s2h_status s2h_exact_tune_info(s2h_ctx *ctx, const uint64_t dims[3], uint32_t r_max, double pitch, int64_t *info,
                               uint64_t n) {
    if (!ctx || !dims || !info) return fail(S2H_ERR_ARG, "null argument");
    if (n != 8) return fail(S2H_ERR_SIZE, "info buffer must hold 8 values");
    S2H_HIP(hipSetDevice(ctx->device));
    ExactPlan *p = nullptr;
    S2H_TRY(exact_plan_get(ctx, dims, r_max, pitch < 1e-9 ? 1e-9 : pitch, &p));
    S2H_TRY(plan_shape(ctx, p));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    info[0] = p->tune.band;
    info[1] = p->tune.waves;
    info[2] = p->tuned;
    info[3] = p->timed_candidates;
    info[4] = (int64_t)p->items.size();
    info[5] = p->ngroups;
    info[6] = (int64_t)p->model_ops;
    info[7] = p->method;
    return S2H_OK;
}

s2h_status s2h_exact_tune_set(s2h_ctx *ctx, const uint64_t dims[3], uint32_t r_max, double pitch, int32_t band,
                              int32_t waves) {
    if (!ctx || !dims) return fail(S2H_ERR_ARG, "null argument");
    if (band < 1 || waves < 1) return fail(S2H_ERR_ARG, "band and waves must be >= 1");
    S2H_HIP(hipSetDevice(ctx->device));
    ExactPlan *p = nullptr;
    S2H_TRY(exact_plan_get(ctx, dims, r_max, pitch < 1e-9 ? 1e-9 : pitch, &p));
    S2H_TRY(apply_tune(ctx, p, ExactTune{band, waves}));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    p->tuned = -1;
    p->timed_candidates = 0;
    return S2H_OK;
}

s2h_status s2h_set_autotune(s2h_ctx *ctx, int32_t enable) {
    if (!ctx) return fail(S2H_ERR_ARG, "null argument");
    ctx->autotune = enable != 0;
    return S2H_OK;
}

s2h_status s2h_exact_method_set(s2h_ctx *ctx, int32_t method) {
    if (!ctx) return fail(S2H_ERR_ARG, "null argument");
    if (method < 0 || method > 2) return fail(S2H_ERR_ARG, "method must be 0 (model), 1 (direct) or 2 (fft64)");
    ctx->exact_method = method;
    return S2H_OK;
}

s2h_status s2h_exact_s2(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, double *s2, uint64_t n) {
    if (!ctx || !vol || !s2) return fail(S2H_ERR_ARG, "null argument");
    if (n != (uint64_t)r_max + 1) return fail(S2H_ERR_SIZE, "s2 buffer must hold r_max+1 values");
    S2H_HIP(hipSetDevice(ctx->device));
    S2H_TRY(inject(kFailLaunch));
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
    FftCube cube;
    S2H_TRY(exact_counts_pick(ctx, vol, p, occupied, &cube));
    // the fold writes the curve straight into page-locked host memory
    void *stage = nullptr;
    S2H_TRY(curve_staging(ctx, p, &stage));
    S2H_TRY(exact_gather_fold(ctx, p, occupied, &cube, false, stage));
    return finish_curve_from_device(ctx, p, vf, s2, stage);
}

}  // extern "C"

extern "C" s2h_status s2h_ordered_sum(s2h_ctx *ctx, const double *terms, uint64_t n, const uint64_t *starts,
                                      uint32_t nseg, double *sums) {
    if (!ctx || !starts || !sums || (n && !terms)) return fail(S2H_ERR_ARG, "null argument");
    for (uint32_t g = 0; g < nseg; ++g)
        if (starts[g] > starts[g + 1] || starts[g + 1] > n) return fail(S2H_ERR_ARG, "segment bounds out of order");
    if (nseg == 0) return S2H_OK;
    S2H_HIP(hipSetDevice(ctx->device));
    DevBuf d_t, d_s, d_o, d_f;
    uint32_t nblk = 0;
    s2h_status st = upload(ctx, d_t, terms, std::max<uint64_t>(n, 1) * 8);
    if (st == S2H_OK) st = upload(ctx, d_s, starts, ((uint64_t)nseg + 1) * 8);
    if (st == S2H_OK) st = dev_alloc(ctx, d_o, (uint64_t)nseg * 8);
    if (st == S2H_OK) st = fold_prepare(ctx, d_f, starts, nseg, &nblk);
    if (st == S2H_OK)
        st = fold_launch(ctx, (const double *)d_t.ptr, (const uint64_t *)d_s.ptr, nseg, fold_view(d_f, nseg, nblk),
                         nullptr, nullptr, (double *)d_o.ptr, false);
    hipError_t e = hipSuccess;
    if (st == S2H_OK) {
        e = hipMemcpyAsync(sums, d_o.ptr, (uint64_t)nseg * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    dev_free(ctx, d_f);
    dev_free(ctx, d_t);
    dev_free(ctx, d_s);
    dev_free(ctx, d_o);
    if (st != S2H_OK) return st;
    if (e != hipSuccess) return fail(S2H_ERR_HIP, std::string("ordered sum: ") + hipGetErrorString(e));
    return S2H_OK;
}
