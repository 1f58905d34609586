// Context, device-memory accounting, packed volumes and the C-ABI glue shared by
// every path of the S2 engine.
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>

#include "s2hip_internal.h"

namespace s2h {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

s2h_status fail(s2h_status code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

// Injected failures (S2HIP_FAIL_AT=alloc|launch|readback[:n]): the n-th (default
// first) pass through that site after the variable last changed returns
// S2H_ERR_HIP before any device work there, so callers' error paths can be
// exercised without a faulting kernel.  The variable is read on each pass;
// the counters are process-wide.
s2h_status inject(FailSite site) {
    static std::mutex mu;
    static std::string cached;
    static std::atomic<int> armed{-1};
    static uint64_t at = 1, seen = 0;
    const char *env = std::getenv("S2HIP_FAIL_AT");
    if (!env && armed < 0) return S2H_OK;
    std::lock_guard<std::mutex> lk(mu);
    const std::string v = env ? env : "";
    if (v != cached) {
        cached = v;
        const size_t colon = v.find(':');
        const std::string name = v.substr(0, colon);
        armed = name == "alloc" ? kFailAlloc : name == "launch" ? kFailLaunch : name == "readback" ? kFailReadback : -1;
        at = colon == std::string::npos ? 1 : std::max<uint64_t>(1, std::strtoull(v.c_str() + colon + 1, nullptr, 10));
        seen = 0;
    }
    if (armed.load() != (int)site || ++seen != at) return S2H_OK;
    static const char *names[] = {"allocation", "launch", "read-back"};
    return fail(S2H_ERR_HIP, std::string("injected ") + names[site] + " failure (S2HIP_FAIL_AT)");
}


// Device memory: blocks released by a context are kept in its cache and
// reused for later requests of a similar size.  hipFree synchronizes the whole
// device, which would drain every other stream (the serving slots) each time a
// request releases its volume; a context has one stream, so a cached block is
// reused in stream order without any synchronization.  The cache is trimmed
// (hipFree) only past kCacheMaxBytes or when the budget needs the room.
static constexpr uint64_t kCacheMaxBytes = 2ull << 30;

static uint64_t round_block(uint64_t bytes) {
    const uint64_t g = bytes <= (1ull << 20) ? 4096 : (64ull << 10);
    return (bytes + g - 1) / g * g;
}

static void cache_trim(s2h_ctx *ctx, uint64_t keep) {
    while (ctx->cached_bytes > keep && !ctx->block_cache.empty()) {
        auto it = std::prev(ctx->block_cache.end());  // largest first
        (void)hipFree(it->second);
        ctx->cached_bytes -= it->first;
        ctx->block_cache.erase(it);
    }
}

s2h_status dev_alloc(s2h_ctx *ctx, DevBuf &b, uint64_t bytes) {
    if (b.ptr && !b.borrowed && b.bytes >= bytes) return S2H_OK;
    S2H_TRY(inject(kFailAlloc));
    dev_free(ctx, b);
    if (bytes == 0) bytes = 8;
    const uint64_t want = round_block(bytes);
    // a cached block no more than a quarter (+64 KiB) larger than asked for
    auto it = ctx->block_cache.lower_bound(want);
    if (it != ctx->block_cache.end() && it->first <= want + want / 4 + (64ull << 10)) {
        if (ctx->budget && ctx->in_use + it->first > ctx->budget)
            return fail(S2H_ERR_BUDGET, "device working set of " + std::to_string(ctx->in_use + it->first) +
                                            " bytes exceeds budget " + std::to_string(ctx->budget));
        b.ptr = it->second;
        b.bytes = it->first;
        ctx->cached_bytes -= it->first;
        ctx->block_cache.erase(it);
        ctx->in_use += b.bytes;
        ctx->held_peak = std::max(ctx->held_peak, ctx->in_use + ctx->cached_bytes);
        return S2H_OK;
    }
    if (ctx->budget && ctx->in_use + want > ctx->budget)
        return fail(S2H_ERR_BUDGET, "device working set of " + std::to_string(ctx->in_use + want) +
                                        " bytes exceeds budget " + std::to_string(ctx->budget));
    if (ctx->budget && ctx->in_use + ctx->cached_bytes + want > ctx->budget)
        cache_trim(ctx, ctx->budget - ctx->in_use - want);
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess && ctx->cached_bytes) {
        (void)hipGetLastError();
        cache_trim(ctx, 0);
        e = hipMalloc(&b.ptr, want);
    }
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return fail(S2H_ERR_HIP, std::string("hipMalloc(") + std::to_string(want) +
                                     ") failed: " + hipGetErrorString(e));
    }
    b.bytes = want;
    ctx->in_use += want;
    ctx->held_peak = std::max(ctx->held_peak, ctx->in_use + ctx->cached_bytes);
    return S2H_OK;
}

void dev_free(s2h_ctx *ctx, DevBuf &b) {
    if (b.ptr && !b.borrowed) {
        ctx->in_use -= b.bytes;
        ctx->block_cache.emplace(b.bytes, b.ptr);
        ctx->cached_bytes += b.bytes;
        if (ctx->cached_bytes > kCacheMaxBytes) cache_trim(ctx, kCacheMaxBytes / 2);
    }
    b.ptr = nullptr;
    b.bytes = 0;
    b.borrowed = false;
}

void dev_cache_release(s2h_ctx *ctx) { cache_trim(ctx, 0); }

void dev_cache_trim(s2h_ctx *ctx, uint64_t keep) { cache_trim(ctx, keep); }

s2h_status upload(s2h_ctx *ctx, DevBuf &b, const void *src, uint64_t bytes) {
    S2H_TRY(dev_alloc(ctx, b, bytes));
    if (bytes) S2H_HIP(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return S2H_OK;
}

// Pack a bool grid into padded rows: one thread per 64-voxel word.
__global__ void pack_bytes_kernel(const uint8_t *__restrict__ occ, uint64_t *__restrict__ words,
                                  uint64_t rows, uint64_t nz, uint64_t row_words) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * row_words) return;
    uint64_t row = i / row_words, w = i % row_words;
    uint64_t z0 = w * 64, v = 0;
    if (z0 < nz) {
        const uint8_t *src = occ + row * nz;
        uint64_t z1 = z0 + 64 < nz ? z0 + 64 : nz;
        for (uint64_t z = z0; z < z1; ++z) v |= (uint64_t)(src[z] != 0) << (z - z0);
    }
    words[i] = v;
}

__global__ void popcount_kernel(const uint64_t *__restrict__ words, uint64_t n,
                                unsigned long long *__restrict__ out) {
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        acc += __popcll(words[i]);
    // wave reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)acc);
}

s2h_status vol_alloc(s2h_ctx *ctx, const uint64_t dims[3], s2h_vol **out) {
    if (!out) return fail(S2H_ERR_ARG, "null output handle");
    for (int a = 0; a < 3; ++a)
        if (dims[a] == 0 || dims[a] > 65535)
            return fail(S2H_ERR_UNSUPPORTED, "grid dimension " + std::to_string(dims[a]) +
                                                 " outside 1..65535");
    auto *v = new s2h_vol();
    for (int a = 0; a < 3; ++a) v->dims[a] = dims[a];
    v->row_words = (dims[2] + 63) / 64 + 1;
    v->nwords = dims[0] * dims[1] * v->row_words;
    v->occupied = std::numeric_limits<uint64_t>::max();
    DevBuf b;
    v->alloc_bytes = v->nwords * 8 + kVolTailBytes;
    s2h_status st = dev_alloc(ctx, b, v->alloc_bytes);
    if (st != S2H_OK) {
        delete v;
        return st;
    }
    v->d_words = (uint64_t *)b.ptr;
    v->alloc_bytes = b.bytes;  // the block as allocated (rounded up)
    hipError_t e = hipMemsetAsync((char *)b.ptr + v->nwords * 8, 0, kVolTailBytes, ctx->stream);
    if (e != hipSuccess) {
        dev_free(ctx, b);
        delete v;
        return fail(S2H_ERR_HIP, std::string("hipMemsetAsync failed: ") + hipGetErrorString(e));
    }
    *out = v;
    return S2H_OK;
}

s2h_status pinned_staging(s2h_ctx *ctx, uint64_t bytes, void **out) {
    if (bytes > ctx->pinned_bytes) {
        if (ctx->pinned) {
            S2H_HIP(hipStreamSynchronize(ctx->stream));
            (void)hipHostFree(ctx->pinned);
            ctx->pinned = nullptr;
            ctx->pinned_bytes = 0;
        }
        const uint64_t want = std::max<uint64_t>(bytes, 4096);
        S2H_HIP(hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
        ctx->pinned_bytes = want;
    }
    *out = ctx->pinned;
    return S2H_OK;
}

s2h_status vol_popcount(s2h_ctx *ctx, s2h_vol *vol, uint64_t *occupied) {
    if (vol->occupied == std::numeric_limits<uint64_t>::max()) {
        S2H_TRY(dev_alloc(ctx, ctx->d_scalar, 64));
        S2H_HIP(hipMemsetAsync(ctx->d_scalar.ptr, 0, 8, ctx->stream));
        uint64_t blocks = (vol->nwords + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        popcount_kernel<<<(unsigned)blocks, 256, 0, ctx->stream>>>(
            vol->d_words, vol->nwords, (unsigned long long *)ctx->d_scalar.ptr);
        S2H_HIP(hipGetLastError());
        uint64_t h = 0;
        S2H_HIP(hipMemcpyAsync(&h, ctx->d_scalar.ptr, 8, hipMemcpyDeviceToHost, ctx->stream));
        S2H_HIP(hipStreamSynchronize(ctx->stream));
        vol->occupied = h;
    }
    *occupied = vol->occupied;
    return S2H_OK;
}

void prof_mark(s2h_ctx *ctx, int k) {
    if (ctx->profiling) (void)hipEventRecord(ctx->ev[k], ctx->stream);
    if (k == 0) ctx->launches = 0;
}

void prof_collect(s2h_ctx *ctx) {
    if (!ctx->profiling) return;
    float a = 0, b = 0, c = 0;
    if (hipEventSynchronize(ctx->ev[2]) != hipSuccess) return;
    (void)hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]);
    (void)hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]);
    (void)hipEventElapsedTime(&c, ctx->ev[0], ctx->ev[2]);
    ctx->prof_ms[0] = a;
    ctx->prof_ms[1] = b;
    ctx->prof_ms[2] = c;
    ctx->prof_ms[3] = ctx->launches;
}

// ---------------------------------------------------------------- shells

// src/geometry/s2.rs:364-385 shell_offset_iter: dx, dy, dz ascending over
// [-lim, lim], keeping d != 0 with low2 <= |d|^2 < high2 (half-open float bounds);
// a radius <= 1e-12 yields only the origin.
void shell_offsets_full(double distance_vox, double half_width_vox, std::vector<Offset> &out) {
    out.clear();
    if (distance_vox <= 1e-12) {
        out.push_back({0, 0, 0});
        return;
    }
    double lo = distance_vox - half_width_vox;
    if (lo < 0.0) lo = 0.0;
    const double low2 = lo * lo;
    const double hi = distance_vox + half_width_vox;
    const double high2 = hi * hi;
    const double limf = std::ceil(distance_vox + half_width_vox + 1.0);
    if (!(limf < 1.0e6)) return;  // radii beyond any supported grid
    const int64_t lim = (int64_t)limf;
    // For each (dx, dy) the admissible dz form at most two runs symmetric about 0;
    // find them from the float bounds, then confirm each candidate with the exact
    // comparison so membership and order equal the reference's cube scan.
    for (int64_t dx = -lim; dx <= lim; ++dx) {
        for (int64_t dy = -lim; dy <= lim; ++dy) {
            const int64_t r2 = dx * dx + dy * dy;
            if ((double)r2 >= high2) continue;
            double zmax_f = std::sqrt(std::max(0.0, high2 - (double)r2));
            int64_t zhi = std::min<int64_t>(lim, (int64_t)zmax_f + 2);
            double zmin_f = std::sqrt(std::max(0.0, low2 - (double)r2));
            int64_t zlo = std::max<int64_t>(0, (int64_t)zmin_f - 2);
            auto ok = [&](int64_t dz) {
                if (dx == 0 && dy == 0 && dz == 0) return false;
                double d2 = (double)(r2 + dz * dz);
                return d2 >= low2 && d2 < high2;
            };
            for (int64_t dz = -zhi; dz <= -zlo; ++dz)
                if (dz <= 0 && ok(dz)) out.push_back({(int32_t)dx, (int32_t)dy, (int32_t)dz});
            for (int64_t dz = std::max<int64_t>(zlo, 1); dz <= zhi; ++dz)
                if (ok(dz)) out.push_back({(int32_t)dx, (int32_t)dy, (int32_t)dz});
            // dz = 0 was emitted by the first loop only when zlo == 0
        }
    }
}

// Full (not domain-filtered) shells of radii 1..r_max, cached per (r_max, pitch):
// the enumeration is the expensive part and many grids share an r_max.
struct ShellSet {
    std::vector<Offset> offs;
    std::vector<uint64_t> start;  // r_max + 2
    int32_t dmax[3] = {0, 0, 0};  // largest |d| per axis
};

static std::shared_ptr<const ShellSet> full_shells(uint32_t r_max, double pitch) {
    static std::mutex mu;
    static std::map<std::pair<uint32_t, uint64_t>, std::shared_ptr<const ShellSet>> cache;
    static std::vector<std::pair<uint32_t, uint64_t>> order;
    uint64_t pb;
    std::memcpy(&pb, &pitch, 8);
    const auto key = std::make_pair(r_max, pb);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    auto set = std::make_shared<ShellSet>();
    set->start.assign(r_max + 2, 0);
    std::vector<Offset> shell;
    for (uint32_t r = 1; r <= r_max; ++r) {
        shell_offsets_full((double)r / pitch, 0.5 / pitch, shell);
        for (const Offset &o : shell) {
            set->offs.push_back(o);
            set->dmax[0] = std::max(set->dmax[0], std::abs(o.dx));
            set->dmax[1] = std::max(set->dmax[1], std::abs(o.dy));
            set->dmax[2] = std::max(set->dmax[2], std::abs(o.dz));
        }
        set->start[r + 1] = set->offs.size();
    }
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    cache[key] = set;
    order.push_back(key);
    if (order.size() > 24) {  // keep the most recent entries
        cache.erase(order.front());
        order.erase(order.begin());
    }
    return set;
}

void shell_order_offsets(const uint64_t dims[3], uint32_t r_max, double pitch,
                         std::vector<Offset> &offs, std::vector<uint64_t> &radius_start) {
    auto set = full_shells(r_max, pitch);
    const bool all_in = (uint64_t)set->dmax[0] < dims[0] && (uint64_t)set->dmax[1] < dims[1] &&
                        (uint64_t)set->dmax[2] < dims[2];
    if (all_in) {
        offs = set->offs;
        radius_start = set->start;
        return;
    }
    offs.clear();
    radius_start.assign(r_max + 2, 0);
    for (uint32_t r = 1; r <= r_max; ++r) {
        for (uint64_t i = set->start[r]; i < set->start[r + 1]; ++i) {
            const Offset &o = set->offs[i];
            if ((uint64_t)std::abs((int64_t)o.dx) < dims[0] && (uint64_t)std::abs((int64_t)o.dy) < dims[1] &&
                (uint64_t)std::abs((int64_t)o.dz) < dims[2])
                offs.push_back(o);
        }
        radius_start[r + 1] = offs.size();
    }
}

// src/geometry/s2.rs:393-501 fill_missing_s2_with_smooth_interpolation.
void fill_missing(double *values, const uint8_t *has_support, uint64_t n, double vf) {
    if (n == 0) return;
    values[0] = vf;
    std::vector<uint64_t> xk;
    std::vector<double> yk;
    for (uint64_t i = 0; i < n; ++i)
        if (i == 0 || has_support[i]) {
            xk.push_back(i);
            yk.push_back(values[i]);
        }
    const uint64_t m = xk.size();
    auto clamp01 = [](double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); };
    if (m < 2) return;
    if (m == 2) {
        uint64_t xl = xk[0], xr = xk[1];
        if (xr <= xl + 1) return;
        double yl = yk[0], yr = yk[1], width = (double)(xr - xl);
        for (uint64_t i = xl + 1; i < xr; ++i) {
            if (has_support[i]) continue;
            double t = (double)(i - xl) / width;
            double s = t * t * (3.0 - 2.0 * t);
            values[i] = clamp01(yl + (yr - yl) * s);
        }
        return;
    }
    std::vector<double> xs(m), h(m - 1), a(m, 0.0), b(m, 0.0), c(m, 0.0), d(m, 0.0), m2(m, 0.0);
    for (uint64_t i = 0; i < m; ++i) xs[i] = (double)xk[i];
    for (uint64_t i = 0; i + 1 < m; ++i) h[i] = std::max(xs[i + 1] - xs[i], 1e-12);
    b[0] = 1.0;
    b[m - 1] = 1.0;
    for (uint64_t i = 1; i + 1 < m; ++i) {
        a[i] = h[i - 1];
        b[i] = 2.0 * (h[i - 1] + h[i]);
        c[i] = h[i];
        d[i] = 6.0 * ((yk[i + 1] - yk[i]) / h[i] - (yk[i] - yk[i - 1]) / h[i - 1]);
    }
    auto guard = [](double v) { return std::fabs(v) < 1e-12 ? 1e-12 : v; };
    for (uint64_t i = 1; i < m; ++i) {
        double w = a[i] / guard(b[i - 1]);
        b[i] -= w * c[i - 1];
        d[i] -= w * d[i - 1];
    }
    m2[m - 1] = d[m - 1] / guard(b[m - 1]);
    for (uint64_t i = m - 1; i-- > 0;) m2[i] = (d[i] - c[i] * m2[i + 1]) / guard(b[i]);
    for (uint64_t seg = 0; seg + 1 < m; ++seg) {
        uint64_t xl = xk[seg], xr = xk[seg + 1];
        if (xr <= xl + 1) continue;
        double x0 = xs[seg], x1 = xs[seg + 1], y0 = yk[seg], y1 = yk[seg + 1];
        double hs = std::max(x1 - x0, 1e-12);
        double m20 = m2[seg], m21 = m2[seg + 1];
        for (uint64_t i = xl + 1; i < xr; ++i) {
            if (has_support[i]) continue;
            double x = (double)i;
            double ac = (x1 - x) / hs, bc = (x - x0) / hs;
            double y = ac * y0 + bc * y1 + ((ac * ac * ac - ac) * m20 + (bc * bc * bc - bc) * m21) * (hs * hs / 6.0);
            values[i] = clamp01(y);
        }
    }
}

}  // namespace s2h

using namespace s2h;

// ================================================================ C ABI

extern "C" {

const char *s2h_last_error(void) { return g_last_error.c_str(); }

s2h_status s2h_create(int device, uint64_t hbm_budget_bytes, s2h_ctx **out) {
    if (!out) return fail(S2H_ERR_ARG, "null output handle");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(S2H_ERR_HIP, std::string("no HIP device: ") + hipGetErrorString(e));
    if (device < 0 || device >= count)
        return fail(S2H_ERR_ARG, "device index " + std::to_string(device) + " out of range");
    S2H_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    S2H_HIP(hipGetDeviceProperties(&prop, device));
    auto *ctx = new s2h_ctx();
    ctx->device = device;
    ctx->budget = hbm_budget_bytes;
    ctx->name = std::string(prop.name) + " " + prop.gcnArchName;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(S2H_ERR_HIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
    }
    const s2h_status st = noalloc_scratch_init(device);
    if (st != S2H_OK) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return st;
    }
    *out = ctx;
    return S2H_OK;
}

void s2h_destroy(s2h_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->exact_plans) exact_plan_release(ctx, kv.second.get());
    ctx->exact_plans.clear();
    shell_devs_release(ctx);
    for (auto &kv : ctx->mc_plans) {
        dev_free(ctx, kv.second->d_shell);
        dev_free(ctx, kv.second->d_start);
        dev_free(ctx, kv.second->d_hv);
    }
    for (auto &kv : ctx->fft_plans) fft_plan_release(ctx, kv.second);
    for (auto &kv : ctx->fftd_plans) fft_plan_release(ctx, kv.second);
    dev_free(ctx, ctx->fft_arena);
    dev_free(ctx, ctx->d_scalar);
    dev_free(ctx, ctx->terms);
    batch_ws_release(ctx);
    dev_cache_release(ctx);
    for (hipEvent_t &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

s2h_status s2h_device_name(s2h_ctx *ctx, char *buf, uint64_t len) {
    if (!ctx || !buf || len == 0) return fail(S2H_ERR_ARG, "null argument");
    std::snprintf(buf, len, "%s", ctx->name.c_str());
    return S2H_OK;
}

s2h_status s2h_profile(s2h_ctx *ctx, int enable) {
    if (!ctx) return fail(S2H_ERR_ARG, "null argument");
    S2H_HIP(hipSetDevice(ctx->device));
    for (hipEvent_t &e : ctx->ev)
        if (!e) S2H_HIP(hipEventCreate(&e));
    ctx->profiling = enable != 0;
    return S2H_OK;
}

s2h_status s2h_profile_read(s2h_ctx *ctx, double *ms, uint64_t n) {
    if (!ctx || !ms) return fail(S2H_ERR_ARG, "null argument");
    if (n != 4) return fail(S2H_ERR_SIZE, "profile buffer must hold 4 values");
    for (int i = 0; i < 4; ++i) ms[i] = ctx->prof_ms[i];
    return S2H_OK;
}

s2h_status s2h_bytes_in_use(s2h_ctx *ctx, uint64_t *bytes) {
    if (!ctx || !bytes) return fail(S2H_ERR_ARG, "null argument");
    *bytes = ctx->in_use;
    return S2H_OK;
}

s2h_status s2h_vol_from_packed(s2h_ctx *ctx, const uint64_t *words, uint64_t nwords,
                               const uint64_t dims[3], s2h_vol **out) {
    if (!ctx || !words || !dims) return fail(S2H_ERR_ARG, "null argument");
    S2H_HIP(hipSetDevice(ctx->device));
    s2h_vol *v = nullptr;
    S2H_TRY(vol_alloc(ctx, dims, &v));
    if (nwords != v->nwords) {
        s2h_vol_free(ctx, v);
        return fail(S2H_ERR_SIZE, "packed volume needs " + std::to_string(v->nwords) + " words, got " +
                                      std::to_string(nwords));
    }
    // Volumes of 256 KiB and more go through the context's page-locked staging
    // buffer: one host copy, then a DMA at full link speed (a copy from
    // pageable memory is staged by the runtime in small synchronous pieces).
    const uint64_t bytes = nwords * 8;
    void *stage = nullptr;
    if (bytes >= (256u << 10)) {
        const s2h_status st = pinned_staging(ctx, bytes, &stage);
        if (st != S2H_OK) {
            s2h_vol_free(ctx, v);
            return st;
        }
        std::memcpy(stage, words, bytes);
    }
    hipError_t e = hipMemcpyAsync(v->d_words, stage ? stage : (const void *)words, bytes, hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        s2h_vol_free(ctx, v);
        return fail(S2H_ERR_HIP, std::string("volume upload failed: ") + hipGetErrorString(e));
    }
    *out = v;
    return S2H_OK;
}

s2h_status s2h_vol_from_bytes(s2h_ctx *ctx, const uint8_t *occ, const uint64_t dims[3], s2h_vol **out) {
    if (!ctx || !occ || !dims) return fail(S2H_ERR_ARG, "null argument");
    S2H_HIP(hipSetDevice(ctx->device));
    s2h_vol *v = nullptr;
    S2H_TRY(vol_alloc(ctx, dims, &v));
    uint64_t cells = dims[0] * dims[1] * dims[2];
    DevBuf tmp;
    s2h_status st = upload(ctx, tmp, occ, cells);
    if (st != S2H_OK) {
        s2h_vol_free(ctx, v);
        return st;
    }
    uint64_t rows = dims[0] * dims[1];
    uint64_t n = rows * v->row_words;
    pack_bytes_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(
        (const uint8_t *)tmp.ptr, v->d_words, rows, dims[2], v->row_words);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, tmp);
    if (e != hipSuccess) {
        s2h_vol_free(ctx, v);
        return fail(S2H_ERR_HIP, std::string("pack kernel failed: ") + hipGetErrorString(e));
    }
    *out = v;
    return S2H_OK;
}

s2h_status s2h_vol_read_packed(s2h_ctx *ctx, const s2h_vol *vol, uint64_t *words, uint64_t nwords) {
    if (!ctx || !vol || !words) return fail(S2H_ERR_ARG, "null argument");
    if (nwords != vol->nwords)
        return fail(S2H_ERR_SIZE, "volume has " + std::to_string(vol->nwords) + " words, buffer " +
                                      std::to_string(nwords));
    S2H_HIP(hipSetDevice(ctx->device));
    S2H_HIP(hipMemcpyAsync(words, vol->d_words, nwords * 8, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    return S2H_OK;
}

s2h_status s2h_vol_dims(const s2h_vol *vol, uint64_t dims[3]) {
    if (!vol || !dims) return fail(S2H_ERR_ARG, "null argument");
    for (int a = 0; a < 3; ++a) dims[a] = vol->dims[a];
    return S2H_OK;
}

s2h_status s2h_vol_occupied(s2h_ctx *ctx, s2h_vol *vol, uint64_t *occupied) {
    if (!ctx || !vol || !occupied) return fail(S2H_ERR_ARG, "null argument");
    S2H_HIP(hipSetDevice(ctx->device));
    return vol_popcount(ctx, vol, occupied);
}

void s2h_vol_free(s2h_ctx *ctx, s2h_vol *vol) {
    if (!ctx || !vol) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    DevBuf b;
    b.ptr = vol->d_words;
    b.bytes = vol->alloc_bytes;
    dev_free(ctx, b);
    delete vol;
}

s2h_status s2h_shell_count(const uint64_t dims[3], uint32_t r_max, double pitch, uint64_t *n_offsets) {
    if (!dims || !n_offsets) return fail(S2H_ERR_ARG, "null argument");
    if (!(pitch > 0.0) || !std::isfinite(pitch)) return fail(S2H_ERR_ARG, "pitch must be finite and > 0");
    std::vector<Offset> offs;
    std::vector<uint64_t> rs;
    shell_order_offsets(dims, r_max, pitch, offs, rs);
    *n_offsets = offs.size();
    return S2H_OK;
}

s2h_status s2h_shell_offsets(const uint64_t dims[3], uint32_t r_max, double pitch, int64_t *out,
                             uint64_t n_offsets, uint64_t *radius_start, uint64_t n_radius_start) {
    if (!dims || !out || !radius_start) return fail(S2H_ERR_ARG, "null argument");
    if (!(pitch > 0.0) || !std::isfinite(pitch)) return fail(S2H_ERR_ARG, "pitch must be finite and > 0");
    std::vector<Offset> offs;
    std::vector<uint64_t> rs;
    shell_order_offsets(dims, r_max, pitch, offs, rs);
    if (offs.size() != n_offsets) return fail(S2H_ERR_SIZE, "offset buffer length mismatch");
    if (rs.size() != n_radius_start) return fail(S2H_ERR_SIZE, "radius_start length mismatch");
    for (uint64_t i = 0; i < offs.size(); ++i) {
        out[3 * i] = offs[i].dx;
        out[3 * i + 1] = offs[i].dy;
        out[3 * i + 2] = offs[i].dz;
    }
    std::memcpy(radius_start, rs.data(), rs.size() * 8);
    return S2H_OK;
}

s2h_status s2h_fill_missing(double *values, const uint8_t *has_support, uint64_t n, double vf) {
    if (!values || !has_support) return fail(S2H_ERR_ARG, "null argument");
    fill_missing(values, has_support, n, vf);
    return S2H_OK;
}

}  // extern "C"
