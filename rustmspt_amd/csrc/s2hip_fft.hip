// FP32 real-FFT autocorrelation S2 (gfx950).
//
// Restates the FFT exact kernel's correlation (src/geometry/s2.rs:965-1012,
// 1219-1261) in FP32: the occupancy is zero-padded to the reference's
// 2,3,5-smooth lengths (padded_fft_dims, src/geometry/s2.rs:554-562), transformed,
// squared in magnitude and transformed back; the correlation at each shell
// offset d (wrapped, d mod f) is the pair count C(d) plus FP32 round-off.  The
// values are NOT rounded: this path reports its error against an FP64 FFT and
// the exact integer counts instead of claiming exactness.
//
// Layout and passes (H = fz/2 + 1, R_a = largest |d_a| of any shell offset;
// spectrum rows are P >= H values apart, P a multiple of 128 bytes, so every
// row starts on a 128-byte line and a tile of 128 bytes of kz columns reads and
// writes whole lines):
//   z-fwd : two real z rows per complex FFT (a + ib) read straight from the
//           packed bits; DIF in LDS; split into the two half spectra
//                                                  -> G1[nx][fy][H]  (y < ny)
//   y-fwd : per (x, tile of kz) DIF along y, rows y >= ny are zero -> G1
//   x     : per (ky, tile of kz) DIF along x (x >= nx zero), |F|^2, DIT back,
//           keep the 2R_x+1 needed x           -> G2[2Rx+1][fy][H]
//   y-inv : per (ix, tile of kz) DIT along y, keep 2R_y+1 y -> G3[2Rx+1][2Ry+1][H]
//   z-inv : two spectra per complex inverse (Hermitian extension), keep 2R_z+1
//           z, scale by 1/(fx fy fz)            -> cube[2Rx+1][2Ry+1][2Rz+1]
//   gather: cube value at every shell offset, in shell order.
// Forward transforms are in-place decimation-in-frequency (output in
// digit-reversed order) and inverses the matching decimation-in-time (input in
// that order, output natural), so no reordering pass is needed; the pointwise
// |F|^2 is order-independent.
//
// Where an axis length is L = A B with in-register DFTs of A and B points
// (x2_split: 96 ... 512), each pass has a two-stage form (fft_z_fwd2,
// fft_y_fwd2, fft_x_power2, fft_y_inv2, fft_z_inv2): stage A, an A-point DFT
// of x[b + B r] straight from global memory (or the staged bits) times
// w^(b q), one LDS exchange, stage B, the B-point DFT of each block, giving
// X[q + A k] in registers.  The two-stage y forward pass leaves natural
// frequency order and the two-stage y inverse reads it, so the y pair is
// switched together; the x pass (per ky line) and the z passes (natural kz at
// both ends) mix freely with either form.  The tile tuner times both forms.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <mutex>
#include <set>

#include "s2hip_internal.h"

namespace s2h {

static constexpr int kMaxFactors = 24;
static constexpr int kTileMax = 16;   // kz columns per workgroup in the y and x passes (max)
static constexpr int kFftThreads = 256;      // z passes
static constexpr int kFftMaxThreads = 1024;  // y and x passes: one first-stage butterfly per thread, up to this

// Values between spectrum rows: H = f/2 + 1 rounded up to a whole number of
// 128-byte lines.
__host__ __device__ constexpr uint64_t spectrum_pitch(uint64_t f, int real_bytes) {
    const uint64_t per_line = 128 / (2 * (uint64_t)real_bytes);
    return (f / 2 + 1 + per_line - 1) / per_line * per_line;
}

struct AxisPlan {
    int L = 1;
    int nf = 0;
    int fac[kMaxFactors];
};

struct FftPlan {
    uint64_t dims[3];
    uint64_t f[3];
    int R[3];
    AxisPlan ax[3];
    int real_bytes = 4;  // 4: FP32 path (float2 spectra), 8: FP64 rounded exact path (double2)
    DevBuf tw[3];    // per axis: the stages' twiddle rows (Fac::tw) in the plan's precision
    DevBuf rev_z;    // int32[fz]: natural frequency held at each DIF output position
    DevBuf pos_z;    // int32[fz]: DIF output position of each natural frequency
    DevBuf g1, g2, g3, cube, fcounts;
    int tile_y = 0;  // kz columns per workgroup in the y passes (0: not chosen yet)
    int tile_x = 0;  // kz columns per workgroup in the x pass
    int sh_z = 3, sh_y = 3, sh_x = 3;  // LDS padding shift per pass kind (3: a slot every 8; 30: none)
    // x pass in two register stages (fft_x_power2) where f_x = A B is one of
    // its lengths: A, B (0: the LDS-staged pass), its block padding in slots,
    // and the stage-A twiddle rows w^(b q) (b < B, q = 1 .. A - 1)
    int x2_a = 0, x2_b = 0, x2_pad = 0;
    bool x2_on = false;
    DevBuf tw_x2;
    // the same for the y passes (forward and inverse together: the two-stage
    // forward leaves natural frequency order, the LDS-staged one DIF order)
    // and for the z forward pass
    int y2_a = 0, y2_b = 0, y2_pad = 0, z2_a = 0, z2_b = 0;
    bool y2_on = false, z2_on = false;
    DevBuf tw_y2, tw_z2;
};

void fft_plan_release(s2h_ctx *ctx, FftPlan *p) {
    if (!p) return;
    for (auto &b : p->tw) dev_free(ctx, b);
    DevBuf *bufs[] = {&p->rev_z, &p->pos_z, &p->g1, &p->g2, &p->g3, &p->cube, &p->fcounts, &p->tw_x2, &p->tw_y2, &p->tw_z2};
    for (DevBuf *b : bufs) dev_free(ctx, *b);
    delete p;
}

namespace {

// t / n for 0 <= t < 2^31 without a division: a shift when n is a power of
// two, else the high word of t * m shifted by sh, with s = ceil(log2 n) and
// m = floor(2^(31+s) / n) + 1 < 2^32.  t * m / 2^(31+s) exceeds t / n by at
// most t / 2^(31+s) < 1 / n, too little to carry past the next multiple of n,
// so the quotient is exact.
struct FastDiv {
    uint32_t n = 1, m = 0;
    int32_t sh = 0;
};

FastDiv fast_div(uint32_t n) {
    FastDiv d;
    d.n = n;
    if ((n & (n - 1)) == 0) {
        d.m = 0;
        d.sh = __builtin_ctz(n);
        return d;
    }
    int s = 0;
    while ((1u << s) < n) ++s;
    d.m = (uint32_t)(((uint64_t)1 << (31 + s)) / n + 1);
    d.sh = s - 1;
    return d;
}

__device__ __forceinline__ void divmod(int t, const FastDiv &d, int &q, int &r) {
    const uint32_t u = (uint32_t)t;
    const uint32_t qq = d.m ? (__umulhi(u, d.m) >> d.sh) : (u >> d.sh);
    q = (int)qq;
    r = t - (int)(qq * d.n);
}

// Stages of one axis: radix f[s] of stage s (forward order; f[kMaxFactors - 1]
// carries the length L), the offset of its twiddle rows in the axis table, the
// butterflies per line nb[s] = L / f[s] and the sub-length quotient
// mp[s] = m_s / f[s] (m_0 = L, m_{s+1} = mp[s]) as fast divisors, and L itself.
// Stage s's row k (k < mp[s]) holds w^(k q L / m_s), q = 1 .. f[s] - 1, with
// w = exp(-2 pi i / L): a butterfly reads its twiddles from one contiguous row.
struct Fac {
    int n;
    int f[kMaxFactors];
    int tw[kMaxFactors];
    FastDiv nb[kMaxFactors];
    FastDiv mp[kMaxFactors];
    FastDiv len;
};

// LDS layout of a line: element j sits at pidx(j, sh) = j + (j >> sh), with
// sh = kPad (one slot every 8) or kNoPad.  The radix-8 stages of a 512-point
// line read with strides 64 and 8 (and 1): unpadded, the lanes of a wave hit
// the same banks 8 ways; padded, 16 consecutive lanes reach 16 different
// 16-byte bank groups.  The padding costs 1/8 more LDS, which for some tiles
// leaves room for one workgroup fewer per CU, so each pass kind has its own
// choice.  Lines are lstride(L, sh) apart (odd, so the load phases, which step
// across lines, spread too).
static constexpr int kPad = 3;
static constexpr int kNoPad = 30;
__host__ __device__ constexpr int pidx(int j, int sh) { return j + (j >> sh); }
__host__ __device__ constexpr int ltw(int L, int sh) { return pidx(L - 1, sh) + 1; }
__host__ __device__ constexpr int lstride(int L, int sh) { return ltw(L, sh) | 1; }

// q = t / n and r = t - q n; by shift and mask when n is a power of two
// (lg = log2 n, else -1; wave-uniform), which is the usual case for the
// stage and tile arithmetic, saving an integer division per use.
__device__ __forceinline__ int log2_if_pow2(int n) { return (n & (n - 1)) == 0 ? __builtin_ctz(n) : -1; }
__device__ __forceinline__ void divmod(int t, int n, int lg, int &q, int &r) {
    if (lg >= 0) {
        q = t >> lg;
        r = t & (n - 1);
    } else {
        q = t / n;
        r = t - q * n;
    }
}

// Complex values of the working precision T (float: the FP32 path; double:
// the rounded exact path).
template <class T> struct Cx;
template <> struct Cx<float> { using type = float2; };
template <> struct Cx<double> { using type = double2; };
template <class T> using C2 = typename Cx<T>::type;

template <class V>
__device__ __forceinline__ V mk(decltype(V::x) re, decltype(V::x) im) {
    V v;
    v.x = re;
    v.y = im;
    return v;
}
template <class V>
__device__ __forceinline__ V cmul(V a, V b) {
    return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class V>
__device__ __forceinline__ V cmulc(V a, V b) {  // a * conj(b)
    return mk<V>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <class V>
__device__ __forceinline__ V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <class V>
__device__ __forceinline__ V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }

// 4-point DFT in registers, natural order in and out.
template <int SIGN, class V>
__device__ __forceinline__ void dft4(V &a0, V &a1, V &a2, V &a3) {
    V s0 = cadd(a0, a2), d0 = csub(a0, a2);
    V s1 = cadd(a1, a3), d1 = csub(a1, a3);
    // multiply d1 by -i (forward) or +i (inverse)
    V jd1 = SIGN < 0 ? mk<V>(d1.y, -d1.x) : mk<V>(-d1.y, d1.x);
    a0 = cadd(s0, s1);
    a2 = csub(s0, s1);
    a1 = cadd(d0, jd1);
    a3 = csub(d0, jd1);
}

// R-point DFT in registers; sign = -1 forward, +1 inverse (unnormalized).
template <int SIGN, class V>
__device__ __forceinline__ void small_dft(V *a, int R) {
    using T = decltype(V::x);
    if (R == 8) {
        // even and odd 4-point DFTs, then X[k] = E[k] + w^k O[k], X[k+4] = E[k] - w^k O[k]
        V e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
        V o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
        dft4<SIGN>(e0, e1, e2, e3);
        dft4<SIGN>(o0, o1, o2, o3);
        const T h = (T)0.70710678118654752440;
        V t1, t2, t3;
        if (SIGN < 0) {  // w = exp(-i pi / 4)
            t1 = mk<V>(h * (o1.x + o1.y), h * (o1.y - o1.x));
            t2 = mk<V>(o2.y, -o2.x);
            t3 = mk<V>(h * (o3.y - o3.x), -h * (o3.x + o3.y));
        } else {         // w = exp(+i pi / 4)
            t1 = mk<V>(h * (o1.x - o1.y), h * (o1.x + o1.y));
            t2 = mk<V>(-o2.y, o2.x);
            t3 = mk<V>(-h * (o3.x + o3.y), h * (o3.x - o3.y));
        }
        a[0] = cadd(e0, o0);
        a[4] = csub(e0, o0);
        a[1] = cadd(e1, t1);
        a[5] = csub(e1, t1);
        a[2] = cadd(e2, t2);
        a[6] = csub(e2, t2);
        a[3] = cadd(e3, t3);
        a[7] = csub(e3, t3);
    } else if (R == 2) {
        V t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if (R == 4) {
        dft4<SIGN>(a[0], a[1], a[2], a[3]);
    } else if (R == 3) {
        const T c = (T)-0.5, s = (T)(SIGN * 0.86602540378443864676);
        V t1 = cadd(a[1], a[2]), t2 = csub(a[1], a[2]);
        V m = mk<V>(a[0].x + c * t1.x, a[0].y + c * t1.y);
        V js = mk<V>(-s * t2.y, s * t2.x);  // i*s*t2
        a[0] = cadd(a[0], t1);
        a[1] = cadd(m, js);
        a[2] = csub(m, js);
    } else {  // R == 5
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)(SIGN * 0.95105651629515357212), s2 = (T)(SIGN * 0.58778525229247312917);
        V b1 = cadd(a[1], a[4]), b4 = csub(a[1], a[4]);
        V b2 = cadd(a[2], a[3]), b3 = csub(a[2], a[3]);
        V x0 = a[0];
        V m1 = mk<V>(x0.x + c1 * b1.x + c2 * b2.x, x0.y + c1 * b1.y + c2 * b2.y);
        V m2 = mk<V>(x0.x + c2 * b1.x + c1 * b2.x, x0.y + c2 * b1.y + c1 * b2.y);
        // i * (s1 b4 + s2 b3) and i * (s2 b4 - s1 b3)
        V n1 = mk<V>(-(s1 * b4.y + s2 * b3.y), s1 * b4.x + s2 * b3.x);
        V n2 = mk<V>(-(s2 * b4.y - s1 * b3.y), s2 * b4.x - s1 * b3.x);
        a[0] = cadd(x0, cadd(b1, b2));
        a[1] = cadd(m1, n1);
        a[4] = csub(m1, n1);
        a[2] = cadd(m2, n2);
        a[3] = csub(m2, n2);
    }
}

// cos and sin of 2 pi k / N, correctly rounded, for the lengths of the
// composite in-register transforms below (k < N).
__host__ __device__ constexpr double unit_root_cos(int N, int k) {
    constexpr double c16[16] = {1.0, 0.923879532511286756128, 0.707106781186547524401, 0.382683432365089771728,
                                0.0, -0.382683432365089771728, -0.707106781186547524401, -0.923879532511286756128,
                                -1.0, -0.923879532511286756128, -0.707106781186547524401, -0.382683432365089771728,
                                0.0, 0.382683432365089771728, 0.707106781186547524401, 0.923879532511286756128};
    constexpr double c20[20] = {1.0, 0.951056516295153572116, 0.809016994374947424102, 0.587785252292473129169,
                                0.309016994374947424102, 0.0, -0.309016994374947424102, -0.587785252292473129169,
                                -0.809016994374947424102, -0.951056516295153572116, -1.0, -0.951056516295153572116,
                                -0.809016994374947424102, -0.587785252292473129169, -0.309016994374947424102, 0.0,
                                0.309016994374947424102, 0.587785252292473129169, 0.809016994374947424102,
                                0.951056516295153572116};
    constexpr double c10[10] = {1.0, 0.809016994374947424102, 0.309016994374947424102, -0.309016994374947424102, -0.809016994374947424102, -1.0, -0.809016994374947424102, -0.309016994374947424102, 0.309016994374947424102, 0.809016994374947424102};
    constexpr double c12[12] = {1.0, 0.866025403784438646764, 0.500000000000000000000, 0.0, -0.500000000000000000000, -0.866025403784438646764, -1.0, -0.866025403784438646764, -0.500000000000000000000, 0.0, 0.500000000000000000000, 0.866025403784438646764};
    constexpr double c15[15] = {1.0, 0.913545457642600895502, 0.669130606358858213826, 0.309016994374947424102, -0.104528463267653471400, -0.500000000000000000000, -0.809016994374947424102, -0.978147600733805637929, -0.978147600733805637929, -0.809016994374947424102, -0.500000000000000000000, -0.104528463267653471400, 0.309016994374947424102, 0.669130606358858213826, 0.913545457642600895502};
    constexpr double c24[24] = {1.0, 0.965925826289068286750, 0.866025403784438646764, 0.707106781186547524401, 0.500000000000000000000, 0.258819045102520762349, 0.0, -0.258819045102520762349, -0.500000000000000000000, -0.707106781186547524401, -0.866025403784438646764, -0.965925826289068286750, -1.0, -0.965925826289068286750, -0.866025403784438646764, -0.707106781186547524401, -0.500000000000000000000, -0.258819045102520762349, 0.0, 0.258819045102520762349, 0.500000000000000000000, 0.707106781186547524401, 0.866025403784438646764, 0.965925826289068286750};
    constexpr double c32[32] = {1.0, 0.980785280403230449126, 0.923879532511286756128, 0.831469612302545237079, 0.707106781186547524401, 0.555570233019602224743, 0.382683432365089771728, 0.195090322016128267848, 0.0, -0.195090322016128267848, -0.382683432365089771728, -0.555570233019602224743, -0.707106781186547524401, -0.831469612302545237079, -0.923879532511286756128, -0.980785280403230449126, -1.0, -0.980785280403230449126, -0.923879532511286756128, -0.831469612302545237079, -0.707106781186547524401, -0.555570233019602224743, -0.382683432365089771728, -0.195090322016128267848, 0.0, 0.195090322016128267848, 0.382683432365089771728, 0.555570233019602224743, 0.707106781186547524401, 0.831469612302545237079, 0.923879532511286756128, 0.980785280403230449126};
    switch (N) {
        case 32: return c32[k % 32];
        case 10: return c10[k % 10];
        case 12: return c12[k % 12];
        case 15: return c15[k % 15];
        case 16: return c16[k % 16];
        case 20: return c20[k % 20];
        default: return c24[k % 24];
    }
}
__host__ __device__ constexpr double unit_root_sin(int N, int k) {
    constexpr double s10[10] = {0.0, 0.587785252292473129169, 0.951056516295153572116, 0.951056516295153572116, 0.587785252292473129169, 0.0, -0.587785252292473129169, -0.951056516295153572116, -0.951056516295153572116, -0.587785252292473129169};
    constexpr double s15[15] = {0.0, 0.406736643075800207754, 0.743144825477394235015, 0.951056516295153572116, 0.994521895368273336923, 0.866025403784438646764, 0.587785252292473129169, 0.207911690817759337102, -0.207911690817759337102, -0.587785252292473129169, -0.866025403784438646764, -0.994521895368273336923, -0.951056516295153572116, -0.743144825477394235015, -0.406736643075800207754};
    if (N == 10) return s10[k % 10];
    if (N == 15) return s15[k % 15];
    return unit_root_cos(N, (k + 3 * N / 4) % N);  // sin t = cos(t - pi/2) where 4 | N
}

// x * exp(SIGN 2 pi i e / N) with e a compile-time constant after unrolling;
// the quarter turns are exact swaps and sign flips.
template <int N, int SIGN, class V>
__device__ __forceinline__ V mul_root(V x, int e) {
    using T = decltype(V::x);
    e %= N;
    if (e == 0) return x;
    if (4 * e == N) return SIGN > 0 ? mk<V>(-x.y, x.x) : mk<V>(x.y, -x.x);
    if (2 * e == N) return mk<V>(-x.x, -x.y);
    if (4 * e == 3 * N) return SIGN > 0 ? mk<V>(x.y, -x.x) : mk<V>(-x.y, x.x);
    const T c = (T)unit_root_cos(N, e), s = (T)(SIGN * unit_root_sin(N, e));
    return mk<V>(x.x * c - x.y * s, x.x * s + x.y * c);
}

// N = N1 N2 point DFT in registers, natural order in and out (sign as
// small_dft): input n = N2 n1 + n2, output k = k1 + N1 k2;
// X[k1 + N1 k2] = sum_n2 w_N2^(n2 k2) w_N^(n2 k1) sum_n1 x[N2 n1 + n2] w_N1^(n1 k1).
// Its round-off is that of its two factor stages and the twiddle product
// between them, as for the same factors run as LDS stages.
template <int N1, int N2, int SIGN, class V>
__device__ __forceinline__ void dft_comp(V *a) {
    constexpr int N = N1 * N2;
    V u[N2][N1];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) u[n2][n1] = a[N2 * n1 + n2];
        small_dft<SIGN>(u[n2], N1);
#pragma unroll
        for (int k1 = 1; k1 < N1; ++k1) u[n2][k1] = mul_root<N, SIGN>(u[n2][k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        V v[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) v[n2] = u[n2][k1];
        small_dft<SIGN>(v, N2);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) a[k1 + N1 * k2] = v[k2];
    }
}

// The in-register DFTs of the two-stage passes: 8 as small_dft, 10 = 2 x 5,
// 12 = 4 x 3, 15 = 3 x 5, 16 = 4 x 4, 20 = 4 x 5, 24 = 8 x 3, 32 = 4 x 8.
template <int N, int SIGN, class V>
__device__ __forceinline__ void dft_reg(V *a) {
    static_assert(N == 8 || N == 10 || N == 12 || N == 15 || N == 16 || N == 20 || N == 24 || N == 32,
                  "in-register DFT length");
    if constexpr (N == 8) small_dft<SIGN>(a, 8);
    else if constexpr (N == 10) dft_comp<2, 5, SIGN>(a);
    else if constexpr (N == 12) dft_comp<4, 3, SIGN>(a);
    else if constexpr (N == 15) dft_comp<3, 5, SIGN>(a);
    else if constexpr (N == 16) dft_comp<4, 4, SIGN>(a);
    else if constexpr (N == 20) dft_comp<4, 5, SIGN>(a);
    else if constexpr (N == 24) dft_comp<8, 3, SIGN>(a);
    else dft_comp<4, 8, SIGN>(a);
}

// One radix-R stage s over all lines: in-place butterflies on sub-length
// m_s = R mp[s]; the butterfly (grp, k) takes its twiddles from row k of the
// stage (none on sub-length R, where mp == 1 and k == 0).
template <int R, bool INV, class V>
__device__ __forceinline__ void fft_stage(V *d, int S, int sh, int lines, const Fac &fc, int s, const V *tw) {
    const FastDiv dnb = fc.nb[s], dmp = fc.mp[s];
    const int nb = (int)dnb.n, mp = (int)dmp.n, m = mp * R;
    const V *trow = tw + fc.tw[s];
    for (int t = threadIdx.x; t < nb * lines; t += blockDim.x) {
        int line, b, grp, k;
        divmod(t, dnb, line, b);
        divmod(b, dmp, grp, k);
        V *base = d + line * S;
        const int j0 = grp * m + k;
        V a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = base[pidx(j0 + r * mp, sh)];
        const V *w = trow + k * (R - 1);
        if (!INV) {
            small_dft<-1>(a, R);
            if (mp > 1) {
#pragma unroll
                for (int q = 1; q < R; ++q) a[q] = cmul(a[q], w[q - 1]);
            }
        } else {
            if (mp > 1) {
#pragma unroll
                for (int q = 1; q < R; ++q) a[q] = cmulc(a[q], w[q - 1]);
            }
            small_dft<1>(a, R);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) base[pidx(j0 + r * mp, sh)] = a[r];
    }
    __syncthreads();
}

template <bool INV, class V>
__device__ __forceinline__ void fft_stage_any(V *d, int S, int sh, int lines, const Fac &fc, int s, const V *tw) {
    switch (fc.f[s]) {
        case 8: fft_stage<8, INV>(d, S, sh, lines, fc, s, tw); break;
        case 4: fft_stage<4, INV>(d, S, sh, lines, fc, s, tw); break;
        case 2: fft_stage<2, INV>(d, S, sh, lines, fc, s, tw); break;
        case 3: fft_stage<3, INV>(d, S, sh, lines, fc, s, tw); break;
        default: fft_stage<5, INV>(d, S, sh, lines, fc, s, tw); break;
    }
}

// In-place mixed-radix transform of `lines` lines of length L held in LDS at
// d[line*S + pidx(pos, sh)].  Forward: decimation in frequency, natural in, digit-reversed
// out.  Inverse: decimation in time with conjugate twiddles, digit-reversed in,
// natural out (unnormalized); inverse stage s works on the same sub-length m_s
// as forward stage s.
template <bool INV, class V>
__device__ void lds_fft(V *d, int S, int sh, int lines, const Fac &fc, const V *tw) {
    if (!INV) {
        for (int s = 0; s < fc.n; ++s) fft_stage_any<false>(d, S, sh, lines, fc, s, tw);
    } else {
        for (int s = fc.n - 1; s >= 0; --s) fft_stage_any<true>(d, S, sh, lines, fc, s, tw);
    }
}

// |F|^2 correlation of `lines` lines in LDS: forward transform, |F|^2,
// inverse transform.  The last forward stage and the first inverse stage both
// work on sub-length R = f_{n-1} (mp = 1, no twiddles) over the same R
// elements, so one thread runs forward DFT, |F|^2 and inverse DFT on them in
// registers: one LDS round trip and two barriers fewer than running
// lds_fft<false>, a |F|^2 pass and lds_fft<true>.
template <int R, class V>
__device__ __forceinline__ void power_stage(V *d, int S, int sh, int lines, const FastDiv dnb) {
    const int nb = (int)dnb.n;
    for (int t = threadIdx.x; t < nb * lines; t += blockDim.x) {
        int line, b;
        divmod(t, dnb, line, b);
        V *base = d + line * S;
        V a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = base[pidx(b * R + r, sh)];
        small_dft<-1>(a, R);
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = mk<V>(a[r].x * a[r].x + a[r].y * a[r].y, 0);
        small_dft<1>(a, R);
#pragma unroll
        for (int r = 0; r < R; ++r) base[pidx(b * R + r, sh)] = a[r];
    }
    __syncthreads();
}

template <class V>
__device__ __forceinline__ void power_stage_any(V *d, int S, int sh, int lines, const Fac &fc) {
    const int last = fc.n - 1;  // sub-length f[last]: mp == 1
    switch (fc.f[last]) {
        case 8: power_stage<8>(d, S, sh, lines, fc.nb[last]); break;
        case 4: power_stage<4>(d, S, sh, lines, fc.nb[last]); break;
        case 2: power_stage<2>(d, S, sh, lines, fc.nb[last]); break;
        case 3: power_stage<3>(d, S, sh, lines, fc.nb[last]); break;
        default: power_stage<5>(d, S, sh, lines, fc.nb[last]); break;
    }
}

template <class V>
__device__ void lds_power(V *d, int S, int sh, int lines, const Fac &fc, const V *tw) {
    if (fc.n == 0) {  // L == 1
        for (int t = threadIdx.x; t < lines; t += blockDim.x) {
            const V v = d[t * S];
            d[t * S] = mk<V>(v.x * v.x + v.y * v.y, 0);
        }
        __syncthreads();
        return;
    }
    for (int s = 0; s + 1 < fc.n; ++s) fft_stage_any<false>(d, S, sh, lines, fc, s, tw);
    power_stage_any(d, S, sh, lines, fc);
    for (int s = fc.n - 2; s >= 0; --s) fft_stage_any<true>(d, S, sh, lines, fc, s, tw);
}

// Outer stages straight between global memory and registers.  Line c's
// element j lives at g[j * rs + c] (rs: the row stride; kz columns c vary
// fastest across lanes, so neighbouring lanes touch one contiguous row
// segment).  Each saves the LDS round trip and barrier of a plain load or
// store loop.
//   load_first_fwd: stage 0 of the forward (DIF) transform (sub-length L,
//     mp = L / R): thread (b, c) loads j = b + mp r (j < jlo, j >= nvalid are zero),
//     runs the butterfly and its twiddles, stores to LDS.
//   store_last_inv: stage 0 of the inverse (DIT) transform: butterfly from
//     LDS, conjugate twiddles, inverse DFT; writes only the kept j
//     (j <= Rk -> row j, and unless half j >= L - Rk -> row j - L + 2Rk + 1).
//   store_last_fwd: the forward transform's last stage (sub-length R, mp = 1,
//     no twiddles) from LDS, every output j = R b + r written to g.
//   load_first_inv: the inverse's first stage (sub-length R, mp = 1) from g.
// tw0 is stage 0's twiddle rows (row b: w^(b q), q = 1 .. R - 1).
template <int R, class V>
__device__ __forceinline__ void load_first_fwd(V *d, int L, int S, int sh, int tile, int lgt, int nk,
                                               const V *__restrict__ g, uint64_t rs, int jlo, int nvalid,
                                               const V *__restrict__ tw) {
    const int mp = L / R;
    for (int t = threadIdx.x; t < mp * tile; t += blockDim.x) {
        int b, c;
        divmod(t, tile, lgt, b, c);
        if (c >= nk) continue;
        V a[R];
        const V *gp = g + (uint64_t)b * rs + c;
        const uint64_t step = (uint64_t)mp * rs;  // rows j = b + mp r: pointer steps, no per-row multiply
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = b + mp * r;
            a[r] = (j >= jlo && j < nvalid) ? *gp : mk<V>(0, 0);
            gp += step;
        }
        small_dft<-1>(a, R);
        if (mp > 1) {
            const V *w = tw + b * (R - 1);
#pragma unroll
            for (int q = 1; q < R; ++q) a[q] = cmul(a[q], w[q - 1]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) d[c * S + pidx(b + mp * r, sh)] = a[r];
    }
    __syncthreads();
}

template <int R, class V>
__device__ __forceinline__ void store_last_inv(const V *d, int L, int S, int sh, int tile, int lgt, int nk,
                                               V *__restrict__ g, uint64_t rs, int Rk, bool half,
                                               const V *__restrict__ tw) {
    const int mp = L / R, nko = 2 * Rk + 1;
    for (int t = threadIdx.x; t < mp * tile; t += blockDim.x) {
        int b, c;
        divmod(t, tile, lgt, b, c);
        if (c >= nk) continue;
        V a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = d[c * S + pidx(b + mp * r, sh)];
        if (mp > 1) {
            const V *w = tw + b * (R - 1);
#pragma unroll
            for (int q = 1; q < R; ++q) a[q] = cmulc(a[q], w[q - 1]);
        }
        small_dft<1>(a, R);
        V *gc = g + c;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = b + mp * r;
            const int o = j <= Rk ? j : (!half && j >= L - Rk ? j - L + nko : -1);
            if (o >= 0) gc[(uint64_t)(uint32_t)o * rs] = a[r];
        }
    }
}

template <int R, class V>
__device__ __forceinline__ void store_last_fwd(const V *d, int L, int S, int sh, int tile, int lgt, int nk,
                                               V *__restrict__ g, uint64_t rs) {
    const int nb = L / R;
    for (int t = threadIdx.x; t < nb * tile; t += blockDim.x) {
        int b, c;
        divmod(t, tile, lgt, b, c);
        if (c >= nk) continue;
        V a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = d[c * S + pidx(b * R + r, sh)];
        small_dft<-1>(a, R);
        V *gp = g + (uint64_t)(b * R) * rs + c;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            *gp = a[r];
            gp += rs;
        }
    }
}

template <int R, class V>
__device__ __forceinline__ void load_first_inv(V *d, int L, int S, int sh, int tile, int lgt, int nk,
                                               const V *__restrict__ g, uint64_t rs) {
    const int nb = L / R;
    for (int t = threadIdx.x; t < nb * tile; t += blockDim.x) {
        int b, c;
        divmod(t, tile, lgt, b, c);
        if (c >= nk) continue;
        V a[R];
        const V *gp = g + (uint64_t)(b * R) * rs + c;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            a[r] = *gp;
            gp += rs;
        }
        small_dft<1>(a, R);
#pragma unroll
        for (int r = 0; r < R; ++r) d[c * S + pidx(b * R + r, sh)] = a[r];
    }
    __syncthreads();
}

#define S2H_RADIX_SWITCH(R, CALL)                    \
    switch (R) {                                     \
        case 8: { constexpr int RR = 8; CALL; } break; \
        case 4: { constexpr int RR = 4; CALL; } break; \
        case 2: { constexpr int RR = 2; CALL; } break; \
        case 3: { constexpr int RR = 3; CALL; } break; \
        default: { constexpr int RR = 5; CALL; } break; \
    }

// Stage 0 of the z forward transform (sub-length L, R = f_0, mp = L / R)
// from the packed rows staged in LDS: line pl holds rows y (real part) and
// y + 1 (imaginary part) of column x, at sw[(2 pl + i) row_words] (a missing
// row y + 1 staged as zeros); z >= nz is zero.
template <int R, class V>
__device__ __forceinline__ void z_first_stage(V *d, int S, int sh, int np, int nz, int row_words,
                                              const uint64_t *sw, const Fac &fc, const V *__restrict__ tw) {
    using T = decltype(V::x);
    const FastDiv dmp = fc.mp[0];
    const int mp = (int)dmp.n;
    for (int t = threadIdx.x; t < mp * np; t += blockDim.x) {
        int pl, b;
        divmod(t, dmp, pl, b);
        const uint64_t *ra = sw + 2 * pl * row_words;
        V a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int z = b + mp * r;
            T re = 0, im = 0;
            if (z < nz) {
                re = (T)((ra[z >> 6] >> (z & 63)) & 1ull);
                im = (T)((ra[row_words + (z >> 6)] >> (z & 63)) & 1ull);
            }
            a[r] = mk<V>(re, im);
        }
        small_dft<-1>(a, R);
        if (mp > 1) {
            const V *w = tw + fc.tw[0] + b * (R - 1);
#pragma unroll
            for (int q = 1; q < R; ++q) a[q] = cmul(a[q], w[q - 1]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) d[pl * S + pidx(b + mp * r, sh)] = a[r];
    }
    __syncthreads();
}

// LDS of a z inverse workgroup: its lines, then the natural frequency of
// every DIF position.
template <class T>
size_t z_inv_lds_bytes(int L, int pairs, int sh) {
    return (size_t)pairs * lstride(L, sh) * sizeof(C2<T>) + (size_t)L * 4;
}

// LDS of a z forward workgroup: its lines, then the packed rows of its pairs
// and the DIF output positions.
template <class T>
size_t z_fwd_lds_bytes(int L, int pairs, int sh, int row_words) {
    return (size_t)pairs * lstride(L, sh) * sizeof(C2<T>) + (size_t)pairs * 2 * row_words * 8 + (size_t)L * 4;
}

// ---- z forward: pairs of real rows (x, y), (x, y+1) -> half spectra in G1
template <class T>
__global__ __launch_bounds__(kFftThreads) void fft_z_fwd(const uint64_t *__restrict__ words, int nx, int ny, int nz,
                                                         int row_words, int fy, int H, int P, Fac fc, int L,
                                                         const C2<T> *__restrict__ tw, const int *__restrict__ pos,
                                                         int pairs_per_block, C2<T> *__restrict__ g1, int sh,
                                                         FastDiv dyp, FastDiv dh) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    V *sm = reinterpret_cast<V *>(sm_raw);
    V *d = sm;
    const int S = lstride(L, sh);
    const int npairs = nx * ((ny + 1) / 2);
    const int p0 = blockIdx.x * pairs_per_block;
    const int np = min(pairs_per_block, npairs - p0);
    // the pairs' packed rows and the DIF output positions, staged in LDS
    uint64_t *sw = reinterpret_cast<uint64_t *>(sm + (size_t)pairs_per_block * S);
    int *spos = reinterpret_cast<int *>(sw + (size_t)pairs_per_block * 2 * row_words);
    for (int t = threadIdx.x; t < np * 2 * row_words; t += blockDim.x) {
        const int pl = t / (2 * row_words), rem = t - pl * 2 * row_words;
        const int i = rem / row_words, wd = rem - i * row_words;
        int x, yp;
        divmod(p0 + pl, dyp, x, yp);
        const int y = 2 * yp + i;
        sw[t] = y < ny ? words[((uint64_t)x * ny + y) * row_words + wd] : 0ull;
    }
    for (int t = threadIdx.x; t < L; t += blockDim.x) spos[t] = pos[t];
    __syncthreads();
    if (fc.n >= 1) {
        // the first DIF stage from the staged bits: thread (line, b) takes
        // z = b + mp r of the two rows
        S2H_RADIX_SWITCH(fc.f[0], (z_first_stage<RR>(d, S, sh, np, nz, row_words, sw, fc, tw)));
        for (int s = 1; s < fc.n; ++s) fft_stage_any<false>(d, S, sh, np, fc, s, tw);
    } else {
        for (int t = threadIdx.x; t < np * L; t += blockDim.x) {
            int pl, z;
            divmod(t, fc.len, pl, z);
            T re = 0, im = 0;
            if (z < nz) {
                const uint64_t *ra = sw + 2 * pl * row_words;
                re = (T)((ra[z >> 6] >> (z & 63)) & 1ull);
                im = (T)((ra[row_words + (z >> 6)] >> (z & 63)) & 1ull);
            }
            d[pl * S + pidx(z, sh)] = mk<V>(re, im);
        }
        __syncthreads();
        lds_fft<false>(d, S, sh, np, fc, tw);
    }
    for (int t = threadIdx.x; t < np * H; t += blockDim.x) {
        int pl, k, x, yp;
        divmod(t, dh, pl, k);
        divmod(p0 + pl, dyp, x, yp);
        const int y = 2 * yp;
        const V zk = d[pl * S + pidx(spos[k], sh)];
        const V zn = d[pl * S + pidx(spos[k == 0 ? 0 : L - k], sh)];
        // A = (Z[k] + conj Z[-k]) / 2,  B = (Z[k] - conj Z[-k]) / (2i)
        const T half = (T)0.5;
        const V A = mk<V>(half * (zk.x + zn.x), half * (zk.y - zn.y));
        const V B = mk<V>(half * (zk.y + zn.y), -half * (zk.x - zn.x));
        g1[((uint64_t)x * fy + y) * P + k] = A;
        if (y + 1 < ny) g1[((uint64_t)x * fy + y + 1) * P + k] = B;
    }
}

// ---- y forward: per (x, kz tile), rows y >= ny are zero; in place in G1
template <class T>
__global__ __launch_bounds__(kFftMaxThreads) void fft_y_fwd(int ny, int fy, int H, int P, int tiles_kz, int tile, Fac fc,
                                                         const C2<T> *__restrict__ tw, C2<T> *__restrict__ g1, int sh,
                                                         const FftdCand *__restrict__ dcands, int Fx, int Rx,
                                                         int Ry) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    V *sm = reinterpret_cast<V *>(sm_raw);
    const int L = fy, S = lstride(L, sh), lgt = log2_if_pow2(tile);
    V *d = sm;
    const int x = blockIdx.x / tiles_kz, kz0 = (blockIdx.x % tiles_kz) * tile;
    int jlo = 0, jhi = ny;  // rows that may be nonzero
    if (dcands) {
        // frames: plane x is field (x / F_x) % 2 of candidate x / (2 F_x); a
        // delta plane outside the box's frame rows R_x .. R_x + |B_x| - 1 is
        // zero and is not read again (the x pass loads only the box rows),
        // and inside them delta is zero outside rows R_y .. R_y + |B_y| - 1
        const int c = x / (2 * Fx), field = (x / Fx) & 1, i = x % Fx;
        if (field == 1) {
            if (i < Rx || i >= Rx + (dcands[c].b1[0] - dcands[c].b0[0])) return;
            jlo = Ry;
            jhi = Ry + (dcands[c].b1[1] - dcands[c].b0[1]);
        }
    }
    const int nk = min(tile, H - kz0);
    V *base = g1 + (uint64_t)x * fy * P + kz0;
    if (fc.n >= 2) {
        S2H_RADIX_SWITCH(fc.f[0], (load_first_fwd<RR>(d, L, S, sh, tile, lgt, nk, base, (uint64_t)P, jlo, jhi,
                                                       tw + fc.tw[0])));
        for (int s = 1; s + 1 < fc.n; ++s) fft_stage_any<false>(d, S, sh, nk, fc, s, tw);
        S2H_RADIX_SWITCH(fc.f[fc.n - 1], (store_last_fwd<RR>(d, L, S, sh, tile, lgt, nk, base, (uint64_t)P)));
        return;
    }
    for (int t = threadIdx.x; t < L * tile; t += blockDim.x) {
        int y, c;
        divmod(t, tile, lgt, y, c);
        V v = mk<V>(0, 0);
        if (y >= jlo && y < jhi && c < nk) v = base[(uint64_t)y * P + c];
        d[c * S + pidx(y, sh)] = v;
    }
    __syncthreads();
    lds_fft<false>(d, S, sh, nk, fc, tw);
    for (int t = threadIdx.x; t < L * tile; t += blockDim.x) {
        int y, c;
        divmod(t, tile, lgt, y, c);
        if (c < nk) base[(uint64_t)y * P + c] = d[c * S + pidx(y, sh)];
    }
}

// Stages 1 .. n-1 of the forward transform, |F|^2 and stages n-1 .. 1 of the
// inverse (lds_power without the outer stages; n >= 2).
template <class V>
__device__ void lds_power_inner(V *d, int S, int sh, int lines, const Fac &fc, const V *tw) {
    for (int s = 1; s + 1 < fc.n; ++s) fft_stage_any<false>(d, S, sh, lines, fc, s, tw);
    power_stage_any(d, S, sh, lines, fc);
    for (int s = fc.n - 2; s >= 1; --s) fft_stage_any<true>(d, S, sh, lines, fc, s, tw);
}

// ---- x: forward, |F|^2, inverse; keep the needed x.  One block per (ky, kz tile).
template <class T>
__global__ __launch_bounds__(kFftMaxThreads) void fft_x_power(int nx, int fy, int H, int P, int tiles_kz, int tile, Fac fc,
                                                           const C2<T> *__restrict__ tw, int Rx,
                                                           const C2<T> *__restrict__ g1, C2<T> *__restrict__ g2, int sh) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    V *sm = reinterpret_cast<V *>(sm_raw);
    const int L = (int)fc.f[kMaxFactors - 1];  // unused slot carries L
    const int S = lstride(L, sh), lgt = log2_if_pow2(tile);
    V *d = sm;
    const int ky = blockIdx.x / tiles_kz, kz0 = (blockIdx.x % tiles_kz) * tile;
    const int nk = min(tile, H - kz0);
    if (fc.n >= 2) {
        const uint64_t rs = (uint64_t)fy * P;
        S2H_RADIX_SWITCH(fc.f[0], (load_first_fwd<RR>(d, L, S, sh, tile, lgt, nk, g1 + (uint64_t)ky * P + kz0, rs,
                                                       0, nx, tw + fc.tw[0])));
        lds_power_inner(d, S, sh, nk, fc, tw);
        S2H_RADIX_SWITCH(fc.f[0], (store_last_inv<RR>(d, L, S, sh, tile, lgt, nk, g2 + (uint64_t)ky * P + kz0, rs,
                                                       Rx, true, tw + fc.tw[0])));
        return;
    }
    for (int t = threadIdx.x; t < L * tile; t += blockDim.x) {
        int x, c;
        divmod(t, tile, lgt, x, c);
        V v = mk<V>(0, 0);
        if (x < nx && c < nk) v = g1[((uint64_t)x * fy + ky) * P + kz0 + c];
        d[c * S + pidx(x, sh)] = v;
    }
    __syncthreads();
    lds_power(d, S, sh, nk, fc, tw);
    for (int t = threadIdx.x; t < (Rx + 1) * tile; t += blockDim.x) {
        int ix, c;
        divmod(t, tile, lgt, ix, c);
        if (c >= nk) continue;
        g2[((uint64_t)ix * fy + ky) * P + kz0 + c] = d[c * S + pidx(ix, sh)];
    }
}

// ---- x in two register stages, for L = A B (A = 16, B = 16 or 20): one block
// per (ky, tile of TILE kz columns, 128 bytes of each row).
//   stage A: thread (b, c), b < B: the forward DIF butterfly of radix A on rows
//     x = b + B r (x >= nx zero) straight from G1, its twiddles w_L^(b q),
//     into LDS;
//   stage B: thread (q, c), q < A: the B-point block q (its sub-length after
//     stage A) forward, |F|^2 and back, all in registers;
//   stage A inverse: thread (b, c): conjugate twiddles, inverse DFT over q,
//     rows dx = b + B r <= Rx to G2.
// Every value makes two LDS round trips (the LDS-staged pass makes four) and
// the block has two barriers.  LDS: element j = B q + k of column c at
// q (B TILE + pad) + k TILE + c, i.e. column-interleaved so that the lanes of
// a wave (c fastest) touch consecutive 128-byte rows of slots, with an
// optional pad-slot gap after each B-element block (pad = 128 bytes puts
// stage B's neighbouring blocks on the other half of the banks).
template <class T>
__host__ __device__ constexpr int x2_tile() { return (int)(128 / (2 * sizeof(T))); }
template <class T>
size_t x2_lds_bytes(int A, int B, int pad) {
    return ((size_t)A * B * x2_tile<T>() + (size_t)(A - 1) * pad) * sizeof(C2<T>);
}
template <class T, int A, int B>
__global__ __launch_bounds__(512) void fft_x_power2(int nx, int fy, int H, int P, int tiles_kz, int Rx, int pad,
                                                    const C2<T> *__restrict__ twA, const C2<T> *__restrict__ g1,
                                                    C2<T> *__restrict__ g2) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    constexpr int TILE = x2_tile<T>();
    V *d = reinterpret_cast<V *>(sm_raw);
    const int ky = blockIdx.x / tiles_kz, kz0 = (blockIdx.x % tiles_kz) * TILE;
    const int nk = min(TILE, H - kz0);
    const uint64_t rs = (uint64_t)fy * P;
    const int t = threadIdx.x, c = t % TILE, grp = t / TILE;
    if (grp < B && c < nk) {
        const int b = grp;
        V a[A];
        const V *gp = g1 + (uint64_t)ky * P + kz0 + c + (uint64_t)b * rs;
        const uint64_t step = (uint64_t)B * rs;
#pragma unroll
        for (int r = 0; r < A; ++r) {
            a[r] = b + B * r < nx ? *gp : mk<V>(0, 0);
            gp += step;
        }
        dft_reg<A, -1>(a);
        const V *w = twA + b * (A - 1);
#pragma unroll
        for (int q = 1; q < A; ++q) a[q] = cmul(a[q], w[q - 1]);
        V *dp = d + b * TILE + c;
#pragma unroll
        for (int q = 0; q < A; ++q) dp[q * (B * TILE + pad)] = a[q];
    }
    __syncthreads();
    if (grp < A && c < nk) {
        V a[B];
        V *dp = d + grp * (B * TILE + pad) + c;
#pragma unroll
        for (int k = 0; k < B; ++k) a[k] = dp[k * TILE];
        dft_reg<B, -1>(a);
#pragma unroll
        for (int k = 0; k < B; ++k) a[k] = mk<V>(a[k].x * a[k].x + a[k].y * a[k].y, 0);
        dft_reg<B, 1>(a);
#pragma unroll
        for (int k = 0; k < B; ++k) dp[k * TILE] = a[k];
    }
    __syncthreads();
    if (grp < B && c < nk) {
        const int b = grp;
        V a[A];
        const V *dp = d + b * TILE + c;
#pragma unroll
        for (int q = 0; q < A; ++q) a[q] = dp[q * (B * TILE + pad)];
        const V *w = twA + b * (A - 1);
#pragma unroll
        for (int q = 1; q < A; ++q) a[q] = cmulc(a[q], w[q - 1]);
        dft_reg<A, 1>(a);
        V *gp = g2 + (uint64_t)ky * P + kz0 + c + (uint64_t)b * rs;
        const uint64_t step = (uint64_t)B * rs;
#pragma unroll
        for (int r = 0; r < A; ++r) {
            if (b + B * r <= Rx) *gp = a[r];
            gp += step;
        }
    }
}

// ---- y in two register stages (f_y = A B, as the x pass above): the
// forward pass writes its spectrum in natural frequency order, which the
// inverse reads back; the x pass in between works per ky line, in any order.
//   forward, per (x, tile): stage A from G1 rows y = b + B r (y >= ny zero)
//     into LDS; stage B: block q's B-point DFT, X[q + A k] to row q + A k.
//   inverse, per (ix, tile): the B-point inverse DFT of rows q + A k, k < B,
//     into block q of LDS; inverse stage A; kept rows |dy| <= Ry to G3.
template <class T, int A, int B>
__global__ __launch_bounds__(512) void fft_y_fwd2(int ny, int fy, int P, int H, int tiles_kz, int pad,
                                                  const C2<T> *__restrict__ twA, C2<T> *__restrict__ g1) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    constexpr int TILE = x2_tile<T>();
    V *d = reinterpret_cast<V *>(sm_raw);
    const int x = blockIdx.x / tiles_kz, kz0 = (blockIdx.x % tiles_kz) * TILE;
    const int nk = min(TILE, H - kz0);
    V *base = g1 + (uint64_t)x * fy * P + kz0;
    const int t = threadIdx.x, c = t % TILE, grp = t / TILE;
    if (grp < B && c < nk) {
        const int b = grp;
        V a[A];
        const V *gp = base + c + (uint64_t)b * P;
#pragma unroll
        for (int r = 0; r < A; ++r) {
            a[r] = b + B * r < ny ? *gp : mk<V>(0, 0);
            gp += (uint64_t)B * P;
        }
        dft_reg<A, -1>(a);
        const V *w = twA + b * (A - 1);
#pragma unroll
        for (int q = 1; q < A; ++q) a[q] = cmul(a[q], w[q - 1]);
        V *dp = d + b * TILE + c;
#pragma unroll
        for (int q = 0; q < A; ++q) dp[q * (B * TILE + pad)] = a[q];
    }
    __syncthreads();
    if (grp < A && c < nk) {
        const int q = grp;
        V a[B];
        const V *dp = d + q * (B * TILE + pad) + c;
#pragma unroll
        for (int k = 0; k < B; ++k) a[k] = dp[k * TILE];
        dft_reg<B, -1>(a);
        V *gp = base + c + (uint64_t)q * P;
#pragma unroll
        for (int k = 0; k < B; ++k) gp[(uint64_t)(A * k) * P] = a[k];
    }
}

template <class T, int A, int B>
__global__ __launch_bounds__(512) void fft_y_inv2(int fy, int P, int H, int tiles_kz, int Ry, int pad,
                                                  const C2<T> *__restrict__ twA, const C2<T> *__restrict__ g2,
                                                  C2<T> *__restrict__ g3) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    constexpr int TILE = x2_tile<T>(), L = A * B;
    V *d = reinterpret_cast<V *>(sm_raw);
    const int ix = blockIdx.x / tiles_kz, kz0 = (blockIdx.x % tiles_kz) * TILE;
    const int nk = min(TILE, H - kz0);
    const int t = threadIdx.x, c = t % TILE, grp = t / TILE;
    if (grp < A && c < nk) {
        const int q = grp;
        V a[B];
        const V *gp = g2 + (uint64_t)ix * fy * P + kz0 + c + (uint64_t)q * P;
#pragma unroll
        for (int k = 0; k < B; ++k) a[k] = gp[(uint64_t)(A * k) * P];
        dft_reg<B, 1>(a);
        V *dp = d + q * (B * TILE + pad) + c;
#pragma unroll
        for (int k = 0; k < B; ++k) dp[k * TILE] = a[k];
    }
    __syncthreads();
    if (grp < B && c < nk) {
        const int b = grp;
        V a[A];
        const V *dp = d + b * TILE + c;
#pragma unroll
        for (int q = 0; q < A; ++q) a[q] = dp[q * (B * TILE + pad)];
        const V *w = twA + b * (A - 1);
#pragma unroll
        for (int q = 1; q < A; ++q) a[q] = cmulc(a[q], w[q - 1]);
        dft_reg<A, 1>(a);
        const int nyo = 2 * Ry + 1;
        V *gp = g3 + (uint64_t)ix * nyo * P + kz0 + c;
#pragma unroll
        for (int r = 0; r < A; ++r) {
            const int y = b + B * r;
            const int o = y <= Ry ? y : (y >= L - Ry ? y - L + nyo : -1);
            if (o >= 0) gp[(uint64_t)(uint32_t)o * P] = a[r];
        }
    }
}

// ---- z forward in two register stages (f_z = A B): lines (two packed rows
// per complex line, as fft_z_fwd) of A blocks of B + 1 slots in LDS.
//   stage A: thread (line, b): the radix-A butterfly of z = b + B r straight
//     from the staged bits, twiddles, into block q at slot b;
//   stage B: thread (line, q): block q's B-point DFT in place, X[q + A k] at
//     slot k of block q;
//   split: thread (line, k), k < H: the two half spectra from X[k] and
//     X[L - k], as fft_z_fwd.
// The odd block stride puts the lanes of stage B and the split on distinct
// 16-byte bank groups.
template <class T>
size_t z2_lds_bytes(int A, int B, int lines, int row_words) {
    return (size_t)lines * A * (B + 1) * sizeof(C2<T>) + (size_t)lines * 2 * row_words * 8;
}
template <class T, int A, int B>
__global__ __launch_bounds__(256) void fft_z_fwd2(const uint64_t *__restrict__ words, int nx, int ny, int nz,
                                                  int row_words, int fy, int H, int P, int lines_per_block,
                                                  const C2<T> *__restrict__ twA, C2<T> *__restrict__ g1,
                                                  FastDiv dyp, FastDiv dh) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    constexpr int L = A * B, BS = B + 1, LS = A * BS;
    V *d = reinterpret_cast<V *>(sm_raw);
    const int npairs = nx * ((ny + 1) / 2);
    const int p0 = blockIdx.x * lines_per_block;
    const int np = min(lines_per_block, npairs - p0);
    uint64_t *sw = reinterpret_cast<uint64_t *>(d + (size_t)lines_per_block * LS);
    for (int t = threadIdx.x; t < np * 2 * row_words; t += blockDim.x) {
        const int pl = t / (2 * row_words), rem = t - pl * 2 * row_words;
        const int i = rem / row_words, wd = rem - i * row_words;
        int x, yp;
        divmod(p0 + pl, dyp, x, yp);
        const int y = 2 * yp + i;
        sw[t] = y < ny ? words[((uint64_t)x * ny + y) * row_words + wd] : 0ull;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < np * B; t += blockDim.x) {
        const int pl = t / B, b = t - pl * B;
        const uint64_t *ra = sw + 2 * pl * row_words;
        using R = decltype(V::x);
        // the two rows shifted right by b once, so that the bit of z = b + B r
        // sits at the compile-time position B r of the shifted words
        constexpr int NW = (A * B + 63) / 64;
        uint64_t sa[NW], sb[NW];
        {
            uint64_t wa[NW + 1], wb[NW + 1];
#pragma unroll
            for (int i = 0; i <= NW; ++i) {
                wa[i] = i < row_words ? ra[i] : 0ull;
                wb[i] = i < row_words ? ra[row_words + i] : 0ull;
            }
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                sa[i] = b ? (wa[i] >> b) | (wa[i + 1] << (64 - b)) : wa[i];
                sb[i] = b ? (wb[i] >> b) | (wb[i + 1] << (64 - b)) : wb[i];
            }
        }
        V a[A];
#pragma unroll
        for (int r = 0; r < A; ++r) {
            const int pos = B * r;
            const bool in = b + pos < nz;
            const R re = in ? (R)((sa[pos >> 6] >> (pos & 63)) & 1ull) : (R)0;
            const R im = in ? (R)((sb[pos >> 6] >> (pos & 63)) & 1ull) : (R)0;
            a[r] = mk<V>(re, im);
        }
        dft_reg<A, -1>(a);
        const V *w = twA + b * (A - 1);
#pragma unroll
        for (int q = 1; q < A; ++q) a[q] = cmul(a[q], w[q - 1]);
        V *dp = d + pl * LS + b;
#pragma unroll
        for (int q = 0; q < A; ++q) dp[q * BS] = a[q];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < np * A; t += blockDim.x) {
        const int pl = t / A, q = t - pl * A;
        V *dp = d + pl * LS + q * BS;
        V a[B];
#pragma unroll
        for (int k = 0; k < B; ++k) a[k] = dp[k];
        dft_reg<B, -1>(a);
#pragma unroll
        for (int k = 0; k < B; ++k) dp[k] = a[k];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < np * H; t += blockDim.x) {
        int pl, k, x, yp;
        divmod(t, dh, pl, k);
        divmod(p0 + pl, dyp, x, yp);
        const int y = 2 * yp;
        const int kn = k == 0 ? 0 : L - k;
        const V zk = d[pl * LS + (k % A) * BS + k / A];
        const V zn = d[pl * LS + (kn % A) * BS + kn / A];
        // A = (Z[k] + conj Z[-k]) / 2,  B = (Z[k] - conj Z[-k]) / (2i)
        using R = decltype(V::x);
        const R half = (R)0.5;
        const V va = mk<V>(half * (zk.x + zn.x), half * (zk.y - zn.y));
        const V vb = mk<V>(half * (zk.y + zn.y), -half * (zk.x - zn.x));
        g1[((uint64_t)x * fy + y) * P + k] = va;
        if (y + 1 < ny) g1[((uint64_t)x * fy + y + 1) * P + k] = vb;
    }
}

// ---- z inverse in two register stages (f_z = A B): two Hermitian half
// spectra per complex inverse, as fft_z_inv, read in natural order.
//   stage B: thread (pair, q): the values k = q + A j, j < B (k >= H from the
//     conjugate of L - k), the B-point inverse DFT into block q of LDS;
//   stage A: thread (pair, b): conjugate twiddles, inverse DFT over q, the
//     kept z = b + B r (|z| <= Rz) scaled to the cube.
template <class T, int A, int B>
__global__ __launch_bounds__(256) void fft_z_inv2(int nlines, int H, int P, int pairs_per_block, int Rz, T scale,
                                                  const C2<T> *__restrict__ twA, const C2<T> *__restrict__ g3,
                                                  T *__restrict__ cube) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    constexpr int L = A * B, BS = B + 1, LS = A * BS;
    V *d = reinterpret_cast<V *>(sm_raw);
    const int p0 = blockIdx.x * pairs_per_block;
    const int npairs = (nlines + 1) / 2;
    const int np = min(pairs_per_block, npairs - p0);
    for (int t = threadIdx.x; t < np * A; t += blockDim.x) {
        const int pl = t / A, q = t - pl * A;
        const int la = 2 * (p0 + pl), lb = la + 1;
        const V *ga = g3 + (uint64_t)la * P;
        const V *gb = g3 + (uint64_t)lb * P;
        V a[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int k = q + A * j;
            const bool upper = k >= H;
            const int kk = upper ? L - k : k;
            V va = ga[kk];
            V vb = lb < nlines ? gb[kk] : mk<V>(0, 0);
            if (upper) {
                va.y = -va.y;
                vb.y = -vb.y;
            }
            a[j] = mk<V>(va.x - vb.y, va.y + vb.x);  // A + iB
        }
        dft_reg<B, 1>(a);
        V *dp = d + pl * LS + q * BS;
#pragma unroll
        for (int j = 0; j < B; ++j) dp[j] = a[j];
    }
    __syncthreads();
    const int nzo = 2 * Rz + 1;
    for (int t = threadIdx.x; t < np * B; t += blockDim.x) {
        const int pl = t / B, b = t - pl * B;
        const int la = 2 * (p0 + pl), lb = la + 1;
        const V *dp = d + pl * LS + b;
        V a[A];
#pragma unroll
        for (int q = 0; q < A; ++q) a[q] = dp[q * BS];
        const V *w = twA + b * (A - 1);
#pragma unroll
        for (int q = 1; q < A; ++q) a[q] = cmulc(a[q], w[q - 1]);
        dft_reg<A, 1>(a);
#pragma unroll
        for (int r = 0; r < A; ++r) {
            const int z = b + B * r;
            const int o = z <= Rz ? z : (z >= L - Rz ? z - L + nzo : -1);
            if (o >= 0) {
                cube[(uint64_t)la * nzo + o] = a[r].x * scale;
                if (lb < nlines) cube[(uint64_t)lb * nzo + o] = a[r].y * scale;
            }
        }
    }
}

// ---- y inverse: per (ix, kz tile), keep the needed y
template <class T>
__global__ __launch_bounds__(kFftMaxThreads) void fft_y_inv(int fy, int H, int P, int tiles_kz, int tile, Fac fc,
                                                         const C2<T> *__restrict__ tw, int Ry,
                                                         const C2<T> *__restrict__ g2, C2<T> *__restrict__ g3, int sh) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    V *sm = reinterpret_cast<V *>(sm_raw);
    const int L = fy, S = lstride(L, sh), lgt = log2_if_pow2(tile);
    V *d = sm;
    const int ix = blockIdx.x / tiles_kz, kz0 = (blockIdx.x % tiles_kz) * tile;
    const int nk = min(tile, H - kz0);
    const V *base = g2 + (uint64_t)ix * fy * P + kz0;
    if (fc.n >= 2) {
        S2H_RADIX_SWITCH(fc.f[fc.n - 1], (load_first_inv<RR>(d, L, S, sh, tile, lgt, nk, base, (uint64_t)P)));
        for (int s = fc.n - 2; s >= 1; --s) fft_stage_any<true>(d, S, sh, nk, fc, s, tw);
        S2H_RADIX_SWITCH(fc.f[0], (store_last_inv<RR>(d, L, S, sh, tile, lgt, nk,
                                                       g3 + (uint64_t)ix * (2 * Ry + 1) * P + kz0, (uint64_t)P, Ry,
                                                       false, tw + fc.tw[0])));
        return;
    }
    for (int t = threadIdx.x; t < L * tile; t += blockDim.x) {
        int y, c;
        divmod(t, tile, lgt, y, c);
        d[c * S + pidx(y, sh)] = c < nk ? base[(uint64_t)y * P + c] : mk<V>(0, 0);
    }
    __syncthreads();
    lds_fft<true>(d, S, sh, nk, fc, tw);
    const int nyo = 2 * Ry + 1;
    for (int t = threadIdx.x; t < nyo * tile; t += blockDim.x) {
        int iy, c;
        divmod(t, tile, lgt, iy, c);
        if (c >= nk) continue;
        const int y = iy <= Ry ? iy : L - nyo + iy;
        g3[((uint64_t)ix * nyo + iy) * P + kz0 + c] = d[c * S + pidx(y, sh)];
    }
}

// ---- z inverse: two Hermitian half spectra per complex inverse; keep needed z
template <class T>
__global__ __launch_bounds__(kFftThreads) void fft_z_inv(int nlines, int H, int P, Fac fc, int L,
                                                         const C2<T> *__restrict__ tw, const int *__restrict__ rev,
                                                         int pairs_per_block, int Rz, T scale,
                                                         const C2<T> *__restrict__ g3, T *__restrict__ cube, int sh,
                                                         FastDiv dzo) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    V *sm = reinterpret_cast<V *>(sm_raw);
    V *d = sm;
    const int S = lstride(L, sh);
    const int p0 = blockIdx.x * pairs_per_block;
    const int npairs = (nlines + 1) / 2;
    const int np = min(pairs_per_block, npairs - p0);
    int *srev = reinterpret_cast<int *>(sm + (size_t)pairs_per_block * S);  // natural frequency per position
    for (int t = threadIdx.x; t < L; t += blockDim.x) srev[t] = rev[t];
    __syncthreads();
    for (int t = threadIdx.x; t < np * L; t += blockDim.x) {
        int pl, pos;
        divmod(t, fc.len, pl, pos);
        const int la = 2 * (p0 + pl), lb = la + 1;
        const int k = srev[pos];
        const bool upper = k >= H;
        const int kk = upper ? L - k : k;
        V A = g3[(uint64_t)la * P + kk];
        V B = lb < nlines ? g3[(uint64_t)lb * P + kk] : mk<V>(0, 0);
        if (upper) {
            A.y = -A.y;
            B.y = -B.y;
        }
        d[pl * S + pidx(pos, sh)] = mk<V>(A.x - B.y, A.y + B.x);  // A + iB
    }
    __syncthreads();
    lds_fft<true>(d, S, sh, np, fc, tw);
    const int nzo = 2 * Rz + 1;
    for (int t = threadIdx.x; t < np * nzo; t += blockDim.x) {
        int pl, iz;
        divmod(t, dzo, pl, iz);
        const int z = iz <= Rz ? iz : L - nzo + iz;
        const int la = 2 * (p0 + pl), lb = la + 1;
        const V v = d[pl * S + pidx(z, sh)];
        cube[(uint64_t)la * nzo + iz] = v.x * scale;
        if (lb < nlines) cube[(uint64_t)lb * nzo + iz] = v.y * scale;
    }
}

template <class T>
__global__ void fft_gather(const Offset *__restrict__ offs, uint64_t n, const T *__restrict__ cube, int Rx,
                           int Ry, int Rz, T *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Offset o = offs[i];
    out[i] = cube[half_cube_index(o, Ry, Rz)];
}

// ---------------------------------------------------------------- delta counts by region correlation
//
// A candidate changes the seed occupancy o to n = o + delta inside its box B
// (delta = A - D in {-1, 0, 1}).  Its pair-count change at offset d is
//   dC(d) = sum_u [o(u) delta(u+d) + delta(u) o(u+d) + delta(u) delta(u+d)],
// whose transform over a frame is 2 Re(conj(O) X) + |X|^2 (O, X: spectra of o
// and delta).  The frame of candidate c starts at B's corner minus R (R: the
// shell extent per axis) and has lengths F >= |B| + 2R: delta lives in
// [R, R + |B|) and every |d| <= R term of the circular correlation is the
// linear one (a wrapped u + d lands in [0, R) or [R + |B|, F), where delta is
// zero), whatever o holds elsewhere in the frame.  The frame reads o from
// the seed volume (zero outside it) and delta from the candidate's (A, D) rows.

// z forward of frame rows: line (c, i, j) is the complex row o + i delta of
// candidate c at frame x = i, y = j; its transform splits into the half
// spectra of o (field 0) and delta (field 1), written to g1 planes
// (2c + field) F_x + i.  The workgroup's seed rows (zero outside the volume)
// and (A, D) rows (zero outside the box) are staged in LDS first.
struct FftdLine {
    int32_t z0;       // volume z of the frame's first z
    int32_t drow;     // 1: the line lies in the box's (x, y) range
    uint64_t row;     // g1 row of the o field
};

template <int R, class V>
__device__ __forceinline__ void fftd_first_stage(V *d, int S, int sh, int np, int nz, int row_words, int nwd,
                                                 const uint64_t *so, const uint2 *sd, const FftdLine *sl,
                                                 const Fac &fc, const V *__restrict__ tw) {
    using T = decltype(V::x);
    const FastDiv dmp = fc.mp[0];
    const int mp = (int)dmp.n;
    for (int t = threadIdx.x; t < mp * np; t += blockDim.x) {
        int pl, b;
        divmod(t, dmp, pl, b);
        const uint64_t *ro = so + (size_t)pl * row_words;
        const uint2 *rd = sd + (size_t)pl * nwd;
        const int z0 = sl[pl].z0;
        V a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int z = z0 + b + mp * r;
            T re = 0, im = 0;
            if (z >= 0 && z < nz) {
                re = (T)((ro[z >> 6] >> (z & 63)) & 1ull);
                const uint2 w = rd[z >> 5];
                im = (T)(int)(((w.x >> (z & 31)) & 1u)) - (T)(int)(((w.y >> (z & 31)) & 1u));
            }
            a[r] = mk<V>(re, im);
        }
        small_dft<-1>(a, R);
        if (mp > 1) {
            const V *w = tw + fc.tw[0] + b * (R - 1);
#pragma unroll
            for (int q = 1; q < R; ++q) a[q] = cmul(a[q], w[q - 1]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) d[pl * S + pidx(b + mp * r, sh)] = a[r];
    }
    __syncthreads();
}

// LDS of a frame z forward workgroup: its lines, the staged seed and (A, D)
// rows, the line records and the DIF output positions.
template <class T>
size_t fftd_z_lds_bytes(int L, int lines, int sh, int row_words, int nwd) {
    return (size_t)lines * lstride(L, sh) * sizeof(C2<T>) + (size_t)lines * ((size_t)row_words * 8 + (size_t)nwd * 8) +
           (size_t)lines * sizeof(FftdLine) + (size_t)L * 4;
}

template <class T>
__global__ __launch_bounds__(kFftThreads) void fftd_z_fwd(const FftdCand *__restrict__ cands, int nlines,
                                                          const uint64_t *__restrict__ words, int nx, int ny, int nz,
                                                          int row_words, const uint2 *__restrict__ ad, int nwd, int Rx,
                                                          int Ry, int Rz, int Fx, int Fy, int H, int P, Fac fc, int L,
                                                          const C2<T> *__restrict__ tw, const int *__restrict__ pos,
                                                          int lines_per_block, C2<T> *__restrict__ g1, int sh,
                                                          FastDiv dplane, FastDiv dFy, FastDiv dh) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    V *d = reinterpret_cast<V *>(sm_raw);
    const int S = lstride(L, sh);
    const int p0 = blockIdx.x * lines_per_block;
    const int np = min(lines_per_block, nlines - p0);
    uint64_t *so = reinterpret_cast<uint64_t *>(d + (size_t)lines_per_block * S);
    uint2 *sdw = reinterpret_cast<uint2 *>(so + (size_t)lines_per_block * row_words);
    FftdLine *sl = reinterpret_cast<FftdLine *>(sdw + (size_t)lines_per_block * nwd);
    int *spos = reinterpret_cast<int *>(sl + lines_per_block);
    for (int pl = threadIdx.x; pl < np; pl += blockDim.x) {
        int c, rem, i, j;
        divmod(p0 + pl, dplane, c, rem);
        divmod(rem, dFy, i, j);
        const FftdCand cd = cands[c];
        const int x = cd.b0[0] - Rx + i, y = cd.b0[1] - Ry + j;
        FftdLine ln;
        ln.z0 = cd.b0[2] - Rz;
        ln.drow = (x >= cd.b0[0] && x < cd.b1[0] && y >= cd.b0[1] && y < cd.b1[1]) ? 1 : 0;
        ln.row = ((uint64_t)(2 * c) * Fx + i) * Fy + j;
        sl[pl] = ln;
    }
    for (int t = threadIdx.x; t < np * row_words; t += blockDim.x) {
        const int pl = t / row_words, wd = t - pl * row_words;
        int c, rem, i, j;
        divmod(p0 + pl, dplane, c, rem);
        divmod(rem, dFy, i, j);
        const FftdCand cd = cands[c];
        const int x = cd.b0[0] - Rx + i, y = cd.b0[1] - Ry + j;
        const bool orow = x >= 0 && x < nx && y >= 0 && y < ny;
        so[t] = orow ? words[((uint64_t)x * ny + y) * row_words + wd] : 0ull;
    }
    for (int t = threadIdx.x; t < np * nwd; t += blockDim.x) {
        const int pl = t / nwd, wd = t - pl * nwd;
        int c, rem, i, j;
        divmod(p0 + pl, dplane, c, rem);
        divmod(rem, dFy, i, j);
        const FftdCand cd = cands[c];
        const int x = cd.b0[0] - Rx + i, y = cd.b0[1] - Ry + j;
        const bool drow = x >= cd.b0[0] && x < cd.b1[0] && y >= cd.b0[1] && y < cd.b1[1];
        sdw[t] = drow ? ad[cd.ad_off + ((uint64_t)(x - cd.b0[0]) * (cd.b1[1] - cd.b0[1]) + (y - cd.b0[1])) * nwd + wd]
                      : make_uint2(0u, 0u);
    }
    for (int t = threadIdx.x; t < L; t += blockDim.x) spos[t] = pos[t];
    __syncthreads();
    S2H_RADIX_SWITCH(fc.f[0], (fftd_first_stage<RR>(d, S, sh, np, nz, row_words, nwd, so, sdw, sl, fc, tw)));
    for (int s = 1; s < fc.n; ++s) fft_stage_any<false>(d, S, sh, np, fc, s, tw);
    for (int t = threadIdx.x; t < np * H; t += blockDim.x) {
        int pl, k;
        divmod(t, dh, pl, k);
        const V zk = d[pl * S + pidx(spos[k], sh)];
        const V zn = d[pl * S + pidx(spos[k == 0 ? 0 : L - k], sh)];
        const T half = (T)0.5;
        const V A = mk<V>(half * (zk.x + zn.x), half * (zk.y - zn.y));   // spectrum of o
        const uint64_t row = sl[pl].row;
        g1[row * P + k] = A;
        // delta rows outside the box are zero: the y and x passes never read them
        if (sl[pl].drow) {
            const V B = mk<V>(half * (zk.y + zn.y), -half * (zk.x - zn.x));  // spectrum of delta
            g1[(row + (uint64_t)Fx * Fy) * P + k] = B;
        }
    }
}

// Last forward stage of both fields' lines (line l: o, line l + lines: delta),
// 2 Re(conj(O) X) + |X|^2, and the first inverse stage, in registers; the
// result replaces the o line.
template <int R, class V>
__device__ __forceinline__ void cross_stage(V *d, int S, int sh, int lines, const FastDiv dnb) {
    const int nb = (int)dnb.n;
    for (int t = threadIdx.x; t < nb * lines; t += blockDim.x) {
        int line, b;
        divmod(t, dnb, line, b);
        V *bo = d + line * S, *bx = d + (line + lines) * S;
        V a[R], x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            a[r] = bo[pidx(b * R + r, sh)];
            x[r] = bx[pidx(b * R + r, sh)];
        }
        small_dft<-1>(a, R);
        small_dft<-1>(x, R);
#pragma unroll
        for (int r = 0; r < R; ++r)
            a[r] = mk<V>(2 * (a[r].x * x[r].x + a[r].y * x[r].y) + (x[r].x * x[r].x + x[r].y * x[r].y), 0);
        small_dft<1>(a, R);
#pragma unroll
        for (int r = 0; r < R; ++r) bo[pidx(b * R + r, sh)] = a[r];
    }
    __syncthreads();
}

// x pass of the frames: one block per (candidate, ky, kz tile); both fields'
// lines forward, the cross spectrum, the inverse keeping the 2R_x + 1 needed
// x (g2 planes c (2R_x + 1) + ix).
template <class T>
__global__ __launch_bounds__(kFftMaxThreads) void fftd_x_cross(const FftdCand *__restrict__ cands, int Fx, int Fy,
                                                            int H, int P, int tiles_kz, int tile, Fac fc,
                                                            const C2<T> *__restrict__ tw, int Rx,
                                                            const C2<T> *__restrict__ g1, C2<T> *__restrict__ g2,
                                                            int sh) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    V *d = reinterpret_cast<V *>(sm_raw);
    const int L = Fx, S = lstride(L, sh), lgt = log2_if_pow2(tile);
    const int per = Fy * tiles_kz;
    const int c = blockIdx.x / per, rem = blockIdx.x % per;
    const int ky = rem / tiles_kz, kz0 = (rem % tiles_kz) * tile;
    const int nk = min(tile, H - kz0);
    const uint64_t rs = (uint64_t)Fy * P, field = (uint64_t)Fx * Fy * P;
    const V *go = g1 + (uint64_t)(2 * c) * field + (uint64_t)ky * P + kz0;
    const int last = fc.n - 1;
    // delta is zero outside the box rows R_x .. R_x + |B_x| - 1 of the frame
    // (those planes are not transformed along y either)
    const int dlo = Rx, dhi = Rx + (cands[c].b1[0] - cands[c].b0[0]);
    S2H_RADIX_SWITCH(fc.f[0], (load_first_fwd<RR>(d, L, S, sh, tile, lgt, nk, go, rs, 0, L, tw + fc.tw[0])));
    S2H_RADIX_SWITCH(fc.f[0], (load_first_fwd<RR>(d + tile * S, L, S, sh, tile, lgt, nk, go + field, rs, dlo, dhi,
                                                   tw + fc.tw[0])));
    for (int s = 1; s < last; ++s) fft_stage_any<false>(d, S, sh, 2 * tile, fc, s, tw);
    switch (fc.f[last]) {
        case 8: cross_stage<8>(d, S, sh, tile, fc.nb[last]); break;
        case 4: cross_stage<4>(d, S, sh, tile, fc.nb[last]); break;
        case 2: cross_stage<2>(d, S, sh, tile, fc.nb[last]); break;
        case 3: cross_stage<3>(d, S, sh, tile, fc.nb[last]); break;
        default: cross_stage<5>(d, S, sh, tile, fc.nb[last]); break;
    }
    for (int s = last - 1; s >= 1; --s) fft_stage_any<true>(d, S, sh, tile, fc, s, tw);
    const int nxo = Rx + 1;  // dx >= 0 only: the delta correlation is even
    S2H_RADIX_SWITCH(fc.f[0], (store_last_inv<RR>(d, L, S, sh, tile, lgt, nk,
                                                   g2 + (uint64_t)c * nxo * rs + (uint64_t)ky * P + kz0, rs, Rx,
                                                   true, tw + fc.tw[0])));
}

// The frames' x pass in two register stages (F_x = A B, as fft_x_power2): stage A
// of both fields (delta only on its box rows), stage B's blocks of both fields
// forward, the cross spectrum 2 Re(conj(O) X) + |X|^2 and its inverse in
// registers, then the inverse stage A of the result keeping dx = 0 .. R_x.
// LDS holds the two fields one after the other, each laid out as in
// fft_x_power2.
template <class T>
size_t xc2_lds_bytes(int A, int B, int pad) { return 2 * x2_lds_bytes<T>(A, B, pad); }
template <class T, int A, int B>
__global__ __launch_bounds__(512) void fftd_x_cross2(const FftdCand *__restrict__ cands, int Fy, int H, int P,
                                                     int tiles_kz, int Rx, int pad, const C2<T> *__restrict__ twA,
                                                     const C2<T> *__restrict__ g1, C2<T> *__restrict__ g2) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    using V = C2<T>;
    constexpr int TILE = x2_tile<T>(), L = A * B;
    V *d = reinterpret_cast<V *>(sm_raw);
    const int fsz = L * TILE + (A - 1) * pad;  // slots of one field
    const int per = Fy * tiles_kz;
    const int c = blockIdx.x / per, rem = blockIdx.x % per;
    const int ky = rem / tiles_kz, kz0 = (rem % tiles_kz) * TILE;
    const int nk = min(TILE, H - kz0);
    const uint64_t rs = (uint64_t)Fy * P, field = (uint64_t)L * Fy * P;
    const V *go = g1 + (uint64_t)(2 * c) * field + (uint64_t)ky * P + kz0;
    // delta is zero outside the box rows R_x .. R_x + |B_x| - 1 of the frame
    const int dlo = Rx, dhi = Rx + (cands[c].b1[0] - cands[c].b0[0]);
    const int t = threadIdx.x, col = t % TILE, grp = t / TILE;
    if (grp < 2 * B && col < nk) {
        const int f = grp / B, b = grp - f * B;
        const int lo = f ? dlo : 0, hi = f ? dhi : L;
        V a[A];
        const V *gp = go + (f ? field : 0) + col + (uint64_t)b * rs;
#pragma unroll
        for (int r = 0; r < A; ++r) {
            const int x = b + B * r;
            a[r] = (x >= lo && x < hi) ? *gp : mk<V>(0, 0);
            gp += (uint64_t)B * rs;
        }
        dft_reg<A, -1>(a);
        const V *w = twA + b * (A - 1);
#pragma unroll
        for (int q = 1; q < A; ++q) a[q] = cmul(a[q], w[q - 1]);
        V *dp = d + f * fsz + b * TILE + col;
#pragma unroll
        for (int q = 0; q < A; ++q) dp[q * (B * TILE + pad)] = a[q];
    }
    __syncthreads();
    if (grp < A && col < nk) {
        V o[B], x[B];
        V *dp = d + grp * (B * TILE + pad) + col;
#pragma unroll
        for (int k = 0; k < B; ++k) {
            o[k] = dp[k * TILE];
            x[k] = dp[fsz + k * TILE];
        }
        dft_reg<B, -1>(o);
        dft_reg<B, -1>(x);
#pragma unroll
        for (int k = 0; k < B; ++k)
            o[k] = mk<V>(2 * (o[k].x * x[k].x + o[k].y * x[k].y) + (x[k].x * x[k].x + x[k].y * x[k].y), 0);
        dft_reg<B, 1>(o);
#pragma unroll
        for (int k = 0; k < B; ++k) dp[k * TILE] = o[k];
    }
    __syncthreads();
    if (grp < B && col < nk) {
        const int b = grp;
        V a[A];
        const V *dp = d + b * TILE + col;
#pragma unroll
        for (int q = 0; q < A; ++q) a[q] = dp[q * (B * TILE + pad)];
        const V *w = twA + b * (A - 1);
#pragma unroll
        for (int q = 1; q < A; ++q) a[q] = cmulc(a[q], w[q - 1]);
        dft_reg<A, 1>(a);
        V *gp = g2 + (uint64_t)c * (Rx + 1) * rs + (uint64_t)ky * P + kz0 + col + (uint64_t)b * rs;
#pragma unroll
        for (int r = 0; r < A; ++r) {
            if (b + B * r <= Rx) *gp = a[r];
            gp += (uint64_t)B * rs;
        }
    }
}

// Delta counts of the candidates: the value at each shell offset rounded to
// the nearest integer, stored at the offset's canonical index (d and -d share
// one; the zero offset has none).
__global__ void fftd_gather(const FftdCand *__restrict__ cands, uint32_t ncand, const Offset *__restrict__ offs,
                            const int32_t *__restrict__ sg, uint64_t noffs, const double *__restrict__ cube, int Rx,
                            int Ry, int Rz, int32_t *__restrict__ delta) {
    const int nyo = 2 * Ry + 1, nzo = 2 * Rz + 1;
    const uint64_t per = (uint64_t)(Rx + 1) * nyo * nzo, total = (uint64_t)ncand * noffs;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t c = t / noffs, i = t - c * noffs;
        const int32_t k = sg[i];
        if (k < 0) continue;
        const Offset o = offs[i];
        const double v = cube[c * per + half_cube_index(o, Ry, Rz)];
        delta[cands[c].delta_off + k] = (int32_t)rint(v);
    }
}

// ---------------------------------------------------------------- host plan

AxisPlan axis_plan(int L) {
    AxisPlan a;
    a.L = L;
    a.nf = 0;
    int n = L;
    while (n % 8 == 0 && a.nf < kMaxFactors - 1) a.fac[a.nf++] = 8, n /= 8;
    while (n % 4 == 0 && a.nf < kMaxFactors - 1) a.fac[a.nf++] = 4, n /= 4;
    while (n % 2 == 0 && a.nf < kMaxFactors - 1) a.fac[a.nf++] = 2, n /= 2;
    while (n % 3 == 0 && a.nf < kMaxFactors - 1) a.fac[a.nf++] = 3, n /= 3;
    while (n % 5 == 0 && a.nf < kMaxFactors - 1) a.fac[a.nf++] = 5, n /= 5;
    if (n != 1) a.nf = -1;
    return a;
}

Fac to_fac(const AxisPlan &a) {
    Fac f;
    f.n = a.nf;
    for (int i = 0; i < kMaxFactors; ++i) f.f[i] = i < a.nf ? a.fac[i] : 0;
    f.f[kMaxFactors - 1] = a.L;
    int m = a.L, off = 0;
    for (int s = 0; s < kMaxFactors; ++s) {
        f.tw[s] = off;
        if (s < a.nf) {
            const int mp = m / a.fac[s];
            f.nb[s] = fast_div((uint32_t)(a.L / a.fac[s]));
            f.mp[s] = fast_div((uint32_t)mp);
            off += mp * (a.fac[s] - 1);
            m = mp;
        }
    }
    f.len = fast_div((uint32_t)std::max(a.L, 1));
    return f;
}

// The twiddle rows of an axis (Fac::tw): stage s, row k < m_s / f_s, entry q
// = 1 .. f_s - 1 holds w^(k q L / m_s), taken from w[j] = exp(-2 pi i j / L).
template <class W>
std::vector<W> stage_twiddles(const AxisPlan &a, const std::vector<W> &w) {
    std::vector<W> out;
    int m = a.L;
    for (int s = 0; s < a.nf; ++s) {
        const int R = a.fac[s], mp = m / R, step = a.L / m;
        for (int k = 0; k < mp; ++k)
            for (int q = 1; q < R; ++q) out.push_back(w[(size_t)k * q * step]);
        m = mp;
    }
    if (out.empty()) out.push_back(w[0]);
    return out;
}

// src/geometry/s2.rs:519-551 smooth_fft_length
uint64_t smooth_length(uint64_t min_len) {
    uint64_t target = std::max<uint64_t>(min_len, 1), best = 0;
    for (uint64_t p5 = 1;; p5 *= 5) {
        for (uint64_t p35 = p5;; p35 *= 3) {
            uint64_t v = p35;
            while (v < target) v *= 2;
            if (best == 0 || v < best) best = v;
            if (p35 >= target) break;
        }
        if (p5 >= target) break;
    }
    return best;
}

// Padded transform length of an axis of n voxels whose needed offsets reach
// |d| <= R.  The circular correlation of length f is the linear one folded by
// f: c_f(d) = sum_k c(d + k f), with c(m) = 0 for |m| >= n.  For |d| <= R every
// folded term has |d + k f| >= f - R, so f >= n + R leaves the needed offsets
// exact (the reference pads to 2n - 1, src/geometry/s2.rs:554-562, which keeps
// every offset).  Among the 2,3,5-smooth lengths from n + R up to the
// reference's, the one with the least f x (stages + 2) -- element passes of
// the LDS stages plus the global load and store -- is taken: 320 for
// n = 256, R = 64 (512 in the reference's padding), 960 = 8 8 3 5 for
// n = 768, R = 128 (1536).
uint64_t fft_length(uint64_t n, uint64_t R) {
    const uint64_t full = smooth_length(2 * n - 1);
    const uint64_t lo = n + std::min<uint64_t>(R, n - 1);
    uint64_t best = full;
    double best_cost = INFINITY;
    for (uint64_t p2 = 1; p2 <= full; p2 *= 2)
        for (uint64_t p23 = p2; p23 <= full; p23 *= 3)
            for (uint64_t v = p23; v <= full; v *= 5) {
                if (v < lo || v > 8192) continue;
                const AxisPlan ap = axis_plan((int)v);
                if (ap.nf < 0) continue;
                const double cost = (double)v * (ap.nf + 2);
                if (cost < best_cost || (cost == best_cost && v < best)) {
                    best_cost = cost;
                    best = v;
                }
            }
    return best;
}

std::tuple<uint64_t, uint64_t, uint64_t> fft_plan_key(const uint64_t dims[3], const int R[3], int real_bytes) {
    return std::make_tuple(dims[0] | ((uint64_t)R[0] << 32) | ((uint64_t)(real_bytes == 8) << 63),
                           dims[1] | ((uint64_t)R[1] << 32), dims[2] | ((uint64_t)R[2] << 32));
}

// The lengths the two-stage passes run: L = A B with their in-register DFTs
// (the instantiations listed in S2H_TWO_STAGE_LENGTHS).
bool x2_split(int L, int *A, int *B) {
    static constexpr int kSplits[][3] = {{96, 8, 12},   {120, 8, 15},  {128, 8, 16},  {160, 16, 10},
                                         {192, 16, 12}, {256, 16, 16}, {320, 16, 20}, {360, 15, 24},
                                         {384, 16, 24}, {400, 20, 20}, {512, 16, 32}};
    for (const auto &e : kSplits)
        if (e[0] == L) {
            *A = e[1];
            *B = e[2];
            return true;
        }
    *A = *B = 0;
    return false;
}

// The per-length tables of a plan whose f[] and ax[] are set: each axis'
// stage twiddle rows in the plan's precision, and the z axis' DIF output order.
s2h_status fft_plan_tables(s2h_ctx *ctx, FftPlan *p) {
    const int real_bytes = p->real_bytes;
    for (int a = 0; a < 3; ++a) {
        const int L = (int)p->f[a];
        std::vector<double2> tw(L);
        std::vector<float2> tw32(L);
        for (int k = 0; k < L; ++k) {
            const double ang = -2.0 * M_PI * (double)k / (double)L;
            tw[k] = make_double2(std::cos(ang), std::sin(ang));
            tw32[k] = make_float2((float)tw[k].x, (float)tw[k].y);
        }
        const std::vector<double2> rows = stage_twiddles(p->ax[a], tw);
        const std::vector<float2> rows32 = stage_twiddles(p->ax[a], tw32);
        S2H_TRY(real_bytes == 8 ? upload(ctx, p->tw[a], rows.data(), rows.size() * sizeof(double2))
                                : upload(ctx, p->tw[a], rows32.data(), rows32.size() * sizeof(float2)));
    }
    // stage-A twiddle rows of the two-stage passes: w^(b q), b < B, q = 1 .. A - 1
    auto two_stage = [&](int L, int *A, int *B, DevBuf &buf) -> s2h_status {
        if (!x2_split(L, A, B)) return S2H_OK;
        std::vector<double2> rows;
        std::vector<float2> rows32;
        for (int b = 0; b < *B; ++b)
            for (int q = 1; q < *A; ++q) {
                const double ang = -2.0 * M_PI * (double)((b * q) % L) / (double)L;
                rows.push_back(make_double2(std::cos(ang), std::sin(ang)));
                rows32.push_back(make_float2((float)rows.back().x, (float)rows.back().y));
            }
        return real_bytes == 8 ? upload(ctx, buf, rows.data(), rows.size() * sizeof(double2))
                               : upload(ctx, buf, rows32.data(), rows32.size() * sizeof(float2));
    };
    S2H_TRY(two_stage((int)p->f[0], &p->x2_a, &p->x2_b, p->tw_x2));
    S2H_TRY(two_stage((int)p->f[1], &p->y2_a, &p->y2_b, p->tw_y2));
    S2H_TRY(two_stage((int)p->f[2], &p->z2_a, &p->z2_b, p->tw_z2));
    const AxisPlan &az = p->ax[2];  // natural frequency held at each DIF output position (z axis)
    std::vector<int> rev(az.L);
    for (int pos = 0; pos < az.L; ++pos) {
        int k = 0, mult = 1, m = az.L, q = pos;
        for (int s = 0; s < az.nf; ++s) {
            const int mp = m / az.fac[s];
            k += (q / mp) * mult;
            q %= mp;
            mult *= az.fac[s];
            m = mp;
        }
        rev[pos] = k;
    }
    std::vector<int> posv(az.L);
    for (int pos = 0; pos < az.L; ++pos) posv[rev[pos]] = pos;
    S2H_TRY(upload(ctx, p->rev_z, rev.data(), rev.size() * sizeof(int)));
    return upload(ctx, p->pos_z, posv.data(), posv.size() * sizeof(int));
}

// Bytes of the plan's big buffers: g1 (n_x planes of f_y rows of the half
// spectrum), g2 (the 2R_x+1 kept x planes), g3 (the kept (x, y) rows) and the
// real cube of kept offsets; rows are spectrum_pitch complex values apart.
void plan_buffer_sizes(const uint64_t dims[3], const int R[3], int real_bytes, uint64_t sizes[4]) {
    const uint64_t fy = fft_length(dims[1], (uint64_t)R[1]);
    const uint64_t H = spectrum_pitch(fft_length(dims[2], (uint64_t)R[2]), real_bytes);
    const uint64_t nxo = (uint64_t)R[0] + 1, nyo = 2 * (uint64_t)R[1] + 1, nzo = 2 * (uint64_t)R[2] + 1;
    const uint64_t cb = 2 * (uint64_t)real_bytes;  // bytes per complex value
    const uint64_t raw[4] = {dims[0] * fy * H * cb, nxo * fy * H * cb, nxo * nyo * H * cb,
                             nxo * nyo * nzo * (uint64_t)real_bytes};
    for (int i = 0; i < 4; ++i) sizes[i] = (raw[i] + 255) / 256 * 256;
}

s2h_status fft_plan_get(s2h_ctx *ctx, const uint64_t dims[3], const int R[3], int real_bytes, FftPlan **out) {
    auto key = fft_plan_key(dims, R, real_bytes);
    auto it = ctx->fft_plans.find(key);
    if (it != ctx->fft_plans.end()) {
        *out = it->second;
        return S2H_OK;
    }
    auto *p = new FftPlan();
    p->real_bytes = real_bytes;
    for (int a = 0; a < 3; ++a) {
        p->dims[a] = dims[a];
        p->f[a] = fft_length(dims[a], (uint64_t)R[a]);
        p->R[a] = R[a];
        p->ax[a] = axis_plan((int)p->f[a]);
        if (p->ax[a].nf < 0 || p->f[a] > 8192) {
            delete p;
            return fail(S2H_ERR_UNSUPPORTED, "FFT length " + std::to_string(p->f[a]) + " unsupported");
        }
    }
    auto cleanup = [&](s2h_status st) {
        fft_plan_release(ctx, p);
        return st;
    };
    {
        const s2h_status st = fft_plan_tables(ctx, p);
        if (st != S2H_OK) return cleanup(st);
    }
    uint64_t sizes[4];
    plan_buffer_sizes(dims, R, real_bytes, sizes);
    DevBuf *big[4] = {&p->g1, &p->g2, &p->g3, &p->cube};
    if (ctx->fft_arena.ptr && sizes[0] + sizes[1] + sizes[2] + sizes[3] <= ctx->fft_arena.bytes) {
        // carve the spectra and cube from the arena; the plan they held before
        // is released first (its kernels are earlier on this stream)
        for (auto &kv : ctx->fft_plans) fft_plan_release(ctx, kv.second);
        ctx->fft_plans.clear();
        uint64_t off = 0;
        for (int i = 0; i < 4; ++i) {
            big[i]->ptr = (unsigned char *)ctx->fft_arena.ptr + off;
            big[i]->bytes = sizes[i];
            big[i]->borrowed = true;
            off += sizes[i];
        }
    } else {
        for (int i = 0; i < 4; ++i) {
            const s2h_status st = dev_alloc(ctx, *big[i], sizes[i]);
            if (st != S2H_OK) return cleanup(st);
        }
    }
    ctx->fft_plans[key] = p;
    *out = p;
    return S2H_OK;
}

template <class T>
size_t lds_bytes(int L, int lines, int sh) {
    return (size_t)lines * lstride(L, sh) * sizeof(C2<T>);
}

// Allow a kernel the whole 160 KB of LDS as dynamic shared memory, once per
// kernel and process: the limit is per function, so setting each launch's own
// size would let a context on another thread lower it under a launch in
// flight.  Occupancy follows each launch's actual size.
s2h_status set_lds(const void *fn, size_t bytes) {
    constexpr size_t kMax = 160 * 1024;
    if (bytes > kMax) return fail(S2H_ERR_UNSUPPORTED, "FFT tile exceeds LDS");
    static std::mutex mu;
    static std::set<const void *> done;
    std::lock_guard<std::mutex> lk(mu);
    if (done.count(fn)) return S2H_OK;
    S2H_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMax));
    done.insert(fn);
    return S2H_OK;
}

// Run every pass; leaves the correlation cube in p->cube.  With ev (six
// events), records ev[4] before the z forward pass, ev[0] before the y forward
// pass, ev[1..3] after it, the x pass and the y inverse pass, and ev[5] after
// the z inverse pass.
// Line pairs per workgroup of the z passes: one radix-8 butterfly per thread
// per stage (2048 points over 256 threads), capped by 48 KB of LDS.
template <class T>
int z_lines_per_block(int fz) {
    const int by_work = std::max(1, 2048 / std::max(fz, 1));
    const int by_lds = std::max(1, (int)(48 * 1024 / ((size_t)fz * sizeof(C2<T>))));
    return std::min(16, std::min(by_work, by_lds));
}

// Threads of a y or x pass workgroup: one butterfly of the first stage per
// thread ((L / f_0) x tile), at least 256, at most kFftMaxThreads.
static int pass_threads(int L, const Fac &fc, int tile) {
    const int mp = fc.n >= 1 ? L / fc.f[0] : L;
    const int t = (mp * tile + 63) / 64 * 64;
    return std::min(kFftMaxThreads, std::max(kFftThreads, t));
}

// The (A, B) instantiations of the two-stage kernels (x2_split's table).
#define S2H_TWO_STAGE_LENGTHS(M) \
    M(8, 12) M(8, 15) M(8, 16) M(16, 10) M(16, 12) M(16, 16) M(16, 20) M(15, 24) M(16, 24) M(20, 20) M(16, 32)

// The two-stage x pass of a plan with x2_a, x2_b set (f_x = A B).
template <class T>
s2h_status fft_x2_launch(FftPlan *p, int nx, int fy, int H, int P, int Rx, hipStream_t s) {
    using V = C2<T>;
    const int A = p->x2_a, B = p->x2_b, tile = x2_tile<T>();
    if (A * B != (int)p->f[0] || nx > A * B || Rx >= A * B) return fail(S2H_ERR_UNSUPPORTED, "two-stage x pass shape");
    const int tiles = (H + tile - 1) / tile;
    const int threads = (std::max(A, B) * tile + 63) / 64 * 64;
    const size_t lds = x2_lds_bytes<T>(A, B, p->x2_pad);
    const V *tw = (const V *)p->tw_x2.ptr;
    const V *g1 = (const V *)p->g1.ptr;
    V *g2 = (V *)p->g2.ptr;
#define S2H_X2(AA, BB)                                                                                          \
    if (A == AA && B == BB) {                                                                                   \
        S2H_TRY(set_lds((const void *)fft_x_power2<T, AA, BB>, lds));                                          \
        fft_x_power2<T, AA, BB><<<fy * tiles, threads, lds, s>>>(nx, fy, H, P, tiles, Rx, p->x2_pad, tw, g1, g2); \
        S2H_HIP(hipGetLastError());                                                                             \
        return S2H_OK;                                                                                          \
    }
    S2H_TWO_STAGE_LENGTHS(S2H_X2)
#undef S2H_X2
    return fail(S2H_ERR_UNSUPPORTED, "two-stage x pass length");
}

// The two-stage y passes (forward: inverse == false) of a plan with y2_a,
// y2_b set (f_y = A B).
template <class T>
s2h_status fft_y2_launch(FftPlan *p, bool inverse, int nplanes, int ny, int H, int P, int Ry, hipStream_t s) {
    using V = C2<T>;
    const int A = p->y2_a, B = p->y2_b, tile = x2_tile<T>(), fy = (int)p->f[1];
    if (A * B != fy || ny > fy || Ry >= fy) return fail(S2H_ERR_UNSUPPORTED, "two-stage y pass shape");
    const int tiles = (H + tile - 1) / tile;
    const int threads = (std::max(A, B) * tile + 63) / 64 * 64;
    const size_t lds = x2_lds_bytes<T>(A, B, p->y2_pad);
    const V *tw = (const V *)p->tw_y2.ptr;
    V *g1 = (V *)p->g1.ptr, *g2 = (V *)p->g2.ptr, *g3 = (V *)p->g3.ptr;
#define S2H_Y2(AA, BB)                                                                                            \
    if (A == AA && B == BB) {                                                                                     \
        if (!inverse) {                                                                                           \
            S2H_TRY(set_lds((const void *)fft_y_fwd2<T, AA, BB>, lds));                                          \
            fft_y_fwd2<T, AA, BB><<<nplanes * tiles, threads, lds, s>>>(ny, fy, P, H, tiles, p->y2_pad, tw, g1);  \
        } else {                                                                                                  \
            S2H_TRY(set_lds((const void *)fft_y_inv2<T, AA, BB>, lds));                                          \
            fft_y_inv2<T, AA, BB><<<nplanes * tiles, threads, lds, s>>>(fy, P, H, tiles, Ry, p->y2_pad, tw, g2,  \
                                                                         g3);                                    \
        }                                                                                                         \
        S2H_HIP(hipGetLastError());                                                                               \
        return S2H_OK;                                                                                            \
    }
    S2H_TWO_STAGE_LENGTHS(S2H_Y2)
#undef S2H_Y2
    return fail(S2H_ERR_UNSUPPORTED, "two-stage y pass length");
}

// Lines (row pairs) per workgroup of the two-stage z forward pass, and line
// pairs per workgroup of the two-stage z inverse, each the fastest of 1-12
// timed at 320^3 (z forward: 4 lines 53 us, 6 lines 58 us; z inverse: 3 pairs
// 11.0 us, 1 pair 16.6 us).
static constexpr int kZ2Lines = 4;
static constexpr int kZ2InvPairs = 3;

// The two-stage z inverse pass of a plan with z2_a, z2_b set.
template <class T>
s2h_status fft_z2inv_launch(FftPlan *p, int nlines, int H, int P, int Rz, T scale, hipStream_t s) {
    using V = C2<T>;
    const int A = p->z2_a, B = p->z2_b, fz = (int)p->f[2];
    if (A * B != fz || Rz >= fz) return fail(S2H_ERR_UNSUPPORTED, "two-stage z inverse shape");
    const int pairs = kZ2InvPairs, threads = (pairs * std::max(A, B) + 63) / 64 * 64;
    const size_t lds = (size_t)pairs * A * (B + 1) * sizeof(V);
    const int npairs = (nlines + 1) / 2;
    const unsigned grid = (unsigned)((npairs + pairs - 1) / pairs);
    const V *tw = (const V *)p->tw_z2.ptr;
#define S2H_Z2I(AA, BB)                                                                                        \
    if (A == AA && B == BB) {                                                                                  \
        S2H_TRY(set_lds((const void *)fft_z_inv2<T, AA, BB>, lds));                                           \
        fft_z_inv2<T, AA, BB><<<grid, threads, lds, s>>>(nlines, H, P, pairs, Rz, scale, tw,                   \
                                                          (const V *)p->g3.ptr, (T *)p->cube.ptr);             \
        S2H_HIP(hipGetLastError());                                                                            \
        return S2H_OK;                                                                                         \
    }
    S2H_TWO_STAGE_LENGTHS(S2H_Z2I)
#undef S2H_Z2I
    return fail(S2H_ERR_UNSUPPORTED, "two-stage z inverse length");
}

// The two-stage z forward pass of a plan with z2_a, z2_b set (f_z = A B).
template <class T>
s2h_status fft_z2_launch(FftPlan *p, s2h_vol *vol, int H, int P, hipStream_t s) {
    using V = C2<T>;
    const int A = p->z2_a, B = p->z2_b, fz = (int)p->f[2], fy = (int)p->f[1];
    const int nx = (int)vol->dims[0], ny = (int)vol->dims[1], nz = (int)vol->dims[2];
    if (A * B != fz || nz > fz) return fail(S2H_ERR_UNSUPPORTED, "two-stage z pass shape");
    const int lines = kZ2Lines, threads = (lines * std::max(A, B) + 63) / 64 * 64;
    const size_t lds = z2_lds_bytes<T>(A, B, lines, (int)vol->row_words);
    const int npairs = nx * ((ny + 1) / 2);
    const unsigned grid = (unsigned)((npairs + lines - 1) / lines);
    const V *tw = (const V *)p->tw_z2.ptr;
    V *g1 = (V *)p->g1.ptr;
    const FastDiv dyp = fast_div((uint32_t)((ny + 1) / 2)), dh = fast_div((uint32_t)H);
#define S2H_Z2(AA, BB)                                                                                        \
    if (A == AA && B == BB) {                                                                                 \
        S2H_TRY(set_lds((const void *)fft_z_fwd2<T, AA, BB>, lds));                                          \
        fft_z_fwd2<T, AA, BB><<<grid, threads, lds, s>>>(vol->d_words, nx, ny, nz, (int)vol->row_words, fy, H, \
                                                          P, lines, tw, g1, dyp, dh);                         \
        S2H_HIP(hipGetLastError());                                                                           \
        return S2H_OK;                                                                                        \
    }
    S2H_TWO_STAGE_LENGTHS(S2H_Z2)
#undef S2H_Z2
    return fail(S2H_ERR_UNSUPPORTED, "two-stage z pass length");
}

template <class T>
s2h_status fft_run(s2h_ctx *ctx, s2h_vol *vol, FftPlan *p, hipEvent_t *ev = nullptr) {
    using V = C2<T>;
    const int nx = (int)vol->dims[0], ny = (int)vol->dims[1], nz = (int)vol->dims[2];
    const int fx = (int)p->f[0], fy = (int)p->f[1], fz = (int)p->f[2];
    const int H = fz / 2 + 1, P = (int)spectrum_pitch((uint64_t)fz, (int)sizeof(T));
    const int Rx = p->R[0], Ry = p->R[1], Rz = p->R[2];
    const int nxo = Rx + 1, nyo = 2 * Ry + 1;  // x planes dx = 0 .. R_x: the correlation is even
    const Fac fcx = to_fac(p->ax[0]), fcy = to_fac(p->ax[1]), fcz = to_fac(p->ax[2]);
    hipStream_t s = ctx->stream;
    const int ty = p->tile_y, tx = p->tile_x;
    const int tiles_y = (H + ty - 1) / ty, tiles_x = (H + tx - 1) / tx;
    if (ev) S2H_HIP(hipEventRecord(ev[4], s));
    // z forward
    if (p->z2_on) {
        S2H_TRY(fft_z2_launch<T>(p, vol, H, P, s));
        ++ctx->launches;
    } else {
        const int ppb = z_lines_per_block<T>(fz);
        const size_t lds = z_fwd_lds_bytes<T>(fz, ppb, p->sh_z, (int)vol->row_words);
        S2H_TRY(set_lds((const void *)fft_z_fwd<T>, lds));
        const int npairs = nx * ((ny + 1) / 2);
        fft_z_fwd<T><<<(npairs + ppb - 1) / ppb, kFftThreads, lds, s>>>(
            vol->d_words, nx, ny, nz, (int)vol->row_words, fy, H, P, fcz, fz, (const V *)p->tw[2].ptr,
            (const int *)p->pos_z.ptr, ppb, (V *)p->g1.ptr, p->sh_z, fast_div((uint32_t)((ny + 1) / 2)),
            fast_div((uint32_t)H));
        S2H_HIP(hipGetLastError());
        ++ctx->launches;
    }
    if (ev) S2H_HIP(hipEventRecord(ev[0], s));
    if (p->y2_on) {
        S2H_TRY(fft_y2_launch<T>(p, false, nx, ny, H, P, Ry, s));
        ++ctx->launches;
    } else {
        const size_t lds = lds_bytes<T>(fy, ty, p->sh_y);
        S2H_TRY(set_lds((const void *)fft_y_fwd<T>, lds));
        fft_y_fwd<T><<<nx * tiles_y, pass_threads(fy, fcy, ty), lds, s>>>(ny, fy, H, P, tiles_y, ty, fcy,
                                                                            (const V *)p->tw[1].ptr,
                                                            (V *)p->g1.ptr, p->sh_y, nullptr, 0, 0, 0);
        S2H_HIP(hipGetLastError());
        ++ctx->launches;
    }
    if (ev) S2H_HIP(hipEventRecord(ev[1], s));
    if (p->x2_on) {
        S2H_TRY(fft_x2_launch<T>(p, nx, fy, H, P, Rx, s));
        ++ctx->launches;
    } else {
        const size_t lds = lds_bytes<T>(fx, tx, p->sh_x);
        S2H_TRY(set_lds((const void *)fft_x_power<T>, lds));
        fft_x_power<T><<<fy * tiles_x, pass_threads(fx, fcx, tx), lds, s>>>(nx, fy, H, P, tiles_x, tx, fcx,
                                                                             (const V *)p->tw[0].ptr,
                                                              Rx, (const V *)p->g1.ptr, (V *)p->g2.ptr, p->sh_x);
        S2H_HIP(hipGetLastError());
        ++ctx->launches;
    }
    if (ev) S2H_HIP(hipEventRecord(ev[2], s));
    if (p->y2_on) {
        S2H_TRY(fft_y2_launch<T>(p, true, nxo, ny, H, P, Ry, s));
        ++ctx->launches;
    } else {
        const size_t lds = lds_bytes<T>(fy, ty, p->sh_y);
        S2H_TRY(set_lds((const void *)fft_y_inv<T>, lds));
        fft_y_inv<T><<<nxo * tiles_y, pass_threads(fy, fcy, ty), lds, s>>>(fy, H, P, tiles_y, ty, fcy,
                                                                             (const V *)p->tw[1].ptr, Ry,
                                                             (const V *)p->g2.ptr, (V *)p->g3.ptr, p->sh_y);
        S2H_HIP(hipGetLastError());
        ++ctx->launches;
    }
    if (ev) S2H_HIP(hipEventRecord(ev[3], s));
    if (p->z2_on) {
        const T scale = (T)(1.0 / ((double)fx * (double)fy * (double)fz));
        S2H_TRY(fft_z2inv_launch<T>(p, nxo * nyo, H, P, Rz, scale, s));
        ++ctx->launches;
    } else {
        const int ppb = z_lines_per_block<T>(fz);
        const size_t lds = z_inv_lds_bytes<T>(fz, ppb, p->sh_z);
        S2H_TRY(set_lds((const void *)fft_z_inv<T>, lds));
        const int nlines = nxo * nyo;
        const int npairs = (nlines + 1) / 2;
        const T scale = (T)(1.0 / ((double)fx * (double)fy * (double)fz));
        fft_z_inv<T><<<(npairs + ppb - 1) / ppb, kFftThreads, lds, s>>>(
            nlines, H, P, fcz, fz, (const V *)p->tw[2].ptr, (const int *)p->rev_z.ptr, ppb, Rz, scale,
            (const V *)p->g3.ptr, (T *)p->cube.ptr, p->sh_z, fast_div((uint32_t)(2 * Rz + 1)));
        S2H_HIP(hipGetLastError());
        ++ctx->launches;
    }
    if (ev) S2H_HIP(hipEventRecord(ev[5], s));
    return S2H_OK;
}

// Workgroups of `lds` bytes that fit one CU's 160 KB of LDS.
static int blocks_per_cu(size_t lds) { return (int)((160 * 1024) / std::max<size_t>(lds, 1)); }

// The padding shift for a pass of `lines` lines of length L: padded unless
// the padding leaves room for fewer workgroups per CU.
template <class T>
int pad_rule(int L, int lines) {
    return blocks_per_cu(lds_bytes<T>(L, lines, kPad)) >= blocks_per_cu(lds_bytes<T>(L, lines, kNoPad)) ? kPad : kNoPad;
}

// kz columns per workgroup of the y and x passes: powers of two up to kTileMax
// whose LDS tile fits -- untimed, the widest for y and for x the narrowest
// whose row segment fills a whole 128-byte line (else the widest up to 4), or,
// on a context that autotunes, for each pass kind the fastest of those
// candidates timed on the device (HIP events around each pass of one run per
// candidate, after a warm-up) among the whole-line tiles.  A segment narrower than a line still fetches
// the whole line, and the neighbouring tile's workgroup seldom finds it in L2.
// The y passes favour wide tiles (longer contiguous kz runs per row), the x
// pass the narrowest whole-line tile (more workgroups per CU).  Every run
// rewrites all of the plan's buffers, so the trial runs leave nothing behind.
// The LDS padding of the y and x passes is timed with the tiles (each tile
// padded and unpadded).  Where the lengths allow, the two-stage passes (block
// gap 0 or 128 bytes) are timed last: the x pass, the y pair and the z pair
// each take the two-stage form if it beats the fastest LDS-staged one; untimed
// plans take it wherever the length allows.
template <class T>
s2h_status fft_pick_tile(s2h_ctx *ctx, s2h_vol *vol, FftPlan *p) {
    const int lmax = (int)std::max(p->f[0], p->f[1]);
    std::vector<int> cands;
    for (int t = kTileMax; t >= 1; t /= 2)
        if (lds_bytes<T>(lmax, t, kNoPad) <= 150 * 1024) cands.push_back(t);
    if (cands.empty()) return fail(S2H_ERR_UNSUPPORTED, "FFT tile exceeds LDS");
    // untimed: the widest tile for y; for x the narrowest whole-line tile,
    // else at most 4 columns; padded where the padding costs no occupancy
    p->tile_y = cands.front();
    p->tile_x = cands.back();
    for (int t : cands)
        if (t <= 4) {
            p->tile_x = t;
            break;
        }
    for (int t : cands)
        if ((size_t)t * sizeof(C2<T>) >= 128) p->tile_x = t;
    const int fz = (int)p->f[2];
    p->sh_z = pad_rule<T>(fz, z_lines_per_block<T>(fz));
    p->sh_y = lds_bytes<T>((int)p->f[1], p->tile_y, kPad) <= 150 * 1024 ? pad_rule<T>((int)p->f[1], p->tile_y) : kNoPad;
    p->sh_x = lds_bytes<T>((int)p->f[0], p->tile_x, kPad) <= 150 * 1024 ? pad_rule<T>((int)p->f[0], p->tile_x) : kNoPad;
    // the two-stage passes where their lengths allow, blocks padded
    p->x2_on = p->x2_a > 0;
    p->x2_pad = (int)(128 / sizeof(C2<T>));
    p->y2_on = p->y2_a > 0;
    p->y2_pad = (int)(128 / sizeof(C2<T>));
    p->z2_on = p->z2_a > 0;
    const char *env = std::getenv("S2HIP_AUTOTUNE");
    const double cells = (double)p->f[0] * (double)p->f[1] * (double)p->f[2];
    // tune only where a run is short (up to a 1536^3 padded grid: about 50 ms
    // per FP64 run there, 12 runs once per plan)
    if (!ctx->autotune || (env && env[0] == '0') || cands.size() < 2 || cells > 3.7e9) return S2H_OK;
    if (cands.size() > 3) cands.resize(3);
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    s2h_status st = S2H_OK;
    for (hipEvent_t &e : ev)
        if (st == S2H_OK && hipEventCreate(&e) != hipSuccess) st = fail(S2H_ERR_HIP, "hipEventCreate failed");
    float best_y = 1e30f, best_x = 1e30f;
    int pick_y = p->tile_y, pick_x = p->tile_x, psh_y = p->sh_y, psh_x = p->sh_x;
    const bool x2_ok = p->x2_on, y2_ok = p->y2_on, z2_ok = p->z2_on;
    int pick_x2 = -1;  // the two-stage x pass' padding if it is fastest, else -1
    int pick_y2 = -1;  // the same for the two-stage y passes
    bool pick_z2 = false;
    float best_z = 1e30f;  // z forward + z inverse of the LDS-staged passes
    p->x2_on = false;
    p->y2_on = false;
    p->z2_on = false;
    for (int t : cands) {
        for (int sh : {kPad, kNoPad}) {
            if (st != S2H_OK) break;
            if (lds_bytes<T>(lmax, t, sh) > 150 * 1024) continue;
            p->tile_y = p->tile_x = t;
            p->sh_y = p->sh_x = sh;
            st = fft_run<T>(ctx, vol, p);  // warm-up
            if (st == S2H_OK) st = fft_run<T>(ctx, vol, p, ev);
            hipError_t e = hipSuccess;
            if (st == S2H_OK) e = hipEventSynchronize(ev[5]);
            float yf = 0, xm = 0, yi = 0, zf = 0, zi = 0;
            if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&yf, ev[0], ev[1]);
            if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&xm, ev[1], ev[2]);
            if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&yi, ev[2], ev[3]);
            if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&zf, ev[4], ev[0]);
            if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&zi, ev[3], ev[5]);
            if (st == S2H_OK && e != hipSuccess)
                st = fail(S2H_ERR_HIP, std::string("FFT tile tuning: ") + hipGetErrorString(e));
            if (st != S2H_OK) break;
            best_z = std::min(best_z, zf + zi);
            // only whole-line tiles are taken: a narrower one reads each
            // 128-byte line once per workgroup that shares it
            const bool whole = (size_t)t * sizeof(C2<T>) >= 128;
            if (yf + yi < best_y && whole) {
                best_y = yf + yi;
                pick_y = t;
                psh_y = sh;
            }
            if (xm < best_x && whole) {
                best_x = xm;
                pick_x = t;
                psh_x = sh;
            }
        }
    }
    // the two-stage x and y passes, unpadded and padded, and the two-stage z
    // passes, against the best LDS-staged ones
    p->tile_y = pick_y;
    p->sh_y = psh_y;
    for (int pad : {0, (int)(128 / sizeof(C2<T>))}) {
        if (st != S2H_OK || !(x2_ok || y2_ok || z2_ok)) break;
        p->x2_on = x2_ok;
        p->x2_pad = pad;
        p->y2_on = y2_ok;
        p->y2_pad = pad;
        p->z2_on = z2_ok;
        st = fft_run<T>(ctx, vol, p);  // warm-up
        if (st == S2H_OK) st = fft_run<T>(ctx, vol, p, ev);
        hipError_t e = hipSuccess;
        float yf = 0, xm = 0, yi = 0, zf = 0, zi = 0;
        if (st == S2H_OK) e = hipEventSynchronize(ev[5]);
        if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&yf, ev[0], ev[1]);
        if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&xm, ev[1], ev[2]);
        if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&yi, ev[2], ev[3]);
        if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&zf, ev[4], ev[0]);
        if (st == S2H_OK && e == hipSuccess) e = hipEventElapsedTime(&zi, ev[3], ev[5]);
        if (st == S2H_OK && e != hipSuccess)
            st = fail(S2H_ERR_HIP, std::string("FFT tile tuning: ") + hipGetErrorString(e));
        if (st == S2H_OK && z2_ok && zf + zi < best_z) {
            best_z = zf + zi;
            pick_z2 = true;
        }
        if (st == S2H_OK && x2_ok && xm < best_x) {
            best_x = xm;
            pick_x2 = pad;
        }
        if (st == S2H_OK && y2_ok && yf + yi < best_y) {
            best_y = yf + yi;
            pick_y2 = pad;
        }
    }
    for (hipEvent_t &e : ev)
        if (e) (void)hipEventDestroy(e);
    p->tile_y = pick_y;
    p->tile_x = pick_x;
    p->sh_y = psh_y;
    p->sh_x = psh_x;
    p->x2_on = pick_x2 >= 0;
    p->x2_pad = pick_x2 >= 0 ? pick_x2 : 0;
    p->y2_on = pick_y2 >= 0;
    p->y2_pad = pick_y2 >= 0 ? pick_y2 : 0;
    p->z2_on = pick_z2;
    return st;
}

// The pruned correlation cube (precision T) of the exact plan's shape in
// fp->cube (device), and with per_offset also the value of every shell offset,
// in shell order, in fp->fcounts.
template <class T>
s2h_status fft_counts_device(s2h_ctx *ctx, s2h_vol *vol, ExactPlan *ep, FftPlan **fp_out, bool per_offset) {
    const int *ext = ep->shells->host->extent;
    int R[3] = {ext[0], ext[1], ext[2]};
    FftPlan *fp = nullptr;
    S2H_TRY(fft_plan_get(ctx, vol->dims, R, (int)sizeof(T), &fp));
    const uint64_t K = ep->offs.size();
    if (per_offset) S2H_TRY(dev_alloc(ctx, fp->fcounts, std::max<uint64_t>(K, 1) * sizeof(T)));
    if (fp->tile_y == 0) S2H_TRY(fft_pick_tile<T>(ctx, vol, fp));
    prof_mark(ctx, 0);
    S2H_TRY(fft_run<T>(ctx, vol, fp));
    if (per_offset && K) {
        fft_gather<T><<<(unsigned)((K + 255) / 256), 256, 0, ctx->stream>>>(
            (const Offset *)ep->d_offs.ptr, K, (const T *)fp->cube.ptr, R[0], R[1], R[2], (T *)fp->fcounts.ptr);
        S2H_HIP(hipGetLastError());
    }
    *fp_out = fp;
    return S2H_OK;
}

}  // namespace

// Worst-case bound on |computed - exact| of every FP64 correlation value.
// The input is 0/1 with K ones (||x||_2^2 = K, |Y_k| <= K for Y = DFT x).  A
// transform over L radix-2-equivalent levels has ||Y^ - Y||_2 <= L eta ||Y||_2
// with eta = mu + gamma_4 (sqrt 2 + mu) <= 8u (u = 2^-53, twiddles within
// mu <= 2u; Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed.,
// Thm 24.2).  Then ||Z^ - Z||_2 <= (2 rho + gamma_2) K ||Y||_2 for Z = |Y|^2,
// and the scaled inverse adds rho again, so every output satisfies
// |c^ - c| <= ||c^ - c||_2 <= (3 rho + 3u) K^{3/2}, rho = L * 8u.  L counts
// each radix-8, -4 and -2 stage as the 3, 2 and 1 radix-2 levels its
// butterfly is built of (its internal twiddles are +-1, +-i and (1 +- i)/sqrt 2),
// a radix-3 stage as 3 levels and a radix-5 stage as 4.  For those two, as
// small_dft codes them: the input twiddle product errs by at most
// tau |x_q|, tau = mu + sqrt 2 gamma_2 (1 + mu) <= 4.9u; each real output
// component is a sum of terms +-k Re x_q, +-k Im x_q with constants k in
// {1, 1/2, cos, sin} (rounded within u), at most d = 3 (radix 3) or 5
// (radix 5) roundings deep, and sum |terms| <= sum_q |x_q| <= sqrt p ||x||_2
// since each x_q enters with a (cos, sin) pair.  So the p outputs err by at
// most sqrt(2p) gamma_{d+1} ||y||_2 (||y||_2 = sqrt p ||x||_2): eta_3 <=
// 4.9u + sqrt 6 gamma_4 < 3 * 8u and eta_5 <= 4.9u + sqrt 10 gamma_6 < 4 * 8u.
// Over the three axes, plus two for the real-to-complex packing and
// unpacking along z.  With
// the bound below 1/2 the rounded values are the exact integer counts; the
// path requires 1/4.
double fft64_error_bound(const uint64_t dims[3], const int R[3], uint64_t occupied) {
    double levels = 2.0;
    for (int a = 0; a < 3; ++a) {
        const AxisPlan ap = axis_plan((int)std::min<uint64_t>(fft_length(dims[a], (uint64_t)R[a]), 1u << 30));
        if (ap.nf < 0) return INFINITY;
        for (int s = 0; s < ap.nf; ++s)
            levels += ap.fac[s] == 8 ? 3 : ap.fac[s] == 4 ? 2 : ap.fac[s] == 2 ? 1 : ap.fac[s] == 3 ? 3 : 4;
    }
    const double u = std::ldexp(1.0, -53);
    const double rho = levels * 8.0 * u;
    const double k = (double)occupied;
    return (3.0 * rho + 3.0 * u) * k * std::sqrt(k);
}

// FP64 correlation of every shell offset of `ep` for `vol` (device array of
// ep->offs.size() doubles, in shell order).  The caller rounds them.
s2h_status fft64_exact_device(s2h_ctx *ctx, s2h_vol *vol, ExactPlan *ep, FftCube *cube) {
    FftPlan *fp = nullptr;
    S2H_TRY(fft_counts_device<double>(ctx, vol, ep, &fp, false));
    *cube = FftCube();
    cube->f64 = (const double *)fp->cube.ptr;
    for (int a = 0; a < 3; ++a) cube->R[a] = fp->R[a];
    return S2H_OK;
}

// ---------------------------------------------------------------- delta counts by region correlation (host)

namespace {

constexpr uint64_t kFftdChunkBytes = 2ull << 30;  // frame buffers of one launch sequence

// Smooth length of least f x (stages + 2) from need (at least 16) to twice it.
int fftd_length(int need) {
    const uint64_t lo = (uint64_t)std::max(need, 16), hi = 2 * lo;
    uint64_t best = 0;
    double best_cost = INFINITY;
    for (uint64_t p2 = 1; p2 <= hi; p2 *= 2)
        for (uint64_t p23 = p2; p23 <= hi; p23 *= 3)
            for (uint64_t v = p23; v <= hi; v *= 5) {
                if (v < lo || v > 8192) continue;
                const AxisPlan ap = axis_plan((int)v);
                if (ap.nf < 2) continue;
                const double cost = (double)v * (ap.nf + 2);
                if (cost < best_cost || (cost == best_cost && v < best)) {
                    best_cost = cost;
                    best = v;
                }
            }
    return (int)best;
}

// Radix-2-equivalent levels of a length (as in fft64_error_bound).
double fft_levels(int L) {
    const AxisPlan a = axis_plan(L);
    double lv = 0;
    for (int s = 0; s < a.nf; ++s)
        lv += a.fac[s] == 8 ? 3 : a.fac[s] == 4 ? 2 : a.fac[s] == 2 ? 1 : a.fac[s] == 3 ? 3 : 4;
    return lv;
}

// The frame plan: twiddles and z order for lengths F, tiles chosen untimed
// (the y passes the widest tile whose LDS fits, the x pass -- two fields'
// lines per column -- the narrowest whole-line tile), and grow-only frame
// buffers.
s2h_status fftd_plan_get(s2h_ctx *ctx, const int F[3], const int R[3], FftPlan **out) {
    const auto key = std::make_tuple((uint64_t)F[0] | ((uint64_t)R[0] << 32), (uint64_t)F[1] | ((uint64_t)R[1] << 32),
                                     (uint64_t)F[2] | ((uint64_t)R[2] << 32));
    auto it = ctx->fftd_plans.find(key);
    if (it != ctx->fftd_plans.end()) {
        *out = it->second;
        return S2H_OK;
    }
    auto *p = new FftPlan();
    p->real_bytes = 8;
    for (int a = 0; a < 3; ++a) {
        p->dims[a] = (uint64_t)F[a];
        p->f[a] = (uint64_t)F[a];
        p->R[a] = R[a];
        p->ax[a] = axis_plan(F[a]);
        if (p->ax[a].nf < 2 || F[a] > 8192) {
            delete p;
            return fail(S2H_ERR_UNSUPPORTED, "frame length " + std::to_string(F[a]) + " unsupported");
        }
    }
    const s2h_status st = fft_plan_tables(ctx, p);
    if (st != S2H_OK) {
        fft_plan_release(ctx, p);
        return st;
    }
    p->tile_y = 1;
    for (int t = kTileMax; t >= 1; t /= 2)
        if (lds_bytes<double>(F[1], t, kNoPad) <= 150 * 1024) {
            p->tile_y = t;
            break;
        }
    p->tile_x = 1;
    for (int t = 1; t <= kTileMax; t *= 2)
        if (lds_bytes<double>(F[0], 2 * t, kNoPad) <= 150 * 1024) {
            p->tile_x = t;
            if ((size_t)t * sizeof(double2) >= 128) break;
        }
    p->sh_z = pad_rule<double>(F[2], z_lines_per_block<double>(F[2]));
    p->sh_y = lds_bytes<double>(F[1], p->tile_y, kPad) <= 150 * 1024 ? pad_rule<double>(F[1], p->tile_y) : kNoPad;
    p->sh_x = lds_bytes<double>(F[0], 2 * p->tile_x, kPad) <= 150 * 1024 ? pad_rule<double>(F[0], 2 * p->tile_x)
                                                                       : kNoPad;
    ctx->fftd_plans[key] = p;
    *out = p;
    return S2H_OK;
}

}  // namespace

bool fft_delta_frame(const int bmax[3], const int R[3], int F[3]) {
    for (int a = 0; a < 3; ++a) {
        F[a] = fftd_length(std::max(bmax[a], 1) + 2 * R[a]);
        if (F[a] <= 0) return false;
    }
    return lds_bytes<double>(std::max(F[0], F[1]), 2, kNoPad) <= 150 * 1024 &&
           lds_bytes<double>(F[2], 1, kNoPad) <= 150 * 1024;
}

// Worst-case bound on |computed - exact| of every frame delta value.  With K_o
// ones of o and K_x nonzeros of delta in the frame (N points), the spectra
// satisfy |O_k| <= K_o, |X_k| <= K_x, ||O||_2 = sqrt(N K_o), ||X||_2 =
// sqrt(N K_x), and the forward passes leave ||O^ - O||_2 <= rho ||O||_2 and
// ||X^ - X||_2 <= rho ||X||_2, rho = 8u L with L the radix-2-equivalent levels
// of the three axes plus two for the packing and splitting along z (as for
// the autocorrelation).  The cross spectrum P = 2 Re(conj(O) X) + |X|^2,
// formed with three roundings per element, then has ||P^ - P||_2 <=
// (2 rho + 2 gamma_3)(K_o sqrt(N K_x)) + 2 rho K_x sqrt(N K_o) + (3 rho +
// gamma_3) K_x sqrt(N K_x) up to second-order terms, and the scaled inverse
// divides the 2-norm by sqrt(N) and adds rho ||P||_2 / sqrt(N) <= rho (2 K_o
// sqrt(K_x) + K_x sqrt(K_x)).  Every value's error is at most the 2-norm:
// (4 rho + 2 gamma_3) (K_o sqrt(K_x) + K_x sqrt(K_o) + K_x^{3/2}), taken here
// with 5 rho + 8u and 2 % slack for the second-order terms, plus one rounding
// of the scaling on a value of magnitude at most 2 K_x.
double fft_delta_error_bound(const int F[3], uint64_t ones, uint64_t changed) {
    constexpr double u = 1.1102230246251565e-16;
    const double L = fft_levels(F[0]) + fft_levels(F[1]) + fft_levels(F[2]) + 2.0;
    const double rho = 8.0 * u * L;
    const double ko = (double)ones, kx = (double)changed;
    return 1.02 * (5.0 * rho + 8.0 * u) * (ko * std::sqrt(kx) + kx * std::sqrt(ko) + kx * std::sqrt(kx)) +
           2.0 * u * kx;
}

// Bytes the frame passes move per candidate (each reads its input and writes
// its output once) at the FP64 passes' modelled rate, plus launch time.
double fft_delta_model_ms(const int F[3], const int R[3], uint32_t ncand) {
    const double fx = F[0], fy = F[1], h = F[2] / 2 + 1;
    const double xo = 2.0 * R[0] + 1, yo = 2.0 * R[1] + 1, zo = 2.0 * R[2] + 1;
    const double bytes = 2 * fx * fy * h * 16 * 3 + xo * fy * h * 16 * 2 + xo * yo * h * 16 * 2 + xo * yo * zo * 16;
    return ncand * bytes / 4.0e12 * 1e3 + 0.04;
}

// The frames' two-stage x pass (fftd_x_cross2) for nc candidates.
static s2h_status fftd_xc2_launch(FftPlan *p, const FftdCand *cc, int nc, int Fy, int H, int P, int Rx,
                                  const double2 *g1, double2 *g2, hipStream_t s) {
    const int A = p->x2_a, B = p->x2_b, tile = x2_tile<double>(), pad = 128 / (int)sizeof(double2);
    const int tiles = (H + tile - 1) / tile;
    const int threads = (std::max(2 * B, A) * tile + 63) / 64 * 64;
    const size_t lds = xc2_lds_bytes<double>(A, B, pad);
    const double2 *tw = (const double2 *)p->tw_x2.ptr;
    const unsigned grid = (unsigned)((uint64_t)nc * Fy * tiles);
#define S2H_XC2(AA, BB)                                                                                           \
    if (A == AA && B == BB) {                                                                                     \
        S2H_TRY(set_lds((const void *)fftd_x_cross2<double, AA, BB>, lds));                                      \
        fftd_x_cross2<double, AA, BB><<<grid, threads, lds, s>>>(cc, Fy, H, P, tiles, Rx, pad, tw, g1, g2);       \
        S2H_HIP(hipGetLastError());                                                                               \
        return S2H_OK;                                                                                            \
    }
    S2H_XC2(8, 12) S2H_XC2(8, 15) S2H_XC2(8, 16) S2H_XC2(16, 10) S2H_XC2(16, 12) S2H_XC2(16, 16) S2H_XC2(16, 20)
#undef S2H_XC2
    return fail(S2H_ERR_UNSUPPORTED, "two-stage frame x pass length");
}

s2h_status fft_delta_run(s2h_ctx *ctx, const FftdSeed &sd, const FftdCand *cands, uint32_t ncand, const int F[3],
                         const uint2 *ad, int nwd, int32_t *delta) {
    using V = double2;
    if (ncand == 0) return S2H_OK;
    FftPlan *p = nullptr;
    S2H_TRY(fftd_plan_get(ctx, F, sd.R, &p));
    const int Fx = F[0], Fy = F[1], Fz = F[2];
    const int H = Fz / 2 + 1, P = (int)spectrum_pitch((uint64_t)Fz, 8);
    const int Rx = sd.R[0], Ry = sd.R[1], Rz = sd.R[2];
    const int nxo = Rx + 1, nyo = 2 * Ry + 1, nzo = 2 * Rz + 1;  // x planes dx = 0 .. R_x
    const uint64_t b1 = 2ull * Fx * Fy * P * 16, b2 = (uint64_t)nxo * Fy * P * 16, b3 = (uint64_t)nxo * nyo * P * 16,
                   b4 = (uint64_t)nxo * nyo * nzo * 8;
    uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(ncand, kFftdChunkBytes / (b1 + b2 + b3 + b4)));
    // every line and block count of a launch stays below 2^31
    const uint64_t per_cand_lines = std::max<uint64_t>({2ull * Fx * Fy, (uint64_t)nxo * nyo, (uint64_t)Fy * H});
    chunk = std::min<uint64_t>(chunk, ((1ull << 31) - 1) / per_cand_lines);
    S2H_TRY(dev_alloc(ctx, p->g1, chunk * b1));
    S2H_TRY(dev_alloc(ctx, p->g2, chunk * b2));
    S2H_TRY(dev_alloc(ctx, p->g3, chunk * b3));
    S2H_TRY(dev_alloc(ctx, p->cube, chunk * b4));
    const Fac fcx = to_fac(p->ax[0]), fcy = to_fac(p->ax[1]), fcz = to_fac(p->ax[2]);
    hipStream_t s = ctx->stream;
    const int ty = p->tile_y, tx = p->tile_x;
    const int tiles_y = (H + ty - 1) / ty, tiles_x = (H + tx - 1) / tx;
    const int ppb = z_lines_per_block<double>(Fz);
    const size_t lds_z = fftd_z_lds_bytes<double>(Fz, ppb, p->sh_z, (int)sd.row_words, nwd),
                 lds_y = lds_bytes<double>(Fy, ty, p->sh_y), lds_x = lds_bytes<double>(Fx, 2 * tx, p->sh_x);
    S2H_TRY(set_lds((const void *)fftd_z_fwd<double>, lds_z));
    S2H_TRY(set_lds((const void *)fft_y_fwd<double>, lds_y));
    S2H_TRY(set_lds((const void *)fftd_x_cross<double>, lds_x));
    S2H_TRY(set_lds((const void *)fft_y_inv<double>, lds_y));
    const size_t lds_zi = z_inv_lds_bytes<double>(Fz, ppb, p->sh_z);
    S2H_TRY(set_lds((const void *)fft_z_inv<double>, lds_zi));
    const double scale = 1.0 / ((double)Fx * (double)Fy * (double)Fz);
    // the x pass in two register stages where F_x allows it (stage B holds
    // both fields' blocks: B <= 20)
    const bool xc2 = p->x2_a > 0 && p->x2_b <= 20;
    const V *tw0 = (const V *)p->tw[0].ptr, *tw1 = (const V *)p->tw[1].ptr, *tw2 = (const V *)p->tw[2].ptr;
    V *g1 = (V *)p->g1.ptr, *g2 = (V *)p->g2.ptr, *g3 = (V *)p->g3.ptr;
    double *cube = (double *)p->cube.ptr;
    for (uint64_t c0 = 0; c0 < ncand; c0 += chunk) {
        const int nc = (int)std::min<uint64_t>(chunk, ncand - c0);
        const FftdCand *cc = cands + c0;
        const int zl = nc * Fx * Fy;
        fftd_z_fwd<double><<<(zl + ppb - 1) / ppb, kFftThreads, lds_z, s>>>(
            cc, zl, sd.words, sd.dims[0], sd.dims[1], sd.dims[2], (int)sd.row_words, ad, nwd, Rx, Ry, Rz, Fx, Fy, H, P,
            fcz, Fz, tw2, (const int *)p->pos_z.ptr, ppb, g1, p->sh_z, fast_div((uint32_t)(Fx * Fy)),
            fast_div((uint32_t)Fy), fast_div((uint32_t)H));
        fft_y_fwd<double><<<nc * 2 * Fx * tiles_y, pass_threads(Fy, fcy, ty), lds_y, s>>>(Fy, Fy, H, P, tiles_y, ty,
                                                                                          fcy, tw1, g1, p->sh_y, cc,
                                                                                          Fx, Rx, Ry);
        if (xc2) {
            S2H_TRY(fftd_xc2_launch(p, cc, nc, Fy, H, P, Rx, g1, g2, s));
        } else {
            fftd_x_cross<double><<<nc * Fy * tiles_x, pass_threads(Fx, fcx, tx), lds_x, s>>>(
                cc, Fx, Fy, H, P, tiles_x, tx, fcx, tw0, Rx, g1, g2, p->sh_x);
        }
        fft_y_inv<double><<<nc * nxo * tiles_y, pass_threads(Fy, fcy, ty), lds_y, s>>>(Fy, H, P, tiles_y, ty, fcy, tw1,
                                                                                       Ry, g2, g3, p->sh_y);
        const int zi = nc * nxo * nyo;
        fft_z_inv<double><<<((zi + 1) / 2 + ppb - 1) / ppb, kFftThreads, lds_zi, s>>>(
            zi, H, P, fcz, Fz, tw2, (const int *)p->rev_z.ptr, ppb, Rz, scale, g3, cube, p->sh_z,
            fast_div((uint32_t)nzo));
        fftd_gather<<<2048, 256, 0, s>>>(cc, (uint32_t)nc, sd.offs, sd.sg, sd.noffs, cube, Rx, Ry, Rz, delta);
        S2H_HIP(hipGetLastError());
        ctx->launches += 6;
    }
    return S2H_OK;
}

// Bytes the pruned passes of one FP64 evaluation move (each reads its input
// and writes its output once), for the method model.
double fft64_model_bytes(const uint64_t dims[3], const int R[3]) {
    const double fy = (double)fft_length(dims[1], (uint64_t)R[1]);
    const double h = (double)(fft_length(dims[2], (uint64_t)R[2]) / 2 + 1);
    const double nx = (double)dims[0], ny = (double)dims[1], nz = (double)dims[2];
    const double xo = R[0] + 1.0, yo = 2.0 * R[1] + 1, zo = 2.0 * R[2] + 1;
    const double c = 16.0;
    return nx * ny * (std::ceil(nz / 64.0) + 1) * 8.0 + nx * ny * h * c + (nx * ny * h + nx * fy * h) * c +
           (nx * fy * h + xo * fy * h) * c + (xo * fy * h + xo * yo * h) * c + xo * yo * h * c + xo * yo * zo * 8.0;
}

// Device bytes of an FP64 plan's spectra, cube and per-offset values (the
// working set the method choice budgets, as fft_working_set_bytes does for the
// reference's planner, src/geometry/s2.rs:695-714).
double fft64_plan_bytes(const uint64_t dims[3], const int R[3]) {
    return (double)fft_plan_buffer_bytes(dims, R, 8);
}

uint64_t fft_plan_buffer_bytes(const uint64_t dims[3], const int R[3], int real_bytes) {
    uint64_t sizes[4];
    plan_buffer_sizes(dims, R, real_bytes, sizes);
    return sizes[0] + sizes[1] + sizes[2] + sizes[3];
}

uint64_t fft_padded_length(uint64_t n, uint64_t R) { return fft_length(n, R); }

bool fft64_plan_cached(s2h_ctx *ctx, const uint64_t dims[3], const int R[3]) {
    return ctx->fft_plans.count(fft_plan_key(dims, R, 8)) != 0;
}

void fft_plans_release(s2h_ctx *ctx) {
    for (auto &kv : ctx->fft_plans) fft_plan_release(ctx, kv.second);
    ctx->fft_plans.clear();
}

// Whether the FP64 passes can run this shape (lengths of 2, 3, 5 factors up
// to 8192 whose LDS tiles fit).
bool fft64_supported(const uint64_t dims[3], const int R[3]) {
    uint64_t f[3];
    for (int a = 0; a < 3; ++a) {
        f[a] = fft_length(dims[a], (uint64_t)R[a]);
        if (f[a] > 8192 || axis_plan((int)f[a]).nf < 0) return false;
    }
    const int lmax = (int)std::max(f[0], f[1]);
    const int lz = (int)f[2];
    return lds_bytes<double>(lmax, 1, kNoPad) <= 150 * 1024 && lds_bytes<double>(lz, 1, kNoPad) <= 150 * 1024;
}

}  // namespace s2h

using namespace s2h;

extern "C" {

s2h_status s2h_fft_padded_dims(const uint64_t dims[3], uint32_t r_max, double pitch, uint64_t padded[3]) {
    if (!dims || !padded) return fail(S2H_ERR_ARG, "null argument");
    if (!(pitch > 0.0) || !std::isfinite(pitch)) return fail(S2H_ERR_ARG, "pitch must be finite and > 0");
    for (int a = 0; a < 3; ++a)
        if (dims[a] == 0 || dims[a] > 65535) return fail(S2H_ERR_ARG, "grid dimension outside 1..65535");
    int R[3];
    shell_extent(dims, r_max, pitch, R);
    for (int a = 0; a < 3; ++a) padded[a] = fft_length(dims[a], (uint64_t)R[a]);
    return S2H_OK;
}

s2h_status s2h_fft32_counts(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, float *counts,
                            uint64_t n_offsets) {
    if (!ctx || !vol || (!counts && n_offsets)) return fail(S2H_ERR_ARG, "null argument");
    if (!(pitch > 0.0) || !std::isfinite(pitch)) return fail(S2H_ERR_ARG, "pitch must be finite and > 0");
    S2H_HIP(hipSetDevice(ctx->device));
    S2H_TRY(inject(kFailLaunch));
    ExactPlan *ep = nullptr;
    S2H_TRY(exact_plan_get(ctx, vol->dims, r_max, pitch, &ep));
    if (ep->offs.size() != n_offsets)
        return fail(S2H_ERR_SIZE, "counts buffer holds " + std::to_string(n_offsets) + " entries, plan has " +
                                      std::to_string(ep->offs.size()));
    FftPlan *fp = nullptr;
    S2H_TRY(fft_counts_device<float>(ctx, vol, ep, &fp, true));
    prof_mark(ctx, 1);
    prof_mark(ctx, 2);
    S2H_TRY(inject(kFailReadback));
    if (n_offsets)
        S2H_HIP(hipMemcpyAsync(counts, fp->fcounts.ptr, n_offsets * 4, hipMemcpyDeviceToHost, ctx->stream));
    S2H_HIP(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    return S2H_OK;
}

s2h_status s2h_fft32_s2(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, double *s2, uint64_t n) {
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
    ExactPlan *ep = nullptr;
    S2H_TRY(exact_plan_get(ctx, vol->dims, r_max, pm, &ep));
    FftPlan *fp = nullptr;
    S2H_TRY(fft_counts_device<float>(ctx, vol, ep, &fp, false));
    FftCube cube;
    cube.f32 = (const float *)fp->cube.ptr;
    for (int a = 0; a < 3; ++a) cube.R[a] = fp->R[a];
    void *stage = nullptr;
    S2H_TRY(curve_staging(ctx, ep, &stage));
    S2H_TRY(exact_gather_fold(ctx, ep, occupied, &cube, false, stage));
    const double vf = (double)occupied / (double)(vol->dims[0] * vol->dims[1] * vol->dims[2]);
    return finish_curve_from_device(ctx, ep, vf, s2, stage);
}

}  // extern "C"
