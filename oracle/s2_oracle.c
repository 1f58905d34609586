/*
 * s2_oracle.c -- CPU restatement of RustMSPT's S2 two-point-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP engine is compared
 * against and the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never links it.
 *
 * Every function restates one reference function; the citation (file:line, in
 * the RustMSPT source tree) is given beside it.  Floating-point expressions keep
 * the reference's operation order and the file is compiled with
 * -ffp-contract=off, so f64 results match Rust's (which never contracts).
 *
 * Pinning (how each part is checked without running the Rust reference):
 *   shells            closed-form lattice counts + the reference's order test
 *   packed counts     brute-force pair enumeration (bool scan) on the reference's
 *                     own tiny grids, C(d)=C(-d), solid/empty closed forms
 *   curve / spline    ordered-sum restatement; spline pinned against the textbook
 *                     natural cubic spline (scipy CubicSpline bc_type='natural')
 *   voxelization      analytic boxes from the reference tests, icospheres vs the
 *                     analytic sphere away from the surface
 *   padded dims       the reference's known answers [7,11,13]->[15,24,25] etc.
 *   FFT f64           brute-force integer pair counts (rounding) and numpy.fft
 *   Philox4x32-10     Random123 known-answer vectors
 *   voxel MC          statistical only (3 sigma against the exact estimand);
 *                     the reference's own SmallRng stream is parity unpinned.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double x, y, z; } v3;

/* src/types.rs:16-43 -- sub, dot ((x*x'+y*y')+z*z'), cross in the reference order. */
static inline v3 v3_sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline double v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 v3_cross(v3 a, v3 b) {
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}

/* src/geometry/s2.rs:14 */
static const v3 RAY_DIR = {0.9428090415820634, 0.2705980500730985, 0.19611613513818402};

/* src/geometry/s2.rs:27-56 ray_intersects_triangle.  Returns 1 and *t on a forward hit. */
int orc_ray_triangle(const double *o, const double *d, const double *a, const double *b,
                     const double *c, double *t_out) {
    const double eps = 1e-10;
    v3 O = {o[0], o[1], o[2]}, D = {d[0], d[1], d[2]};
    v3 A = {a[0], a[1], a[2]}, B = {b[0], b[1], b[2]}, C = {c[0], c[1], c[2]};
    v3 e1 = v3_sub(B, A), e2 = v3_sub(C, A);
    v3 h = v3_cross(D, e2);
    double det = v3_dot(e1, h);
    if (fabs(det) <= eps) return 0;
    double inv = 1.0 / det;
    v3 s = v3_sub(O, A);
    double u = inv * v3_dot(s, h);
    if (!(u >= 0.0 - eps && u <= 1.0 + eps)) return 0;
    v3 q = v3_cross(s, e1);
    double v = inv * v3_dot(D, q);
    if (v < -eps || (u + v) > 1.0 + eps) return 0;
    double t = inv * v3_dot(e2, q);
    if (t > eps) { *t_out = t; return 1; }
    return 0;
}

static int cmp_f64(const void *pa, const void *pb) {
    double a = *(const double *)pa, b = *(const double *)pb;
    return (a < b) ? -1 : (a > b) ? 1 : 0;
}

/* Granule bbox (src/geometry/bbox.rs:8-25) over the vertices its faces reference. */
typedef struct { double mn[3], mx[3]; int valid; } bbox3;

/* src/geometry/mesh_query.rs:73-116 contains_point, without the BVH (the BVH only
 * prunes triangles its conservative slab test proves unreachable, so the hit
 * multiset -- and hence the sorted, anchored-dedup parity -- is the full scan's).
 * tris: granule's triangles as 9 doubles each. */
static int granule_contains(const double *tris, size_t ntri, const bbox3 *bb, v3 p,
                            double *scratch) {
    const double eps = 1e-9;
    if (!bb->valid) return 0;
    if (p.x < bb->mn[0] - eps || p.x > bb->mx[0] + eps || p.y < bb->mn[1] - eps ||
        p.y > bb->mx[1] + eps || p.z < bb->mn[2] - eps || p.z > bb->mx[2] + eps)
        return 0;
    double o[3] = {p.x, p.y, p.z}, d[3] = {RAY_DIR.x, RAY_DIR.y, RAY_DIR.z};
    size_t n = 0;
    for (size_t i = 0; i < ntri; i++) {
        double t;
        if (orc_ray_triangle(o, d, tris + 9 * i, tris + 9 * i + 3, tris + 9 * i + 6, &t))
            scratch[n++] = t;
    }
    if (n == 0) return 0;
    qsort(scratch, n, sizeof(double), cmp_f64);
    size_t unique = 0;
    double last = -INFINITY;
    for (size_t i = 0; i < n; i++) {
        if (fabs(scratch[i] - last) > 1e-8) { unique++; last = scratch[i]; }
    }
    return (int)(unique & 1);
}

/* src/geometry/s2.rs:110-117 occupancy_dims */
void orc_occupancy_dims(const double *bmin, const double *bmax, double pitch, uint64_t *dims) {
    for (int a = 0; a < 3; a++) {
        double n = ceil((bmax[a] - bmin[a]) / pitch);
        uint64_t u = (n != n || n <= 0.0) ? 0 : (n >= 1.8446744073709552e19 ? UINT64_MAX : (uint64_t)n);
        dims[a] = u < 1 ? 1 : u;
    }
}

/* Rust `f as isize` saturates and maps NaN to 0. */
static int64_t f2i_sat(double f) {
    if (f != f) return 0;
    if (f >= 9.2233720368547758e18) return INT64_MAX;
    if (f <= -9.2233720368547758e18) return INT64_MIN;
    return (int64_t)f;
}

/* src/geometry/mesh_ops.rs:48-~120 split_mesh_into_granules: connected components
 * over shared vertex indices; component k is the one whose BFS starts at the k-th
 * smallest face index not yet visited.  Writes the component id per face and
 * returns the number of components. */
uint32_t orc_granules(const uint32_t *tri, uint64_t ntri, uint64_t nverts, uint32_t *comp) {
    if (ntri == 0 || nverts == 0) return 0;
    uint64_t *off = calloc(nverts + 1, sizeof(uint64_t));
    for (uint64_t f = 0; f < ntri; f++)
        for (int k = 0; k < 3; k++) off[tri[3 * f + k] + 1]++;
    for (uint64_t v = 0; v < nverts; v++) off[v + 1] += off[v];
    uint64_t *cur = malloc(nverts * sizeof(uint64_t));
    memcpy(cur, off, nverts * sizeof(uint64_t));
    uint64_t *adj = malloc(off[nverts] * sizeof(uint64_t) + 1);
    for (uint64_t f = 0; f < ntri; f++)
        for (int k = 0; k < 3; k++) adj[cur[tri[3 * f + k]]++] = f;
    uint8_t *seen = calloc(ntri, 1);
    uint64_t *queue = malloc(ntri * sizeof(uint64_t));
    uint32_t ncomp = 0;
    for (uint64_t s = 0; s < ntri; s++) {
        if (seen[s]) continue;
        uint64_t head = 0, tail = 0;
        seen[s] = 1;
        queue[tail++] = s;
        while (head < tail) {
            uint64_t f = queue[head++];
            comp[f] = ncomp;
            for (int k = 0; k < 3; k++) {
                uint32_t v = tri[3 * f + k];
                for (uint64_t j = off[v]; j < off[v + 1]; j++) {
                    uint64_t g = adj[j];
                    if (!seen[g]) { seen[g] = 1; queue[tail++] = g; }
                }
            }
        }
        ncomp++;
    }
    free(off); free(cur); free(adj); free(seen); free(queue);
    return ncomp;
}

/* src/geometry/s2.rs:160-201 build_bbox_occupancy (with 126-147 part_voxel_ranges,
 * 150-152 voxel_centre, 17-19 index order).  verts: nverts*3 doubles (welded mesh),
 * tri: ntri*3 indices, comp: component id per face from orc_granules.  occ is
 * nx*ny*nz bytes, x-major, z fastest.  Returns the occupied count. */
uint64_t orc_voxelize(const double *verts, uint64_t nverts, const uint32_t *tri, uint64_t ntri,
                      const uint32_t *comp, uint32_t ncomp, const double *bmin, double pitch,
                      const uint64_t *dims, uint8_t *occ) {
    (void)nverts;
    uint64_t nx = dims[0], ny = dims[1], nz = dims[2];
    memset(occ, 0, nx * ny * nz);
    if (ncomp == 0) return 0;
    /* gather triangles per component (face order within a component does not affect
     * the hit multiset) and the component bbox over referenced vertices */
    uint64_t *cstart = calloc(ncomp + 1, sizeof(uint64_t));
    for (uint64_t f = 0; f < ntri; f++) cstart[comp[f] + 1]++;
    for (uint32_t c = 0; c < ncomp; c++) cstart[c + 1] += cstart[c];
    uint64_t *fill = malloc(ncomp * sizeof(uint64_t));
    memcpy(fill, cstart, ncomp * sizeof(uint64_t));
    double *ctri = malloc(ntri * 9 * sizeof(double));
    bbox3 *cbb = calloc(ncomp, sizeof(bbox3));
    for (uint64_t f = 0; f < ntri; f++) {
        uint32_t c = comp[f];
        double *dst = ctri + 9 * fill[c]++;
        for (int k = 0; k < 3; k++) {
            const double *v = verts + 3 * (uint64_t)tri[3 * f + k];
            for (int a = 0; a < 3; a++) {
                dst[3 * k + a] = v[a];
                if (!cbb[c].valid) { cbb[c].mn[a] = cbb[c].mx[a] = v[a]; }
                else {
                    /* f64::min / f64::max ignore a NaN operand */
                    if (v[a] < cbb[c].mn[a] || cbb[c].mn[a] != cbb[c].mn[a]) cbb[c].mn[a] = v[a];
                    if (v[a] > cbb[c].mx[a] || cbb[c].mx[a] != cbb[c].mx[a]) cbb[c].mx[a] = v[a];
                }
            }
            cbb[c].valid = 1;
        }
    }
    uint64_t occupied = 0;
    int64_t *rng = malloc(ncomp * 6 * sizeof(int64_t));
    for (uint32_t c = 0; c < ncomp; c++) {
        int64_t *r = rng + 6 * c;
        for (int a = 0; a < 3; a++) {
            int64_t lo = f2i_sat(floor((cbb[c].mn[a] - bmin[a]) / pitch));
            int64_t hi = f2i_sat(ceil((cbb[c].mx[a] - bmin[a]) / pitch));
            int64_t n = (int64_t)dims[a];
            r[2 * a] = lo < 0 ? 0 : lo;
            r[2 * a + 1] = hi < 0 ? 0 : (hi > n ? n : hi);
        }
    }
    uint64_t maxtri = 0;
    for (uint32_t c = 0; c < ncomp; c++)
        if (cstart[c + 1] - cstart[c] > maxtri) maxtri = cstart[c + 1] - cstart[c];
#pragma omp parallel reduction(+ : occupied)
    {
        double *scratch = malloc((maxtri + 1) * sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (uint64_t x = 0; x < nx; x++) {
            double cx = bmin[0] + ((double)x + 0.5) * pitch;
            uint8_t *slab = occ + x * ny * nz;
            for (uint32_t c = 0; c < ncomp; c++) {
                const int64_t *r = rng + 6 * c;
                if (!(r[0] < r[1] && r[2] < r[3] && r[4] < r[5])) continue;
                if ((int64_t)x < r[0] || (int64_t)x >= r[1]) continue;
                for (int64_t y = r[2]; y < r[3]; y++) {
                    double cy = bmin[1] + ((double)y + 0.5) * pitch;
                    for (int64_t z = r[4]; z < r[5]; z++) {
                        uint64_t idx = (uint64_t)y * nz + (uint64_t)z;
                        if (slab[idx]) continue;
                        v3 p = {cx, cy, bmin[2] + ((double)z + 0.5) * pitch};
                        if (granule_contains(ctri + 9 * cstart[c], cstart[c + 1] - cstart[c], &cbb[c], p, scratch)) {
                            slab[idx] = 1;
                            occupied++;
                        }
                    }
                }
            }
        }
        free(scratch);
    }
    free(cstart); free(fill); free(ctri); free(cbb); free(rng);
    return occupied;
}

/* ---------------------------------------------------------------- shells */

/* src/geometry/s2.rs:336-361 shell_offsets_for_distance (== 364-385 shell_offset_iter
 * for every radius the tests use).  Writes up to cap offsets (dx,dy,dz) and returns
 * the total count (call with cap=0 to size). */
uint64_t orc_shell_offsets(double distance_vox, double half_width_vox, int64_t *out, uint64_t cap) {
    if (distance_vox <= 1e-12) {
        if (cap >= 1) { out[0] = 0; out[1] = 0; out[2] = 0; }
        return 1;
    }
    double lo = distance_vox - half_width_vox;
    if (lo < 0.0) lo = 0.0;
    double low2 = lo * lo; /* powi(2) is a single multiplication */
    double hi = distance_vox + half_width_vox;
    double high2 = hi * hi;
    int64_t lim = f2i_sat(ceil(distance_vox + half_width_vox + 1.0));
    uint64_t n = 0;
    for (int64_t dx = -lim; dx <= lim; dx++)
        for (int64_t dy = -lim; dy <= lim; dy++)
            for (int64_t dz = -lim; dz <= lim; dz++) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                double d2 = (double)(dx * dx + dy * dy + dz * dz);
                if (d2 >= low2 && d2 < high2) {
                    if (n < cap) { out[3 * n] = dx; out[3 * n + 1] = dy; out[3 * n + 2] = dz; }
                    n++;
                }
            }
    return n;
}

/* src/geometry/s2.rs:1028-1030 offset_in_domain */
static inline int in_domain(const int64_t *d, const uint64_t *dims) {
    for (int a = 0; a < 3; a++) {
        uint64_t m = (uint64_t)(d[a] < 0 ? -d[a] : d[a]);
        if (m >= dims[a]) return 0;
    }
    return 1;
}

/* All in-domain offsets of radii 1..r_max in shell order (the curve's summation
 * order, src/geometry/s2.rs:1190-1193).  radius_start has r_max+2 entries:
 * offsets of radius r are [radius_start[r], radius_start[r+1]); radius 0 is empty.
 * Returns the total; call with cap=0 to size. */
uint64_t orc_domain_offsets(const uint64_t *dims, uint32_t r_max, double pitch, int64_t *out,
                            uint64_t cap, uint64_t *radius_start) {
    uint64_t total = 0;
    if (radius_start) { radius_start[0] = 0; radius_start[1] = 0; }
    for (uint32_t r = 1; r <= r_max; r++) {
        double dist = (double)r / pitch, hw = 0.5 / pitch;
        uint64_t n = orc_shell_offsets(dist, hw, NULL, 0);
        int64_t *tmp = malloc((n ? n : 1) * 3 * sizeof(int64_t));
        orc_shell_offsets(dist, hw, tmp, n);
        for (uint64_t i = 0; i < n; i++) {
            if (!in_domain(tmp + 3 * i, dims)) continue;
            if (total < cap) memcpy(out + 3 * total, tmp + 3 * i, 3 * sizeof(int64_t));
            total++;
        }
        free(tmp);
        if (radius_start) radius_start[r + 1] = total;
    }
    return total;
}

/* ---------------------------------------------------------------- packed counts */

/* src/geometry/s2.rs:1035-1054 OccupancyBits::new: row (x,y) holds ceil(nz/64)+1 u64
 * words, bit z%64 of word z/64, trailing pad word zero. */
uint64_t orc_row_words(uint64_t nz) { return (nz + 63) / 64 + 1; }

void orc_pack(const uint8_t *occ, const uint64_t *dims, uint64_t *words) {
    uint64_t ny = dims[1], nz = dims[2], w = orc_row_words(nz);
    uint64_t rows = dims[0] * ny;
    memset(words, 0, rows * w * sizeof(uint64_t));
    for (uint64_t r = 0; r < rows; r++)
        for (uint64_t z = 0; z < nz; z++)
            if (occ[r * nz + z]) words[r * w + z / 64] |= 1ull << (z % 64);
}

/* src/geometry/s2.rs:1065-1088 shifted_and_count */
static uint64_t shifted_and_count(const uint64_t *a, const uint64_t *b, uint64_t shift, uint64_t len) {
    if (len == 0) return 0;
    uint64_t ws = shift / 64, bs = shift % 64, full = len / 64, hits = 0;
#define FETCH(i) (bs == 0 ? b[(i) + ws] : ((b[(i) + ws] >> bs) | (b[(i) + ws + 1] << (64 - bs))))
    for (uint64_t i = 0; i < full; i++) hits += (uint64_t)__builtin_popcountll(a[i] & FETCH(i));
    uint64_t tail = len % 64;
    if (tail) hits += (uint64_t)__builtin_popcountll(a[full] & FETCH(full) & ((1ull << tail) - 1));
#undef FETCH
    return hits;
}

/* src/geometry/s2.rs:1092-1120 direct_pair_counts_bits (in-domain offset required). */
void orc_pair_counts_bits(const uint64_t *words, const uint64_t *dims, const int64_t *off,
                          uint64_t *hits_out, uint64_t *valid_out) {
    uint64_t nx = dims[0], ny = dims[1], nz = dims[2], w = orc_row_words(nz);
    int64_t dx = off[0], dy = off[1], dz = off[2];
    uint64_t x0 = dx < 0 ? (uint64_t)(-dx) : 0, x1 = dx < 0 ? nx : nx - (uint64_t)dx;
    uint64_t y0 = dy < 0 ? (uint64_t)(-dy) : 0, y1 = dy < 0 ? ny : ny - (uint64_t)dy;
    uint64_t shift = (uint64_t)(dz < 0 ? -dz : dz), run = nz - shift, hits = 0;
    for (uint64_t x = x0; x < x1; x++) {
        uint64_t x2 = (uint64_t)((int64_t)x + dx);
        for (uint64_t y = y0; y < y1; y++) {
            uint64_t y2 = (uint64_t)((int64_t)y + dy);
            const uint64_t *first = words + (x * ny + y) * w, *second = words + (x2 * ny + y2) * w;
            hits += dz >= 0 ? shifted_and_count(first, second, shift, run)
                            : shifted_and_count(second, first, shift, run);
        }
    }
    *hits_out = hits;
    *valid_out = (x1 - x0) * (y1 - y0) * run;
}

/* src/geometry/s2.rs:1125-1154 direct_pair_counts (bool scan, the reference's own
 * test oracle for the packed kernel). */
void orc_pair_counts_bool(const uint8_t *occ, const uint64_t *dims, const int64_t *off,
                          uint64_t *hits_out, uint64_t *valid_out) {
    uint64_t nx = dims[0], ny = dims[1], nz = dims[2];
    int64_t dx = off[0], dy = off[1], dz = off[2];
    uint64_t x0 = dx < 0 ? (uint64_t)(-dx) : 0, x1 = dx < 0 ? nx : nx - (uint64_t)dx;
    uint64_t y0 = dy < 0 ? (uint64_t)(-dy) : 0, y1 = dy < 0 ? ny : ny - (uint64_t)dy;
    uint64_t z0 = dz < 0 ? (uint64_t)(-dz) : 0, z1 = dz < 0 ? nz : nz - (uint64_t)dz;
    uint64_t run = z1 - z0, hits = 0;
    for (uint64_t x = x0; x < x1; x++)
        for (uint64_t y = y0; y < y1; y++) {
            const uint8_t *a = occ + (x * ny + y) * nz + z0;
            const uint8_t *b = occ + ((uint64_t)((int64_t)x + dx) * ny + (uint64_t)((int64_t)y + dy)) * nz +
                               (uint64_t)((int64_t)z0 + dz);
            for (uint64_t z = 0; z < run; z++) hits += (a[z] && b[z]);
        }
    *hits_out = hits;
    *valid_out = (x1 - x0) * (y1 - y0) * run;
}

/* Per-offset hits for a list of offsets (the direct kernel's inner map,
 * src/geometry/s2.rs:1197-1201), OpenMP over offsets. */
void orc_hits_list(const uint64_t *words, const uint64_t *dims, const int64_t *offs, uint64_t n,
                   uint64_t *hits) {
#pragma omp parallel for schedule(dynamic, 16)
    for (uint64_t i = 0; i < n; i++) {
        uint64_t v;
        orc_pair_counts_bits(words, dims, offs + 3 * i, hits + i, &v);
    }
}

/* ---------------------------------------------------------------- curve */

/* src/geometry/s2.rs:393-501 fill_missing_s2_with_smooth_interpolation */
void orc_fill_missing(double *values, const uint8_t *has_support, uint64_t n, double vf) {
    if (n == 0) return;
    values[0] = vf;
    uint64_t *xk = malloc(n * sizeof(uint64_t));
    double *yk = malloc(n * sizeof(double));
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; i++)
        if (i == 0 || has_support[i]) { xk[m] = i; yk[m] = values[i]; m++; }
    if (m < 2) goto done;
    if (m == 2) {
        uint64_t xl = xk[0], xr = xk[1];
        if (xr <= xl + 1) goto done;
        double yl = yk[0], yr = yk[1], width = (double)(xr - xl);
        for (uint64_t i = xl + 1; i < xr; i++) {
            if (has_support[i]) continue;
            double t = (double)(i - xl) / width;
            double s = t * t * (3.0 - 2.0 * t);
            double v = yl + (yr - yl) * s;
            values[i] = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
        }
        goto done;
    }
    {
        double *xs = malloc(m * sizeof(double)), *h = malloc(m * sizeof(double));
        double *a = calloc(m, sizeof(double)), *b = calloc(m, sizeof(double));
        double *c = calloc(m, sizeof(double)), *d = calloc(m, sizeof(double));
        double *m2 = calloc(m, sizeof(double));
        for (uint64_t i = 0; i < m; i++) xs[i] = (double)xk[i];
        for (uint64_t i = 0; i + 1 < m; i++) { double hh = xs[i + 1] - xs[i]; h[i] = hh > 1e-12 ? hh : 1e-12; }
        b[0] = 1.0;
        b[m - 1] = 1.0;
        for (uint64_t i = 1; i + 1 < m; i++) {
            a[i] = h[i - 1];
            b[i] = 2.0 * (h[i - 1] + h[i]);
            c[i] = h[i];
            d[i] = 6.0 * ((yk[i + 1] - yk[i]) / h[i] - (yk[i] - yk[i - 1]) / h[i - 1]);
        }
        for (uint64_t i = 1; i < m; i++) {
            double den = fabs(b[i - 1]) < 1e-12 ? 1e-12 : b[i - 1];
            double w = a[i] / den;
            b[i] -= w * c[i - 1];
            d[i] -= w * d[i - 1];
        }
        double last = fabs(b[m - 1]) < 1e-12 ? 1e-12 : b[m - 1];
        m2[m - 1] = d[m - 1] / last;
        for (uint64_t ii = m - 1; ii-- > 0;) {
            double den = fabs(b[ii]) < 1e-12 ? 1e-12 : b[ii];
            m2[ii] = (d[ii] - c[ii] * m2[ii + 1]) / den;
        }
        for (uint64_t seg = 0; seg + 1 < m; seg++) {
            uint64_t xl = xk[seg], xr = xk[seg + 1];
            if (xr <= xl + 1) continue;
            double x0 = xs[seg], x1 = xs[seg + 1], y0 = yk[seg], y1 = yk[seg + 1];
            double hs = x1 - x0;
            if (!(hs > 1e-12)) hs = 1e-12;
            double m20 = m2[seg], m21 = m2[seg + 1];
            for (uint64_t i = xl + 1; i < xr; i++) {
                if (has_support[i]) continue;
                double x = (double)i;
                double ac = (x1 - x) / hs, bc = (x - x0) / hs;
                double y = ac * y0 + bc * y1 +
                           ((ac * ac * ac - ac) * m20 + (bc * bc * bc - bc) * m21) * (hs * hs / 6.0);
                values[i] = y < 0.0 ? 0.0 : (y > 1.0 ? 1.0 : y);
            }
        }
        free(xs); free(h); free(a); free(b); free(c); free(d); free(m2);
    }
done:
    free(xk);
    free(yk);
}

/* src/geometry/s2.rs:1187-1210 (per-radius ordered f64 fold) + 1157-1167
 * finish_exact_curve.  hits: per in-domain offset in shell order; valid is
 * analytic.  curve has r_max+1 entries. */
void orc_curve_from_hits(const uint64_t *dims, uint32_t r_max, const int64_t *offs,
                         const uint64_t *radius_start, const uint64_t *hits, double vf,
                         double *curve) {
    uint8_t *sup = calloc(r_max + 1, 1);
    curve[0] = vf;
    sup[0] = 1;
    for (uint32_t r = 1; r <= r_max; r++) {
        uint64_t s = radius_start[r], e = radius_start[r + 1];
        if (s == e) { curve[r] = 0.0; continue; }
        double sum = 0.0;
        for (uint64_t i = s; i < e; i++) {
            const int64_t *d = offs + 3 * i;
            uint64_t valid = 1;
            for (int a = 0; a < 3; a++) valid *= dims[a] - (uint64_t)(d[a] < 0 ? -d[a] : d[a]);
            sum = sum + (double)hits[i] / (double)valid;
        }
        curve[r] = sum / (double)(e - s);
        sup[r] = 1;
    }
    orc_fill_missing(curve, sup, r_max + 1, vf);
    curve[0] = vf;
    free(sup);
}

/* src/geometry/s2.rs:1482-1495 calculate_exact_with(Direct) on a packed grid:
 * occupied==0 gives zeros; vf = occupied/cells. */
void orc_exact_s2(const uint64_t *words, const uint64_t *dims, uint32_t r_max, double pitch,
                  uint64_t occupied, double *curve) {
    uint64_t cells = dims[0] * dims[1] * dims[2];
    if (occupied == 0) { memset(curve, 0, (r_max + 1) * sizeof(double)); return; }
    double vf = (double)occupied / (double)(cells ? cells : 1);
    if (pitch < 1e-9) pitch = 1e-9;
    uint64_t *rs = malloc((r_max + 2) * sizeof(uint64_t));
    uint64_t k = orc_domain_offsets(dims, r_max, pitch, NULL, 0, NULL);
    int64_t *offs = malloc((k ? k : 1) * 3 * sizeof(int64_t));
    orc_domain_offsets(dims, r_max, pitch, offs, k, rs);
    uint64_t *hits = malloc((k ? k : 1) * sizeof(uint64_t));
    orc_hits_list(words, dims, offs, k, hits);
    orc_curve_from_hits(dims, r_max, offs, rs, hits, vf, curve);
    free(rs); free(offs); free(hits);
}

/* ---------------------------------------------------------------- planner */

/* src/geometry/s2.rs:519-551 smooth_fft_length: smallest 2,3,5-smooth >= min_len. */
uint64_t orc_smooth_fft_length(uint64_t min_len) {
    uint64_t target = min_len < 1 ? 1 : min_len, best = 0;
    for (uint64_t p5 = 1;; ) {
        for (uint64_t p35 = p5;; ) {
            uint64_t v = p35;
            while (v < target) { if (v > UINT64_MAX / 2) break; v *= 2; }
            if (v >= target && (best == 0 || v < best)) best = v;
            if (p35 >= target || p35 > UINT64_MAX / 3) break;
            p35 *= 3;
        }
        if (p5 >= target || p5 > UINT64_MAX / 5) break;
        p5 *= 5;
    }
    return best;
}

/* src/geometry/s2.rs:554-562 padded_fft_dims: each axis to smooth(max(1, 2n-1)). */
void orc_padded_fft_dims(const uint64_t *dims, uint64_t *out) {
    for (int a = 0; a < 3; a++) {
        uint64_t need = dims[a] * 2;
        need = need >= 1 ? need - 1 : 0;
        out[a] = orc_smooth_fft_length(need < 1 ? 1 : need);
    }
}

/* src/geometry/s2.rs:623-692 exact_shell_work: in-domain offsets K, pair work W and
 * the largest per-radius count, by enumerating every shell (the reference's
 * octant enumeration is tested equal to this enumeration, 1899-1917). */
void orc_exact_shell_work(const uint64_t *dims, uint32_t r_max, double pitch, uint64_t *k_out,
                          double *w_out, uint64_t *maxk_out) {
    uint64_t K = 0, maxk = 0;
    __uint128_t W = 0;
    for (uint32_t r = 1; r <= r_max; r++) {
        double dist = (double)r / pitch, hw = 0.5 / pitch;
        uint64_t n = orc_shell_offsets(dist, hw, NULL, 0);
        int64_t *tmp = malloc((n ? n : 1) * 3 * sizeof(int64_t));
        orc_shell_offsets(dist, hw, tmp, n);
        uint64_t shell = 0;
        for (uint64_t i = 0; i < n; i++) {
            if (!in_domain(tmp + 3 * i, dims)) continue;
            shell++;
            __uint128_t w = 1;
            for (int a = 0; a < 3; a++) w *= dims[a] - (uint64_t)(tmp[3 * i + a] < 0 ? -tmp[3 * i + a] : tmp[3 * i + a]);
            W += w;
        }
        free(tmp);
        K += shell;
        if (shell > maxk) maxk = shell;
    }
    *k_out = K;
    *w_out = (double)W;
    *maxk_out = maxk;
}

/* src/geometry/s2.rs:695-714 fft_working_set_bytes (no overflow handling needed
 * for the sizes this is called with). */
double orc_fft_working_set_bytes(const uint64_t *dims, const uint64_t *padded, uint32_t r_max,
                                 uint64_t workers) {
    double cells = (double)dims[0] * dims[1] * dims[2];
    double pc = (double)padded[0] * padded[1] * padded[2];
    uint64_t w = workers < 1 ? 1 : workers;
    uint64_t t = (uint64_t)ceil(pc / 65536.0);
    uint64_t tasks = w < t ? w : t;
    if (tasks < 1) tasks = 1;
    double cx = 16.0, grid = pc * cx, band = 0.0;
    if (tasks > 1) {
        double plane = (double)padded[0] * padded[2] * cx;
        if (plane < 1) plane = 1;
        double planes = floor(8.0 * 1024 * 1024 / plane);
        double py = (double)(padded[1] < 1 ? 1 : padded[1]);
        if (planes < 1) planes = 1;
        if (planes > py) planes = py;
        band = planes * plane;
    }
    uint64_t mx = padded[0];
    if (padded[1] > mx) mx = padded[1];
    if (padded[2] > mx) mx = padded[2];
    double scratch = (double)w * 4.0 * mx * cx, plans = 12.0 * mx * cx;
    double output = ((double)r_max + 1) * 32.0;
    return cells + grid + band + scratch + plans + output;
}

/* src/geometry/s2.rs:717-725 direct_working_set_bytes */
double orc_direct_working_set_bytes(const uint64_t *dims, uint32_t r_max, uint64_t max_shell,
                                    uint64_t workers) {
    double cells = (double)dims[0] * dims[1] * dims[2];
    uint64_t w = workers < 1 ? 1 : workers, rm = r_max < 1 ? 1 : r_max;
    double concurrent = (double)(w < rm ? w : rm);
    double per_offset = 24.0 + 16.0;
    double packed = (double)dims[0] * dims[1] * (double)((dims[2] + 63) / 64 + 1) * 8.0;
    return cells + packed + (double)max_shell * per_offset * concurrent + ((double)r_max + 1) * 32.0;
}

/* src/geometry/s2.rs:733-766 plan_exact_cpu (constants 512-516).  method: 0=fft,
 * 1=direct.  out[0..4] = fft_seconds, direct_seconds, fft_bytes, direct_bytes,
 * pair_work; returns the method and the K offsets in *k_out. */
int orc_plan_exact_cpu(const uint64_t *dims, uint32_t r_max, double pitch, uint64_t workers,
                       double budget_bytes, double *out, uint64_t *k_out) {
    const double NS_FFT = 7.0, FFT_EFF = 0.16, NS_DIRECT = 0.115, DIRECT_EFF = 0.3;
    uint64_t w = workers < 1 ? 1 : workers, padded[3], K, maxk;
    double W;
    orc_padded_fft_dims(dims, padded);
    orc_exact_shell_work(dims, r_max, pitch, &K, &W, &maxk);
    double fb = orc_fft_working_set_bytes(dims, padded, r_max, w);
    double db = orc_direct_working_set_bytes(dims, r_max, maxk, w);
    double pc = (double)padded[0] * padded[1] * padded[2];
    double fs = NS_FFT * pc * log2(pc < 2 ? 2 : pc) / (1.0 + FFT_EFF * ((double)w - 1.0)) * 1e-9;
    double dt = (double)w < (double)(K < 1 ? 1 : K) ? (double)w : (double)(K < 1 ? 1 : K);
    double ds = NS_DIRECT * W / (1.0 + DIRECT_EFF * (dt - 1.0)) * 1e-9;
    int ffit = fb <= budget_bytes, dfit = db <= budget_bytes, method;
    if (ffit && dfit) method = fs <= ds ? 0 : 1;
    else if (ffit) method = 0;
    else method = 1;
    out[0] = fs; out[1] = ds; out[2] = fb; out[3] = db; out[4] = W;
    *k_out = K;
    return method;
}

/* ---------------------------------------------------------------- FFT (f64) */

/* Mixed-radix (2,3,5) complex FFT, recursive decimation in time.  Used for the
 * FFT exact kernel (src/geometry/s2.rs:965-1012, 1219-1261), whose results are
 * rounded to integers, and as the FP64 reference for the FP32 FFT path. */
static void fft_rec(double complex *x, uint64_t n, uint64_t stride, double complex *out,
                    double complex *tw, uint64_t tw_stride) {
    if (n == 1) { out[0] = x[0]; return; }
    uint64_t p = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : 5;
    uint64_t m = n / p;
    for (uint64_t j = 0; j < p; j++) fft_rec(x + j * stride, m, stride * p, out + j * m, tw, tw_stride * p);
    double complex tmp[5];
    for (uint64_t k = 0; k < m; k++) {
        for (uint64_t j = 0; j < p; j++) tmp[j] = out[j * m + k] * tw[(j * k * tw_stride)];
        for (uint64_t q = 0; q < p; q++) {
            double complex acc = 0;
            for (uint64_t j = 0; j < p; j++) acc += tmp[j] * tw[((j * q * m) % n) * tw_stride];
            out[q * m + k] = acc;
        }
    }
}

static void fft_line(double complex *line, uint64_t n, double complex *tw, double complex *buf,
                     double complex *buf2) {
    memcpy(buf2, line, n * sizeof(double complex));
    fft_rec(buf2, n, 1, buf, tw, 1);
    memcpy(line, buf, n * sizeof(double complex));
}

static double complex *twiddles(uint64_t n, int inverse) {
    double complex *tw = malloc(n * sizeof(double complex));
    for (uint64_t k = 0; k < n; k++) {
        double ang = (inverse ? 2.0 : -2.0) * M_PI * (double)k / (double)n;
        tw[k] = cos(ang) + I * sin(ang);
    }
    return tw;
}

static void fft3d(double complex *g, const uint64_t *f, int inverse) {
    uint64_t fx = f[0], fy = f[1], fz = f[2], mx = fx;
    if (fy > mx) mx = fy;
    if (fz > mx) mx = fz;
    double complex *tws[3] = {twiddles(fx, inverse), twiddles(fy, inverse), twiddles(fz, inverse)};
#pragma omp parallel
    {
        double complex *line = malloc(mx * sizeof(double complex));
        double complex *b1 = malloc(mx * sizeof(double complex)), *b2 = malloc(mx * sizeof(double complex));
#pragma omp for
        for (uint64_t r = 0; r < fx * fy; r++) fft_line(g + r * fz, fz, tws[2], b1, b2);
#pragma omp for
        for (uint64_t xz = 0; xz < fx * fz; xz++) {
            uint64_t x = xz / fz, z = xz % fz;
            for (uint64_t y = 0; y < fy; y++) line[y] = g[(x * fy + y) * fz + z];
            fft_line(line, fy, tws[1], b1, b2);
            for (uint64_t y = 0; y < fy; y++) g[(x * fy + y) * fz + z] = line[y];
        }
#pragma omp for
        for (uint64_t yz = 0; yz < fy * fz; yz++) {
            for (uint64_t x = 0; x < fx; x++) line[x] = g[x * fy * fz + yz];
            fft_line(line, fx, tws[0], b1, b2);
            for (uint64_t x = 0; x < fx; x++) g[x * fy * fz + yz] = line[x];
        }
        free(line); free(b1); free(b2);
    }
    if (inverse) {
        double norm = (double)(fx * fy * fz);
#pragma omp parallel for
        for (uint64_t i = 0; i < fx * fy * fz; i++) g[i] /= norm;
    }
    for (int a = 0; a < 3; a++) free(tws[a]);
}

/* Autocorrelation on the padded grid (src/geometry/s2.rs:965-1012): corr has
 * padded[0]*padded[1]*padded[2] doubles (real part after the inverse). */
void orc_fft_autocorr(const uint8_t *occ, const uint64_t *dims, const uint64_t *padded, double *corr) {
    uint64_t fx = padded[0], fy = padded[1], fz = padded[2], n = fx * fy * fz;
    double complex *g = calloc(n, sizeof(double complex));
    for (uint64_t x = 0; x < dims[0]; x++)
        for (uint64_t y = 0; y < dims[1]; y++)
            for (uint64_t z = 0; z < dims[2]; z++)
                g[(x * fy + y) * fz + z] = occ[(x * dims[1] + y) * dims[2] + z] ? 1.0 : 0.0;
    fft3d(g, padded, 0);
#pragma omp parallel for
    for (uint64_t i = 0; i < n; i++) g[i] = g[i] * conj(g[i]);
    fft3d(g, padded, 1);
    for (uint64_t i = 0; i < n; i++) corr[i] = creal(g[i]);
    free(g);
}

/* src/geometry/s2.rs:1219-1261 calculate_s2_exact_fft hits: correlation at the
 * wrapped offset, rounded and clamped at zero. */
void orc_fft_hits(const double *corr, const uint64_t *padded, const int64_t *offs, uint64_t n,
                  uint64_t *hits) {
    for (uint64_t i = 0; i < n; i++) {
        uint64_t idx[3];
        for (int a = 0; a < 3; a++) {
            int64_t d = offs[3 * i + a];
            idx[a] = d >= 0 ? (uint64_t)d : padded[a] - (uint64_t)(-d);
        }
        double v = round(corr[(idx[0] * padded[1] + idx[1]) * padded[2] + idx[2]]);
        hits[i] = v > 0.0 ? (uint64_t)v : 0;
    }
}

/* ---------------------------------------------------------------- Philox + voxel MC */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
 * 1, 2, 3", SC11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl increments
 * 0x9E3779B9 / 0xBB67AE85, ten rounds. */
void orc_philox4x32_10(const uint32_t *ctr_in, const uint32_t *key_in, uint32_t *out) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Voxel Monte Carlo estimator of src/geometry/s2.rs:1412-1464 (uniform voxel,
 * uniform offset of the full -- not domain-filtered -- shell, out-of-grid partner
 * rejected, hits over valid), with this engine's Philox stream: sample s of
 * radius r draws philox(ctr={s_lo, s_hi, r, 0x53324D43}, key={seed_lo, seed_hi});
 * words 0..2 pick x,y,z and word 3 the shell offset, each as (u32 * n) >> 32.
 * hv: 2*(r_max+1) u64, (hits, valid) per radius, radius 0 left zero. */
void orc_mc_counts(const uint64_t *words, const uint64_t *dims, uint32_t r_max, double pitch,
                   uint64_t samples, uint64_t seed, uint64_t *hv) {
    uint64_t nx = dims[0], ny = dims[1], nz = dims[2], w = orc_row_words(nz);
    memset(hv, 0, 2 * (r_max + 1) * sizeof(uint64_t));
    if (pitch < 1e-9) pitch = 1e-9;
    for (uint32_t r = 1; r <= r_max; r++) {
        double dist = (double)r / pitch, hw = 0.5 / pitch;
        uint64_t n = orc_shell_offsets(dist, hw, NULL, 0);
        int64_t *sh = malloc((n ? n : 1) * 3 * sizeof(int64_t));
        orc_shell_offsets(dist, hw, sh, n);
        uint64_t hits = 0, valid = 0;
#pragma omp parallel for reduction(+ : hits, valid) schedule(static)
        for (uint64_t s = 0; s < samples; s++) {
            uint32_t ctr[4] = {(uint32_t)s, (uint32_t)(s >> 32), r, 0x53324D43u};
            uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
            orc_philox4x32_10(ctr, key, o);
            uint64_t x = ((uint64_t)o[0] * nx) >> 32, y = ((uint64_t)o[1] * ny) >> 32;
            uint64_t z = ((uint64_t)o[2] * nz) >> 32, k = ((uint64_t)o[3] * n) >> 32;
            int64_t x2 = (int64_t)x + sh[3 * k], y2 = (int64_t)y + sh[3 * k + 1], z2 = (int64_t)z + sh[3 * k + 2];
            if (x2 < 0 || y2 < 0 || z2 < 0 || x2 >= (int64_t)nx || y2 >= (int64_t)ny || z2 >= (int64_t)nz) continue;
            valid++;
            uint64_t a = (words[(x * ny + y) * w + z / 64] >> (z % 64)) & 1;
            uint64_t b = (words[((uint64_t)x2 * ny + (uint64_t)y2) * w + (uint64_t)z2 / 64] >> ((uint64_t)z2 % 64)) & 1;
            hits += a & b;
        }
        hv[2 * r] = hits;
        hv[2 * r + 1] = valid;
        free(sh);
    }
}

/* Continuous mesh Monte Carlo estimator of src/geometry/s2.rs:1273-1320 with this
 * engine's Philox stream: sample i of radius r draws
 * philox(ctr={i_lo, i_hi, r, 0x4D455348 + k}, key={seed_lo, seed_hi}) for call k.
 * Calls 0-1 give p (53-bit uniforms u = ((a >> 5) * 2^26 + (b >> 6)) / 2^53 from
 * word pairs (0,1), (2,3) of call 0 and (0,1) of call 1; coordinate
 * min + (max - min) u); direction attempt a uses calls 2+2a and 3+2a the same way
 * for x, y, z in [-1, 1), accepted when 1e-12 < |d|^2 <= 1 and normalized by
 * 1/sqrt(|d|^2); after 64 rejections the direction is +x.  q = p + r d; a q
 * outside the box is rejected, else valid += 1, and hits += 1 when both points
 * are inside the mesh: the whole-mesh contains_point of
 * src/geometry/mesh_query.rs:73-116 (bbox of all vertices +-1e-9, full scan).
 * max(samples, 200) samples per radius.  hv: 2*(r_max+1) u64. */
static double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

static inline double rmin(double a, double b) { return isnan(a) ? b : (isnan(b) ? a : (a < b ? a : b)); }
static inline double rmax(double a, double b) { return isnan(a) ? b : (isnan(b) ? a : (a > b ? a : b)); }

void orc_mesh_mc_counts(const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntri,
                        const double *bmin, const double *bmax, uint32_t r_max, uint64_t samples,
                        uint64_t seed, uint64_t *hv) {
    memset(hv, 0, 2 * (r_max + 1) * sizeof(uint64_t));
    const uint64_t spr = samples < 200 ? 200 : samples;
    bbox3 bb;
    bb.valid = nverts > 0;
    for (uint64_t v = 0; v < nverts; v++)
        for (int a = 0; a < 3; a++) {
            double c = xyz[3 * v + a];
            bb.mn[a] = v ? rmin(bb.mn[a], c) : c;
            bb.mx[a] = v ? rmax(bb.mx[a], c) : c;
        }
    double *tris = malloc((ntri ? ntri : 1) * 9 * sizeof(double));
    for (uint64_t f = 0; f < ntri; f++)
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) tris[9 * f + 3 * k + a] = xyz[3 * (uint64_t)tri[3 * f + k] + a];
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t r = 1; r <= r_max; r++) {
        uint64_t hits = 0, valid = 0;
        const double rr = (double)r;
#pragma omp parallel reduction(+ : hits, valid)
        {
            double *scratch = malloc((ntri ? ntri : 1) * sizeof(double));
#pragma omp for schedule(static)
            for (uint64_t i = 0; i < spr; i++) {
                uint32_t a[4], b[4];
                uint32_t c0[4] = {(uint32_t)i, (uint32_t)(i >> 32), r, 0x4D455348u};
                uint32_t c1[4] = {(uint32_t)i, (uint32_t)(i >> 32), r, 0x4D455348u + 1u};
                orc_philox4x32_10(c0, key, a);
                orc_philox4x32_10(c1, key, b);
                v3 p = {bmin[0] + (bmax[0] - bmin[0]) * u53(a[0], a[1]),
                        bmin[1] + (bmax[1] - bmin[1]) * u53(a[2], a[3]),
                        bmin[2] + (bmax[2] - bmin[2]) * u53(b[0], b[1])};
                v3 d = {1.0, 0.0, 0.0};
                for (uint32_t t = 0; t < 64; t++) {
                    uint32_t e0[4] = {(uint32_t)i, (uint32_t)(i >> 32), r, 0x4D455348u + 2u + 2u * t};
                    uint32_t e1[4] = {(uint32_t)i, (uint32_t)(i >> 32), r, 0x4D455348u + 3u + 2u * t};
                    orc_philox4x32_10(e0, key, a);
                    orc_philox4x32_10(e1, key, b);
                    double x = -1.0 + 2.0 * u53(a[0], a[1]), y = -1.0 + 2.0 * u53(a[2], a[3]);
                    double z = -1.0 + 2.0 * u53(b[0], b[1]);
                    double n2 = x * x + y * y + z * z;
                    if (n2 > 1e-12 && n2 <= 1.0) {
                        double inv = 1.0 / sqrt(n2);
                        d.x = x * inv;
                        d.y = y * inv;
                        d.z = z * inv;
                        break;
                    }
                }
                v3 q = {p.x + d.x * rr, p.y + d.y * rr, p.z + d.z * rr};
                if (q.x < bmin[0] || q.x > bmax[0] || q.y < bmin[1] || q.y > bmax[1] || q.z < bmin[2] ||
                    q.z > bmax[2])
                    continue;
                valid++;
                if (granule_contains(tris, ntri, &bb, p, scratch) && granule_contains(tris, ntri, &bb, q, scratch))
                    hits++;
            }
            free(scratch);
        }
        hv[2 * r] = hits;
        hv[2 * r + 1] = valid;
    }
    free(tris);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
