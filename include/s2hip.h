/*
 * s2hip.h -- C ABI of the MI355X (gfx950) S2 two-point-correlation engine.
 *
 * Drop-in boundary for RustMSPT's S2 path.  Each entry point names the reference
 * interface it replaces (file:line in the RustMSPT source tree).  All functions
 * are extern "C", take plain pointers and sizes, never throw, and report failure
 * through a status code plus a thread-local message (s2h_last_error), which the
 * caller maps to its own Result<_, String> and fallback policy exactly as the
 * reference callers do for the wgpu path (src/pipeline/measure.rs:225-246).
 *
 * Ownership: the context owns the device, streams, its device-memory accounting
 * and the per-(grid, r_max, pitch) shell plans.  A volume handle is owned by the
 * context until s2h_vol_free.  Host pointers are borrowed for the call only.
 * Output arrays are caller-allocated with explicit lengths; a length mismatch is
 * an error (S2H_ERR_SIZE), never a truncation.
 *
 * Threading: a context is not thread-safe (the reference serializes its GPU
 * objects behind a Mutex, src/pipeline/optimize_execution.rs:147); use one context
 * per host thread or lock around it.  The serve_* calls are meant for one owner
 * thread.
 *
 * Packed volume layout (shared with the reference's OccupancyBits,
 * src/geometry/s2.rs:1035-1054): row (x, y) is (ceil(nz/64)+1) little-endian
 * uint64 words at word index (x*ny + y)*row_words; voxel z is bit z%64 of word
 * z/64; the trailing pad word and every bit at z >= nz are zero.
 *
 * Offset order: "shell order" is the reference's curve summation order -- radii
 * 1..r_max, each radius' offsets in shell_offset_iter order (dx, then dy, then dz
 * ascending; src/geometry/s2.rs:364-385) -- keeping only in-domain offsets
 * (|d_i| < n_i, src/geometry/s2.rs:1028-1030).
 *
 * Environment (read by the library):
 *   S2HIP_AUTOTUNE=0            launch shapes and FFT tiles from the cost model only
 *                               (no device timing on a plan's first run)
 *   S2HIP_BATCH_DELTA=direct|fft  force the batched candidates' delta kernel
 *                               (default: chosen per seed by a cost model)
 *   S2HIP_FAIL_AT=alloc|launch|readback[:n]  the n-th (default first) pass through
 *                               that site after the variable last changed returns
 *                               S2H_ERR_HIP before any device work there, to exercise
 *                               callers' error paths (no kernel is affected)
 */
#ifndef S2HIP_H
#define S2HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct s2h_ctx s2h_ctx;
typedef struct s2h_vol s2h_vol;
typedef int32_t s2h_status;

#define S2H_OK 0
#define S2H_ERR_ARG 1         /* invalid argument (message says which) */
#define S2H_ERR_HIP 2         /* HIP runtime error */
#define S2H_ERR_BUDGET 3      /* working set exceeds the context's device budget */
#define S2H_ERR_SIZE 4        /* caller buffer length does not match */
#define S2H_ERR_UNSUPPORTED 5 /* shape outside what the kernels handle */
#define S2H_ERR_IO 6          /* file read failure */

/* ------------------------------------------------------------------ context */

/* Replaces the shared wgpu device (src/gpu/context.rs:115-119, shared_device) and
 * the budget of AccelerationConfig.gpu_memory_limit_mb
 * (src/config/acceleration.rs:15).  hbm_budget_bytes = 0 means no budget. */
s2h_status s2h_create(int device, uint64_t hbm_budget_bytes, s2h_ctx **out);
void s2h_destroy(s2h_ctx *ctx);
/* Thread-local message for the last failing call on this thread. */
const char *s2h_last_error(void);
/* Device name for "Compute Backend: exact=hip/<name>" (src/compute/backend.rs:89-93). */
s2h_status s2h_device_name(s2h_ctx *ctx, char *buf, uint64_t len);
/* Device bytes currently held by the context (volumes, plans, scratch). */
s2h_status s2h_bytes_in_use(s2h_ctx *ctx, uint64_t *bytes);

/* Kernel timing on the context's stream: when enabled, each exact / FFT / MC call
 * records HIP events around its phases; s2h_profile_read returns, in ms, for the
 * last call: [0] count kernels, [1] gather + fold, [2] whole device work, and
 * [3] the number of kernel launches in [0].  No counterpart in the reference
 * (it has no GPU kernel timestamps, AGENTS.md:638); used by the benchmark. */
s2h_status s2h_profile(s2h_ctx *ctx, int enable);
s2h_status s2h_profile_read(s2h_ctx *ctx, double *ms, uint64_t n);
/* Measured integer-VALU issue rate of the device, in 32-bit lane-operations per
 * second, for the exact counting loop's instruction pair (v_and_b32 with a
 * wave-uniform operand + v_bcnt_u32_b32 accumulate) with every SIMD occupied:
 * the roofline peak of the exact count kernel.  No counterpart in the reference. */
s2h_status s2h_valu_probe(s2h_ctx *ctx, double *lane_ops_per_s);

/* ------------------------------------------------------------------ volumes */

/* Upload a packed volume (layout above); the pad word and bits z >= nz must be 0. */
s2h_status s2h_vol_from_packed(s2h_ctx *ctx, const uint64_t *words, uint64_t nwords,
                               const uint64_t dims[3], s2h_vol **out);
/* Upload a bool grid (x-major, z fastest, one byte per voxel) and pack it on the
 * device (replaces OccupancyBits::new, src/geometry/s2.rs:1042-1054). */
s2h_status s2h_vol_from_bytes(s2h_ctx *ctx, const uint8_t *occ, const uint64_t dims[3],
                              s2h_vol **out);
s2h_status s2h_vol_read_packed(s2h_ctx *ctx, const s2h_vol *vol, uint64_t *words, uint64_t nwords);
s2h_status s2h_vol_dims(const s2h_vol *vol, uint64_t dims[3]);
/* Occupied voxel count (the numerator of vf, src/geometry/s2.rs:1486-1491). */
s2h_status s2h_vol_occupied(s2h_ctx *ctx, s2h_vol *vol, uint64_t *occupied);
void s2h_vol_free(s2h_ctx *ctx, s2h_vol *vol);

/* ------------------------------------------------------------ volume ingest */

/* CT-style sample volumes: a folder of RAW slice files or a TIFF stack.
 * Samples are z-major as Volume3D (idx = z*width*height + y*width + x,
 * src/io/volume.rs:45-53).  The occupancy grid maps file axes one to one:
 * S2 (x, y, z) = file (x, y, slice), so dims = {width, height, depth} and
 * each S2 row (x, y) runs along the slices. */
#define S2H_VOLUME_RAW 0
#define S2H_VOLUME_TIFF 1
typedef struct {
    const char *path;       /* RAW: folder of slice files; TIFF: a multi-page file or a folder of TIFF files */
    int32_t kind;           /* S2H_VOLUME_RAW or S2H_VOLUME_TIFF */
    int32_t reserved;
    int64_t slice_start;    /* inclusive; -1 = first (resolve_slice_range, src/io/volume.rs:258-285) */
    int64_t slice_end;      /* inclusive; -1 = last */
    uint64_t width, height; /* RAW only (RawFolderSpec, src/io/volume.rs:193-203) */
    uint32_t bits;          /* RAW only: 8, 16 or 32 */
    int32_t is_signed;      /* RAW only */
    int32_t big_endian;     /* RAW only */
} s2h_volume_source;
typedef struct {
    uint64_t width, height, depth;
    uint32_t bits;       /* 8, 16 or 32 */
    int32_t is_signed;
    int32_t big_endian;  /* byte order of the staged samples */
    int32_t reserved;
} s2h_volume_info;

/* Which samples are occupied. */
#define S2H_PHASE_RANGE 0           /* lo <= v <= hi (a threshold window) */
#define S2H_PHASE_FOREGROUND 1      /* v != background */
#define S2H_PHASE_FOREGROUND_AUTO 2 /* v != the modal boundary value, smallest on ties
                                       (detect_background_mode, src/pipeline/crop.rs:338-373) */
typedef struct {
    int32_t mode;
    int32_t reserved;
    int64_t lo, hi, background;
} s2h_phase;

/* Host only (no device): the volume's dimensions and staged layout after the
 * slice range (load_raw_folder_typed, src/io/volume.rs:344-403;
 * load_tiff_or_folder_typed_with_range, src/io/volume.rs:521-581). */
s2h_status s2h_volume_probe(const s2h_volume_source *src, s2h_volume_info *info);
/* Host only: decode every slice into samples (width*height*depth*bits/8 bytes,
 * in the staged layout). */
s2h_status s2h_volume_read(const s2h_volume_source *src, void *samples, uint64_t nbytes, s2h_volume_info *info);
/* Classify and bit-pack a z-major sample buffer on the device.  background_used
 * (optional) receives the background of the FOREGROUND modes; occupied
 * (optional) the occupied voxel count. */
s2h_status s2h_vol_from_samples(s2h_ctx *ctx, const void *samples, uint64_t nbytes, const s2h_volume_info *layout,
                                const s2h_phase *phase, s2h_vol **out, int64_t *background_used,
                                uint64_t *occupied);
/* Read a RAW folder or TIFF stack, streaming its slices through pinned staging
 * to the device, then classify and bit-pack there. */
s2h_status s2h_vol_load(s2h_ctx *ctx, const s2h_volume_source *src, const s2h_phase *phase, s2h_vol **out,
                        s2h_volume_info *info, int64_t *background_used, uint64_t *occupied);

/* --------------------------------------------------------------- voxelization */

/* Replaces build_bbox_occupancy / VoxelS2::new (src/geometry/s2.rs:160-201, 1475)
 * and GpuVoxelPipeline::voxelize (src/gpu/voxel.rs:285).  The host passes the
 * welded mesh (vertices as f64 triples, src/io/stl.rs:55-80) and the granule id of
 * every triangle (split_mesh_into_granules, src/geometry/mesh_ops.rs:48); the
 * device evaluates the reference's f64 parity predicate per granule at every
 * voxel centre of the granule's range and ORs granules, bit-exactly.  pitch must
 * be > 0 (the caller applies the reference's max(1e-9) / fallback-to-1.0 rules). */
s2h_status s2h_voxelize(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri,
                        uint64_t ntris, const uint32_t *granule_of_tri, uint32_t ngranules,
                        const double bbox_min[3], double pitch, const uint64_t dims[3],
                        s2h_vol **out, uint64_t *occupied);

/* ------------------------------------------------------- device-resident meshes */

/* A welded, granule-split mesh kept on the device (vertices f64, faces u32,
 * granule id per face), so that voxelization needs no host preparation. */
typedef struct s2h_mesh s2h_mesh;
/* Replaces load_stl / load_stl_or_merge_folder (src/io/stl.rs:214-239, 282-302)
 * followed by split_mesh_into_granules (src/geometry/mesh_ops.rs:48): one STL
 * byte stream per file (ASCII sniffed and parsed on the host, binary decoded on
 * the device), welded on the device with the reference's 1e-6 key, the first
 * occurrence keeping its index, never across files; files concatenate in the
 * order given.  The vertices, faces and granule ids equal s2h_stl_load +
 * s2h_mesh_granules on the same bytes. */
s2h_status s2h_mesh_from_stl(s2h_ctx *ctx, const void *const *files, const uint64_t *sizes, uint32_t nfiles,
                             s2h_mesh **out);
/* As s2h_mesh_from_stl, reading the files (in the order given). */
s2h_status s2h_mesh_load_stl(s2h_ctx *ctx, const char *const *paths, uint32_t npaths, s2h_mesh **out);
/* An already welded mesh; granules are split on the device. */
s2h_status s2h_mesh_from_arrays(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri,
                                uint64_t ntris, s2h_mesh **out);
/* Sizes and the vertex bbox (mesh_bbox, src/geometry/bbox.rs:8-25: f64::min/max,
 * NaN ignored) of a resident mesh. */
s2h_status s2h_mesh_info(const s2h_mesh *mesh, uint64_t *nverts, uint64_t *ntris, uint32_t *ngranules,
                         double bbox_min[3], double bbox_max[3]);
/* Copy the mesh back: xyz 3*nverts, tri 3*ntris, gid ntris (any may be null). */
s2h_status s2h_mesh_read(s2h_ctx *ctx, const s2h_mesh *mesh, double *xyz, uint64_t nxyz, uint32_t *tri,
                         uint64_t ntri, uint32_t *gid, uint64_t ngid);
/* Overwrite vertices [vstart, vstart+n) of a resident mesh (xyz: 3n doubles):
 * a rigid particle move in the optimize loop, whose topology and granules do
 * not change.  Replaces GpuS2Pipeline::update_mesh's partial upload of changed
 * faces (src/gpu/s2.rs:480-560); the bbox is recomputed. */
s2h_status s2h_mesh_set_vertices(s2h_ctx *ctx, s2h_mesh *mesh, uint64_t vstart, const double *xyz, uint64_t n);
/* s2h_voxelize on the resident mesh (same result bits). */
s2h_status s2h_mesh_voxelize(s2h_ctx *ctx, const s2h_mesh *mesh, const double bbox_min[3], double pitch,
                             const uint64_t dims[3], s2h_vol **out, uint64_t *occupied);
void s2h_mesh_free(s2h_ctx *ctx, s2h_mesh *mesh);

/* ------------------------------------------------------------------ shells */

/* Number of in-domain offsets of radii 1..r_max in shell order (sizes the outputs
 * of s2h_exact_counts / s2h_fft32_counts; exact_shell_work's K,
 * src/geometry/s2.rs:623-692). */
s2h_status s2h_shell_count(const uint64_t dims[3], uint32_t r_max, double pitch, uint64_t *n_offsets);
/* The offsets themselves (3 int64 each) and radius_start[r_max+2] (offsets of
 * radius r are [radius_start[r], radius_start[r+1])). */
s2h_status s2h_shell_offsets(const uint64_t dims[3], uint32_t r_max, double pitch, int64_t *offs,
                             uint64_t n_offsets, uint64_t *radius_start, uint64_t n_radius_start);

/* The reference's ordered f64 fold, sum = sum + t_k left to right
 * (src/geometry/s2.rs:1187-1201), of each segment [starts[g], starts[g+1]) of
 * terms, computed on the device by the same code as the exact curve's fold
 * (64 terms per step on the exact binade grid; bit-identical to the sequential
 * sum).  starts has nseg+1 entries. */
s2h_status s2h_ordered_sum(s2h_ctx *ctx, const double *terms, uint64_t n, const uint64_t *starts, uint32_t nseg,
                           double *sums);

/* ------------------------------------------------------------------ exact */

/* Exact pair counts ("hits") of every in-domain offset in shell order (replaces
 * direct_pair_counts_bits over the shells, src/geometry/s2.rs:1092-1120,
 * 1187-1201, and GpuShellS2Pipeline::compute_shell_input, src/gpu/s2_shell.rs:459).
 * valid is analytic (prod n_i-|d_i|) and is not returned. */
s2h_status s2h_exact_counts(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch,
                            uint64_t *hits, uint64_t n_offsets);
/* Finished exact S2 curve r = 0..r_max (replaces VoxelS2::calculate_exact_with,
 * src/geometry/s2.rs:1482-1495, and try_calculate_s2_gpu_exact's result,
 * src/geometry/s2.rs:1616-1678): per-radius ordered f64 fold of hits/valid in
 * shell order, smooth fill of unsupported radii, S2(0)=vf; all-empty -> zeros.
 * Bit-identical to the reference CPU exact path. */
s2h_status s2h_exact_s2(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, double *s2,
                        uint64_t n);
/* Which exact kernel s2h_exact_counts / s2h_exact_s2 run (replaces the method
 * choice of plan_exact_cpu and calculate_exact_with's forced method,
 * src/geometry/s2.rs:733-766, 1482-1495).  Both kernels give the same integer
 * counts, so the choice never changes a result: 1 direct counting; 2 the FP64
 * FFT correlation rounded to the nearest integer, `.re.round().max(0.0)` of
 * calculate_s2_exact_fft (src/geometry/s2.rs:1219-1261), allowed only when the
 * worst-case round-off bound for the volume, (3 rho + 3u) K^(3/2) with K
 * occupied voxels, is below 1/4 (else S2H_ERR_UNSUPPORTED); 0 (default) the
 * faster modelled kernel among the eligible ones.  s2h_exact_tune_info's
 * info[7] reports the kernel of the plan's last evaluation (1 or 2). */
s2h_status s2h_exact_method_set(s2h_ctx *ctx, int32_t method);

/* Shape of the exact count launch for (dims, r_max, pitch) -- no counterpart in
 * the reference, whose direct kernel has no launch shape.  The half-ball columns
 * are grouped 64 to a wave in bands of `band` consecutive dz extents, and the
 * launch is split into about `waves` waves.  The first exact call on a new shape
 * ranks every candidate with a cost model and, when a run is short, times the
 * best few on the device and keeps the fastest (S2HIP_AUTOTUNE=0: model only).
 * Counts never depend on the shape.
 * info (8 values): [0] band, [1] target waves, [2] 1 picked by device timing,
 * -1 by the cost model, 0 not run yet, [3] candidates timed, [4] work items,
 * [5] column groups, [6] modelled VALU wave-instructions per run, [7] kernel of
 * the last evaluation (1 direct, 2 FP64 FFT, 0 none yet). */
s2h_status s2h_exact_tune_info(s2h_ctx *ctx, const uint64_t dims[3], uint32_t r_max, double pitch,
                               int64_t *info, uint64_t n);
/* Force a shape (band >= 1, waves >= 1) for this context; the autotuner is
 * skipped for it. */
s2h_status s2h_exact_tune_set(s2h_ctx *ctx, const uint64_t dims[3], uint32_t r_max, double pitch,
                              int32_t band, int32_t waves);
/* enable != 0: a new exact plan's first run times the best-modelled shapes
 * (default); 0: the cost model's pick is used as is (for callers that see many
 * distinct shapes once each, such as the serving loop).  A shape timed once is
 * reused process-wide by every context. */
s2h_status s2h_set_autotune(s2h_ctx *ctx, int32_t enable);

/* ------------------------------------------------------------------ FFT (FP32) */

/* Padded transform lengths of the FFT kernels (FP64 exact and FP32) for this
 * shape -- no counterpart in the reference, which pads each axis to the
 * smallest 2,3,5-smooth length >= 2n - 1 (padded_fft_dims, src/geometry/s2.rs:554-562).
 * Only offsets |d_a| <= R_a (the largest of the shells) are read, and a length
 * f >= n + R leaves those exact, so each axis takes the 2,3,5-smooth length in
 * [n + R, the reference's] with the least modelled work.  Host only. */
s2h_status s2h_fft_padded_dims(const uint64_t dims[3], uint32_t r_max, double pitch, uint64_t padded[3]);

/* FP32 real-FFT autocorrelation counts of every in-domain offset in shell order
 * (the FFT exact kernel's correlation, src/geometry/s2.rs:965-1012, 1219-1261, in
 * FP32 and NOT rounded: values carry FP32 round-off). */
s2h_status s2h_fft32_counts(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch,
                            float *counts, uint64_t n_offsets);
/* S2 curve from the FP32 correlation (same fold and fill as s2h_exact_s2). */
s2h_status s2h_fft32_s2(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, double *s2,
                        uint64_t n);

/* ------------------------------------------------------------------ Monte Carlo */

/* Voxel Monte Carlo counts (estimator of voxel_mc_radii, src/geometry/s2.rs:1412-1464)
 * with a Philox4x32-10 counter-based stream keyed by seed: hits_valid has
 * 2*(r_max+1) entries, (hits, valid) per radius, radius 0 zero. */
s2h_status s2h_mc_counts(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch,
                         uint64_t samples, uint64_t seed, uint64_t *hits_valid, uint64_t n);
/* s2h_mc_counts without any device allocation -- no counterpart in the reference; for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and the same counts bit for bit.  Everything the kernels need beyond
 * the caller's volume (the shell table, built on the device, the radius starts and the counters)
 * lives in one 32 MiB file-scope device array per GPU, held until the process ends.  The table
 * holds about 4.19 million offsets: r_max / pitch up to about 96 voxels; larger shells (and
 * r_max > 4096) are refused with S2H_ERR_BUDGET before any launch.  Calls on one GPU share the
 * array and keep the last call's contents: calls that overlap on one GPU race. */
s2h_status s2h_mc_counts_noalloc(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch,
                                 uint64_t samples, uint64_t seed, uint64_t *hits_valid, uint64_t n);
/* Finished MC curve (VoxelS2::calculate_seeded "monte_carlo", src/geometry/s2.rs:1521-1537):
 * hits/valid per radius, unsupported radii filled, S2(0)=vf. samples is per radius
 * and is raised to 200 as in the reference. */
s2h_status s2h_mc_s2(s2h_ctx *ctx, s2h_vol *vol, uint32_t r_max, double pitch, uint64_t samples,
                     uint64_t seed, double *s2, uint64_t n);

/* Continuous mesh Monte Carlo (calculate_s2_mesh_mc_seeded, src/geometry/s2.rs:1273-1320;
 * the path the reference takes for a non-exact method with voxel_pitch <= 0,
 * src/geometry/s2.rs:1351-1356): per radius, max(samples, 200) point pairs p, p + r d
 * with p uniform in the box [bbox_min, bbox_max] and d uniform on the sphere; pairs
 * leaving the box are rejected; hits count pairs with both points inside the mesh
 * by its whole-mesh parity query (PreparedMeshQuery::contains_point,
 * src/geometry/mesh_query.rs:73-116), bit-exactly.  Philox4x32-10 stream keyed by
 * seed (the reference's ChaCha12 streams are not reproduced; the estimand is).
 * hits_valid: 2*(r_max+1) values, (hits, valid) per radius, radius 0 zero. */
s2h_status s2h_mesh_mc_counts(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri,
                              uint64_t ntris, const double bbox_min[3], const double bbox_max[3],
                              uint32_t r_max, uint64_t samples, uint64_t seed, uint64_t *hits_valid,
                              uint64_t n);
/* s2h_mesh_mc_counts on a resident mesh (GpuS2Pipeline::calculate_s2_gpu_seeded,
 * src/gpu/s2.rs:640-670, after update_mesh): the same counts as the host-array
 * entry on the same vertices and faces. */
s2h_status s2h_mesh_mc_counts_resident(s2h_ctx *ctx, const s2h_mesh *mesh, const double bbox_min[3],
                                       const double bbox_max[3], uint32_t r_max, uint64_t samples, uint64_t seed,
                                       uint64_t *hits_valid, uint64_t n);
/* The curve: s2[0] = vf (the caller's geometric volume fraction in the box,
 * volume_fraction_in_bbox, src/geometry/volume.rs:437-469), s2[r] = hits/valid
 * (0 when no pair was valid), no fill -- as the reference. */
s2h_status s2h_mesh_mc_s2(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri,
                          uint64_t ntris, const double bbox_min[3], const double bbox_max[3],
                          uint32_t r_max, uint64_t samples, uint64_t seed, double vf, double *s2,
                          uint64_t n);

/* ------------------------------------------------------------ batched candidates */

/* A resident seed packing for batched simulated-annealing evaluation: packed
 * occupancy, per-voxel granule coverage counts (VoxelCoverage,
 * src/geometry/s2.rs:266-296), the seed's exact counts, and every particle's
 * geometry.  Arguments as for s2h_voxelize, plus the particles -- particle p owns
 * vertices [particle_vert_start[p], particle_vert_start[p+1]) and triangles
 * [particle_tri_start[p], particle_tri_start[p+1]), and its triangles use only
 * its own vertices (the per-particle meshes of the packing, concatenated) --
 * and the r_max every candidate of this seed is evaluated at. */
typedef struct s2h_seed s2h_seed;
s2h_status s2h_seed_create(s2h_ctx *ctx, const double *xyz, uint64_t nverts, const uint32_t *tri,
                           uint64_t ntris, const uint32_t *granule_of_tri, uint32_t ngranules,
                           const uint32_t *particle_vert_start, const uint32_t *particle_tri_start,
                           uint32_t nparticles, const double bbox_min[3], double pitch,
                           const uint64_t dims[3], uint32_t r_max, s2h_seed **out);
void s2h_seed_free(s2h_ctx *ctx, s2h_seed *seed);
/* The seed's packed volume (owned by the seed; valid until s2h_seed_free). */
s2h_status s2h_seed_volume(s2h_seed *seed, s2h_vol **vol);

/* One candidate: particle `particle` of seed `seed` moved to new vertex
 * positions (VoxelCoverage::replace, src/geometry/s2.rs:314-317).  new_xyz holds
 * 3 doubles per vertex of that particle, in the seed's vertex order. */
typedef struct {
    uint32_t seed;
    uint32_t particle;
    const double *new_xyz;
} s2h_candidate;

/* Exact S2 curves of every candidate packing in one pass (replaces the per-
 * candidate island_s2 -> evaluate_voxel_grid of run_sa_island,
 * src/pipeline/optimize.rs:719-739): curves receives, per candidate in order,
 * r_max(seed)+1 values, bit-identical to a full evaluation of that packing.
 * The device work is one replay of a HIP graph captured once per context. */
s2h_status s2h_batch_eval(s2h_ctx *ctx, s2h_seed *const *seeds, uint32_t nseeds,
                          const s2h_candidate *cands, uint32_t ncand, double *curves,
                          uint64_t ncurve);
/* Commit a move into the seed (run_sa_island's accept path,
 * src/pipeline/optimize.rs:740-771, with VoxelCoverage::replace): coverage,
 * occupancy rows, exact counts and the particle's vertices are updated on the
 * device; curve (r_max+1 values, may be null) receives the new seed's curve. */
s2h_status s2h_seed_accept(s2h_ctx *ctx, s2h_seed *seed, uint32_t particle, const double *new_xyz,
                           double *curve, uint64_t n);

/* ------------------------------------------------------------------ serving */

/* A request loop with admission by device-memory footprint (no counterpart in the
 * reference, whose GPU objects serve one call at a time behind a Mutex,
 * src/pipeline/optimize_execution.rs:274-288).  nslots worker contexts, each on
 * its own stream, run admitted requests concurrently; a waiting request is
 * admitted as soon as a slot is free and its footprint fits the budget, in
 * arrival order (a request that does not fit is passed over and counted as
 * deferred).  Volumes stay resident in HBM while their request runs.  With a
 * budget, a quarter of it is left to the slots' cached shell plans, a quarter
 * to per-slot FFT arenas reserved at create (an FFT plan whose spectra fit its
 * slot's arena is carved from it, so no request allocates spectra), and
 * requests are admitted against the other half. */
typedef struct s2h_server s2h_server;
typedef struct {
    const uint64_t *words;  /* packed volume (copied at submit) */
    uint64_t nwords;
    uint64_t dims[3];
    uint32_t r_max;
    int32_t method;         /* 0 exact, 1 FP32 FFT */
    double pitch;
    uint64_t tag;           /* caller's id, echoed in the result */
} s2h_request;
typedef struct {
    int32_t ready;          /* 0 while the request is waiting or running */
    int32_t status;
    uint32_t deferrals;     /* admission passes that skipped this request */
    uint32_t reserved;
    uint64_t tag;
    uint64_t footprint;     /* device bytes reserved while admitted */
    double t_submit, t_start, t_done;  /* seconds since the server started */
} s2h_result;
s2h_status s2h_serve_create(int device, uint64_t hbm_budget_bytes, uint32_t nslots, s2h_server **out);
void s2h_serve_destroy(s2h_server *server);
s2h_status s2h_serve_submit(s2h_server *server, const s2h_request *req, uint64_t *ticket);
/* Result of a ticket (wait != 0 blocks until it is done).  Once a finished
 * result has been returned its ticket is released. */
s2h_status s2h_serve_poll(s2h_server *server, uint64_t ticket, int32_t wait, s2h_result *res,
                          double *s2, uint64_t n);
/* n = 8: [admitted, deferred, completed, failed, waiting, bytes_in_use, peak_bytes, budget], where
 * bytes_in_use / peak_bytes sum the admitted footprints; n = 10 adds [8] the device bytes the slot
 * contexts hold between requests (in use + cached blocks) and [9] the number of requests during
 * which a slot held more device memory than its share (the request's footprint plus the slot's
 * plan allowance; 0 unless an estimate was exceeded).  Under a budget each slot's device
 * memory, cached blocks included, is held to that share. */
s2h_status s2h_serve_stats(s2h_server *server, uint64_t *stats, uint64_t n);

/* --------------------------------------------------------------- host helpers */

/* Fill unsupported radii (fill_missing_s2_with_smooth_interpolation,
 * src/geometry/s2.rs:393-501). */
s2h_status s2h_fill_missing(double *values, const uint8_t *has_support, uint64_t n, double vf);
/* Binary/ASCII STL load with the reference's f32->f64 conversion and 1e-6 vertex
 * weld (src/io/stl.rs:55-80, 157-189).  Arrays are allocated by the library and
 * released with s2h_free. */
s2h_status s2h_stl_load(const char *path, double **xyz, uint64_t *nverts, uint32_t **tri,
                        uint64_t *ntris);
void s2h_free(void *p);
/* Volume of a closed mesh inside a box (particle_volume_in_bbox,
 * src/geometry/volume.rs:424-430): the mesh volume (mesh_volume, 6-15) when every
 * vertex lies in the box, else the volume of the mesh clipped by the six box
 * planes with each cut capped (clip_mesh_by_bbox, 352-421).  Host only. */
s2h_status s2h_mesh_volume_in_bbox(const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntris,
                                   const double bbox_min[3], const double bbox_max[3], double *volume);
/* Geometric volume fraction in the box (volume_fraction_in_bbox,
 * src/geometry/volume.rs:437-469): the granules (reference BFS order) each
 * measured by s2h_mesh_volume_in_bbox, summed in granule order, over the box
 * volume (at least 1e-12), clamped to [0, 1].  Bit-identical steps, so the
 * measure report's "Volume Fraction:" and S2(0)-VF lines print the same.
 * Host only. */
s2h_status s2h_volume_fraction_in_bbox(const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntris,
                                       const double bbox_min[3], const double bbox_max[3], double *vf);
/* Connected components over shared vertices (split_mesh_into_granules,
 * src/geometry/mesh_ops.rs:48-~120): granule id per triangle, in the reference's
 * component order. */
s2h_status s2h_mesh_granules(const uint32_t *tri, uint64_t ntris, uint64_t nverts,
                             uint32_t *granule_of_tri, uint32_t *ngranules);

#ifdef __cplusplus
}
#endif
#endif /* S2HIP_H */
