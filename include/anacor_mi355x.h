/*
 * anacor_mi355x.h -- C-ABI of the MI355X (gfx950) absorption-correction ray
 * tracer, libanacor_mi355x.so.
 *
 * The first four entry points keep the exact signatures of the reference's
 * ctypes targets, so AnACor's Python drivers can load this library in place of
 * src/ray_tracing_cpu.so / src/ray_tracing_gpu_{card}.so.  The rest is the
 * library's own extended interface (resident volume, diagnostics).
 *
 * Error convention: no entry point aborts.  A reflection on which the
 * reference would abort (cube_face exit(), cal_coord assert) yields NaN;
 * anacor_last_error() then returns a nonzero code and a message.  Results of
 * the drop-in entry points are allocated with libc malloc(), so callers free
 * them with libc free() exactly as utils/utils_mp.py:212 does.
 */
#ifndef ANACOR_MI355X_H
#define ANACOR_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define ANACOR_OK            0
#define ANACOR_ERR_ARG       1   /* invalid argument */
#define ANACOR_ERR_NO_DEVICE 2   /* no HIP device visible */
#define ANACOR_ERR_HIP       3   /* HIP runtime failure */
#define ANACOR_ERR_GEOMETRY  4   /* some reflection hit a reference abort case (result NaN) */

/* per-ray error bits in anacor_path_t.err */
#define ANACOR_RAY_ERR_FACE   1u  /* cube_face found no exit plane */
#define ANACOR_RAY_ERR_ASSERT 2u  /* exit face inconsistent with the beam angles */

/* flags of anacor_ctx_run */
#define ANACOR_FLAG_F64_MEAN 1   /* FP64 exp and accumulation instead of FP32 exp + compensated sum */
#define ANACOR_FLAG_NO_EXP   2   /* average sum(mu*L) instead of exp(-sum(mu*L)) (IsExp = 0) */

/* ------------------------------------------------------------------ drop-in */

/* Replaces ray_tracing_overall, AnACor/src/ray_tracing.c:365-374 (ctypes
 * binding utils/utils_mp.py:59-79, call :197-204).  label_list is the int8
 * row-pointer table of python_2_c_3d (utils/utils_os.py:45-68); low, up and
 * store_paths are accepted and unused as in the reference; num_workers sets the
 * host threads of the O(R) setup.  Returns malloc'd double[len_result]. */
double *ray_tracing_overall(int64_t low, int64_t up, int64_t *coord_list, int64_t len_coord_list,
                            const double *scattering_vector_list, const double *omega_list,
                            const double *raw_xray, const double *omega_axis,
                            const double *kp_rotation_matrix, int64_t len_result,
                            double *voxel_size, double *coefficients, int8_t ***label_list,
                            int64_t *shape, int full_iteration, int store_paths, int num_workers,
                            int IsExp);

/* Replaces ray_tracing_gpu_overall, AnACor/src/ray_tracing_gpu.c:33-40
 * (ctypes binding utils/utils_mp.py:128-148, call :166-187).  float inputs are
 * widened to double and traced with the same FP64 rules; label_list_1d is
 * used when non-null; gpumethod is accepted and ignored.  Returns malloc'd
 * float[len_result]. */
float *ray_tracing_gpu_overall(size_t low, size_t up, int *coord_list, size_t len_coord_list,
                               const float *scattering_vector_list, const float *omega_list,
                               const float *raw_xray, const float *omega_axis,
                               const float *kp_rotation_matrix, size_t len_result,
                               float *voxel_size, float *coefficients, int8_t ***label_list,
                               int8_t *label_list_1d, int *shape, int full_iteration,
                               int store_paths, int gpumethod);

/* Replaces ray_tracing_single, AnACor/src/ray_tracing.c:166-172 (ctypes
 * binding utils/utils_mp.py:82-94): one reflection given already rotated
 * vectors; returns the mean absorption factor. */
double ray_tracing_single(int64_t *coord_list, int64_t len_coord_list, const double *rotated_s1,
                          const double *xray, double *voxel_size, double *coefficients,
                          int8_t ***label_list, int64_t *shape, int full_iteration,
                          int64_t store_paths, int IsExp);

/* Replaces ray_tracing_single_mp, AnACor/src/ray_tracing.c:294-300 (called by
 * verification/main_verification.py:213-248); num_workers is ignored. */
double ray_tracing_single_mp(int64_t *coord_list, int64_t len_coord_list,
                             const double *rotated_s1, const double *xray, double *voxel_size,
                             double *coefficients, int8_t ***label_list, int64_t *shape,
                             int full_iteration, int64_t store_paths, int IsExp, int num_workers);

/* ------------------------------------------------------- extended interface */

/* Last error of the calling thread's most recent call: returns the code and
 * copies the message into buf (no counterpart in the reference, which aborts;
 * replaces the printf/exit of ray_tracing_basic.c:150-155). */
int anacor_last_error(char *buf, size_t n);

/* Number of visible HIP devices. */
int anacor_device_count(void);

/* A device-resident problem: the label volume [z][y][x] (uint8, contiguous)
 * and the crystal-voxel list (int64 [n][3], z y x) are uploaded once;
 * voxel_size (z, y, x) and coefficients [li, lo, cr, bu] as in
 * ray_tracing_overall.  The device holds the volume inside a one-voxel shell
 * of zeros: each extent at most 2^20 and (Z+2)(Y+2)(X+2) < 2^31.  Returns
 * NULL on failure (see anacor_last_error). */
typedef struct anacor_ctx anacor_ctx;
anacor_ctx *anacor_ctx_create(const uint8_t *volume, const int64_t shape[3],
                              const int64_t *coords, int64_t n_coords,
                              const double voxel_size[3], const double coefficients[4],
                              int full_iteration, int device);
void anacor_ctx_destroy(anacor_ctx *ctx);

/* Voxels per workgroup tile (0 = automatic: at most 1,024 in the raw
 * layout, 384 in the bricked ones). */
int anacor_ctx_set_tile(anacor_ctx *ctx, int voxels_per_tile);

/* Launch order of the trace kernel (speed only, never results): reflections
 * are sorted by the directions of their rays, led by the incident ray's
 * (direction_major 0) or the diffracted ray's (1; -1: automatic, the
 * diffracted one in the bricked layouts, else the incident one), and launched
 * in groups of refl_group consecutive reflections, tile by tile within a
 * group, so that the workgroups resident on an XCD share the bricks their
 * rays cross (0: automatic, 2,048 in the bricked layouts and none in the raw
 * one; 1: no groups, every tile of a reflection in turn).  refl_group at most
 * 65,536. */
int anacor_ctx_set_order(anacor_ctx *ctx, int refl_group, int direction_major);

/* The work shape in use: voxel block (bz, by, bx; 0, 0, 0 = Morton), tile
 * setting (0 = automatic), reflections per launch group for the current
 * layout (1 = none) and the leading ray of the direction sort.  Any output
 * may be NULL. */
int anacor_ctx_get_work_shape(const anacor_ctx *ctx, int block[3], int *tile, int *refl_group,
                              int *direction_major);

/* Volume layout of the trace kernel.  Never changes a result: every layout
 * yields the same labels along every ray (bit-identical results).  Bricks are
 * 8x8x8 voxels; 4-bit bricks store labels 0..4 as they are and any other value
 * as 5, which the boundary rule cannot tell apart (only 0, "changed" and
 * 0..4 matter to it).  The packed copy is built on the GPU on first use; the
 * raw volume stays resident (path records and the batched path use it). */
#define ANACOR_LAYOUT_AUTO   (-1)  /* bricks above 64 MB; ANACOR_VOLUME_LAYOUT=raw|brick overrides */
#define ANACOR_LAYOUT_RAW    0     /* uint8 [z][y][x] */
#define ANACOR_LAYOUT_BRICK  1     /* 2-bit bricks if every label is 0..3, else 4-bit */
#define ANACOR_LAYOUT_BRICK2 2     /* 2-bit bricks (error unless every label is 0..3) */
#define ANACOR_LAYOUT_BRICK4 4     /* 4-bit bricks */
int anacor_ctx_set_layout(anacor_ctx *ctx, int layout);
/* Layout used by the most recent trace: ANACOR_LAYOUT_RAW, _BRICK2 or _BRICK4. */
int anacor_ctx_get_layout(const anacor_ctx *ctx);

/* Grouping of the crystal voxels into waves: blocks of bz x by x bx voxels in
 * z, y, x order, lanes along x inside a block; (0, 0, 0) selects Morton
 * order.  Affects speed only, never results beyond the FP32 summation order. */
int anacor_ctx_set_voxel_block(anacor_ctx *ctx, int bz, int by, int bx);

/* Mean absorption factor of R reflections (s1 [R][3], omega [R], DIALS frame)
 * into out[R] (host memory).  The same computation as ray_tracing_overall on
 * a resident volume. */
int anacor_ctx_run(anacor_ctx *ctx, const double *s1, const double *omega, int64_t R,
                   const double raw_xray[3], const double omega_axis[3], const double F[9],
                   int flags, double *out);

/* The same as anacor_ctx_run, allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * per-call device buffer (rules, processing order, tile partials, error words,
 * means, step counter) lives in one static device array of
 * ANACOR_NOALLOC_SCRATCH_BYTES that each process holds per GPU until it
 * exits; host staging is static too and the host setup runs on the calling
 * thread.  At most ANACOR_NOALLOC_MAX_REFLECTIONS reflections per call
 * (ANACOR_ERR_ARG beyond); the voxel tile is enlarged so that reflections x
 * tiles fit ANACOR_NOALLOC_MAX_PARTIALS.  The context must already hold the
 * volume layout the trace uses (a bricked copy is built by
 * anacor_ctx_set_layout or an earlier anacor_ctx_run; ANACOR_ERR_ARG
 * otherwise).  Calls in one process are serialised by a lock and every call
 * rewrites the array (its error words and step counter are cleared first);
 * any other use of the array by overlapping work on the same GPU would race. */
#define ANACOR_NOALLOC_MAX_REFLECTIONS 16384
#define ANACOR_NOALLOC_MAX_PARTIALS    (1 << 21)
#define ANACOR_NOALLOC_SCRATCH_BYTES   25428224   /* 24.25 MiB */
int anacor_ctx_run_noalloc(anacor_ctx *ctx, const double *s1, const double *omega, int64_t R,
                           const double raw_xray[3], const double omega_axis[3],
                           const double F[9], int flags, double *out);

/* Per (reflection, voxel) path record, both rays ([0] incident, [1]
 * diffracted): exit face, len_path_2, boundary count, last path coordinate
 * (z, y, x) and the four lengths [li, lo, cr, bu] of cal_path2_plus. */
typedef struct {
    int32_t face[2];
    int32_t len[2];
    int32_t nseg[2];
    int32_t err;
    int32_t pad;
    int32_t last[2][3];
    double  L[2][4];
} anacor_path_t;

/* Path records of R reflections into out[R * n_coords] (voxels in the order
 * given to anacor_ctx_create). */
int anacor_ctx_paths(anacor_ctx *ctx, const double *s1, const double *omega, int64_t R,
                     const double raw_xray[3], const double omega_axis[3], const double F[9],
                     anacor_path_t *out);

/* Statistics of the most recent run on ctx. */
typedef struct {
    double   kernel_ms;      /* trace kernel duration (HIP events on the ctx stream) */
    uint64_t voxel_steps;    /* sum of len_path_2 over all rays */
    int64_t  workgroups;     /* workgroups of the trace launch */
    int64_t  n_voxels;
    double   host_ms;        /* host setup of the reflections (rotation, angles, face rules, order) */
    uint64_t outside_reads;  /* bisection runs: label reads outside the volume (read as vacuum) */
} anacor_stats_t;
int anacor_ctx_stats(const anacor_ctx *ctx, anacor_stats_t *stats);

/* ----------------------------------------------------------- batched path */

/* One independent sample of a batch: its label volume, crystal voxels,
 * voxel size and coefficients, as for anacor_ctx_create. */
typedef struct {
    const uint8_t *volume;
    int64_t        shape[3];
    const int64_t *coords;
    int64_t        n_coords;
    double         voxel_size[3];
    double         coefficients[4];
} anacor_sample_t;

/* Several samples traced as a clear kernel, per sample its host setup, rule
 * upload and trace launch (the setup of each sample overlapping the trace of
 * the one before), then one finish kernel over all of them: launched from
 * the host, or (anacor_batch_set_graph) as one hipGraph with the setups as
 * host nodes, captured on the first run and relaunched while the reflection
 * counts and buffers stay the same.
 * No counterpart in the reference, whose driver runs samples one after the
 * other (main_lite.py:276-341).  All samples share full_iteration. */
typedef struct anacor_batch anacor_batch;
anacor_batch *anacor_batch_create(const anacor_sample_t *samples, int n_samples,
                                  int full_iteration, int device);
void anacor_batch_destroy(anacor_batch *batch);

/* Borrowed handle of sample i (for anacor_ctx_set_voxel_block /
 * anacor_ctx_set_tile / anacor_ctx_run, e.g. by an autotuner); never pass it
 * to anacor_ctx_destroy. */
anacor_ctx *anacor_batch_sample(anacor_batch *batch, int i);

/* Trace R[i] reflections of every sample i (s1[i] [R[i]][3], omega[i] [R[i]])
 * with one goniometer (raw_xray, omega_axis, F); out[i] receives R[i] means. */
int anacor_batch_run(anacor_batch *batch, const double *const *s1, const double *const *omega,
                     const int64_t *R, const double raw_xray[3], const double omega_axis[3],
                     const double F[9], int flags, double *const *out);

/* Statistics of the most recent batch run (kernel_ms covers the whole
 * sequence, the first sample's host setup included). */
int anacor_batch_stats(const anacor_batch *batch, anacor_stats_t *stats);

/* Run the batch as one hipGraph (use_graph 1) or launch its sequence from
 * the host (0, the default: faster where the runtime runs the graph's host
 * nodes in line with its kernels).  Results are the same bit for bit. */
int anacor_batch_set_graph(anacor_batch *batch, int use_graph);

/* 1 when the most recent batch run launched the batch graph, 0 when it ran
 * the sequence from the host (graph off, or rays that need the guarded
 * kernel). */
int anacor_batch_last_graph(const anacor_batch *batch);

/* ------------------------------------------------------------ gridding path */

/* Replaces ray_tracing_single_gridding, AnACor/src/ray_tracing.c:484-534
 * (ctypes binding utils/utils_gridding.py:104-118, per-voxel loop :162-231):
 * sum of mu*L of one ray from coord along rotated_s1 walked with the given
 * angles (IsExp = 1: its exp(-.)).  NaN (with ANACOR_ERR_GEOMETRY) where the
 * reference's asserts abort.  One ray per call; anacor_grid_create builds a
 * whole map in one launch. */
double ray_tracing_single_gridding(int64_t *coord, const double *rotated_s1, double theta,
                                   double phi, double *voxel_size, double *coefficients,
                                   int8_t ***label_list, int64_t *shape, int full_iteration,
                                   int index, int IsExp);

/* Replaces nearest_neighbor_interpolate_overall, AnACor/src/gridding_interpolation.c:
 * 98-161 (binding utils/utils_gridding.py:567-591, call :627-641): per
 * reflection the mean over voxels of exp(-(m1 + m2)) where m1, m2 are the
 * padded map gridding_data[N][ny][nx] looked up at the incident and diffracted
 * (theta, phi): interpolation_method 1 = the reference's truncated
 * nearest-neighbour index, 2 = bilinear as GSL's interp2d (the GSL path of
 * gridding_interpolation.c:53-78).  Returns malloc'd double[len_result]. */
double *nearest_neighbor_interpolate_overall(
    int64_t low, int64_t up, int64_t *coord_list, int64_t len_coord_list,
    const double *scattering_vector_list, const double *omega_list, const double *raw_xray,
    const double *omega_axis, const double *kp_rotation_matrix, int64_t len_result,
    double *voxel_size, double *coefficients, int num_workers, int IsExp, double *theta_list,
    double *phi_list, double *gridding_data, size_t nx, size_t ny, double theta_min,
    double theta_max, double phi_min, double phi_max, int interpolation_method);

/* A resident absorption map of the crystal voxels of ctx for n_theta x n_phi
 * directions (phi-major, as utils_gridding.py:254-259 orders them):
 * rotated_s1[D][3] and their (theta, phi)[D][2].  Built by one kernel that
 * walks every (direction, voxel) ray with the tracer's exact rules; entries
 * are floats as the reference's .npy maps.  The map is [D][N] (voxels in
 * the order given to anacor_ctx_create), the reference's absorption_map. */
typedef struct anacor_grid anacor_grid;
anacor_grid *anacor_grid_create(anacor_ctx *ctx, int64_t n_theta, int64_t n_phi,
                                const double *rotated_s1, const double *theta_phi, int is_exp);
void anacor_grid_destroy(anacor_grid *grid);
int anacor_grid_map(const anacor_grid *grid, float *out);
/* The correction of R reflections by look-up (method 1 nearest, 2 bilinear)
 * on the padded grid the lists describe (theta_list[nx], phi_list[ny]: the
 * n_theta x n_phi grid extended by (nx - n_theta) / 2 columns and
 * (ny - n_phi) / 2 rows from the opposite ends, utils_gridding.py:406-445). */
int anacor_grid_run(anacor_grid *grid, const double *s1, const double *omega, int64_t R,
                    const double raw_xray[3], const double omega_axis[3], const double F[9],
                    const double *theta_list, int64_t nx, const double *phi_list, int64_t ny,
                    int method, double *out);
/* Duration of the map build and of the latest look-up (HIP events). */
int anacor_grid_stats(const anacor_grid *grid, double *build_ms, double *lookup_ms);

/* ------------------------------------------------- iterative bisection path */

/* Replaces ib_test, AnACor/src/ray_tracing.c:25-160 with bisection.c:21-471
 * (ctypes binding mp_worker.py:390-404, call :509-514; the binding's extra
 * trailing IsExp argument is ignored, as by the reference): the reference's
 * alternative traversal, which bisects each ray for its air, crystal, loop
 * and (num_cls > 3) bubble boundaries instead of walking it; mean over the
 * voxels of exp(-sum mu*L) for one reflection given its rotated vectors.
 * Labels outside the volume read as vacuum (the reference reads its row table
 * unchecked there).  NaN (ANACOR_ERR_GEOMETRY) where the reference asserts. */
double ib_test(int64_t *coord_list, int64_t len_coord_list, double *rotated_s1, double *xray,
               double *voxel_size, double *coefficients, int8_t ***label_list, int64_t *shape,
               int full_iteration, int64_t store_paths, int num_cls);

/* The same for R reflections in one launch on a resident volume: s1 [R][3]
 * and omega [R] rotated as ray_tracing_overall rotates them (the loop of
 * mp_worker.py:491-514).  flags: ANACOR_FLAG_F64_MEAN, ANACOR_FLAG_NO_EXP. */
int anacor_ctx_bisection(anacor_ctx *ctx, const double *s1, const double *omega, int64_t R,
                         const double raw_xray[3], const double omega_axis[3], const double F[9],
                         int num_cls, int flags, double *out);

/* On already rotated vectors (rotated_s1 [R][3], xray [R][3]), as ib_test
 * takes them.  lengths (optional, host) receives [R][n_coords][2][4]: per
 * voxel (in the order given to anacor_ctx_create) and ray ([0] incident, [1]
 * diffracted) the lengths [li, lo, cr, bu] of cal_path_bisection. */
int anacor_ctx_bisection_rotated(anacor_ctx *ctx, const double *rotated_s1, const double *xray,
                                 int64_t R, int num_cls, int flags, double *out,
                                 double *lengths);

/* Host-side geometry of R reflections (diagnostics): theta_phi[R][4] =
 * (theta, phi) of the incident then the diffracted ray; d_table[R][2][6][3] =
 * per-face per-increment deltas (z, y, x).  Either output may be NULL. */
int anacor_reflection_geometry(const double *s1, const double *omega, int64_t R,
                               const double raw_xray[3], const double omega_axis[3],
                               const double F[9], int64_t max_dim, double *theta_phi,
                               double *d_table);

/* -------------------------------------------------------- learned surrogate */

/* An MLP 60 -> 256 -> 256 -> 256 -> 256 -> 1 (SiLU) from features of the
 * incident and diffracted unit vectors in the crystal frame to log A_h for
 * one sample, trained on the GPU on absorption factors the tracer computes.
 * No counterpart in the reference; its nearest analogue is the gridding
 * approximation (src/gridding_interpolation.c:98-161).  The 256x256 layers
 * use BF16 weights on MFMA with FP32 accumulation, the input layer BF16
 * hi/lo pairs.
 *
 * A row of network inputs (float[60]): [0:3] x (incident), [3:6] s
 * (diffracted), [6:12] six more by feature set, [12:60] resonance inputs
 * (ANACOR_MLP_FEATURES_RESONANCE; zero in the other sets): for each ray
 * (x, then s) and component pair (0, 1), (0, 2), (1, 2), k = 3 ray + pair,
 * [12 + k] the ratio t = min(|v_i|, |v_j|) / max(|v_i|, |v_j|) (0 when both
 * are 0) and [18 + 7 k + q - 2] = 2 |q t - rint(q t)| for q = 2..8, all in
 * FP32 from the FP32 components. */
#define ANACOR_MLP_INPUTS 60
#define ANACOR_MLP_WIDTH  256
#define ANACOR_MLP_PARAMS 213249   /* FP32 master parameters, layout in mlp.hip */

/* Network input sets (anacor_mlp_set_features): inputs [6:12] after the two
 * unit vectors x (incident) and s (diffracted), crystal frame. */
#define ANACOR_MLP_FEATURES_HARMONIC  0  /* per vector: 2 v0 v1, v0^2 - v1^2, 2 v2^2 - 1 */
#define ANACOR_MLP_FEATURES_CHAMBER   1  /* +-1 sides of the planes where the tracer's
                                            discrete rules switch (sign of theta, diagonals) */
#define ANACOR_MLP_FEATURES_RESONANCE 2  /* the chamber signs, and the resonance inputs
                                            [12:60]: kinks where two components of a ray
                                            stand in a ratio p / q, q <= 8 */

/* Where the network may answer.  Invariants of a reflection: axial = x . axis,
 * two_theta = angle(x, s), alpha = atan2(x . e2, x . e1) (azimuth of the
 * incident direction about the axis), beta = azimuth of s about x (basis
 * axis x x, x (axis x x)).  A reflection is inside when axial and two_theta
 * are in range and its (alpha, two_theta, beta) cell is trusted. */
typedef struct {
    double  axis[3];
    double  e1[3], e2[3];
    double  axial[2];
    double  two_theta[2];     /* radians */
    int32_t cells[3];         /* grid over alpha [-pi, pi], two_theta range, beta [-pi, pi] */
    int32_t pad;
} anacor_envelope_t;

typedef struct anacor_mlp anacor_mlp;

/* Seeded uniform fan-in initialisation on `device`; NULL on failure. */
anacor_mlp *anacor_mlp_create(uint64_t seed, int device);
void anacor_mlp_destroy(anacor_mlp *mlp);
size_t anacor_mlp_num_params(void);
int anacor_mlp_get_params(anacor_mlp *mlp, float *out, size_t n);
/* Also resets the Adam state. */
int anacor_mlp_set_params(anacor_mlp *mlp, const float *in, size_t n);
/* log A_h = y_mean + y_std * network output. */
int anacor_mlp_set_scale(anacor_mlp *mlp, double y_mean, double y_std);

int anacor_mlp_set_features(anacor_mlp *mlp, int feature_set);

/* Envelope and trust table (uint8[cells[0] * cells[1] * cells[2]], cell
 * index (ia * cells[1] + it) * cells[2] + ib; NULL = every cell trusted).
 * env NULL removes the envelope. */
int anacor_mlp_set_envelope(anacor_mlp *mlp, const anacor_envelope_t *env, const uint8_t *trust);

/* Resonance bands: near a small rational ratio of two components of a ray's
 * crystal-frame direction, the voxel steps of all crystal voxels fall into
 * step together and the traced absorption factor oscillates (by up to a few
 * 1e-3 relative, over a few 0.01 rad) faster than 200k training reflections
 * sample it.  A reflection whose ray `ray` (0 incident, 1 diffracted) has
 * |min(|v_i|, |v_j|) / max(|v_i|, |v_j|) - ratio| < width for some band lies
 * outside the envelope and goes to the tracer.  n = 0 removes them. */
typedef struct {
    int32_t ray;
    int32_t i, j;             /* components (0, 1, 2) of the unit vector, i != j */
    int32_t pad;
    double  ratio;            /* in [0, 1] */
    double  width;
} anacor_band_t;
#define ANACOR_MLP_MAX_BANDS 32
int anacor_mlp_set_bands(anacor_mlp *mlp, const anacor_band_t *bands, int n);

/* Risk trees: a boosted ensemble of complete binary decision trees over 60
 * resonance features of a reflection, computed in FP32 from the network
 * inputs x = features[0..2], s = features[3..5]:
 *   [0, 6)    x0 x1 x2 s0 s1 s2
 *   [6, 12)   t = min(|v_i|, |v_j|) / max(|v_i|, |v_j|) (0 when both are 0) for
 *             (v, (i, j)) = (x, (0,1)), (x, (0,2)), (x, (1,2)), (s, (0,1)), (s, (0,2)),
 *             (s, (1,2)), in that order
 *   [12, 60)  for each of those six ratios k and q = 1..8:
 *             |q t - rint(q t)| * (1 / q) at index 12 + 8 k + q - 1
 *             (the distance of t to the nearest fraction p / q)
 * Tree `t` of depth D has 2^D - 1 internal nodes in level order and 2^D
 * leaves: from node 0, node = 2 node + 1 + (f[feat[node]] > thr[node]),
 * and after D levels the leaf is node - (2^D - 1).  score = base + the sum of
 * the trees' leaves in tree order (FP32).  A reflection whose score is at
 * least `cut` lies outside the envelope and goes to the tracer: the trees
 * estimate where the network's error exceeds the surrogate's bound (the
 * moire of the voxel steps near rational direction ratios).  No counterpart
 * in the reference. */
#define ANACOR_RISK_FEATURES  60
#define ANACOR_RISK_MAX_DEPTH 8
#define ANACOR_RISK_MAX_TREES 1024

/* Arrays feat[n_trees][2^depth - 1] (feature indices), thr[n_trees][2^depth - 1],
 * leaf[n_trees][2^depth]; n_trees = 0 removes the trees. */
int anacor_mlp_set_risk(anacor_mlp *mlp, int n_trees, int depth, const uint8_t *feat,
                        const float *thr, const float *leaf, float base, float cut);

/* The 60 resonance features (out[n][60]) and the trees' scores (scores[n],
 * the trees set by anacor_mlp_set_risk; nullable) of n rows of network
 * inputs x[n][ANACOR_MLP_INPUTS] (only x[:, 0:6] is read), on the GPU. */
int anacor_mlp_risk(anacor_mlp *mlp, const float *x, int64_t n, float *features,
                    float *scores);

/* Fits risk trees on the host (no GPU): Newton-boosted trees of the logistic
 * loss of labels[n] (0/1) over features[n][n_features] (row-major), n_bins
 * (2..256) quantile bins per feature, L2 penalty l2 on the leaf values,
 * splits whose children both carry a Hessian of at least min_hess, leaves
 * scaled by learning_rate; optional per-row weights (NULL: 1).  Writes
 * feat, thr, leaf (layouts of anacor_mlp_set_risk) and base (the log-odds of
 * the weighted label mean).  A node without a gainful split sends every row
 * left (thr = +inf).  Deterministic: the same inputs give the same trees on
 * any number of host threads (threads <= 0: the machine's). */
int anacor_risk_fit(const float *features, const uint8_t *labels, const float *weights,
                    int64_t n, int n_features, int n_trees, int depth, float learning_rate,
                    float l2, float min_hess, int n_bins, int threads, uint8_t *feat,
                    float *thr, float *leaf, float *base);

/* Network inputs of R reflections (features[R][ANACOR_MLP_INPUTS]) and invariants[R][4] =
 * (axial, two_theta, alpha, beta); with an envelope set also the cell index
 * and the in-envelope flag.  Every output is nullable; without an envelope
 * the axis is F^T omega_axis. */
int anacor_mlp_features(anacor_mlp *mlp, const double *s1, const double *omega, int64_t R,
                        const double raw_xray[3], const double omega_axis[3],
                        const double F[9], float *features, double *invariants,
                        int32_t *cells, uint8_t *in_envelope);

/* Network outputs (normalised) for n rows of features. */
int anacor_mlp_forward(anacor_mlp *mlp, const float *x, int64_t n, float *y);

/* Gradient of the mean squared error of n <= 65536 rows against targets t
 * (normalised) into grad[ANACOR_MLP_PARAMS]; no update. */
int anacor_mlp_gradients(anacor_mlp *mlp, const float *x, const float *t, int64_t n,
                         float *grad, float *loss);

/* Adam (beta 0.9 / 0.999, eps 1e-8) over `epochs` shuffled passes of
 * mini-batches, learning rate cosine-decayed from lr to lr_min;
 * epoch_loss[epochs] (nullable) receives the mean batch loss per epoch. */
int anacor_mlp_train(anacor_mlp *mlp, const float *x, const float *t, int64_t n, int epochs,
                     int batch, float lr, float lr_min, uint64_t seed, float *epoch_loss);

/* A_h of R reflections: features, envelope test and forward pass on the GPU.
 * Reflections outside the envelope get NaN and in_envelope 0; the caller
 * traces them.  Needs an envelope (anacor_mlp_set_envelope). */
int anacor_mlp_predict(anacor_mlp *mlp, const double *s1, const double *omega, int64_t R,
                       const double raw_xray[3], const double omega_axis[3], const double F[9],
                       double *A, uint8_t *in_envelope);

typedef struct {
    double  mlp_ms;       /* forward pass (HIP events on the model's stream) */
    double  total_ms;     /* the whole anacor_mlp_predict call */
    int64_t rows;
    int64_t in_envelope;
} anacor_mlp_stats_t;
int anacor_mlp_stats(const anacor_mlp *mlp, anacor_mlp_stats_t *stats);

/* ------------------------------------------------------------------------
 * Measurement helper (no counterpart in the reference): the device's
 * one-byte gather rate against a working set of 2^log2_bytes bytes
 * (6..30), with no arithmetic between the loads.  It bounds the ray walk,
 * which reads one uint8 label per voxel step, from above at the volume's own
 * size.  About `loads` loads (capped at 2^36) per timed launch, five timed
 * launches after one untimed one; allocates and frees its own buffers.
 * ---------------------------------------------------------------------- */
#define ANACOR_GATHER_LANES 0   /* every lane its own pseudo-random byte */
#define ANACOR_GATHER_ROW   1   /* a wave's 64 lanes one 64-byte row */
int anacor_probe_gather(int device, int log2_bytes, int pattern, int64_t loads,
                        double *loads_per_s);

#ifdef __cplusplus
}
#endif
#endif
