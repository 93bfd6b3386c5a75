/*
 * anacor_oracle.h -- CPU restatement (FP64) of AnACor's analytical
 * absorption-correction ray tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: it is built and called by tests/, by __graft_entry__.smoke() and by the
 * CPU-baseline leg of bench.py, never by the product library.
 *
 * Every function follows the reference C implementation op for op (same FP64
 * operation order, no FMA contraction, glibc libm), citing the reference
 * file:line it restates.  See anacor_oracle.c for the pinning status of each
 * function.
 */
#ifndef ANACOR_ORACLE_H
#define ANACOR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes (the reference aborts the process in each case). */
#define ORACLE_OK          0
#define ORACLE_ERR_FACE    1   /* cube_face found no exit plane: exit(EXIT_FAILURE), ray_tracing_basic.c:150-155 */
#define ORACLE_ERR_ASSERT  2   /* cal_coord precondition (assert) fails, ray_tracing_basic.c:420,517,618,717,807,901,1106,1284 */

/* Per-ray diagnostic record. */
typedef struct {
    int32_t face;      /* exit face 1..6 (0 when cube_face failed) */
    int32_t err;       /* ORACLE_* error code */
    int64_t len;       /* len_path_2 */
    int64_t nseg;      /* number of recorded boundaries (len_classes_posi) */
    int64_t first[3];  /* path coordinate at increment 0, (z, y, x) */
    int64_t last[3];   /* last path coordinate, (z, y, x) */
    double  L[4];      /* per-material lengths [li, lo, cr, bu] */
} oracle_ray_t;

/* Per-reflection geometry (ray_tracing.c:393-407 then :182-197). */
typedef struct {
    double xray[3];       /* rotated incident beam, DIALS frame */
    double s1[3];         /* rotated diffracted beam, DIALS frame */
    double theta[2];      /* [0] incident (L1=1), [1] diffracted */
    double phi[2];
    double dir[2][3];     /* myframe (z, y, x) directions used by cube_face */
} oracle_refl_t;

/* Per-(ray, face) stepping rule (ray_tracing_basic.c:418-1377). */
typedef struct {
    double  d[3];        /* per-increment delta, (z, y, x): coord + inc*d */
    int32_t rounding;    /* 0 floor, 1 round-half-away */
    int32_t axis;        /* bounding axis 0 z, 1 y, 2 x */
    int32_t to_max;      /* 1 when the ray leaves through the upper plane of axis */
    int32_t inclusive;   /* loop bound inclusive */
    int32_t valid;       /* the reference's assert for this face/branch holds */
    int32_t vacuum_run;  /* full_iteration vacuum-run limit (50, or 10 for face 1) */
} oracle_rule_t;

void    oracle_kp_rotation(const double axis[3], double theta, double out[9]);
void    oracle_thetaphi(const double v[3], int64_t L1, double *theta, double *phi);
void    oracle_myframe(const double v[3], double out[3]);
int64_t oracle_cube_face(const int64_t origin[3], const double dir[3],
                         const int64_t shape[3], int L1);
void    oracle_face_rule(double theta, double phi, int64_t face, oracle_rule_t *rule);
void    oracle_reflection_setup(const double *s1, const double *omega, int64_t R,
                                const double *raw_xray, const double *omega_axis,
                                const double *F, oracle_refl_t *out);

/*
 * Full path: equivalent of ray_tracing_overall (ray_tracing.c:365-425) on a
 * contiguous int8/uint8 volume [z][y][x].  result[R] receives the per
 * reflection mean (NaN for a reflection on which the reference would abort);
 * rays (optional, R*N*2 records, index (r*N + i)*2 + k with k=0 incident and
 * k=1 diffracted) receives per-ray diagnostics; *steps (optional) the sum of
 * len_path_2 over all rays.  Returns the number of failed reflections.
 */
int64_t oracle_trace(const int64_t *coords, int64_t n_coords,
                     const double *s1, const double *omega, int64_t R,
                     const double *raw_xray, const double *omega_axis, const double *F,
                     const double *voxel_size, const double *coefficients,
                     const uint8_t *volume, const int64_t shape[3],
                     int full_iteration, int is_exp, int n_threads,
                     double *result, oracle_ray_t *rays, int64_t *steps);

/* Equivalent of ray_tracing_single (ray_tracing.c:166-292): already rotated
 * vectors; returns the mean (NaN on an error). */
double oracle_single_ray(const int64_t coord[3], const double rotated_s1[3], double theta,
                         double phi, const double *voxel_size, const double *coefficients,
                         const uint8_t *volume, const int64_t shape[3], int full_iteration,
                         int is_exp, oracle_ray_t *ray);
void   oracle_gridding_map(const int64_t *coords, int64_t n_coords, const double *rotated_s1,
                           const double *theta_phi, int64_t D, const double *voxel_size,
                           const double *coefficients, const uint8_t *volume,
                           const int64_t shape[3], int full_iteration, int is_exp,
                           int n_threads, float *map);
double oracle_single(const int64_t *coords, int64_t n_coords,
                     const double rotated_s1[3], const double xray[3],
                     const double *voxel_size, const double *coefficients,
                     const uint8_t *volume, const int64_t shape[3],
                     int full_iteration, int is_exp, oracle_ray_t *rays, int64_t *steps);

/*
 * Iterative bisection (ib_test, ray_tracing.c:25-160; bisection.c:21-471).
 * Labels outside the volume read as 0.  Diagnostics: label reads made, reads
 * outside the volume, and those of them whose value in the reference (its
 * unchecked row-pointer table) would not be 0 or is not defined.
 */
typedef struct {
    int64_t reads;
    int64_t oob;
    int64_t undefined;
} oracle_ib_diag_t;

int     oracle_ib_ratios(int64_t face, double theta, double phi, double r[3]);
/* ib_test on already rotated vectors; lengths (optional) [n_coords][2][4]. */
double  oracle_ib_test(const int64_t *coords, int64_t n_coords, const double rotated_s1[3],
                       const double xray[3], const double *voxel_size,
                       const double *coefficients, const uint8_t *volume,
                       const int64_t shape[3], int num_cls, int is_exp, double *lengths,
                       oracle_ib_diag_t *diag);
/* Rotation (ray_tracing.c:393-407) then ib_test per reflection; returns the
 * number of reflections on which the reference aborts (result NaN). */
int64_t oracle_ib_overall(const int64_t *coords, int64_t n_coords, const double *s1,
                          const double *omega, int64_t R, const double *raw_xray,
                          const double *omega_axis, const double *F, const double *voxel_size,
                          const double *coefficients, const uint8_t *volume,
                          const int64_t shape[3], int num_cls, int is_exp, int n_threads,
                          double *result, oracle_ib_diag_t *diag);

/* The GPU's quotient sequence (two fused corrections of RN(a * RN(1 / b)))
 * against IEEE division on n seeded cases of `kind` (0: integer / double,
 * 1: integer span / integer D); returns the mismatches, *one_step those of a
 * single correction. */
int64_t oracle_quot_rn_check(int kind, int64_t n, uint64_t seed, int64_t *one_step);

#ifdef __cplusplus
}
#endif
#endif
