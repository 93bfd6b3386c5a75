/*
 * anacor_oracle.c -- CPU restatement (FP64) of AnACor's analytical
 * absorption-correction ray tracer.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only tests/,
 * __graft_entry__.smoke() and the CPU-baseline leg of bench.py load it.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp (oracle/Makefile).  No FMA
 * contraction, glibc libm, so every FP64 operation rounds exactly as in the
 * reference binary built with AnACor/src/Makefile:36,47-60 (gcc -O3, x86-64
 * baseline ISA, which has no FMA instruction to contract into).
 *
 * Arithmetic details that matter for bit-exactness, all checked against the
 * reference build (oracle/_ref) through the golden fixtures it writes
 * (tests/golden/make_golden.py), compared in tests/test_oracle_golden.py:
 *   - pow(v, 2) in the reference is folded to v*v by gcc at -O3; glibc pow(v, 2)
 *     differs from v*v in ~0.1% of inputs, so the restatement writes v*v.
 *   - double -> int64 conversion of an out-of-range or NaN value yields
 *     INT64_MIN on x86-64 (cvttsd2si); it is then clamped to 0.
 *   - "x + -1*inc*r" (int64 -1*inc, then *r) equals x + inc*(-r) and
 *     "y - inc*r" equals y + inc*(-r) in IEEE round-to-nearest, so every axis
 *     is evaluated as  coord + (double)inc * d  with the sign folded into d.
 *
 * Pinning status (what checks each function besides the reference build):
 *   oracle_kp_rotation   closed form (Rodrigues) + golden bits
 *   oracle_thetaphi      golden bits from the reference build
 *   oracle_myframe       closed form (permutation, signed-zero behaviour)
 *   oracle_cube_face     brute-force slab test + golden faces
 *   oracle_face_rule     golden per-ray paths (no closed form exists)
 *   walk / lengths       golden per-ray (face, len, last coord, lengths);
 *                        single-material cube closed form
 *   oracle_trace         golden A_h bits; mu=0 => 1; analytic sphere/cylinder
 *                        tables at percent level
 */
#include "anacor_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* The reference's own literal (ray_tracing.h:10). */
#define ORACLE_PI 3.14159265358979323846

/* ---------------------------------------------------------------- rotation */

/* matrices.h:49-68 (Rodrigues rotation, same expression trees). */
void oracle_kp_rotation(const double axis[3], double theta, double out[9])
{
    const double x = axis[0], y = axis[1], z = axis[2];
    const double c = cos(theta);
    const double s = sin(theta);
    out[0] = c + (x * x) * (1 - c);
    out[1] = x * y * (1 - c) - z * s;
    out[2] = y * s + x * z * (1 - c);
    out[3] = z * s + x * y * (1 - c);
    out[4] = c + (y * y) * (1 - c);
    out[5] = -x * s + y * z * (1 - c);
    out[6] = -y * s + x * z * (1 - c);
    out[7] = x * s + y * z * (1 - c);
    out[8] = c + (z * z) * (1 - c);
}

/* matrices.h:14-27: C[m x p] = A[m x n] B[n x p], each entry accumulated from
 * 0.0 in k order. */
static void mat_mul(const double *A, const double *B, double *C, int m, int n, int p)
{
    for (int i = 0; i < m; i++)
        for (int j = 0; j < p; j++) {
            double acc = 0.0;
            for (int k = 0; k < n; k++)
                acc += A[i * n + k] * B[k * p + j];
            C[i * p + j] = acc;
        }
}

/* ------------------------------------------------------------------ angles */

/* ray_tracing_basic.c:13-50 (dials_2_thetaphi_22).  The incident ray (L1) is
 * negated first.  pow(v,2) is written v*v (see header comment). */
void oracle_thetaphi(const double v_in[3], int64_t L1, double *theta, double *phi)
{
    double v[3] = {v_in[0], v_in[1], v_in[2]};
    if (L1 == 1) {
        v[0] = -v[0];
        v[1] = -v[1];
        v[2] = -v[2];
    }
    if (v[1] == 0) {
        *theta = atan(-v[2] / (-sqrt(v[0] * v[0] + v[1] * v[1]) + 0.001));
        *phi = atan(-v[0] / (v[1] + 0.001));
    } else if (v[1] < 0) {
        *theta = atan(-v[2] / sqrt(v[0] * v[0] + v[1] * v[1]));
        *phi = atan(-v[0] / (v[1]));
    } else {
        const double a = atan(-v[2] / sqrt(v[0] * v[0] + v[1] * v[1]));
        *theta = (v[2] < 0) ? ORACLE_PI - a : -ORACLE_PI - a;
        *phi = -atan(-v[0] / (-v[1]));
    }
}

/* ray_tracing_basic.c:52-67: permutation (v0, v2, v1) evaluated as a 3x3
 * product accumulated from 0.0, which turns -0.0 into +0.0 and 0*inf into
 * NaN exactly as the reference does. */
void oracle_myframe(const double v[3], double out[3])
{
    static const double P[3][3] = {{1, 0, 0}, {0, 0, 1}, {0, 1, 0}};
    for (int i = 0; i < 3; i++) {
        double acc = 0.0;
        for (int j = 0; j < 3; j++)
            acc += P[i][j] * v[j];
        out[i] = acc;
    }
}

/* ------------------------------------------------------------- exit plane */

/* ray_tracing_basic.c:69-156 (cube_face).  Slab distances to the six planes
 * 0 and shape (not shape-1); negated for the incident ray; the smallest
 * non-negative distance wins, ties resolved in the order
 * x_min(6), x_max(1), y_min(4), y_max(5), z_min(2), z_max(3).
 * Returns 0 where the reference calls exit(EXIT_FAILURE). */
int64_t oracle_cube_face(const int64_t origin[3], const double dir[3],
                         const int64_t shape[3], int L1)
{
    double t[6];
    /* storage order of the reference's candidate list: x_min y_min z_min x_max y_max z_max */
    t[0] = (0 - origin[2]) / dir[2];
    t[1] = (0 - origin[1]) / dir[1];
    t[2] = (0 - origin[0]) / dir[0];
    t[3] = (shape[2] - origin[2]) / dir[2];
    t[4] = (shape[1] - origin[1]) / dir[1];
    t[5] = (shape[0] - origin[0]) / dir[0];
    if (L1)
        for (int i = 0; i < 6; i++)
            t[i] = -t[i];

    int have = 0;
    double tmin = 0.0;
    for (int i = 0; i < 6; i++) {
        if (t[i] >= 0) {
            if (!have || t[i] < tmin)
                tmin = t[i];
            have = 1;
        }
    }
    if (!have)
        return 0; /* reference reads an uninitialised value and exits */
    if (tmin == t[0]) return 6;
    if (tmin == t[3]) return 1;
    if (tmin == t[1]) return 4;
    if (tmin == t[4]) return 5;
    if (tmin == t[2]) return 2;
    if (tmin == t[5]) return 3;
    return 0;
}

/* ------------------------------------------------------------ face rules */

/* The per-face stepping rules of cal_coord, ray_tracing_basic.c:418-1377.
 * Each axis advances as coord + inc*d (d carries the reference's sign), the
 * rounding is floor or round-half-away, and the loop runs to the exit plane
 * with the reference's inclusive/exclusive bound. */
void oracle_face_rule(double theta, double phi, int64_t face, oracle_rule_t *r)
{
    const double half_pi = ORACLE_PI / 2;
    const int narrow = fabs(theta) < half_pi;   /* |theta| < pi/2 */
    memset(r, 0, sizeof(*r));
    r->vacuum_run = 50;                          /* va_threshold, :12 */
    r->rounding = (theta > 0) ? 0 : 1;           /* floor when theta > 0 */
    switch (face) {
    case 6: { /* x -> 0, :418-511 */
        const double ry = tan(theta) / cos(phi);
        const double rz = tan(phi);
        r->valid = fabs(theta) <= half_pi;
        r->d[0] = rz; r->d[1] = -ry; r->d[2] = -1.0;
        r->axis = 2; r->to_max = 0; r->inclusive = 1;
        break;
    }
    case 1: { /* x -> x_max, :1280-1377 */
        const double ry = tan(ORACLE_PI - theta) / cos(fabs(phi));
        const double rz = tan(phi);
        r->valid = fabs(theta) >= half_pi;
        r->d[0] = -rz; r->d[1] = -ry; r->d[2] = 1.0;
        r->axis = 2; r->to_max = 1; r->inclusive = 0;
        r->vacuum_run = 10;                      /* :1352 */
        break;
    }
    case 2: { /* z -> 0, :512-710 */
        const double rx = 1 / tan(fabs(phi));
        if (narrow) {
            const double ry = tan(theta) / sin(fabs(phi));
            r->valid = phi < 0;
            r->d[0] = -1.0; r->d[1] = -ry; r->d[2] = -rx;
        } else {
            const double ry = tan(ORACLE_PI - theta) / sin(fabs(phi));
            r->valid = phi > 0;
            r->d[0] = -1.0; r->d[1] = -ry; r->d[2] = rx;
        }
        r->axis = 0; r->to_max = 0; r->inclusive = 1;
        break;
    }
    case 3: { /* z -> z_max, :712-897 */
        const double rx = 1 / tan(fabs(phi));
        if (narrow) {
            const double ry = tan(theta) / sin(fabs(phi));
            r->valid = phi > 0;
            r->d[0] = 1.0; r->d[1] = -ry; r->d[2] = -rx;
            r->inclusive = 0;
        } else {
            const double ry = tan(ORACLE_PI - theta) / sin(fabs(phi));
            r->valid = phi < 0;
            r->d[0] = 1.0; r->d[1] = -ry; r->d[2] = rx;
            r->inclusive = 1;
        }
        r->axis = 0; r->to_max = 1;
        break;
    }
    case 4: { /* y -> 0, always floor, :899-1102 */
        r->valid = theta > 0;
        r->rounding = 0;
        if (narrow) {
            const double rx = cos(fabs(phi)) / tan(fabs(theta));
            const double rz = sin(phi) / tan(fabs(theta));
            r->d[0] = rz; r->d[1] = -1.0; r->d[2] = -rx;
        } else {
            const double rx = cos(fabs(phi)) / (tan((ORACLE_PI - fabs(theta))));
            const double rz = sin(-phi) / (tan((ORACLE_PI - fabs(theta))));
            r->d[0] = rz; r->d[1] = -1.0; r->d[2] = rx;
        }
        r->axis = 1; r->to_max = 0; r->inclusive = 1;
        break;
    }
    case 5: { /* y -> y_max, always round, :1104-1278 */
        r->valid = theta < 0;
        r->rounding = 1;
        if (narrow) {
            const double rx = cos(fabs(phi)) / (tan(fabs(theta)));
            const double rz = sin(phi) / (tan(fabs(theta)));
            r->d[0] = rz; r->d[1] = 1.0; r->d[2] = -rx;
            r->inclusive = 0;
        } else {
            const double rx = cos(fabs(phi)) / (tan(ORACLE_PI - fabs(theta)));
            const double rz = sin(phi) / (tan(ORACLE_PI - fabs(theta)));
            r->d[0] = -rz; r->d[1] = 1.0; r->d[2] = rx;
            r->inclusive = 1;
        }
        r->axis = 1; r->to_max = 1;
        break;
    }
    default:
        r->valid = 0;
        break;
    }
}

/* ------------------------------------------------------------------- walk */

/* double -> int64 as x86-64 cvttsd2si converts it (INT64_MIN when the value
 * does not fit), then the clamp to [0, max-1] of every cal_coord branch. */
static inline int64_t clamp_index(double w, int64_t max)
{
    int64_t v;
    if (!(w >= -0x1p63 && w < 0x1p63))
        v = INT64_MIN;
    else
        v = (int64_t)w;
    if (v >= max)
        v = max - 1;
    else if (v < 0)
        v = 0;
    return v;
}

typedef struct {
    int64_t *posi;
    int8_t *cls;
    int64_t cap;
} seg_buf;

static int seg_reserve(seg_buf *b, int64_t need)
{
    if (need <= b->cap)
        return 1;
    int64_t cap = b->cap ? b->cap : 64;
    while (cap < need)
        cap *= 2;
    int64_t *p = (int64_t *)realloc(b->posi, (size_t)cap * sizeof(int64_t));
    if (!p) return 0;
    b->posi = p;
    int8_t *c = (int8_t *)realloc(b->cls, (size_t)cap);
    if (!c) return 0;
    b->cls = c;
    b->cap = cap;
    return 1;
}

/* One ray: cal_coord (ray_tracing_basic.c:396-1460, with appending :343-394)
 * followed by cal_path2_plus (:1462-1552).  Writes the four lengths. */
static void trace_ray(const int64_t c0[3], const oracle_rule_t *rule, int64_t face,
                      const int8_t *vol, const int64_t shape[3],
                      const double *voxel_size, int full_iteration,
                      seg_buf *segs, oracle_ray_t *out)
{
    const int64_t Z = shape[0], Y = shape[1], X = shape[2];
    const int64_t ext[3] = {Z, Y, X};
    const int64_t ax = rule->axis;
    const int64_t limit = rule->to_max ? (ext[ax] - c0[ax]) + rule->inclusive
                                       : c0[ax] + 1;
    const double base[3] = {(double)c0[0], (double)c0[1], (double)c0[2]};

    int64_t len = 1, nseg = 1, va = 0;
    int64_t first[3] = {0, 0, 0}, last[3] = {0, 0, 0};
    int prev = 0;

    seg_reserve(segs, limit > 0 ? limit + 2 : 2);
    segs->posi[0] = 0;      /* classes_posi[0] = 0, :407-408 */
    segs->cls[0] = 3;       /* classes[0] = 3 */

    for (int64_t inc = 0; inc < limit; inc++) {
        int64_t c[3];
        for (int a = 0; a < 3; a++) {
            double w = base[a] + (double)inc * rule->d[a];
            w = rule->rounding ? round(w) : floor(w);
            c[a] = clamp_index(w, ext[a]);
        }
        const int lab = (int)vol[(c[0] * Y + c[1]) * X + c[2]];
        if (!full_iteration) {
            if (lab == 0)
                break;
        } else {
            va = (lab == 0) ? va + 1 : 0;
            if (va > rule->vacuum_run)
                break;
        }
        if (inc == 0) {
            first[0] = c[0]; first[1] = c[1]; first[2] = c[2];
            last[0] = c[0]; last[1] = c[1]; last[2] = c[2];
            prev = lab;
            continue;
        }
        if (lab != prev && lab >= 0 && lab <= 4) {
            segs->posi[nseg] = inc;
            segs->cls[nseg] = (int8_t)lab;
            nseg++;
        }
        last[0] = c[0]; last[1] = c[1]; last[2] = c[2];
        prev = lab;
        len++;
    }

    double L[4] = {0.0, 0.0, 0.0, 0.0}; /* li, lo, cr, bu */
    if (len != 1) {
        const double dy = (double)(last[1] - first[1]) * voxel_size[1];
        const double dz = (double)(last[0] - first[0]) * voxel_size[0];
        const double dx = (double)(last[2] - first[2]) * voxel_size[2];
        const double T = sqrt(dy * dy + dz * dz + dx * dx);
        for (int64_t j = 0; j < nseg; j++) {
            const int64_t span = (j < nseg - 1) ? segs->posi[j + 1] - segs->posi[j]
                                                : len - segs->posi[j];
            const double term = T * ((double)span / (double)(len - 1));
            switch (segs->cls[j]) {
            case 1: L[0] += term; break;
            case 2: L[1] += term; break;
            case 3: L[2] += term; break;
            case 4: L[3] += term; break;
            default: break;
            }
        }
    }
    if (out) {
        out->face = (int32_t)face;
        out->len = len;
        out->nseg = nseg;
        memcpy(out->first, first, sizeof(first));
        memcpy(out->last, last, sizeof(last));
        memcpy(out->L, L, sizeof(L));
    }
}

/* cal_rate, ray_tracing_basic.c:1554-1583 (same association order). */
static double rate(const double *n1, const double *n2, const double *mu, int is_exp)
{
    double r = (mu[0] * (n1[0] + n2[0]) +
                mu[1] * (n1[1] + n2[1]) +
                mu[2] * (n1[2] + n2[2]) +
                mu[3] * (n1[3] + n2[3]));
    if ((char)is_exp == 1)
        r = exp(-r);
    return r;
}

/* ray_tracing_single, ray_tracing.c:166-292, on already computed angles and
 * directions.  Returns the mean; *err receives the first error code. */
static double single_core(const int64_t *coords, int64_t n, const double theta[2],
                          const double phi[2], const double dir[2][3],
                          const double *voxel_size, const double *mu,
                          const int8_t *vol, const int64_t shape[3],
                          int full_iteration, int is_exp, seg_buf *segs,
                          oracle_ray_t *rays, int64_t *steps, int *err)
{
    double sum = 0.0;
    int64_t st = 0;
    *err = ORACLE_OK;
    for (int64_t i = 0; i < n; i++) {
        const int64_t c[3] = {coords[i * 3], coords[i * 3 + 1], coords[i * 3 + 2]};
        oracle_ray_t rr[2];
        memset(rr, 0, sizeof(rr));
        /* cube_face for both rays runs before either walk (:206-207) */
        const int64_t face[2] = {oracle_cube_face(c, dir[0], shape, 1),
                                 oracle_cube_face(c, dir[1], shape, 0)};
        if (face[0] == 0 || face[1] == 0) {
            *err = ORACLE_ERR_FACE;
            if (rays) {
                rays[i * 2].err = rays[i * 2 + 1].err = ORACLE_ERR_FACE;
            }
            return NAN;
        }
        for (int k = 0; k < 2; k++) {
            oracle_rule_t rule;
            oracle_face_rule(theta[k], phi[k], face[k], &rule);
            if (!rule.valid) {
                *err = ORACLE_ERR_ASSERT;
                if (rays) rays[i * 2 + k].err = ORACLE_ERR_ASSERT;
                return NAN;
            }
            trace_ray(c, &rule, face[k], vol, shape, voxel_size, full_iteration, segs, &rr[k]);
            st += rr[k].len;
        }
        sum += rate(rr[0].L, rr[1].L, mu, is_exp);
        if (rays) {
            rays[i * 2] = rr[0];
            rays[i * 2 + 1] = rr[1];
        }
    }
    if (steps) *steps = st;
    return sum / n;
}

/* ray_tracing.c:393-407 (rotation) then :182-197 (angles, directions). */
static void setup_one(const double s1_in[3], double omega, const double *raw_xray,
                      const double *omega_axis, const double *F, oracle_refl_t *o)
{
    double Rw[9], M[9], T[9];
    oracle_kp_rotation(omega_axis, omega, Rw);
    mat_mul(Rw, F, M, 3, 3, 3);
    for (int i = 0; i < 3; i++)          /* transpose, matrices.h:6-12 */
        for (int j = 0; j < 3; j++)
            T[j * 3 + i] = M[i * 3 + j];
    mat_mul(T, raw_xray, o->xray, 3, 3, 1);
    mat_mul(T, s1_in, o->s1, 3, 3, 1);
    oracle_thetaphi(o->xray, 1, &o->theta[0], &o->phi[0]);
    oracle_thetaphi(o->s1, 0, &o->theta[1], &o->phi[1]);
    oracle_myframe(o->xray, o->dir[0]);
    oracle_myframe(o->s1, o->dir[1]);
}

void oracle_reflection_setup(const double *s1, const double *omega, int64_t R,
                             const double *raw_xray, const double *omega_axis,
                             const double *F, oracle_refl_t *out)
{
    for (int64_t i = 0; i < R; i++)
        setup_one(&s1[i * 3], omega[i], raw_xray, omega_axis, F, &out[i]);
}

int64_t oracle_trace(const int64_t *coords, int64_t n_coords,
                     const double *s1, const double *omega, int64_t R,
                     const double *raw_xray, const double *omega_axis, const double *F,
                     const double *voxel_size, const double *coefficients,
                     const uint8_t *volume, const int64_t shape[3],
                     int full_iteration, int is_exp, int n_threads,
                     double *result, oracle_ray_t *rays, int64_t *steps)
{
    int64_t failed = 0, total = 0;
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel num_threads(n_threads) reduction(+ : failed, total)
    {
        seg_buf segs = {NULL, NULL, 0};
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < R; i++) {
            oracle_refl_t g;
            setup_one(&s1[i * 3], omega[i], raw_xray, omega_axis, F, &g);
            int err = 0;
            int64_t st = 0;
            result[i] = single_core(coords, n_coords, g.theta, g.phi,
                                    (const double(*)[3])g.dir, voxel_size, coefficients,
                                    (const int8_t *)volume, shape, full_iteration, is_exp,
                                    &segs, rays ? rays + i * n_coords * 2 : NULL, &st, &err);
            if (err) failed++;
            total += st;
        }
        free(segs.posi);
        free(segs.cls);
    }
    if (steps) *steps = total;
    return failed;
}

double oracle_single(const int64_t *coords, int64_t n_coords,
                     const double rotated_s1[3], const double xray[3],
                     const double *voxel_size, const double *coefficients,
                     const uint8_t *volume, const int64_t shape[3],
                     int full_iteration, int is_exp, oracle_ray_t *rays, int64_t *steps)
{
    double theta[2], phi[2], dir[2][3];
    oracle_thetaphi(xray, 1, &theta[0], &phi[0]);
    oracle_thetaphi(rotated_s1, 0, &theta[1], &phi[1]);
    oracle_myframe(xray, dir[0]);
    oracle_myframe(rotated_s1, dir[1]);
    seg_buf segs = {NULL, NULL, 0};
    int err = 0;
    double m = single_core(coords, n_coords, theta, phi, (const double(*)[3])dir,
                           voxel_size, coefficients, (const int8_t *)volume, shape,
                           full_iteration, is_exp, &segs, rays, steps, &err);
    free(segs.posi);
    free(segs.cls);
    return m;
}

/* ray_tracing_single_gridding, ray_tracing.c:484-534: one ray from `coord`
 * along rotated_s1 (the diffracted-ray convention, L1 = 0) walked with the
 * given angles; cal_rate with zero incident lengths.  NaN where the reference
 * would abort. */
double oracle_single_ray(const int64_t coord[3], const double rotated_s1[3], double theta,
                         double phi, const double *voxel_size, const double *coefficients,
                         const uint8_t *volume, const int64_t shape[3], int full_iteration,
                         int is_exp, oracle_ray_t *ray)
{
    double dir[3];
    oracle_myframe(rotated_s1, dir);
    const int64_t face = oracle_cube_face(coord, dir, shape, 0);
    if (face == 0) return NAN;
    oracle_rule_t rule;
    oracle_face_rule(theta, phi, face, &rule);
    if (!rule.valid) return NAN;
    seg_buf segs = {NULL, NULL, 0};
    oracle_ray_t rr;
    memset(&rr, 0, sizeof(rr));
    trace_ray(coord, &rule, face, (const int8_t *)volume, shape, voxel_size, full_iteration,
              &segs, &rr);
    free(segs.posi);
    free(segs.cls);
    if (ray) *ray = rr;
    const double zero[4] = {0.0, 0.0, 0.0, 0.0};
    return rate(zero, rr.L, coefficients, is_exp);
}

/* The absorption map of utils/utils_gridding.py:144-231: for D directions
 * (rotated_s1[D][3] with their angles theta_phi[D][2]) and every crystal
 * voxel, oracle_single_ray stored as float, map[D][n_coords]. */
void oracle_gridding_map(const int64_t *coords, int64_t n_coords, const double *rotated_s1,
                         const double *theta_phi, int64_t D, const double *voxel_size,
                         const double *coefficients, const uint8_t *volume,
                         const int64_t shape[3], int full_iteration, int is_exp,
                         int n_threads, float *map)
{
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 1)
    for (int64_t d = 0; d < D; d++) {
        for (int64_t k = 0; k < n_coords; k++) {
            map[d * n_coords + k] = (float)oracle_single_ray(
                coords + 3 * k, rotated_s1 + 3 * d, theta_phi[2 * d], theta_phi[2 * d + 1],
                voxel_size, coefficients, volume, shape, full_iteration, is_exp, NULL);
        }
    }
}


/* ------------------------------------------------------ iterative bisection
 *
 * ib_test (ray_tracing.c:25-160) with increments, bisection,
 * iterative_bisection and cal_path_bisection (bisection.c:21-471), resolution
 * 1.0 (ray_tracing.c:77).  Positions are held as doubles: every value is an
 * integer (floor) that a double holds exactly, so the differences and lengths
 * equal the reference's int64 arithmetic.
 *
 * The reference indexes its row-pointer table without a bounds check; a label
 * outside the volume reads as 0 (vacuum) here.  diag->oob counts such reads,
 * diag->undefined those for which the reference's read is not vacuum: either
 * its row-major wrap-around (label_list[z][y][x] == flat[(z*Y + y)*X + x]
 * while z and z*Y + y index the tables) lands on a non-zero label, or it
 * leaves the tables.  Where undefined == 0 the two agree.
 */

/* increments, bisection.c:21-148: the ratios (z, y, x) with which bisection
 * steps ("z + m*rz", "y - m*ry", "x + m*rx") for exit face `face`.  Returns 0
 * where the reference's assert fails or the face is not 1..6 (it aborts, or
 * goes on with uninitialised ratios). */
int oracle_ib_ratios(int64_t face, double theta, double phi, double r[3])
{
    const double at = fabs(theta), ap = fabs(phi);
    const int narrow = at < ORACLE_PI / 2;
    switch (face) {
    case 4:                                   /* :39-61, y -> 0 */
        if (!(theta > 0)) return 0;
        if (narrow) {
            r[2] = -cos(ap) / tan(at);
            r[0] = sin(phi) / tan(at);
        } else {
            r[2] = cos(ap) / tan(ORACLE_PI - at);
            r[0] = sin(-phi) / tan(ORACLE_PI - at);
        }
        r[1] = 1;
        return 1;
    case 5:                                   /* :62-84, y -> y_max */
        if (!(theta < 0)) return 0;
        if (narrow) {
            r[2] = -cos(ap) / tan(at);
            r[0] = sin(phi) / tan(at);
        } else {
            r[2] = cos(ap) / tan(ORACLE_PI - at);
            r[0] = -sin(phi) / tan(ORACLE_PI - at);
        }
        r[1] = -1;
        return 1;
    case 6:                                   /* :85-95, x -> 0 */
        if (!(at <= ORACLE_PI / 2)) return 0;
        r[2] = -1;
        r[1] = tan(theta) / cos(phi);
        r[0] = tan(phi);
        return 1;
    case 1:                                   /* :96-106, x -> x_max */
        if (!(at >= ORACLE_PI / 2)) return 0;
        r[2] = 1;
        r[1] = tan(ORACLE_PI - theta) / cos(ap);
        r[0] = -tan(phi);
        return 1;
    case 2:                                   /* :107-124, z -> 0 */
    case 3:                                   /* :125-142, z -> z_max */
        if (narrow) {
            if (face == 2 ? !(phi < 0) : !(phi > 0)) return 0;
            r[2] = -1 / tan(ap);
            r[1] = tan(theta) / sin(ap);
        } else {
            if (face == 2 ? !(phi > 0) : !(phi < 0)) return 0;
            r[2] = 1 / tan(ap);
            r[1] = tan(ORACLE_PI - theta) / sin(ap);
        }
        r[0] = face == 2 ? -1 : 1;
        return 1;
    default:
        return 0;
    }
}

/* AirLongest of increments: the distance, in steps of the face axis, from
 * the voxel to the last plane before the face (extent - 1 at the top). */
static double ib_air_longest(int64_t face, const int64_t c[3], const int64_t shape[3])
{
    switch (face) {
    case 4: return (double)c[1];
    case 5: return (double)(shape[1] - 1 - c[1]);
    case 6: return (double)c[2];
    case 1: return (double)(shape[2] - 1 - c[2]);
    case 2: return (double)c[0];
    default: return (double)(shape[0] - 1 - c[0]);
    }
}

static int ib_label(const int8_t *vol, const int64_t shape[3], double z, double y, double x,
                    oracle_ib_diag_t *diag)
{
    const double Z = (double)shape[0], Y = (double)shape[1], X = (double)shape[2];
    if (z >= 0 && z < Z && y >= 0 && y < Y && x >= 0 && x < X)
        return vol[((int64_t)z * shape[1] + (int64_t)y) * shape[2] + (int64_t)x];
    if (diag) {
        diag->oob++;
        int defined = 0, v = 0;
        if (z >= 0 && z < Z) {
            const double row = z * Y + y;
            const double flat = row * X + x;
            if (row >= 0 && row < Z * Y && flat >= 0 && flat < Z * Y * X) {
                defined = 1;
                v = vol[(int64_t)flat];
            }
        }
        if (!defined || v != 0) diag->undefined++;
    }
    return 0;
}

typedef struct { double p[3]; double mid; } ib_point_t;

/* bisection, bisection.c:156-235: halve [shortest, longest] until it is at
 * most one step wide.  `outer` = 0 moves `longest` in while the sample point
 * holds class `cls` (the point where cls begins, seen from outside); outer = 1
 * moves `shortest` out while it does (where cls ends, seen from the voxel).
 * The point returned is the voxel plus floor(mid) steps. */
static ib_point_t ib_search(double longest, double shortest, const double r[3],
                            const int64_t c[3], const int8_t *vol, const int64_t shape[3],
                            int outer, int cls, oracle_ib_diag_t *diag)
{
    const double z = (double)c[0], y = (double)c[1], x = (double)c[2];
    double diff = longest - shortest;
    double mid = (longest + shortest) / 2.0;
    while (diff > 1.0) {
        mid = (longest + shortest) / 2.0;
        const int label = ib_label(vol, shape, floor(z + mid * r[0]), floor(y - mid * r[1]),
                                   floor(x + mid * r[2]), diag);
        const int hit = (int8_t)label == cls;
        if (hit == outer)
            shortest = mid;
        else
            longest = mid;
        diff = longest - shortest;
        if (diag) diag->reads++;
    }
    const double m = floor(mid);
    ib_point_t out = {{floor(z + m * r[0]), floor(y - m * r[1]), floor(x + m * r[2])}, mid};
    return out;
}

/* |a - b|_1 summed z, y, x as the reference's three fabs terms. */
static double ib_l1(const double *a, const double *b)
{
    return fabs(a[0] - b[0]) + fabs(a[1] - b[1]) + fabs(a[2] - b[2]);
}

/* difference_length, bisection.c:244-253 (pow(v, 2) folded to v*v). */
static double ib_dist(const double *a, const double *b, const double *vs)
{
    const double dz = (a[0] - b[0]) * vs[0];
    const double dy = (a[1] - b[1]) * vs[1];
    const double dx = (a[2] - b[2]) * vs[2];
    return sqrt(dz * dz + dy * dy + dx * dx);
}

/* iterative_bisection (bisection.c:298-471) then cal_path_bisection
 * (:255-296) for one ray: L = [li, lo, cr, bu].  Returns an ORACLE_* code. */
static int ib_ray(double theta, double phi, const int64_t c[3], int64_t face,
                  const int8_t *vol, const int64_t shape[3], int num_cls,
                  const double *vs, double L[4], oracle_ib_diag_t *diag)
{
    double r[3];
    if (face == 0) return ORACLE_ERR_FACE;
    if (!oracle_ib_ratios(face, theta, phi, r)) return ORACLE_ERR_ASSERT;
    double pts[7][3];
    int n = 1;
    pts[0][0] = (double)c[0];
    pts[0][1] = (double)c[1];
    pts[0][2] = (double)c[2];
    const ib_point_t A = ib_search(ib_air_longest(face, c, shape), 0, r, c, vol, shape, 0, 0,
                                   diag);
    const ib_point_t C = ib_search(A.mid, 0, r, c, vol, shape, 1, 3, diag);
    memcpy(pts[n++], A.p, sizeof(A.p));
    memcpy(pts[n++], C.p, sizeof(C.p));
    do {
        /* loop (class 2): inner boundary, then the outer one; a pair closer
         * than two voxels per axis on average ends the search (:355-402) */
        const ib_point_t Li = ib_search(A.mid, C.mid, r, c, vol, shape, 0, 2, diag);
        if (ib_l1(Li.p, A.p) / 3 > 1) {
            const ib_point_t Lo = ib_search(A.mid, Li.mid, r, c, vol, shape, 1, 2, diag);
            if (ib_l1(Li.p, Lo.p) / 3 <= 2) break;
            memcpy(pts[n++], Li.p, sizeof(Li.p));
            memcpy(pts[n++], Lo.p, sizeof(Lo.p));
        }
        /* bubble (class 4), num_cls > 3 only; its first test has no /3
         * (:405-448) */
        if (num_cls > 3) {
            const ib_point_t Bi = ib_search(A.mid, C.mid, r, c, vol, shape, 0, 4, diag);
            if (ib_l1(Bi.p, A.p) > 1) {
                const ib_point_t Bo = ib_search(A.mid, Bi.mid, r, c, vol, shape, 1, 4, diag);
                if (ib_l1(Bi.p, Bo.p) / 3 <= 2) break;
                memcpy(pts[n++], Bi.p, sizeof(Bi.p));
                memcpy(pts[n++], Bo.p, sizeof(Bo.p));
            }
        }
    } while (0);
    /* cal_path_bisection: with seven points the loop length stays 0 and the
     * bubble pair is measured; with five, whichever pair was found is taken
     * as the loop */
    const double total = ib_dist(pts[0], pts[1], vs);
    const double cr = ib_dist(pts[0], pts[2], vs);
    double lo = 0, bu = 0;
    const double air = 0;
    if (n == 5)
        lo = ib_dist(pts[3], pts[4], vs);
    else if (n == 7)
        bu = ib_dist(pts[5], pts[6], vs);
    L[0] = total - lo - bu - cr - air;
    L[1] = lo;
    L[2] = cr;
    L[3] = bu;
    return ORACLE_OK;
}

static double ib_core(const int64_t *coords, int64_t n, const double theta[2],
                      const double phi[2], const double dir[2][3], const double *vs,
                      const double *mu, const int8_t *vol, const int64_t shape[3],
                      int num_cls, int is_exp, double *lengths, oracle_ib_diag_t *diag,
                      int *err)
{
    double sum = 0.0;
    *err = ORACLE_OK;
    for (int64_t i = 0; i < n; i++) {
        const int64_t c[3] = {coords[i * 3], coords[i * 3 + 1], coords[i * 3 + 2]};
        const int64_t face[2] = {oracle_cube_face(c, dir[0], shape, 1),
                                 oracle_cube_face(c, dir[1], shape, 0)};
        double L[2][4];
        for (int k = 0; k < 2; k++) {
            const int e = ib_ray(theta[k], phi[k], c, face[k], vol, shape, num_cls, vs, L[k],
                                 diag);
            if (e) {
                *err = e;
                return NAN;
            }
        }
        if (lengths) memcpy(lengths + i * 8, L, sizeof(L));
        sum += rate(L[0], L[1], mu, is_exp);
    }
    return sum / n;
}

double oracle_ib_test(const int64_t *coords, int64_t n_coords, const double rotated_s1[3],
                      const double xray[3], const double *voxel_size,
                      const double *coefficients, const uint8_t *volume,
                      const int64_t shape[3], int num_cls, int is_exp, double *lengths,
                      oracle_ib_diag_t *diag)
{
    double theta[2], phi[2], dir[2][3];
    oracle_thetaphi(xray, 1, &theta[0], &phi[0]);
    oracle_thetaphi(rotated_s1, 0, &theta[1], &phi[1]);
    oracle_myframe(xray, dir[0]);
    oracle_myframe(rotated_s1, dir[1]);
    int err = 0;
    return ib_core(coords, n_coords, theta, phi, (const double(*)[3])dir, voxel_size,
                   coefficients, (const int8_t *)volume, shape, num_cls, is_exp, lengths, diag,
                   &err);
}

int64_t oracle_ib_overall(const int64_t *coords, int64_t n_coords, const double *s1,
                          const double *omega, int64_t R, const double *raw_xray,
                          const double *omega_axis, const double *F, const double *voxel_size,
                          const double *coefficients, const uint8_t *volume,
                          const int64_t shape[3], int num_cls, int is_exp, int n_threads,
                          double *result, oracle_ib_diag_t *diag)
{
    int64_t failed = 0, oob = 0, undef = 0, reads = 0;
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 1) \
    reduction(+ : failed, oob, undef, reads)
    for (int64_t i = 0; i < R; i++) {
        oracle_refl_t g;
        setup_one(&s1[i * 3], omega[i], raw_xray, omega_axis, F, &g);
        oracle_ib_diag_t d = {0, 0, 0};
        int err = 0;
        result[i] = ib_core(coords, n_coords, g.theta, g.phi, (const double(*)[3])g.dir,
                            voxel_size, coefficients, (const int8_t *)volume, shape, num_cls,
                            is_exp, NULL, &d, &err);
        if (err) failed++;
        oob += d.oob;
        undef += d.undefined;
        reads += d.reads;
    }
    if (diag) {
        diag->oob = oob;
        diag->undefined = undef;
        diag->reads = reads;
    }
    return failed;
}

/* ------------------------------------------- the device's quotient sequence */

/* The GPU kernels compute RN(a / b) from a per-ray reciprocal y = RN(1 / b)
 * by two fused corrections of q0 = RN(a * y) (quot_rn in csrc/trace.hip).
 * This replays that sequence with C99 fma() (correctly rounded, as the
 * device's v_fma_f64) against IEEE division on n seeded cases of each kind:
 *   kind 0: a an integer with |a| < 2^21, b a double of any sign with
 *           exponent in [-900, 900] (cube_face's (bound - origin) / dir);
 *   kind 1: b an integer in [1, 2^21), a an integer in [0, b + 1]
 *           (cal_path2_plus' span / (len - 1)).
 * Significands are drawn near 1 and near 2 as well as uniformly, where the
 * error of q0 is largest.  Returns the number of mismatches of the
 * two-correction sequence; *one_step receives those of a single correction
 * (which Markstein's theorem does not cover, since q0 need not be faithful). */
static uint64_t qc_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double qc_significand(uint64_t *s)   /* in [1, 2) */
{
    const uint64_t r = qc_next(s);
    uint64_t frac = r & ((1ull << 52) - 1);
    switch ((r >> 60) & 3) {
    case 0: frac &= 0xFFFFull; break;                          /* just above 1 */
    case 1: frac = ((1ull << 52) - 1) - (frac & 0xFFFFull); break; /* just below 2 */
    default: break;
    }
    double m;
    const uint64_t bits = (1023ull << 52) | frac;
    memcpy(&m, &bits, sizeof m);
    return m;
}

int64_t oracle_quot_rn_check(int kind, int64_t n, uint64_t seed, int64_t *one_step)
{
    uint64_t s = seed;
    int64_t bad = 0, bad1 = 0;
    for (int64_t i = 0; i < n; ++i) {
        double a, b;
        if (kind == 0) {
            const int e = (int)(qc_next(&s) % 1801) - 900;
            b = ldexp(qc_significand(&s), e);
            if (qc_next(&s) & 1) b = -b;
            /* |a| < 2^21: powers of two and their neighbours often */
            int64_t k;
            const uint64_t r = qc_next(&s);
            if ((r & 3) == 0) k = ((int64_t)1 << (r >> 2) % 21) - (int64_t)((r >> 8) % 3);
            else k = (int64_t)((r >> 2) % ((1u << 21) - 1));
            a = (double)((r >> 40) & 1 ? -k : k);
        } else {
            const uint64_t r = qc_next(&s);
            int64_t D = (r & 1) ? (int64_t)((r >> 1) % ((1u << 21) - 1)) + 1
                                : ((int64_t)1 << (r >> 1) % 21) + (int64_t)((r >> 8) % 3) - 1;
            if (D < 1) D = 1;
            const int64_t span = (int64_t)(qc_next(&s) % (uint64_t)(D + 2));
            a = (double)span;
            b = (double)D;
        }
        const double y = 1.0 / b;
        const double q0 = a * y;
        const double q1 = fma(fma(-q0, b, a), y, q0);
        const double q2 = fma(fma(-q1, b, a), y, q1);
        const double ref = a / b;
        if (memcmp(&q2, &ref, sizeof ref) != 0) ++bad;
        if (memcmp(&q1, &ref, sizeof ref) != 0) ++bad1;
    }
    if (one_step) *one_step = bad1;
    return bad;
}
