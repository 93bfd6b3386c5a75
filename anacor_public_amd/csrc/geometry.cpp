// Host-side FP64 setup of each reflection: goniometer rotation, the two beam
// angles and the six per-face stepping rules of each ray.
//
// This is O(R) work done once per reflection with glibc libm, so the angles
// and per-increment ratios carry exactly the bits the reference computes
// (AnACor/src/ray_tracing.c:393-407, ray_tracing_basic.c:13-67 and the ratio
// expressions of cal_coord, :418-1377).  Built with g++ -ffp-contract=off
// (no FMA contraction) -fno-fast-math; gcc folds the reference's pow(v, 2)
// into v*v, which is written out here.
#include "geometry.h"
#include "host_pool.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace anacor {
namespace {

constexpr double kPi = 3.14159265358979323846;  // ray_tracing.h:10

// Rodrigues rotation about `axis` (matrices.h:49-68).
void rodrigues(const double* axis, double w, double* m) {
  const double x = axis[0], y = axis[1], z = axis[2];
  const double c = std::cos(w), s = std::sin(w);
  m[0] = c + (x * x) * (1 - c);
  m[1] = x * y * (1 - c) - z * s;
  m[2] = y * s + x * z * (1 - c);
  m[3] = z * s + x * y * (1 - c);
  m[4] = c + (y * y) * (1 - c);
  m[5] = -x * s + y * z * (1 - c);
  m[6] = -y * s + x * z * (1 - c);
  m[7] = x * s + y * z * (1 - c);
  m[8] = c + (z * z) * (1 - c);
}

// Row-major 3x3 x 3xP product, each entry summed from 0.0 in k order
// (matrices.h:14-27).
void mul3(const double* a, const double* b, double* c, int p) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < p; ++j) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k) acc += a[i * 3 + k] * b[k * p + j];
      c[i * p + j] = acc;
    }
}

// dials_2_thetaphi_22 (ray_tracing_basic.c:13-50).
void beam_angles(const double* in, bool incident, double* theta, double* phi) {
  double v0 = in[0], v1 = in[1], v2 = in[2];
  if (incident) { v0 = -v0; v1 = -v1; v2 = -v2; }
  if (v1 == 0) {
    *theta = std::atan(-v2 / (-std::sqrt(v0 * v0 + v1 * v1) + 0.001));
    *phi = std::atan(-v0 / (v1 + 0.001));
  } else if (v1 < 0) {
    *theta = std::atan(-v2 / std::sqrt(v0 * v0 + v1 * v1));
    *phi = std::atan(-v0 / v1);
  } else {
    const double a = std::atan(-v2 / std::sqrt(v0 * v0 + v1 * v1));
    *theta = (v2 < 0) ? kPi - a : -kPi - a;
    *phi = -std::atan(-v0 / (-v1));
  }
}

// dials_2_myframe (ray_tracing_basic.c:52-67): (v0, v2, v1) through a 0/1
// matrix product accumulated from +0.0 (keeps its signed-zero and 0*inf
// behaviour, which cube_face's divisions see).
void to_myframe(const double* v, double* out) {
  static const double P[9] = {1, 0, 0, 0, 0, 1, 0, 1, 0};
  for (int i = 0; i < 3; ++i) {
    double acc = 0.0;
    for (int j = 0; j < 3; ++j) acc += P[i * 3 + j] * v[j];
    out[i] = acc;
  }
}

uint32_t pack_meta(bool valid, bool round_half, bool inclusive, bool to_max, int axis, int vac) {
  return (valid ? kMetaValid : 0u) | (round_half ? kMetaRound : 0u) |
         (inclusive ? kMetaInclusive : 0u) | (to_max ? kMetaToMax : 0u) |
         (uint32_t(axis) << kMetaAxisShift) | (uint32_t(vac) << kMetaVacShift);
}

void set_d(RayRule& r, int face, double dz, double dy, double dx) {
  r.d[face - 1][0] = dz;
  r.d[face - 1][1] = dy;
  r.d[face - 1][2] = dx;
  r.d[face - 1][3] = 0.0;
}

static bool same_bits(double a, double b) {
  uint64_t x, y;
  std::memcpy(&x, &a, 8);
  std::memcpy(&y, &b, 8);
  return x == y;
}

// The six exit-face rules of cal_coord for one ray with angles (t, p).
// Axis deltas carry the sign with which the reference combines them
// ("c - inc*r" == "c + inc*(-r)" in IEEE arithmetic).  Each libm value is
// computed once per distinct argument (bit-identical arguments give
// bit-identical results), e.g. cos(|p|) is cos(p) when p >= +0.
void face_rules(double t, double p, RayRule& r) {
  const double half_pi = kPi / 2;
  const bool narrow = std::fabs(t) < half_pi;
  const bool pos = t > 0;  // floor when theta > 0, else round (faces 1,2,3,6)
  const double ap = std::fabs(p), at = std::fabs(t);
  const bool p_is_ap = same_bits(p, ap), t_is_at = same_bits(t, at);

  const double tan_p = std::tan(p);
  const double cos_p = std::cos(p);
  const double tan_t = std::tan(t);
  const double tan_pi_t = std::tan(kPi - t);
  const double cos_ap = p_is_ap ? cos_p : std::cos(ap);
  const double tan_ap = p_is_ap ? tan_p : std::tan(ap);
  const double sin_ap = std::sin(ap);
  const double sin_p = p_is_ap ? sin_ap : std::sin(p);

  // face 6: x -> 0 (:418-511)
  set_d(r, 6, tan_p, -(tan_t / cos_p), -1.0);
  r.meta[5] = pack_meta(at <= half_pi, !pos, true, false, 2, 50);

  // face 1: x -> x_max, exclusive, vacuum run 10 (:1280-1377)
  set_d(r, 1, -tan_p, -(tan_pi_t / cos_ap), 1.0);
  r.meta[0] = pack_meta(at >= half_pi, !pos, false, true, 2, 10);

  // faces 2 (z -> 0, inclusive) and 3 (z -> z_max) (:512-897)
  const double rx = 1 / tan_ap;
  if (narrow) {
    const double ry = tan_t / sin_ap;
    set_d(r, 2, -1.0, -ry, -rx);
    r.meta[1] = pack_meta(p < 0, !pos, true, false, 0, 50);
    set_d(r, 3, 1.0, -ry, -rx);
    r.meta[2] = pack_meta(p > 0, !pos, false, true, 0, 50);
  } else {
    const double ry = tan_pi_t / sin_ap;
    set_d(r, 2, -1.0, -ry, rx);
    r.meta[1] = pack_meta(p > 0, !pos, true, false, 0, 50);
    set_d(r, 3, 1.0, -ry, rx);
    r.meta[2] = pack_meta(p < 0, !pos, true, true, 0, 50);
  }

  // face 4: y -> 0, always floor, inclusive (:899-1102)
  // face 5: y -> y_max, always round (:1104-1278)
  if (narrow) {
    const double tt = t_is_at ? tan_t : std::tan(at);
    const double rx4 = cos_ap / tt;
    set_d(r, 4, sin_p / tt, -1.0, -rx4);
    set_d(r, 5, sin_p / tt, 1.0, -rx4);
    r.meta[4] = pack_meta(t < 0, true, false, true, 1, 50);
  } else {
    const double tt = t_is_at ? tan_pi_t : std::tan(kPi - at);
    const double rx4 = cos_ap / tt;
    const double mp = -p;
    const double sin_mp = same_bits(mp, ap) ? sin_ap : std::sin(mp);
    set_d(r, 4, sin_mp / tt, -1.0, rx4);
    set_d(r, 5, -(sin_p / tt), 1.0, rx4);
    r.meta[4] = pack_meta(t < 0, true, true, true, 1, 50);
  }
  r.meta[3] = pack_meta(t > 0, false, true, false, 1, 50);
}

// True when, on some face the ray may walk (its precondition holds), a
// clamped coordinate could come from a value outside the int64 range (or a
// non-finite one), which the kernels then map to 0 as the reference's
// double->int64 conversion does.  Faces whose precondition fails never walk.
bool needs_guard(const RayRule& r, int64_t max_dim) {
  const double lim = 0x1p62;
  const double m = double(max_dim) + 1.0;
  for (int f = 0; f < 6; ++f) {
    if (!(r.meta[f] & kMetaValid)) continue;
    for (int a = 0; a < 3; ++a) {
      const double d = r.d[f][a];
      if (!std::isfinite(d)) return true;
      if (std::fabs(d) * m + m >= lim) return true;
    }
  }
  return false;
}

// RN(1 / dir) per axis into the spare fourth delta slots of faces 1-3: the
// GPU's cube_face divides by the direction through these reciprocals.
static void set_reciprocals(RayRule& r) {
  for (int a = 0; a < 3; ++a) r.d[a][3] = 1.0 / r.dir[a];
}

void fill_ray(const double* v, bool incident, int64_t max_dim, RayRule& r, double* theta,
              double* phi, double* dir) {
  std::memset(&r, 0, sizeof(r));
  double t, p;
  beam_angles(v, incident, &t, &p);
  double d3[3];
  to_myframe(v, d3);
  r.dir[0] = d3[0];
  r.dir[1] = d3[1];
  r.dir[2] = d3[2];
  r.dir[3] = 0.0;
  face_rules(t, p, r);
  set_reciprocals(r);
  r.neg = incident ? 1u : 0u;
  r.guard = needs_guard(r, max_dim) ? 1u : 0u;
  *theta = t;
  *phi = p;
  std::memcpy(dir, d3, sizeof(d3));
}

void setup_range(const double* s1, const double* omega, int64_t lo, int64_t hi,
                 const double* raw_xray, const double* omega_axis, const double* F,
                 int64_t max_dim, ReflRule* rules, ReflGeom* geom) {
  for (int64_t i = lo; i < hi; ++i) {
    double rw[9], m[9], tr[9], xr[3], sv[3];
    rodrigues(omega_axis, omega[i], rw);
    mul3(rw, F, m, 3);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) tr[b * 3 + a] = m[a * 3 + b];
    mul3(tr, raw_xray, xr, 1);
    mul3(tr, s1 + 3 * i, sv, 1);
    ReflGeom g;
    std::memcpy(g.xray, xr, sizeof(xr));
    std::memcpy(g.s1, sv, sizeof(sv));
    fill_ray(xr, true, max_dim, rules[i].ray[0], &g.theta[0], &g.phi[0], g.dir[0]);
    fill_ray(sv, false, max_dim, rules[i].ray[1], &g.theta[1], &g.phi[1], g.dir[1]);
    if (geom) geom[i] = g;
  }
}

// The bisection ray of one beam vector (increments, bisection.c:21-148):
// angles and direction as for the walk, then each ratio expression as the
// reference writes it.  Each libm value is computed once per distinct
// argument (bit-identical arguments give bit-identical results).  A face
// whose assert fails keeps zero ratios and no valid bit.
void ib_ray(const double* v, bool incident, IbRay& ray) {
  std::memset(&ray, 0, sizeof(ray));
  double t, p, d3[3];
  beam_angles(v, incident, &t, &p);
  to_myframe(v, d3);
  ray.dir[0] = d3[0];
  ray.dir[1] = d3[1];
  ray.dir[2] = d3[2];
  ray.neg = incident ? 1u : 0u;
  const double at = std::fabs(t), ap = std::fabs(p);
  const bool narrow = at < kPi / 2;
  const bool p_is_ap = same_bits(p, ap), t_is_at = same_bits(t, at);
  auto set = [&](int face, bool ok, double rz, double ry, double rx) {
    if (!ok) return;
    ray.r[face - 1][0] = rz;
    ray.r[face - 1][1] = ry;
    ray.r[face - 1][2] = rx;
    ray.valid |= 1u << (face - 1);
  };
  const double tan_p = std::tan(p), cos_p = std::cos(p), sin_p = std::sin(p);
  const double tan_t = std::tan(t), tan_pi_t = std::tan(kPi - t);
  const double cos_ap = p_is_ap ? cos_p : std::cos(ap);
  const double tan_ap = p_is_ap ? tan_p : std::tan(ap);
  const double sin_ap = p_is_ap ? sin_p : std::sin(ap);
  // faces 4 (y -> 0) and 5 (y -> y_max), :39-84
  if (narrow) {
    const double tt = t_is_at ? tan_t : std::tan(at);
    set(4, t > 0, sin_p / tt, 1.0, -cos_ap / tt);
    set(5, t < 0, sin_p / tt, -1.0, -cos_ap / tt);
  } else {
    const double tt = t_is_at ? tan_pi_t : std::tan(kPi - at);
    const double mp = -p;
    const double sin_mp = same_bits(mp, ap) ? sin_ap : std::sin(mp);
    set(4, t > 0, sin_mp / tt, 1.0, cos_ap / tt);
    set(5, t < 0, -sin_p / tt, -1.0, cos_ap / tt);
  }
  // faces 6 (x -> 0) and 1 (x -> x_max), :85-106
  set(6, at <= kPi / 2, tan_p, tan_t / cos_p, -1.0);
  set(1, at >= kPi / 2, -tan_p, tan_pi_t / cos_ap, 1.0);
  // faces 2 (z -> 0) and 3 (z -> z_max), :107-142
  if (narrow) {
    const double rx = -1 / tan_ap, ry = tan_t / sin_ap;
    set(2, p < 0, -1.0, ry, rx);
    set(3, p > 0, 1.0, ry, rx);
  } else {
    const double rx = 1 / tan_ap, ry = tan_pi_t / sin_ap;
    set(2, p > 0, -1.0, ry, rx);
    set(3, p < 0, 1.0, ry, rx);
  }
}

void rotate_one(const double* s1, double omega, const double* raw_xray,
                const double* omega_axis, const double* F, double* xr, double* sv) {
  double rw[9], m[9], tr[9];
  rodrigues(omega_axis, omega, rw);
  mul3(rw, F, m, 3);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) tr[b * 3 + a] = m[a * 3 + b];
  mul3(tr, raw_xray, xr, 1);
  mul3(tr, s1, sv, 1);
}

// Run body(lo, hi) over [0, R) on up to n_threads threads (at most 64).
template <class Body>
void parallel_range(int64_t R, int n_threads, Body body) {
  int nt = n_threads > 0 ? n_threads : 1;
  if (nt > 64) nt = 64;
  if (R < 512) nt = 1;
  if (nt == 1) {
    body((int64_t)0, R);
    return;
  }
  host_pool().run(R, nt, WorkerPool::Body(body));
}

}  // namespace

void host_setup(const double* s1, const double* omega, int64_t R, const double* raw_xray,
                const double* omega_axis, const double* F, int64_t max_dim, ReflRule* rules,
                ReflGeom* geom, int n_threads) {
  if (R <= 0) return;
  parallel_range(R, n_threads, [=](int64_t lo, int64_t hi) {
    setup_range(s1, omega, lo, hi, raw_xray, omega_axis, F, max_dim, rules, geom);
  });
}

void ib_setup(const double* s1, const double* omega, int64_t R, const double* raw_xray,
              const double* omega_axis, const double* F, IbRule* rules, int n_threads) {
  if (R <= 0) return;
  parallel_range(R, n_threads, [=](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      double xr[3], sv[3];
      rotate_one(s1 + 3 * i, omega[i], raw_xray, omega_axis, F, xr, sv);
      ib_ray(xr, true, rules[i].ray[0]);
      ib_ray(sv, false, rules[i].ray[1]);
    }
  });
}

void ib_setup_rotated(const double* rotated_s1, const double* xray, int64_t R, IbRule* rules,
                      int n_threads) {
  if (R <= 0) return;
  parallel_range(R, n_threads, [=](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      ib_ray(xray + 3 * i, true, rules[i].ray[0]);
      ib_ray(rotated_s1 + 3 * i, false, rules[i].ray[1]);
    }
  });
}

void host_setup_directions(const double* rotated_s1, const double* theta_phi, int64_t D,
                           int64_t max_dim, RayRule* rays) {
  for (int64_t i = 0; i < D; ++i) {
    RayRule& r = rays[i];
    std::memset(&r, 0, sizeof(r));
    double d3[3];
    to_myframe(rotated_s1 + 3 * i, d3);
    r.dir[0] = d3[0];
    r.dir[1] = d3[1];
    r.dir[2] = d3[2];
    face_rules(theta_phi[2 * i], theta_phi[2 * i + 1], r);
    set_reciprocals(r);
    r.neg = 0u;
    r.guard = needs_guard(r, max_dim) ? 1u : 0u;
  }
}

void host_setup_rotated(const double rotated_s1[3], const double xray[3], int64_t max_dim,
                        ReflRule* rule, ReflGeom* geom) {
  ReflGeom g;
  std::memcpy(g.xray, xray, 3 * sizeof(double));
  std::memcpy(g.s1, rotated_s1, 3 * sizeof(double));
  fill_ray(xray, true, max_dim, rule->ray[0], &g.theta[0], &g.phi[0], g.dir[0]);
  fill_ray(rotated_s1, false, max_dim, rule->ray[1], &g.theta[1], &g.phi[1], g.dir[1]);
  if (geom) *geom = g;
}

}  // namespace anacor
