// Per-reflection FP64 geometry shared by the host setup (geometry.cpp, built
// with g++ against glibc libm) and the HIP kernels (trace.hip).
#pragma once
#include <cstddef>
#include <cstdint>

namespace anacor {

// Bits of RayRule::meta[f] (one word per exit face f = 1..6, index f-1).
enum : uint32_t {
  kMetaValid = 1u << 0,      // the reference's precondition for this face holds
  kMetaRound = 1u << 1,      // round-half-away (else floor)
  kMetaInclusive = 1u << 2,  // loop bound inclusive
  kMetaToMax = 1u << 3,      // leaves through the upper plane of `axis`
  kMetaAxisShift = 4,        // bits 4-5: bounding axis 0 z, 1 y, 2 x
  kMetaVacShift = 8,         // bits 8-15: full_iteration vacuum-run limit
};

// One ray direction (incident or diffracted) of one reflection: the
// cube_face direction and the stepping rule for each possible exit face.
// 256 bytes, so a reflection's two rays are one 512-byte record.
struct alignas(16) RayRule {
  double dir[4];        // myframe (z, y, x) direction, [3] unused
  double d[6][4];       // per face: per-increment delta (z, y, x); d[a][3] for
                        // a = 0, 1, 2: RN(1 / dir[a]), the rest unused
  uint32_t meta[6];     // per face flags, see above
  uint32_t neg;         // 1 for the incident ray (cube_face's L1)
  uint32_t guard;       // 1 when a coordinate can leave the int64 range
};
static_assert(sizeof(RayRule) == 256, "RayRule layout");

struct ReflRule {
  RayRule ray[2];       // [0] incident (traced backwards), [1] diffracted
};

// Angles and rotated vectors of one reflection (diagnostics / tests).
struct ReflGeom {
  double xray[3], s1[3];
  double theta[2], phi[2];
  double dir[2][3];
};

// One ray of the bisection traversal (increments, bisection.c:21-148): the
// cube_face direction, and per exit face the ratios with which the sample
// point at distance m along the face axis is (z + m*r[0], y - m*r[1],
// x + m*r[2]).  240 bytes.
struct alignas(16) IbRay {
  double dir[4];        // myframe (z, y, x) direction, [3] unused
  double r[6][4];       // per face f = 1..6 (index f-1): (rz, ry, rx), [3] unused
  uint32_t valid;       // bit f-1: the reference's assert for face f holds
  uint32_t neg;         // 1 for the incident ray (cube_face's L1)
  uint32_t pad[2];
};
static_assert(sizeof(IbRay) == 240, "IbRay layout");

struct IbRule {
  IbRay ray[2];         // [0] incident, [1] diffracted
};

// Bisection rules of R reflections: rotation as host_setup, then angles,
// directions and the per-face ratios of both rays; up to n_threads threads.
void ib_setup(const double* s1, const double* omega, int64_t R, const double* raw_xray,
              const double* omega_axis, const double* F, IbRule* rules, int n_threads);

// The same for already rotated vectors (rotated_s1[R][3], xray[R][3]), as
// ib_test takes them (ray_tracing.c:65-80).
void ib_setup_rotated(const double* rotated_s1, const double* xray, int64_t R, IbRule* rules,
                      int n_threads);

// Fill rules[R] (and geom[R] when non-null) for R reflections, using up to
// n_threads host threads.  max_dim is the largest volume extent, used to set
// RayRule::guard.  Thread-safe, allocates nothing.
void host_setup(const double* s1, const double* omega, int64_t R,
                const double* raw_xray, const double* omega_axis, const double* F,
                int64_t max_dim, ReflRule* rules, ReflGeom* geom, int n_threads);

// One ray per direction, as ray_tracing_single_gridding walks it
// (ray_tracing.c:484-534): the exit-face direction from rotated_s1[D][3], the
// stepping rules from the given angles theta_phi[D][2] (the diffracted-ray
// convention, not negated).
void host_setup_directions(const double* rotated_s1, const double* theta_phi, int64_t D,
                           int64_t max_dim, RayRule* rays);

// Same for already-rotated vectors (ray_tracing_single).
void host_setup_rotated(const double rotated_s1[3], const double xray[3], int64_t max_dim,
                        ReflRule* rule, ReflGeom* geom);

}  // namespace anacor
