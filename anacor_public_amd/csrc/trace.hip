// MI355X (gfx950) kernels of the analytical absorption-correction ray tracer
// and the C-ABI around them.
//
// Work item: one (reflection, crystal voxel) pair per lane, both rays of the
// pair (incident traced backwards, diffracted forwards) on the same lane so
// cal_rate needs no exchange.  A workgroup covers one reflection and a tile of
// the Morton-ordered voxel list, so every wave's 64 rays are parallel and start
// from neighbouring voxels: their label gathers hit the same cache lines.
//
// Numerics (bit-exact path lengths against AnACor/src/ray_tracing_basic.c):
//   * cube_face: correctly rounded FP64 quotients (one per axis, two for a
//     non-finite component), the reference's candidate order and ties.
//   * stepping: coord + (double)inc * d as a separately rounded multiply and
//     add (this file is built with -ffp-contract=off), then floor or
//     round-half-away, the clamp to [0, extent-1] and the double->int64 quirk
//     (values outside int64 become 0) when the host flags a ray as needing it.
//   * lengths: per-segment terms T * (span / (len-1)) summed in path order
//     from a per-lane segment list in LDS (exact re-walk when it overflows).
//   * cal_rate in FP64 with the reference's association; exp in FP32 after an
//     FP64 range reduction; per-reflection mean by compensated FP32 summation
//     (Kahan per lane, TwoSum merges across lanes, waves and tiles), or FP64
//     accumulation when ANACOR_FLAG_F64_MEAN is set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <utility>
#include <atomic>
#include <chrono>
#include <thread>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <vector>

#include "errors.h"
#include "geometry.h"
#include "../../include/anacor_mi355x.h"

namespace anacor {
struct Context;
}

// The public handle types of include/anacor_mi355x.h.
struct anacor_ctx {
  anacor::Context* c;
};

namespace anacor {
struct Grid;
}

struct anacor_grid {
  anacor::Grid* g;
};

namespace anacor {
struct Batch;
}

struct anacor_batch {
  anacor::Batch* b;
};

namespace anacor {

// ------------------------------------------------------------ error state

static thread_local char g_err[512];
static std::atomic<int> g_err_code{0};

void set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  g_err_code.store(code);
  if (std::getenv("ANACOR_VERBOSE")) fprintf(stderr, "anacor: %s\n", g_err);
}

void clear_error() {
  g_err[0] = 0;
  g_err_code.store(0);
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error(ANACOR_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
      return ANACOR_ERR_HIP;                                                       \
    }                                                                              \
  } while (0)

// ------------------------------------------------------------ device code

constexpr int kBlock = 256;
constexpr int kWave = 64;

struct PathRecord {           // per (reflection, voxel): both rays
  int32_t face[2];
  int32_t len[2];
  int32_t nseg[2];
  int32_t err;
  int32_t pad;
  int32_t last[2][3];
  double L[2][4];
};
static_assert(sizeof(PathRecord) == 120, "PathRecord layout");

struct TraceArgs {
  const uint8_t* vol;          // raw volume or, for the bricked layouts, the packed one
  int vol_bytes;               // bytes behind vol (range check of the loads)
  int cz, cy;                  // brick constants (bricked layouts)
  int Z, Y, X;
  const int4* coords;          // (z, y, x, original index)
  int N;
  const ReflRule* rules;
  int R;
  int tiles;                   // tiles per reflection
  int vox_per_tile;
  int n_blocks;                // logical blocks = R * tiles
  double vs[3];                // voxel size (z, y, x)
  double mu[4];                // [li, lo, cr, bu]
  int is_exp;
  float2* partial32;           // [R * tiles] (sum, compensation)
  double* partial64;           // [R * tiles]
  unsigned* err;               // [R]
  unsigned long long* steps;   // sum of len_path_2
  PathRecord* paths;           // [R * N] in original voxel order (paths kernel)
  const int* order;            // [R] caller's index of the reflection at each processing position
  const double* incd;          // (double)i for i = 0 .. incd_last
  int incd_last;
  int group;                   // reflections per group of the launch order (<= 1: none)
  int pos0;                    // processing position of this launch's first reflection
};

// Map hardware workgroup id to a logical one so that consecutive logical
// blocks land on the same XCD: the dispatcher deals workgroups round-robin
// over the 8 XCDs, each with its own L2.
__device__ __forceinline__ int xcd_logical_block(int hw, int n_hw) {
  const int per = n_hw >> 3;  // n_hw is a multiple of 8
  return (hw & 7) * per + (hw >> 3);
}

// Logical block lb -> (processing position, voxel tile).  Without groups the
// tiles of one reflection are consecutive.  With groups of G consecutive
// processing positions (reflections of similar directions), a group runs
// tile by tile, each tile for all G reflections, so the workgroups resident
// on an XCD trace nearly parallel rays from one compact set of voxels and
// share the bricks they cross in that XCD's L2.  Every group but the last
// has G reflections.
__device__ __forceinline__ void block_work(int lb, int tiles, int R, int group, int& pos,
                                           int& tile) {
  if (group <= 1) {
    pos = lb / tiles;
    tile = lb - pos * tiles;
    return;
  }
  const int span = group * tiles;
  const int g = lb / span;
  const int w = lb - g * span;
  const int gsize = min(group, R - g * group);
  tile = w / gsize;
  pos = g * group + (w - tile * gsize);
}

// The two planes of one axis, t = (0 - c) / e and (n - c) / e with the
// direction e already negated for the incident ray (-(k / d) and k / (-d) are
// the same double).  With 0 <= c < n only one quotient can be >= 0 unless
// c == 0: for e with its sign bit clear (> 0 or +0) the upper plane's
// (positive, or +inf), for e with the sign bit set the lower plane's
// (positive, +inf, -0 when c == 0, NaN for 0 / -0); the other is negative,
// -inf or NaN — except (0 - 0) / e = +0 for e > 0 — so it is not computed
// (a direction component that is not finite gets both quotients).
// Non-candidates are returned as -1, which neither passes the >= 0 test nor
// equals a minimum.  One division per axis instead of two.
//
// RN(a / b) from y = RN(1 / b) by two fused corrections of q0 = RN(a * y).
// q0 can be up to about 1.5 ulp from a / b (the error of y, scaled by a, plus
// the rounding of the product).  The first correction computes
// q0 + RN(a - b q0) y exactly before rounding it; that value lies within
// about 4 u^2 |a / b| (u = 2^-53) of a / b, so q1 is faithful (one of the two
// doubles around a / b).  Markstein's theorem (P. Markstein, IBM J. Res.
// Dev. 34(1), 1990; Muller et al., Handbook of Floating-Point Arithmetic,
// 2nd ed., section 4.8) then makes the second correction's
// RN(q1 + (a - b q1) y), with the exact remainder a - b q1, the correctly
// rounded quotient.  Callers keep a, b and the results away from overflow
// and underflow.
__device__ __forceinline__ double quot_rn(double a, double b, double y) {
  const double q0 = a * y;
  const double q1 = __builtin_fma(__builtin_fma(-q0, b, a), y, q0);
  return __builtin_fma(__builtin_fma(-q1, b, a), y, q1);
}

// k / e for an integer k (|k| < 2^21): with y = RN(1 / e) from the host (the
// direction is the reflection's, so this is a per-ray value), quot_rn.
// Components outside [2^-900, 2^900] in magnitude, where k * y could overflow
// or y lose precision, divide.
template <bool INV>
__device__ __forceinline__ double plane_quot(int k, double e, double y) {
  const double a = (double)k;
  const double m = __builtin_fabs(e);
  if (INV && m >= 0x1p-900 && m <= 0x1p900) return quot_rn(a, e, y);
  return a / e;
}

template <bool INV>
__device__ __forceinline__ void plane_pair(int c, int n, double e, double y, double& tlo,
                                           double& thi) {
  const double no = -1.0;
  if (!__builtin_isfinite(e)) {   // k / +-inf is a signed zero, k / NaN NaN: both as they are
    tlo = (double)(0 - c) / e;
    thi = (double)(n - c) / e;
  } else if (__builtin_signbit(e)) {
    tlo = plane_quot<INV>(0 - c, e, y);
    thi = no;
  } else {
    thi = plane_quot<INV>(n - c, e, y);
    tlo = (c == 0 && e > 0.0) ? 0.0 : no;   // 0 / +0 and 0 / NaN are NaN
  }
}

// cube_face (ray_tracing_basic.c:69-156).  0 when no exit plane qualifies.
// INV: iz, iy, ix are RN(1 / dz), RN(1 / dy), RN(1 / dx).
template <bool INV>
__device__ __forceinline__ int exit_face(int z, int y, int x, double dz, double dy, double dx,
                                         double iz, double iy, double ix, int Z, int Y, int X,
                                         bool neg) {
  double t0, t1, t2, t3, t4, t5;      // x_min, y_min, z_min, x_max, y_max, z_max
  // -(k / d) == k / (-d), and RN(1 / -d) == -RN(1 / d)
  plane_pair<INV>(x, X, neg ? -dx : dx, neg ? -ix : ix, t0, t3);
  plane_pair<INV>(y, Y, neg ? -dy : dy, neg ? -iy : iy, t1, t4);
  plane_pair<INV>(z, Z, neg ? -dz : dz, neg ? -iz : iz, t2, t5);
  const double big = __builtin_huge_val();
  // minimum over the non-negative candidates (NaN never qualifies)
  double m = big;
  bool have = false;
  if (t0 >= 0) { m = t0; have = true; }
  if (t1 >= 0) { m = (!have || t1 < m) ? t1 : m; have = true; }
  if (t2 >= 0) { m = (!have || t2 < m) ? t2 : m; have = true; }
  if (t3 >= 0) { m = (!have || t3 < m) ? t3 : m; have = true; }
  if (t4 >= 0) { m = (!have || t4 < m) ? t4 : m; have = true; }
  if (t5 >= 0) { m = (!have || t5 < m) ? t5 : m; have = true; }
  if (!have) return 0;
  if (m == t0) return 6;
  if (m == t3) return 1;
  if (m == t1) return 4;
  if (m == t4) return 5;
  if (m == t2) return 2;
  if (m == t5) return 3;
  return 0;
}

// ---- one step of a ray
//
// The reference evaluates each axis as  idx = clamp(rnd(c + inc*d), 0, n-1)
// with rnd = floor or round-half-away and a double->int64 conversion that
// maps values outside int64 to INT64_MIN (then clamped to 0).  Here:
//   * the exit face's own axis has d = +-1, so it is an exact integer walk;
//   * for the other two axes, w = c + inc*d is evaluated exactly as the
//     reference does (separately rounded multiply and add), then
//       rnd(w) clamped == clamp(trunc_sat(w + bias)),
//     bias = 0 for floor and 0.5 - 2^-54 for round-half-away: negative values
//     clamp to 0 either way, and for w >= 0, floor(w + (0.5 - 2^-54)) equals
//     round-half-away(w) for every double (0.5 itself would round
//     0.49999999999999994 up).  v_cvt_i32_f64 truncates, saturates and maps
//     NaN to 0; values >= 2^63 (only possible when the host sets GUARD) are
//     mapped to 0 as the reference's conversion does.
constexpr double kRoundBias = 0x1.fffffffffffffp-2;  // 0.5 - 2^-54

__device__ __forceinline__ int cvt_trunc_sat(double x) {
  int r;
  asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// a * b + c on the low 24 bits of a and b (signed), one v_mad_i32_i24
__device__ __forceinline__ int mad24(int a, int b, int c) {
  int r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// The same for 0 <= a < 2^23, written so that the backend selects the
// v_mad_i32_i24 itself and can take a from a scalar register (the walk's
// increment is the same in every lane) without a copy to a vector register.
__device__ __forceinline__ int mad24_small(int a, int b, int c) {
  __builtin_assume(a >= 0 && a < (1 << 23));
  return ((a << 8) >> 8) * ((b << 8) >> 8) + c;
}

// median(v, 0, hi) == clamp to [0, hi] for hi >= 0, one v_med3_i32
__device__ __forceinline__ int clamp_int(int v, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(v), "v"(hi));
  return r;
}

// BIAS = false drops the add of a zero bias (floor rounding): w + 0.0
// converts to the same integer as w, -0.0 included.  CLAMP = false drops the
// clamp where the caller has shown that the index stays in [0, hi].
template <bool GUARD, bool BIAS = true, bool CLAMP = true>
__device__ __forceinline__ int axis_index(double base, double incd, double d, double bias,
                                          int hi) {
  const double w = base + incd * d;
  int v = cvt_trunc_sat(BIAS ? w + bias : w);
  if (GUARD && !(w < 0x1p63)) v = 0;
  return CLAMP ? clamp_int(v, hi) : v;
}

// A ray of one lane, with its axes permuted so that `p` is the exit face's
// axis and a, b are the other two (in z, y, x order).  Origins are inside the
// volume (checked on the host), so the face axis walks p0 + pdir*inc without
// leaving [0, extent-1] except on the last step of an inclusive bound, where
// the reference's clamp keeps it on the last plane: that is a clamp of the
// increment at incmax.
struct RayLane {
  double ba, bb;         // origins of the stepped axes a, b
  double da, db;         // per-increment deltas of a, b
  double bias;           // 0 (floor) or kRoundBias (round-half-away)
  int sa, sb;            // strides (elements) of a, b in the volume
  int ha, hb;            // extent - 1 of a, b
  int pbase, pstep;      // p0 * stride_p and pdir * stride_p
  int incmax;            // increment at which the face axis reaches its last plane
  int p0, pdir, paxis;   // face axis origin, direction (+1/-1), index (0 z, 1 y, 2 x)
  int limit;             // increments tried: inc in [0, limit)
  int vac;               // full_iteration vacuum-run limit
  int ca, cb, cp, sp;    // bricked layouts: brick constants of a, b, p and stride of p
};

// SHELL: the face coordinate is not held on its last plane; past it the
// offset lands in the zero shell of the raw layouts.
template <bool GUARD, bool WIDE, bool BIAS = true, bool CLAMP = true, bool SHELL = false>
__device__ __forceinline__ int lane_offset(const RayLane& L, int inc, double incd) {
  const int ia = axis_index<GUARD, BIAS, CLAMP>(L.ba, incd, L.da, L.bias, L.ha);
  const int ib = axis_index<GUARD, BIAS, CLAMP>(L.bb, incd, L.db, L.bias, L.hb);
  const int ip = SHELL ? inc : min(inc, L.incmax);
  if (WIDE) return ia * L.sa + ib * L.sb + ip * L.pstep + L.pbase;
  // every factor fits 24 bits (checked on the host): full-rate v_mad_i32_i24
  return mad24(ia, L.sa, mad24(ib, L.sb, mad24_small(ip, L.pstep, L.pbase)));
}

template <bool GUARD>
__device__ __forceinline__ void lane_zyx(const RayLane& L, int inc, int& z, int& y, int& x) {
  const double incd = (double)inc;
  const int ia = axis_index<GUARD>(L.ba, incd, L.da, L.bias, L.ha);
  const int ib = axis_index<GUARD>(L.bb, incd, L.db, L.bias, L.hb);
  const int ip = L.p0 + L.pdir * min(inc, L.incmax);
  if (L.paxis == 0) { z = ip; y = ia; x = ib; }
  else if (L.paxis == 1) { z = ia; y = ip; x = ib; }
  else { z = ia; y = ib; x = ip; }
}

// Coordinates at increment 0: the origin itself (c + 0 * d rounds to the
// integer c, and the origin lies inside the volume).
__device__ __forceinline__ void lane_origin(const RayLane& L, int& z, int& y, int& x) {
  const int ia = (int)L.ba, ib = (int)L.bb, ip = L.p0;
  if (L.paxis == 0) { z = ip; y = ia; x = ib; }
  else if (L.paxis == 1) { z = ia; y = ip; x = ib; }
  else { z = ia; y = ib; x = ip; }
}

// Volume layouts (the LAYOUT template parameter of the walk):
//   kLayoutRaw    uint8 [z][y][x], 24-bit address arithmetic
//   kLayoutWide   uint8 [z][y][x], 32-bit address arithmetic
//   kLayoutBrick2 8x8x8 bricks of 2-bit labels (one 128-byte line per brick)
//   kLayoutBrick4 8x8x8 bricks of 4-bit codes: labels 0..4 as they are, any
//                 other value as 5 (the walk only asks whether a label is 0,
//                 whether it differs from the previous one and whether it is
//                 in 0..4, so the boundaries are unchanged)
// Brick b = (bz * NBY + by) * NBX + bx holds voxel (z&7, y&7, x&7) at index
// 512 b + 64 (z&7) + 8 (y&7) + (x&7); per axis that index is
// i * s + (i >> 3) * c with (s, c) = (64, 512 NBX NBY - 512),
// (8, 512 NBX - 64), (1, 504) for z, y, x.
enum { kLayoutRaw = 0, kLayoutWide = 1, kLayoutBrick2 = 2, kLayoutBrick4 = 3 };

// The raw layouts hold the volume inside a one-voxel shell of label 0:
// voxel (z, y, x) is at (z + 1) * sz + (y + 1) * sy + (x + 1), sy = X + 2,
// sz = (Y + 2) * sy.  A ray whose face coordinate steps one plane past the
// volume reads the shell's 0 and stops, so the walk needs no loop bound.
typedef const __attribute__((address_space(4))) double* const_f64_ptr;

struct Volume {
  __amdgpu_buffer_rsrc_t rsrc;  // range-checked: offsets past the end read 0
  int Z, Y, X;
  int sz, sy, base;             // raw layouts: strides of z and y, offset of voxel (0, 0, 0)
  int cz, cy;                   // brick constants c of z and y (bricked layouts)
  // (double)i for i = 0 .. incd_last, read through the scalar cache: the
  // walk's increment is the same in every lane, so its conversion is a
  // scalar load instead of a vector instruction per step
  const_f64_ptr incd;
  int incd_last;
};

__device__ __forceinline__ int load_label(const Volume& V, int off) {
  return __builtin_amdgcn_raw_buffer_load_b8(V.rsrc, off, 0, 0);
}

// Label of brick index f in a packed layout: one dword load and a bit-field
// extract (16 2-bit or 8 4-bit labels per dword).
template <int LAYOUT>
__device__ __forceinline__ int load_packed(const Volume& V, int f) {
  constexpr int BITS = LAYOUT == kLayoutBrick2 ? 2 : 4;
  constexpr int SH = LAYOUT == kLayoutBrick2 ? 4 : 3;
  const int off = (f >> SH) << 2;
  const unsigned dw = __builtin_amdgcn_raw_buffer_load_b32(V.rsrc, off, 0, 0);
  return (int)__builtin_amdgcn_ubfe(dw, (unsigned)(f * BITS), BITS);
}

// Label at increment inc of the lane's ray in the given layout (SHELL: raw
// layouts only, see lane_offset).
// incd: (double)inc.
template <bool GUARD, int LAYOUT, bool BIAS = true, bool CLAMP = true, bool SHELL = false>
__device__ __forceinline__ int label_at(const RayLane& L, const Volume& V, int inc, double incd) {
  if (LAYOUT == kLayoutRaw || LAYOUT == kLayoutWide)
    return load_label(V, lane_offset<GUARD, LAYOUT == kLayoutWide, BIAS, CLAMP, SHELL>(L, inc,
                                                                                        incd));
  const int ia = axis_index<GUARD, BIAS, CLAMP>(L.ba, incd, L.da, L.bias, L.ha);
  const int ib = axis_index<GUARD, BIAS, CLAMP>(L.bb, incd, L.db, L.bias, L.hb);
  const int ip = mad24(L.pdir, min(inc, L.incmax), L.p0);
  // every factor fits 24 bits and every partial sum 31 (checked on the host)
  const int f = mad24(ia, L.sa, mad24(ia >> 3, L.ca, mad24(ib, L.sb, mad24(ib >> 3, L.cb,
                mad24(ip, L.sp, mad24(ip >> 3, L.cp, 0))))));
  return load_packed<LAYOUT>(V, f);
}

// Walk one ray (cal_coord + cal_path2_plus).  Boundaries beyond the implicit
// first one (class 3 at increment 0) are kept in this lane's column of `seg`
// (lane-interleaved, CAP + U slots: every step stores at the write slot,
// which advances only when a boundary opens).  A group of U steps opens at
// most U boundaries, so the bound is tested once per group: a group starts
// with the write slot at CAP or below, or the list has overflowed and the
// group writes into the spare slots above CAP again.  len is len_path_2; out
// receives [li, lo, cr, bu].
typedef __attribute__((address_space(3))) uint32_t lds_u32;

constexpr int seg_slots(int cap, int u) { return cap + u; }

// L04: every label of the volume is in 0..4 (scanned on the host side), so
// any change of label opens a boundary and the walk skips that test.
template <int CAP, int U, bool GUARD, bool FULL, int LAYOUT, bool BIAS, bool CLAMP, bool L04>
__device__ __forceinline__ void walk_ray_impl(const RayLane& L, const Volume& V,
                                              const double vs[3], lds_u32* seg, int& len_out,
                                              int& nseg_out, int last_zyx[3], double out[4]) {
  // raw layouts without full_iteration: the face coordinate runs into the
  // zero shell at increment incmax + 1, which ends the ray where the loop
  // bound of an exclusive bound does (limit = incmax + 1); an inclusive bound
  // (limit = incmax + 2) takes that increment once more after the loop, with
  // the face coordinate held on its last plane
  constexpr bool SHELL = !FULL && LAYOUT < kLayoutBrick2;
  // int8 labels: only values 0..4 open a boundary (appending, :343-394);
  // 2-bit bricks hold 0..3 only
  constexpr bool kAll04 = L04 || LAYOUT == kLayoutBrick2;
  // taken steps (increments 0 .. cnt-1 passed), last label, vacuum run
  int cnt = 0, prev = 0, va = 0;
  bool done = L.limit <= 0;
  bool overflow = false;
  lds_u32* wp = seg;                          // write slot (32-bit LDS address)
  lds_u32* const cap_slot = seg + CAP * kBlock;
  if (!done) {  // increment 0: its label opens no boundary
    const int lab = label_at<GUARD, LAYOUT>(L, V, 0, 0.0);
    bool stop;
    if (FULL) {
      va = (lab == 0) ? 1 : 0;
      stop = va > L.vac;
    } else {
      stop = lab == 0;
    }
    if (stop) done = true;
    else { prev = lab; cnt = 1; }
  }
  // With the shell every walk ends on a label 0, and that step records too
  // (0 differs from any taken label): the last record holds the step count,
  // so the loop keeps no count per step.
  const bool ran = !done;
  for (int inc0 = 1; !done; inc0 += U) {
    if (wp > cap_slot) {  // more than CAP boundaries so far: re-walk at the end
      overflow = true;
      wp = cap_slot;
    }
    int lab[U];
    // the table covers every increment a walk reaches (extent + 2U); the
    // bound only keeps the group's scalar loads inside it
    const const_f64_ptr tbl = V.incd + min(inc0, V.incd_last - (U - 1));
#pragma unroll
    for (int u = 0; u < U; ++u)  // range-checked loads; steps past the end are not taken
      lab[u] = label_at<GUARD, LAYOUT, BIAS, CLAMP, SHELL>(L, V, inc0 + u, tbl[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int inc = inc0 + u;
      const int lb = lab[u];
      bool stop;
      if (FULL) {
        va = (lb == 0) ? va + 1 : 0;
        stop = va > L.vac;
      } else {
        stop = lb == 0;
      }
      // once a lane stops it never takes another step: prev/va need no select
      const bool take = SHELL ? !done && !stop : !done && inc < L.limit && !stop;
      bool chg = SHELL ? !done && lb != prev : take && lb != prev;
      if (!kAll04) chg = chg && (unsigned)lb <= 4u;
      *wp = ((uint32_t)inc << 3) | (uint32_t)lb;
      wp = chg ? wp + kBlock : wp;
      if (!SHELL) cnt += take ? 1 : 0;
      prev = lb;
      done = done || !take;
    }
  }
  bool stop_record = false;
  if (SHELL && ran) {
    // the last record is the stop, written in the lane's last group (after
    // any overflow reset): its increment is the number of steps taken
    cnt = (int)(*(wp - kBlock) >> 3);
    stop_record = true;
  }
  if (SHELL && cnt == L.incmax + 1 && L.limit == L.incmax + 2) {
    const int inc = L.incmax + 1;
    const int lb = label_at<GUARD, LAYOUT, BIAS, CLAMP>(L, V, inc, (double)inc);
    if (lb != 0) {
      // the ray takes this step: drop the shell's stop record; prev holds
      // the label of the group's last step, taken or not, so read the last
      // taken step's label again
      wp -= kBlock;
      stop_record = false;
      const int before = label_at<GUARD, LAYOUT, BIAS, CLAMP>(L, V, L.incmax, (double)L.incmax);
      bool chg = lb != before;
      if (!kAll04) chg = chg && (unsigned)lb <= 4u;
      *wp = ((uint32_t)inc << 3) | (uint32_t)lb;
      wp = chg ? wp + kBlock : wp;
      cnt += 1;
    }
  }

  // exact unless the list overflowed (<= CAP + U records, the stop included)
  int nseg = (int)(wp - seg) / kBlock - (stop_record ? 1 : 0);
  const int last = cnt - 1;
  const int len = cnt > 0 ? cnt : 1;
  len_out = len;
  nseg_out = nseg + 1;
  out[0] = out[1] = out[2] = out[3] = 0.0;
  last_zyx[0] = last_zyx[1] = last_zyx[2] = 0;
  if (last < 0) return;
  int fz, fy, fx, lz, ly, lx;
  lane_origin(L, fz, fy, fx);
  lane_zyx<GUARD>(L, last, lz, ly, lx);
  last_zyx[0] = lz; last_zyx[1] = ly; last_zyx[2] = lx;
  if (len == 1) return;
  // total length first -> last (y, z, x order as in :1480-1482)
  const double ey = (double)(ly - fy) * vs[1];
  const double ez = (double)(lz - fz) * vs[0];
  const double ex = (double)(lx - fx) * vs[2];
  const double T = __builtin_sqrt(ey * ey + ez * ez + ex * ex);
  const double D = (double)(len - 1);
  // RN(span / D) for every segment from one division per ray (quot_rn with
  // y = RN(1 / D); 1 <= D < 2^21, 0 <= span <= D + 1)
  const double yD = 1.0 / D;
  double li = 0.0, lo = 0.0, cr = 0.0, bu = 0.0;
  // With a finite term, acc + (cls == k ? term : 0) equals fma(term, m, acc)
  // for m = 1.0 or 0.0: the product is exact, and term * 0 = +0 leaves acc
  // (which starts at +0 and only grows) unchanged.  A non-finite T (absurd
  // voxel sizes) keeps the selects.
  const bool finite_T = __builtin_isfinite(T);
  auto add = [&](int cls, int span) {
    const double term = T * quot_rn((double)span, D, yD);
    if (finite_T) {
      li = __builtin_fma(term, cls == 1 ? 1.0 : 0.0, li);
      lo = __builtin_fma(term, cls == 2 ? 1.0 : 0.0, lo);
      cr = __builtin_fma(term, cls == 3 ? 1.0 : 0.0, cr);
      bu = __builtin_fma(term, cls == 4 ? 1.0 : 0.0, bu);
    } else {
      li += (cls == 1) ? term : 0.0;
      lo += (cls == 2) ? term : 0.0;
      cr += (cls == 3) ? term : 0.0;
      bu += (cls == 4) ? term : 0.0;
    }
  };
  int cls = 3, start = 0;
  if (!overflow) {
    for (int k = 0; k < nseg; ++k) {
      const uint32_t e = seg[k * kBlock];
      const int pos = (int)(e >> 3);
      add(cls, pos - start);
      cls = (int)(e & 7u);
      start = pos;
    }
  } else {
    // the boundary list overflowed: walk the (now known) path again in order
    int pl = label_at<GUARD, LAYOUT>(L, V, 0, 0.0);
    nseg = 0;
    for (int inc = 1; inc <= last; ++inc) {
      const int lb = label_at<GUARD, LAYOUT>(L, V, inc, (double)inc);
      if (lb != pl && (unsigned)lb <= 4u) {
        add(cls, inc - start);
        cls = lb;
        start = inc;
        ++nseg;
      }
      pl = lb;
    }
    nseg_out = nseg + 1;
  }
  add(cls, len - start);
  out[0] = li; out[1] = lo; out[2] = cr; out[3] = bu;
}

// True when neither stepped coordinate of the lane's ray needs the clamp to
// [0, extent-1] on increments 0 .. last: each is monotone in the increment,
// so its unclamped values at the two ends bound all the others.  Increments
// past the last one (the unrolled tail) are never taken, and their loads stay
// range-checked.
template <bool BIAS>
__device__ __forceinline__ bool lane_in_range(const RayLane& L, int last) {
  const double i0 = 0.0, i1 = (double)last;
  const int a0 = axis_index<false, BIAS, false>(L.ba, i0, L.da, L.bias, L.ha);
  const int a1 = axis_index<false, BIAS, false>(L.ba, i1, L.da, L.bias, L.ha);
  const int b0 = axis_index<false, BIAS, false>(L.bb, i0, L.db, L.bias, L.hb);
  const int b1 = axis_index<false, BIAS, false>(L.bb, i1, L.db, L.bias, L.hb);
  return (unsigned)a0 <= (unsigned)L.ha && (unsigned)a1 <= (unsigned)L.ha &&
         (unsigned)b0 <= (unsigned)L.hb && (unsigned)b1 <= (unsigned)L.hb;
}

// (trace_kernel and grid_map_kernel keep five waves per SIMD in the raw
// layouts, at most 96 VGPRs, which the four inlined forms would otherwise
// exceed.)
// The stepping loop in four forms: floor rounding (zero bias, one FP64 add
// per axis and step fewer) or round-half-away, with or without the clamp of
// the stepped coordinates.  The rounding is fixed per ray and almost always
// the same across a wave; the clamp is dropped only when no lane of the
// wave needs it (a wave vote), so that branch never diverges.
// BRICK_FREE: the bricked layouts take the clamp-free forms too (steps past a
// ray's end, the unrolled tail, are not taken and their loads stay
// range-checked).  The FP64-mean and absorption-map kernels leave it off: in
// the FP64-mean kernel the two more inlined forms made the launch 50% slower
// (1,294 against 864 ms on 1000x1000x800), while the FP32 one gained 1%.
template <int CAP, int U, bool GUARD, bool FULL, int LAYOUT, bool L04, bool BRICK_FREE = false>
__device__ __forceinline__ void walk_ray(const RayLane& L, const Volume& V, const double vs[3],
                                         lds_u32* seg, int& len_out, int& nseg_out,
                                         int last_zyx[3], double out[4]) {
  const bool round = L.bias != 0.0;
  bool unclamped = false;
  if (!GUARD && (BRICK_FREE || LAYOUT < kLayoutBrick2)) {
    // with the zero shell the walk reads increment incmax + 1 of every ray
    // (limit - 1 for an inclusive bound, limit for an exclusive one); without
    // it (full_iteration, bricked layouts) it takes increments below limit
    const int last = (FULL || LAYOUT >= kLayoutBrick2) ? L.limit - 1 : L.incmax + 1;
    const bool in_range = L.limit <= 0 || (round ? lane_in_range<true>(L, last)
                                                 : lane_in_range<false>(L, last));
    unclamped = __all(in_range) != 0;
  }
  if (unclamped) {
    if (round)
      walk_ray_impl<CAP, U, GUARD, FULL, LAYOUT, true, false, L04>(L, V, vs, seg, len_out, nseg_out,
                                                              last_zyx, out);
    else
      walk_ray_impl<CAP, U, GUARD, FULL, LAYOUT, false, false, L04>(L, V, vs, seg, len_out,
                                                               nseg_out, last_zyx, out);
  } else {
    if (round)
      walk_ray_impl<CAP, U, GUARD, FULL, LAYOUT, true, true, L04>(L, V, vs, seg, len_out, nseg_out,
                                                             last_zyx, out);
    else
      walk_ray_impl<CAP, U, GUARD, FULL, LAYOUT, false, true, L04>(L, V, vs, seg, len_out, nseg_out,
                                                              last_zyx, out);
  }
}

// Set up the lane's ray for exit face `face` (1..6) of rule `rr`; false when
// the reference's precondition for that face fails.  Strides are those of the
// layout: element strides of [z][y][x], or in-brick strides plus brick
// constants for the bricked layouts.
template <int LAYOUT>
__device__ __forceinline__ bool make_lane(const RayRule* __restrict__ rr, int face, int z, int y,
                                          int x, const Volume& V, RayLane& L) {
  constexpr bool BRICK = LAYOUT == kLayoutBrick2 || LAYOUT == kLayoutBrick4;
  const uint32_t meta = rr->meta[face - 1];
  const double* d = rr->d[face - 1];
  const int axis = (int)((meta >> kMetaAxisShift) & 3u);
  const int sz = BRICK ? 64 : V.sz, sy = BRICK ? 8 : V.sy;
  const int kz = V.cz, ky = V.cy, kx = 504;
  L.bias = (meta & kMetaRound) ? kRoundBias : 0.0;
  L.vac = (int)((meta >> kMetaVacShift) & 0xffu);
  L.paxis = axis;
  // extents as values first: selecting among the fields of V inside the
  // branches let the compiler merge them into one indexed load, which put V
  // in scratch memory
  const int VZ = V.Z, VY = V.Y, VX = V.X;
  int c, ext, sp;
  if (axis == 0) {          // z: a = y, b = x
    L.ba = (double)y; L.da = d[1]; L.sa = sy; L.ha = VY - 1; L.ca = ky;
    L.bb = (double)x; L.db = d[2]; L.sb = 1;  L.hb = VX - 1; L.cb = kx;
    sp = sz; c = z; ext = VZ; L.cp = kz;
  } else if (axis == 1) {   // y: a = z, b = x
    L.ba = (double)z; L.da = d[0]; L.sa = sz; L.ha = VZ - 1; L.ca = kz;
    L.bb = (double)x; L.db = d[2]; L.sb = 1;  L.hb = VX - 1; L.cb = kx;
    sp = sy; c = y; ext = VY; L.cp = ky;
  } else {                  // x: a = z, b = y
    L.ba = (double)z; L.da = d[0]; L.sa = sz; L.ha = VZ - 1; L.ca = kz;
    L.bb = (double)y; L.db = d[1]; L.sb = sy; L.hb = VY - 1; L.cb = ky;
    sp = 1; c = x; ext = VX; L.cp = kx;
  }
  L.sp = sp;
  L.p0 = c;
  L.pbase = c * sp + (BRICK ? 0 : V.base);
  if (meta & kMetaToMax) {
    L.pdir = 1;
    L.limit = (ext - c) + ((meta & kMetaInclusive) ? 1 : 0);
    L.incmax = ext - 1 - c;
  } else {
    L.pdir = -1;
    L.limit = c + 1;
    L.incmax = c;
  }
  L.pstep = L.pdir * sp;
  return (meta & kMetaValid) != 0;
}

// exp(-r) in FP32 after an FP64 range reduction r = k ln2 + f, |f| <= ln2/2.
__device__ __forceinline__ float exp_neg(double r) {
  const double t = -r;
  const double k = __builtin_rint(t * 1.4426950408889634);
  const double f = (t - k * 0.6931471805599453) - k * 2.3190468138462996e-17;
  const double kc = __builtin_fmin(__builtin_fmax(k, -300.0), 300.0);
  return __builtin_ldexpf(expf((float)f), (int)kc);
}

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Kahan pair (s, c) with true value s - c; merge o into it.
__device__ __forceinline__ void merge_pair(float& s, float& c, float os, float oc) {
  float t, e;
  two_sum(s, os, t, e);
  s = t;
  c = (c + oc) - e;
}

// One sample as the kernels see it.
struct SampleView {
  Volume V;
  const int4* coords;          // (z, y, x, caller index), wave-grouped order
  int N;
  const ReflRule* rules;       // in processing order
  double vs[3];
  double mu[4];
};

// Accumulators of one lane over its voxels of a tile.
struct TileAcc {
  float ks, kc;                // Kahan sum, value ks - kc
  double acc64;
  unsigned long long steps;
  unsigned err;
};

// Trace the voxels [v_begin, v_end) of reflection `rule` with this lane's
// stride; accumulate exp(-sum mu*L) (or write path records).
template <int CAP, int U, bool GUARD, bool FULL, int LAYOUT, bool F64, bool PATHS, bool L04>
__device__ __forceinline__ void trace_tile(const SampleView& S, const ReflRule* __restrict__ rule,
                                           int v_begin, int v_end, int is_exp, lds_u32* seg,
                                           TileAcc& acc, PathRecord* __restrict__ paths) {
  for (int v = v_begin + (int)threadIdx.x; v < v_end; v += kBlock) {
    const int4 co = S.coords[v];
    double Lr[2][4];
    int lens[2], nsegs[2], faces[2], lasts[2][3];
    unsigned verr = 0;
    // cube_face for both rays first (ray_tracing.c:206-207)
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const RayRule* rr = &rule->ray[k];
      faces[k] = exit_face<true>(co.x, co.y, co.z, rr->dir[0], rr->dir[1], rr->dir[2],
                                 rr->d[0][3], rr->d[1][3], rr->d[2][3], S.V.Z, S.V.Y,
                                 S.V.X, rr->neg != 0);
    }
    if (faces[0] == 0 || faces[1] == 0) verr = ANACOR_RAY_ERR_FACE;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      Lr[k][0] = Lr[k][1] = Lr[k][2] = Lr[k][3] = 0.0;
      lens[k] = 0; nsegs[k] = 0;
      lasts[k][0] = lasts[k][1] = lasts[k][2] = 0;
      if (verr) continue;
      RayLane L;
      if (!make_lane<LAYOUT>(&rule->ray[k], faces[k], co.x, co.y, co.z, S.V, L)) {
        verr = ANACOR_RAY_ERR_ASSERT;
        continue;
      }
      walk_ray<CAP, U, GUARD, FULL, LAYOUT, L04, !F64>(L, S.V, S.vs, seg, lens[k], nsegs[k],
                                                      lasts[k], Lr[k]);
      acc.steps += (unsigned long long)lens[k];
    }
    acc.err |= verr;
    if (PATHS) {
      PathRecord& p = paths[co.w];
      p.err = (int)verr;
      p.pad = 0;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        p.face[k] = faces[k];
        p.len[k] = lens[k];
        p.nseg[k] = nsegs[k];
        p.last[k][0] = lasts[k][0]; p.last[k][1] = lasts[k][1]; p.last[k][2] = lasts[k][2];
        p.L[k][0] = Lr[k][0]; p.L[k][1] = Lr[k][1]; p.L[k][2] = Lr[k][2]; p.L[k][3] = Lr[k][3];
      }
      continue;
    }
    // cal_rate (ray_tracing_basic.c:1554-1583), same association order
    const double r = S.mu[0] * (Lr[0][0] + Lr[1][0]) + S.mu[1] * (Lr[0][1] + Lr[1][1]) +
                     S.mu[2] * (Lr[0][2] + Lr[1][2]) + S.mu[3] * (Lr[0][3] + Lr[1][3]);
    if (F64) {
      acc.acc64 += is_exp ? exp(-r) : r;
    } else {
      const float val = is_exp ? exp_neg(r) : (float)r;
      const float y = val - acc.kc;
      const float t = acc.ks + y;
      acc.kc = (t - acc.ks) - y;
      acc.ks = t;
    }
  }
}

struct ReduceLds {
  double r64[kBlock / kWave];
  float r32[2][kBlock / kWave];
  unsigned long long steps[kBlock / kWave];
};

// Workgroup reduction of the lanes' accumulators (wave shuffles, one LDS
// round, fixed order, so results are deterministic); thread 0 publishes.
template <bool F64, bool PATHS>
__device__ __forceinline__ void reduce_tile(TileAcc& acc, ReduceLds& red, unsigned* err_word,
                                            unsigned long long* steps_total, float2* p32,
                                            double* p64) {
  const int tid = (int)threadIdx.x;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    acc.steps += __shfl_xor(acc.steps, off);
    acc.err |= __shfl_xor(acc.err, off);
    if (!PATHS) {
      if (F64) {
        acc.acc64 += __shfl_xor(acc.acc64, off);
      } else {
        const float os = __shfl_xor(acc.ks, off), oc = __shfl_xor(acc.kc, off);
        merge_pair(acc.ks, acc.kc, os, oc);
      }
    }
  }
  if (lane == 0) {
    red.steps[wave] = acc.steps;
    red.r64[wave] = acc.acc64;
    red.r32[0][wave] = acc.ks;
    red.r32[1][wave] = acc.kc;
    if (acc.err) atomicOr(err_word, acc.err);
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long st = red.steps[0];
    double s64 = red.r64[0];
    float s = red.r32[0][0], c = red.r32[1][0];
    for (int w = 1; w < kBlock / kWave; ++w) {
      st += red.steps[w];
      s64 += red.r64[w];
      merge_pair(s, c, red.r32[0][w], red.r32[1][w]);
    }
    atomicAdd(steps_total, st);
    if (!PATHS) {
      if (F64) *p64 = s64;
      else *p32 = make_float2(s, c);
    }
  }
}

// nbytes: size of the buffer behind vol (the raw volume, or the packed one).
__device__ __forceinline__ Volume make_volume(const uint8_t* vol, int Z, int Y, int X, int nbytes,
                                              int cz, int cy, const double* incd = nullptr,
                                              int incd_last = 0) {
  Volume V;
  V.incd = (const_f64_ptr)incd;
  V.incd_last = incd_last;
  V.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)vol, (short)0, nbytes, 0x00020000);
  V.Z = Z; V.Y = Y; V.X = X;
  V.sy = X + 2;
  V.sz = (Y + 2) * V.sy;
  V.base = V.sz + V.sy + 1;
  V.cz = cz; V.cy = cy;
  return V;
}

// Raw layouts: at least five waves per SIMD (the walk is bound by VALU
// issue).  Bricked layouts: no bound (1), so the compiler keeps the registers
// for eight gathers in flight per lane (the walk waits on the L2 and the
// Infinity Cache there): 110 VGPRs, four waves.
template <int CAP, int U, bool GUARD, bool FULL, int LAYOUT, bool F64, bool PATHS, bool L04>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(LAYOUT >= kLayoutBrick2 ? 1 : 5)))
void trace_kernel(TraceArgs a) {
  __shared__ uint32_t seg_lds[seg_slots(CAP, U) * kBlock];
  __shared__ ReduceLds red;

  const int lb = xcd_logical_block((int)blockIdx.x, (int)gridDim.x);
  if (lb >= a.n_blocks) return;
  int pos, tile;                         // processing position, voxel tile
  block_work(lb, a.tiles, a.R, a.group, pos, tile);
  pos += a.pos0;
  const int refl = a.order[pos];         // the caller's reflection index
  SampleView S;
  S.V = make_volume(a.vol, a.Z, a.Y, a.X, a.vol_bytes, a.cz, a.cy, a.incd, a.incd_last);
  S.coords = a.coords;
  S.N = a.N;
  S.rules = a.rules;
  for (int i = 0; i < 3; ++i) S.vs[i] = a.vs[i];
  for (int i = 0; i < 4; ++i) S.mu[i] = a.mu[i];
  lds_u32* seg = (lds_u32*)(seg_lds + threadIdx.x);

  TileAcc acc = {0.f, 0.f, 0.0, 0ull, 0u};
  const int v_begin = tile * a.vox_per_tile;
  const int v_end = min(a.N, v_begin + a.vox_per_tile);
  trace_tile<CAP, U, GUARD, FULL, LAYOUT, F64, PATHS, L04>(
      S, a.rules + refl, v_begin, v_end, a.is_exp, seg, acc,
      PATHS ? a.paths + (size_t)refl * a.N : nullptr);
  const int slot = pos * a.tiles + tile;   // finish_kernel reads [position][tile]
  reduce_tile<F64, PATHS>(acc, red, &a.err[refl], a.steps, a.partial32 + slot,
                          a.partial64 + slot);
}

// Per reflection: combine the tile partials in tile order and divide by N.
__global__ void finish_kernel(const float2* __restrict__ p32, const double* __restrict__ p64,
                              const unsigned* __restrict__ err, const int* __restrict__ order,
                              int R, int tiles, int N, int f64, double* __restrict__ out) {
  const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= R) return;
  double sum = 0.0;
  if (f64) {
    for (int t = 0; t < tiles; ++t) sum += p64[(size_t)r * tiles + t];
  } else {
    float s = 0.f, c = 0.f;
    for (int t = 0; t < tiles; ++t) {
      const float2 q = p32[(size_t)r * tiles + t];
      merge_pair(s, c, q.x, q.y);
    }
    sum = (double)s - (double)c;
  }
  const int refl = order[r];
  out[refl] = err[refl] ? __builtin_nan("") : sum / (double)N;
}

// ------------------------------------------------------------ batched path
//
// Several independent samples (own volume, voxel list, reflections) traced
// as a clear kernel, one trace_kernel launch per sample into its context's
// buffers (each launch fills the GPU: thousands of workgroups), and one
// finish kernel over every sample's reflections.  The host setup of sample
// i + 1 runs while the GPU traces sample i; uploads run on a second stream
// ahead of the launches.  The same sequence can run as one hipGraph (host
// setups as host nodes on a branch of their own that each sample's launch
// waits for), captured once per set of reflection counts and buffers and
// then relaunched; on the eight-sample workload it measured 72.8 ms per step
// against 64.0 ms for the launches from the host, because the graph ran its
// host nodes in line with the kernels (18.5 ms of setups + the traces), so
// the host launches are the default.  (A persistent work-queue kernel over
// all samples measured 96 ms against 77 ms for these launches: its state
// held 124 VGPRs and kept it from the walk's faster forms.)

// What the clear and finish kernels need of each sample.
struct BatchSample {
  const int* order;            // processing position -> reflection
  int N, R, tiles, vpt;
  float2* p32;                 // tile partials [R][tiles]
  double* p64;
  unsigned* err;               // [R]
  double* out;                 // [R]
};

// Clears the step counter and every sample's error words (a kernel rather
// than hipMemsetAsync: the batch graph holds no memset node).
__global__ void batch_clear_kernel(const BatchSample* __restrict__ samples, int S,
                                   unsigned long long* steps) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t == 0) *steps = 0ull;
  for (int s = 0; s < S; ++s) {
    const BatchSample& B = samples[s];
    for (int r = t; r < B.R; r += (int)(gridDim.x * blockDim.x)) B.err[r] = 0u;
  }
}

// Per (sample, reflection): merge the tile partials in order, divide by N.
__global__ void batch_finish_kernel(const BatchSample* __restrict__ samples, int S, int f64,
                                    long long total_refl) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_refl) return;
  long long base = 0;
  int s = 0;
  while (s < S && g >= base + samples[s].R) { base += samples[s].R; ++s; }
  const BatchSample& B = samples[s];
  const int r = (int)(g - base);
  double sum = 0.0;
  if (f64) {
    for (int t = 0; t < B.tiles; ++t) sum += B.p64[(size_t)r * B.tiles + t];
  } else {
    float a = 0.f, c = 0.f;
    for (int t = 0; t < B.tiles; ++t) {
      const float2 q = B.p32[(size_t)r * B.tiles + t];
      merge_pair(a, c, q.x, q.y);
    }
    sum = (double)a - (double)c;
  }
  const int refl = B.order[r];
  B.out[refl] = B.err[refl] ? __builtin_nan("") : sum / (double)B.N;
}

// ------------------------------------------------------------- gridding
//
// The reference's approximate path (utils/utils_gridding.py,
// src/gridding_interpolation.c): an absorption map holding, for every grid
// direction and crystal voxel, the sum of mu*L of one ray from the voxel
// along that direction (ray_tracing_single_gridding, ray_tracing.c:484-534,
// IsExp = 0, stored as float), then per reflection a look-up of the incident
// and the diffracted direction and the mean of exp(-(m1 + m2)).

struct GridArgs {
  const uint8_t* vol;
  int vol_bytes, cz, cy;
  int Z, Y, X;
  const int4* coords;          // (z, y, x, caller index)
  int N;
  const RayRule* rays;         // [D] one ray per direction
  int D, tiles, vox_per_tile, n_blocks;
  double vs[3];
  double mu[4];
  int is_exp;
  float* map;                  // [D][N], voxels in the caller's order
  double* map64;               // same, unrounded (nullable)
  unsigned* err;               // [D]
  const double* incd;          // (double)i for i = 0 .. incd_last
  int incd_last;
};

template <int CAP, int U, bool GUARD, bool FULL, int LAYOUT>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(LAYOUT >= kLayoutBrick2 ? 1 : 5)))
void grid_map_kernel(GridArgs a) {
  __shared__ uint32_t seg_lds[seg_slots(CAP, U) * kBlock];
  const int lb = xcd_logical_block((int)blockIdx.x, (int)gridDim.x);
  if (lb >= a.n_blocks) return;
  const int d = lb / a.tiles;
  const int tile = lb - d * a.tiles;
  const Volume V = make_volume(a.vol, a.Z, a.Y, a.X, a.vol_bytes, a.cz, a.cy, a.incd,
                               a.incd_last);
  lds_u32* seg = (lds_u32*)(seg_lds + threadIdx.x);
  const RayRule* rr = a.rays + d;
  unsigned err = 0;
  const int v_end = min(a.N, (tile + 1) * a.vox_per_tile);
  for (int v = tile * a.vox_per_tile + (int)threadIdx.x; v < v_end; v += kBlock) {
    const int4 co = a.coords[v];
    double val = __builtin_nan("");
    const int face = exit_face<true>(co.x, co.y, co.z, rr->dir[0], rr->dir[1], rr->dir[2],
                                     rr->d[0][3], rr->d[1][3], rr->d[2][3], a.Z, a.Y,
                                     a.X, false);
    RayLane L;
    if (face == 0) {
      err |= ANACOR_RAY_ERR_FACE;
    } else if (!make_lane<LAYOUT>(rr, face, co.x, co.y, co.z, V, L)) {
      err |= ANACOR_RAY_ERR_ASSERT;
    } else {
      double Lr[4];
      int len, nseg, last[3];
      walk_ray<CAP, U, GUARD, FULL, LAYOUT, false>(L, V, a.vs, seg, len, nseg, last, Lr);
      // cal_rate with zero incident lengths, same association (:1554-1583)
      const double r = a.mu[0] * (0.0 + Lr[0]) + a.mu[1] * (0.0 + Lr[1]) +
                       a.mu[2] * (0.0 + Lr[2]) + a.mu[3] * (0.0 + Lr[3]);
      val = a.is_exp ? exp(-r) : r;
    }
    a.map[(size_t)d * a.N + co.w] = (float)val;
    if (a.map64) a.map64[(size_t)d * a.N + co.w] = val;
  }
  if (err) atomicOr(&a.err[d], err);
}

// Look-up geometry of one reflection on the padded grid the reference builds
// (utils_gridding.py:406-445: theta_extra columns and phi_extra rows copied
// from the opposite ends, theta_list / phi_list over the padded extent).
struct GridLookup {
  const float* map;            // resident: [n_phi][n_theta][N] (the map above)
  const double* map64;         // drop-in: [N][ny][nx], padded (reference layout)
  int N, n_theta, n_phi, theta_extra, phi_extra;
  const double* theta_list;    // [nx]
  const double* phi_list;      // [ny]
  int nx, ny;
  double theta_min, theta_max, phi_min, phi_max;
  const double* angles;        // [R][4]: theta_1, phi_1 (incident), theta, phi (diffracted)
  const int* order;            // [R] processing position -> caller's index
  int R;
  double* out;                 // [R]
  unsigned* err;               // [R]
};

// nearest_neighbor_custom (gridding_interpolation.c:16-32): truncated index
// on the padded grid, clamped.
__device__ __forceinline__ void nearest_index(const GridLookup& g, double theta, double phi,
                                              int& it, int& ip) {
  const double theta_step = (g.theta_max - g.theta_min) / (g.nx - 1);
  const double phi_step = (g.phi_max - g.phi_min) / (g.ny - 1);
  it = (int)((theta - g.theta_min) / theta_step);
  ip = (int)((phi - g.phi_min) / phi_step);
  it = it < 0 ? 0 : (it >= g.nx ? g.nx - 1 : it);
  ip = ip < 0 ? 0 : (ip >= g.ny ? g.ny - 1 : ip);
}

// The interval index of GSL's bilinear interp2d (binary search of gsl_interp,
// x_array[i] <= x < x_array[i+1]); false outside the grid, where GSL reports
// a domain error.
__device__ __forceinline__ bool interval(const double* xs, int n, double x, int& i) {
  if (!(x >= xs[0] && x <= xs[n - 1])) return false;
  int lo = 0, hi = n - 1;
  while (hi > lo + 1) {
    const int m = (hi + lo) / 2;
    if (xs[m] > x) hi = m;
    else lo = m;
  }
  i = lo;
  return true;
}

template <bool RESIDENT>
__device__ __forceinline__ double grid_value(const GridLookup& g, int v, int it, int ip) {
  if (RESIDENT) {
    int ut = (it - g.theta_extra) % g.n_theta;
    int up = (ip - g.phi_extra) % g.n_phi;
    ut += ut < 0 ? g.n_theta : 0;
    up += up < 0 ? g.n_phi : 0;
    return (double)g.map[((size_t)up * g.n_theta + ut) * g.N + v];
  }
  return g.map64[((size_t)v * g.ny + ip) * g.nx + it];
}

// One workgroup per reflection; METHOD 1 nearest, 2 bilinear.
template <int METHOD, bool RESIDENT>
__global__ __launch_bounds__(kBlock) void grid_lookup_kernel(GridLookup g) {
  __shared__ double part[kBlock / kWave];
  const int pos = (int)blockIdx.x;
  if (pos >= g.R) return;
  const int r = g.order ? g.order[pos] : pos;
  const double* ang = g.angles + 4 * (size_t)r;
  int it[2], ip[2];
  double wt[2][4];
  bool ok = true;
  for (int k = 0; k < 2; ++k) {
    const double th = ang[2 * k], ph = ang[2 * k + 1];
    if (METHOD == 1) {
      nearest_index(g, th, ph, it[k], ip[k]);
    } else {
      int xi = 0, yi = 0;
      ok = ok && interval(g.theta_list, g.nx, th, xi) && interval(g.phi_list, g.ny, ph, yi);
      if (!ok) { it[k] = ip[k] = 0; continue; }
      it[k] = xi;
      ip[k] = yi;
      const double xmin = g.theta_list[xi], xmax = g.theta_list[xi + 1];
      const double ymin = g.phi_list[yi], ymax = g.phi_list[yi + 1];
      const double t = (th - xmin) / (xmax - xmin), u = (ph - ymin) / (ymax - ymin);
      wt[k][0] = (1. - t) * (1. - u);   // zminmin
      wt[k][1] = t * (1. - u);          // zmaxmin
      wt[k][2] = (1. - t) * u;          // zminmax
      wt[k][3] = t * u;                 // zmaxmax
    }
  }
  double sum = 0.0;
  if (ok) {
    for (int v = (int)threadIdx.x; v < g.N; v += kBlock) {
      double m[2];
      for (int k = 0; k < 2; ++k) {
        if (METHOD == 1) {
          m[k] = grid_value<RESIDENT>(g, v, it[k], ip[k]);
        } else {
          const double zmm = grid_value<RESIDENT>(g, v, it[k], ip[k]);
          const double zMm = grid_value<RESIDENT>(g, v, it[k] + 1, ip[k]);
          const double zmM = grid_value<RESIDENT>(g, v, it[k], ip[k] + 1);
          const double zMM = grid_value<RESIDENT>(g, v, it[k] + 1, ip[k] + 1);
          m[k] = wt[k][0] * zmm + wt[k][1] * zMm + wt[k][2] * zmM + wt[k][3] * zMM;
        }
      }
      sum += exp(-(m[0] + m[1]));
    }
  }
  for (int off = kWave / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
  if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) s += part[w];
    g.out[r] = ok ? s / (double)g.N : __builtin_nan("");
    if (!ok) g.err[r] = ANACOR_RAY_ERR_FACE;
  }
}

// ------------------------------------------------------- volume packing

// flags |= 1 when some label is outside 0..3 (2-bit bricks impossible),
// |= 2 when some label is outside 0..4 (the walk must test which labels open
// a boundary).
__global__ void label_scan_kernel(const uint8_t* __restrict__ vol, long long n,
                                  unsigned* __restrict__ flags) {
  bool over3 = false, over4 = false;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    over3 = over3 || vol[i] > 3;
    over4 = over4 || vol[i] > 4;
  }
  const unsigned f = (__any(over3) ? 1u : 0u) | (__any(over4) ? 2u : 0u);
  if (f && (threadIdx.x & (kWave - 1)) == 0) atomicOr(flags, f);
}

// The caller's [Z][Y][X] volume into the raw layout [Z+2][Y+2][X+2] with its
// one-voxel shell of label 0 (n = padded voxel count).
__global__ void shell_volume_kernel(const uint8_t* __restrict__ in, int Z, int Y, int X,
                                    long long n, uint8_t* __restrict__ out) {
  const int sx = X + 2, sy = Y + 2;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(i % sx) - 1;
    const long long t = i / sx;
    const int y = (int)(t % sy) - 1;
    const int z = (int)(t / sy) - 1;
    uint8_t v = 0;
    if ((unsigned)z < (unsigned)Z && (unsigned)y < (unsigned)Y && (unsigned)x < (unsigned)X)
      v = in[((size_t)z * Y + y) * X + x];
    out[i] = v;
  }
}

// One thread per dword of the bricked volume (16 2-bit or 8 4-bit labels);
// voxels of the padding beyond the volume are 0.
__global__ void pack_volume_kernel(const uint8_t* __restrict__ vol, int Z, int Y, int X, int NBX,
                                   int NBY, int bits, long long ndw, uint32_t* __restrict__ out) {
  const long long d = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndw) return;
  const int per = 32 / bits;
  uint32_t word = 0;
  for (int j = 0; j < per; ++j) {
    const long long f = d * per + j;
    const long long b = f >> 9;
    const int w = (int)(f & 511);
    const int bx = (int)(b % NBX);
    const long long t = b / NBX;
    const int by = (int)(t % NBY), bz = (int)(t / NBY);
    const int z = bz * 8 + (w >> 6), y = by * 8 + ((w >> 3) & 7), x = bx * 8 + (w & 7);
    uint32_t lab = 0;   // vol: the raw layout with its zero shell
    if (z < Z && y < Y && x < X) lab = vol[((size_t)(z + 1) * (Y + 2) + (y + 1)) * (X + 2) + x + 1];
    const uint32_t code = bits == 2 ? lab : (lab <= 4u ? lab : 5u);
    word |= code << (j * bits);
  }
  out[d] = word;
}

// ------------------------------------------------------------ host side

// Grow-only page-locked host buffer (hipHostMalloc) for uploads that should
// run as asynchronous DMA.  Callers synchronise their stream before it grows.
template <class T>
struct Pinned {
  T* p = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t n) {
    if (n <= cap) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const hipError_t e = hipHostMalloc((void**)&p, n * sizeof(T), hipHostMallocDefault);
    if (e == hipSuccess) cap = n;
    return e;
  }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  T& operator[](size_t i) { return p[i]; }
  T* data() { return p; }
  const T* data() const { return p; }
};

// Device allocation owned by one scope: freed on every way out of it.
struct DeviceTemp {
  void* p = nullptr;
  DeviceTemp() = default;
  DeviceTemp(const DeviceTemp&) = delete;
  DeviceTemp& operator=(const DeviceTemp&) = delete;
  ~DeviceTemp() {
    if (p) (void)hipFree(p);
  }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  uint8_t* d_vol = nullptr;              // raw layout with its zero shell
  long long vol_bytes = 0;               // (Z + 2)(Y + 2)(X + 2)
  double* d_incd = nullptr;              // (double)i, i = 0 .. incd_last (walk increments)
  int incd_last = 0;
  int Z = 0, Y = 0, X = 0;
  int layout_req = ANACOR_LAYOUT_AUTO;   // requested layout of the trace kernel
  int layout = kLayoutRaw;               // layout in use (kLayout*)
  uint8_t* d_pack = nullptr;             // bricked copy of the volume
  int pack_bits = 0;                     // 2 or 4 once built
  long long pack_bytes = 0;
  int NBX = 0, NBY = 0, NBZ = 0;
  bool labels_over3 = true;              // some label outside 0..3 (scanned at creation)
  bool labels_over4 = true;              // some label outside 0..4
  int4* d_coords = nullptr;
  int N = 0;
  double vs[3] = {1, 1, 1};
  double mu[4] = {0, 0, 0, 0};
  int full = 0;
  int tile_voxels = 0;          // 0: automatic
  int refl_group = 0;           // reflections per launch group (0: automatic, 1: none)
  int sort_major = -1;          // direction sort key led by ray 0 (incident) or 1 (diffracted);
                                // -1: automatic (sort_major_for)
  std::vector<int4> h_coords;   // (z, y, x, index) in the caller's order
  int block[3] = {0, 0, 0};     // voxel grouping in use (0,0,0: Morton)
  // reusable per-run buffers
  ReflRule* d_rules = nullptr;
  size_t rules_cap = 0;
  void* d_partial = nullptr;
  size_t partial_cap = 0;
  unsigned* d_err = nullptr;
  double* d_out = nullptr;
  size_t refl_cap = 0;
  unsigned long long* d_steps = nullptr;
  Pinned<ReflRule> h_rules;         // rules in the caller's order (upload staging)
  std::vector<int> h_order;         // processing position -> caller's index
  Pinned<int> h_order_pin;          // its upload staging
  int* d_order = nullptr;
  // bisection traversal
  IbRule* d_ib = nullptr;
  size_t ib_cap = 0;
  Pinned<IbRule> h_ib;
  unsigned long long* d_oob = nullptr;
  unsigned long long last_oob = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // uploads of a call's second part, beside the first part's trace
  hipStream_t copy_stream = nullptr;
  hipEvent_t copied = nullptr;
  float last_kernel_ms = 0.f;
  unsigned long long last_steps = 0;
  int64_t last_blocks = 0;
  double last_host_ms = 0.0;
  std::mutex mu_run;
};

static uint64_t spread_bits(uint64_t v) {   // 21 bits -> every third bit
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

static int ensure_run_buffers(Context* c, int64_t R, int64_t n_partials, bool f64) {
  if ((size_t)R > c->rules_cap) {
    if (c->d_rules) HIP_TRY(hipFree(c->d_rules));
    c->d_rules = nullptr;
    HIP_TRY(hipMalloc(&c->d_rules, (size_t)R * sizeof(ReflRule)));
    c->rules_cap = (size_t)R;
  }
  const size_t pbytes = (size_t)n_partials * (f64 ? sizeof(double) : sizeof(float2));
  if (pbytes > c->partial_cap) {
    if (c->d_partial) HIP_TRY(hipFree(c->d_partial));
    c->d_partial = nullptr;
    HIP_TRY(hipMalloc(&c->d_partial, pbytes));
    c->partial_cap = pbytes;
  }
  if ((size_t)R > c->refl_cap) {
    if (c->d_err) HIP_TRY(hipFree(c->d_err));
    if (c->d_out) HIP_TRY(hipFree(c->d_out));
    if (c->d_order) HIP_TRY(hipFree(c->d_order));
    c->d_err = nullptr;
    c->d_out = nullptr;
    c->d_order = nullptr;
    HIP_TRY(hipMalloc(&c->d_err, (size_t)R * sizeof(unsigned)));
    HIP_TRY(hipMalloc(&c->d_out, (size_t)R * sizeof(double)));
    HIP_TRY(hipMalloc(&c->d_order, (size_t)R * sizeof(int)));
    c->refl_cap = (size_t)R;
  }
  if (!c->d_steps) HIP_TRY(hipMalloc(&c->d_steps, sizeof(unsigned long long)));
  return ANACOR_OK;
}

// Voxel tile per workgroup: enough workgroups to fill 256 CUs several times
// over, at least one voxel per lane, at most 64 voxels per lane.
// A dispatch holds at most 2^32 - 1 work-items: 2^24 - 8 workgroups of 256
// (a multiple of 8, for the XCD mapping of xcd_logical_block).
constexpr int64_t kMaxBlocks = (1ll << 24) - 8;

// The tile (voxels per workgroup) for R rows of work over n voxels: vpt,
// enlarged until R * tiles fits one dispatch; 0 when even one tile per row
// does not fit (more than kMaxBlocks rows).
static int fit_tile(int64_t n, int64_t R, int vpt) {
  if (n <= 0 || R <= 0) return vpt;
  if (R * ((n + vpt - 1) / vpt) <= kMaxBlocks) return vpt;
  const int64_t max_tiles = kMaxBlocks / R;
  if (max_tiles < 1) return 0;
  return (int)((n + max_tiles - 1) / max_tiles);
}

// Automatic: at most 4 voxels per lane (1,024 per tile; measured on the 256^3
// workload: 80 ms at 1,024, 84 ms at 6,144, 108 ms from 8,192 up), 1.5 in the
// bricked layouts, where a smaller tile narrows the bundle of rays whose
// bricks a launch group shares (1000x1000x800, 10k reflections, groups of
// 2,048, sort led by the diffracted ray, three workgroups per CU: 939 ms at
// 512, 864 ms at 384, 894 ms at 256; the same volume with seed 7, 6k
// reflections: 580, 527, 544 ms; 700x900x900, 80k voxels: 398, 378, 378 ms).
static bool is_bricked(int layout);
static int choose_tile(const Context* c, int64_t R, int layout = -1) {
  if (c->tile_voxels > 0) return c->tile_voxels;
  const int64_t cap = is_bricked(layout) ? kBlock + kBlock / 2 : (int64_t)kBlock * 4;
  const int64_t target_blocks = 256 * 16;
  int64_t tiles = (target_blocks + R - 1) / std::max<int64_t>(R, 1);
  tiles = std::max<int64_t>(1, std::min<int64_t>(tiles, (c->N + kBlock - 1) / kBlock));
  int64_t vpt = (c->N + tiles - 1) / tiles;
  vpt = std::min<int64_t>(std::max<int64_t>(vpt, kBlock), cap);
  return (int)vpt;
}

constexpr int kCap = 16;     // boundaries kept per lane before the re-walk fallback
constexpr int kUnrollRaw = 6;  // label gathers in flight per lane, raw layouts

// Label gathers in flight per lane.  Raw layouts (256^3 workload, clamp-free
// forms, 1,024-voxel tiles): 2: 91.9 ms, 3: 83.8, 4: 80.1, 5: 78.8, 6: 78.4,
// 8: 78.4.  The bricked walk waits on L2/Infinity Cache latency with VALU to
// spare (1000x1000x800, 1,000 reflections: 164 ms with 4, 140 ms with 8;
// 10k reflections at three workgroups per CU: 892.8 ms with 8, 904.0 with
// 12, 1,697 with 16, whose code the compiler serialises).
constexpr int unroll_for(int layout) { return layout >= kLayoutBrick2 ? 8 : kUnrollRaw; }

// Workgroups of the bricked walk resident per CU.  On a volume far larger
// than the caches the bricked walk is bound by how much of the bricks that its
// rays in flight cross the L2 holds, and fewer resident workgroups keep that
// working set smaller (1000x1000x800 volume, 10k reflections, 256-voxel
// tiles, groups of 2,048: 996.7 ms with 4 workgroups per CU, the register
// limit; 892.8 ms with 3; 891.7 ms with 2.  The 256^3 volume in bricks: 146,
// 150 and 170 ms).  The limit is set by the launch's dynamic LDS, sized from
// the device's LDS per CU so that no further workgroup fits beside them.  The
// raw walk, bound by VALU issue on cache-resident volumes, keeps its register
// limit of five (256^3 workload: 62.4 ms; 65.7 at four, 73.3 at three).
constexpr int kBrickWgPerCu = 3;

// Dynamic LDS bytes that hold launches of `kernel` to `per_cu` workgroups per
// CU on the current device; 0 (no limit) when the device attributes are not
// available or the limit does not fit a workgroup's LDS.
// (a fixed table, so that no call allocates, not even host memory)
static int lds_pad_for(const void* kernel, int per_cu) {
  struct Entry {
    int dev;
    const void* kernel;
    int pad;
  };
  static std::mutex mu;
  static Entry cache[256];
  static int used = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  std::lock_guard<std::mutex> lk(mu);
  for (int i = 0; i < used; ++i)
    if (cache[i].dev == dev && cache[i].kernel == kernel) return cache[i].pad;
  int pad = 0, lds_cu = 0, lds_block = 0;
  hipFuncAttributes fa{};
  if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) ==
          hipSuccess &&
      hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
          hipSuccess &&
      hipFuncGetAttributes(&fa, kernel) == hipSuccess) {
    // just over 1 / (per_cu + 1) of the CU's LDS per workgroup, whole KiB: a
    // further workgroup cannot fit, and the allocation granularity cannot
    // push out one of the per_cu
    const int target = ((lds_cu / (per_cu + 1)) & ~1023) + 1024;
    const int stat = (int)fa.sharedSizeBytes;
    if (target > stat && target <= lds_block && (int64_t)per_cu * target <= lds_cu)
      pad = target - stat;
  } else {
    (void)hipGetLastError();
  }
  if (used < 256) cache[used++] = Entry{dev, kernel, pad};
  return pad;
}

template <bool GUARD, bool FULL, int LAYOUT, bool F64, bool PATHS, bool L04>
static void launch_trace(const TraceArgs& a, int grid, hipStream_t s) {
  auto* k = &trace_kernel<kCap, unroll_for(LAYOUT), GUARD, FULL, LAYOUT, F64, PATHS, L04>;
  const int pad = LAYOUT >= kLayoutBrick2 ? lds_pad_for((const void*)k, kBrickWgPerCu) : 0;
  hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), (unsigned)pad, s, a);
}

using LaunchFn = void (*)(const TraceArgs&, int, hipStream_t);

// Index bits: guard 1, full 2, f64 4, paths 8, layout 16 * (0..3), labels in
// 0..4 64.  Path records are written from the raw layouts only; the 0..4
// variant exists for the raw layouts' A_h kernels without the guard.
template <int I>
static constexpr LaunchFn launcher() {
  constexpr int layout = (I >> 4) & 3;
  constexpr bool guard = (I & 1) != 0, paths = (I & 8) != 0, l04 = (I & 64) != 0;
  if constexpr (paths && layout >= kLayoutBrick2) {
    return nullptr;
  } else if constexpr (l04 && (guard || paths || layout >= kLayoutBrick2)) {
    return nullptr;
  } else {
    return &launch_trace<guard, (I & 2) != 0, layout, (I & 4) != 0, paths, l04>;
  }
}

template <int... I>
static constexpr std::array<LaunchFn, sizeof...(I)> launch_table(std::integer_sequence<int, I...>) {
  return {launcher<I>()...};
}

static const std::array<LaunchFn, 128> kLaunch =
    launch_table(std::make_integer_sequence<int, 128>{});

// labels04: every label of the volume is in 0..4.
static void dispatch_trace(const TraceArgs& a, int grid, bool guard, bool full, int layout,
                           bool f64, bool paths, bool labels04, hipStream_t s) {
  const bool l04 = labels04 && !guard && !paths && layout < kLayoutBrick2;
  const int i = (guard ? 1 : 0) | (full ? 2 : 0) | (f64 ? 4 : 0) | (paths ? 8 : 0) |
                (layout << 4) | (l04 ? 64 : 0);
  kLaunch[i](a, grid, s);
}

// 24-bit address arithmetic unless a z-plane or the volume is too large for it.
static bool raw_needs_wide(const Context* c) {
  return (int64_t)(c->Y + 2) * (c->X + 2) >= (1 << 23) || c->Z + 2 >= (1 << 23);
}

// The bricked layouts need every factor of the address chain in 24 bits and
// every partial sum in 31: brick constants below 2^23, padded volume below 2^30.
static bool bricks_fit(const Context* c) {
  const int64_t nbx = (c->X + 7) / 8, nby = (c->Y + 7) / 8, nbz = (c->Z + 7) / 8;
  return 512 * nbx * nby < (1 << 23) && 512 * nbx * nby * nbz < (1 << 30) &&
         c->X < (1 << 23) && c->Y < (1 << 23) && c->Z < (1 << 23);
}

// Choose (and if needed build) the volume layout of the trace kernel.
// Automatic: bricks for volumes above 64 MB, which exceed the L2s and most of
// the Infinity Cache in the raw layout; raw otherwise, where the kernel is
// bound by VALU issue and the brick address arithmetic would only add to it.
// may_build false: an error instead of allocating a bricked copy that is not
// there yet.
static int ensure_layout(Context* c, bool may_build = true) {
  int want = c->layout_req;
  if (want == ANACOR_LAYOUT_AUTO) {
    const int64_t nvox = (int64_t)c->Z * c->Y * c->X;
    const char* env = std::getenv("ANACOR_VOLUME_LAYOUT");
    if (env && std::strcmp(env, "raw") == 0) want = ANACOR_LAYOUT_RAW;
    else if (env && std::strcmp(env, "brick") == 0) want = ANACOR_LAYOUT_BRICK;
    else want = nvox > (64ll << 20) ? ANACOR_LAYOUT_BRICK : ANACOR_LAYOUT_RAW;
    if (want != ANACOR_LAYOUT_RAW && !bricks_fit(c)) want = ANACOR_LAYOUT_RAW;
  }
  if (want == ANACOR_LAYOUT_RAW) {
    c->layout = raw_needs_wide(c) ? kLayoutWide : kLayoutRaw;
    return ANACOR_OK;
  }
  if (!bricks_fit(c)) {
    set_error(ANACOR_ERR_ARG, "volume %dx%dx%d too large for the bricked layout", c->Z, c->Y,
              c->X);
    return ANACOR_ERR_ARG;
  }
  int bits = (want == ANACOR_LAYOUT_BRICK4 || c->labels_over3) ? 4 : 2;
  if (want == ANACOR_LAYOUT_BRICK2 && c->labels_over3) {
    set_error(ANACOR_ERR_ARG, "2-bit bricks need labels in 0..3");
    return ANACOR_ERR_ARG;
  }
  if (c->pack_bits != bits && !may_build) {
    set_error(ANACOR_ERR_ARG, "the %d-bit bricked copy of the volume is not built yet "
              "(anacor_ctx_set_layout builds it)", bits);
    return ANACOR_ERR_ARG;
  }
  if (c->pack_bits != bits) {
    c->NBX = (c->X + 7) / 8;
    c->NBY = (c->Y + 7) / 8;
    c->NBZ = (c->Z + 7) / 8;
    const long long nb = (long long)c->NBX * c->NBY * c->NBZ;
    const long long bytes = nb * 512 * bits / 8;
    if (c->d_pack) HIP_TRY(hipFree(c->d_pack));
    c->d_pack = nullptr;
    c->pack_bits = 0;
    HIP_TRY(hipMalloc(&c->d_pack, (size_t)bytes));
    const long long ndw = bytes / 4;
    hipLaunchKernelGGL(pack_volume_kernel, dim3((unsigned)((ndw + 255) / 256)), dim3(256), 0,
                       c->stream, c->d_vol, c->Z, c->Y, c->X, c->NBX, c->NBY, bits, ndw,
                       (uint32_t*)c->d_pack);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->pack_bits = bits;
    c->pack_bytes = bytes;
  }
  c->layout = bits == 2 ? kLayoutBrick2 : kLayoutBrick4;
  return ANACOR_OK;
}

// Processing order of the reflections: sorted by the directions of their two
// rays (7-bit Morton code of each unit vector, ray `major` first), so the
// workgroups resident on an XCD at one time trace nearly parallel rays whose
// label gathers share that XCD's L2.  Results are written back through the
// inverse permutation, so the order is invisible to callers.
// order[R] receives the processing order; key[R] and tmp[R] are work space.
//
// Stable LSD radix sort of v[0..R) by bits [lo, 64), 11 bits per pass (tmp:
// R values of work space).  The keys below carry the reflection index in
// their low bits, so every key is distinct and the result is the one order
// a comparison sort gives them; on 20k reflections it takes 0.26 ms against
// std::sort's 1.17 ms.
static void radix_sort_high(uint64_t* v, uint64_t* tmp, int64_t R, int lo) {
  uint32_t cnt[2048];
  uint64_t* a = v;
  uint64_t* b = tmp;
  for (int sh = lo; sh < 64; sh += 11) {
    std::memset(cnt, 0, sizeof cnt);
    for (int64_t i = 0; i < R; ++i) ++cnt[(a[i] >> sh) & 2047u];
    uint32_t sum = 0;
    for (int d = 0; d < 2048; ++d) {
      const uint32_t n = cnt[d];
      cnt[d] = sum;
      sum += n;
    }
    for (int64_t i = 0; i < R; ++i) b[cnt[(a[i] >> sh) & 2047u]++] = a[i];
    std::swap(a, b);
  }
  if (a != v) std::memcpy(v, a, (size_t)R * sizeof(uint64_t));
}

template <class DirOf>   // DirOf(i, k): direction (z, y, x) of ray k of reflection i
static void direction_order(int64_t R, DirOf dir_of, int* order, uint64_t* key, uint64_t* tmp,
                            int major = 0) {
  for (int64_t i = 0; i < R; ++i) order[i] = (int)i;
  const char* env = std::getenv("ANACOR_SORT_REFLECTIONS");
  if (env && env[0] == '0') return;
  auto q = [](double v) {
    double t = (v + 1.0) * 64.0;
    if (!(t >= 0.0)) t = 0.0;
    if (t > 127.0) t = 127.0;
    return (uint64_t)t;
  };
  for (int64_t i = 0; i < R; ++i) {
    const double* a = dir_of(i, major);
    const double* b = dir_of(i, 1 - major);
    const uint64_t ka = (spread_bits(q(a[0])) << 2) | (spread_bits(q(a[1])) << 1) | spread_bits(q(a[2]));
    const uint64_t kb = (spread_bits(q(b[0])) << 2) | (spread_bits(q(b[1])) << 1) | spread_bits(q(b[2]));
    key[i] = (ka << 21) | kb;
  }
  if (R <= (int64_t(1) << 22)) {
    // 42-bit key above a 22-bit index, sorted by the key: the input is in
    // index order and the radix sort is stable, so ties keep the index order
    for (int64_t i = 0; i < R; ++i) key[i] = (key[i] << 22) | (uint64_t)i;
    radix_sort_high(key, tmp, R, 22);
    for (int64_t i = 0; i < R; ++i) order[i] = (int)(key[i] & ((1u << 22) - 1));
    return;
  }
  std::stable_sort(order, order + R, [&](int x, int y) { return key[x] < key[y]; });
}

template <class DirOf>
static void order_by_direction(Context* c, int64_t R, DirOf dir_of, int major) {
  c->h_order.resize((size_t)R);
  std::vector<uint64_t> key((size_t)R), tmp((size_t)R);
  direction_order(R, dir_of, c->h_order.data(), key.data(), tmp.data(), major);
}

static bool is_bricked(int layout) { return layout == kLayoutBrick2 || layout == kLayoutBrick4; }

// The ray whose direction leads the sort key.  Automatic: the incident one in
// the raw layouts; the diffracted one in the bricked layouts, whose launch
// groups then share the diffracted rays' bricks (1000x1000x800, 10k
// reflections, groups of 128, 512-voxel tiles: 1,156 ms led by the incident
// ray, 1,090 ms by the diffracted one).
static int sort_major_for(const Context* c, int layout) {
  if (c->sort_major >= 0) return c->sort_major;
  return is_bricked(layout) ? 1 : 0;
}

static void order_reflections(Context* c, const ReflRule* rules, int64_t R, int layout) {
  order_by_direction(c, R, [rules](int64_t i, int k) { return rules[i].ray[k].dir; },
                     sort_major_for(c, layout));
}

// Reflections per group of the trace launch order (block_work).  Automatic:
// none in the raw layouts, whose volumes the L2s and the Infinity Cache
// already hold; kBrickGroup in the bricked ones (1000x1000x800, 10k
// reflections, 256-voxel tiles: 1,350 ms without groups and the incident
// ray leading the sort, 1,025 ms with groups of 64, 1,014 ms of 512,
// 995 ms of 2,048).
constexpr int kBrickGroup = 2048;
static int launch_group(const Context* c, int layout) {
  if (c->refl_group > 0) return c->refl_group;
  return (layout == kLayoutBrick2 || layout == kLayoutBrick4) ? kBrickGroup : 1;
}

// Device buffers of one trace call: the context's grow-only ones
// (ensure_run_buffers) or, for anacor_ctx_run_noalloc, parts of the static
// scratch array.
struct RunBuffers {
  ReflRule* rules = nullptr;           // [R]
  int* order = nullptr;                // [R] processing position -> caller's index
  unsigned* err = nullptr;             // [R] per-reflection error bits
  double* out = nullptr;               // [R] means
  void* partial = nullptr;             // [R * tiles] float2 or double
  unsigned long long* steps = nullptr; // voxel-step counter
};

// Clears what a trace call accumulates into (a fill kernel rather than
// hipMemsetAsync, so the sequence holds no memset node if it is captured).
__global__ void clear_run_kernel(unsigned* __restrict__ err, int R,
                                 unsigned long long* __restrict__ steps) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < R) err[i] = 0u;
  if (i == 0) *steps = 0ull;
}

// One trace of R reflections from their host rules h_rules and processing
// order h_order (both uploaded), with vpt voxels per tile in `layout`: the
// trace kernel, then (without path records) the finish kernel and the
// download of the means into out.  h_err (R words, host) receives the error
// words when some mean is NaN; nullptr: a temporary.  Synchronous.
//
// With first in (0, R) and prepare_rest set, the call runs in two parts: the
// rules and order of positions [0, first) (reflections [0, first)) go up and
// their trace is launched, then prepare_rest() fills h_rules and h_order for
// [first, R) on the host while the GPU traces the first part, and those go
// up and are traced behind it.
static int trace_call(Context* c, const RunBuffers& B, const ReflRule* h_rules,
                      const int* h_order, int64_t R, int vpt, int layout, int flags, double* out,
                      PathRecord* d_paths, PathRecord* paths_out, unsigned* h_err,
                      int64_t first = -1, const std::function<int()>& prepare_rest = {}) {
  const bool f64 = (flags & ANACOR_FLAG_F64_MEAN) != 0;
  const bool paths = paths_out != nullptr;
  const int64_t tiles = (c->N + vpt - 1) / vpt;
  const int64_t nblk = R * tiles;
  const int64_t RA = (prepare_rest && first > 0 && first < R) ? first : R;
  auto guard_of = [&](int64_t lo, int64_t hi) {
    bool g = false;
    for (int64_t i = lo; i < hi; ++i)
      g = g || h_rules[i].ray[0].guard || h_rules[i].ray[1].guard;
    return g;
  };

  HIP_TRY(hipMemcpyAsync(B.rules, h_rules, (size_t)RA * sizeof(ReflRule),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(B.order, h_order, (size_t)RA * sizeof(int), hipMemcpyHostToDevice,
                         c->stream));
  hipLaunchKernelGGL(clear_run_kernel, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, c->stream,
                     B.err, (int)R, B.steps);
  HIP_TRY(hipGetLastError());

  TraceArgs a;
  const bool bricked = layout == kLayoutBrick2 || layout == kLayoutBrick4;
  a.vol = bricked ? c->d_pack : c->d_vol;
  a.vol_bytes = bricked ? (int)c->pack_bytes : (int)c->vol_bytes;
  a.cz = 512 * c->NBX * c->NBY - 512;
  a.cy = 512 * c->NBX - 64;
  a.Z = c->Z; a.Y = c->Y; a.X = c->X;
  a.coords = c->d_coords;
  a.N = c->N;
  a.rules = B.rules;
  a.R = 0;                     // per launch, below
  a.tiles = (int)tiles;
  a.vox_per_tile = vpt;
  a.n_blocks = 0;
  for (int i = 0; i < 3; ++i) a.vs[i] = c->vs[i];
  for (int i = 0; i < 4; ++i) a.mu[i] = c->mu[i];
  a.is_exp = (flags & ANACOR_FLAG_NO_EXP) ? 0 : 1;
  a.partial32 = (float2*)B.partial;
  a.partial64 = (double*)B.partial;
  a.err = B.err;
  a.steps = B.steps;
  a.paths = d_paths;
  a.order = B.order;
  a.incd = c->d_incd;
  a.incd_last = c->incd_last;
  a.group = launch_group(c, layout);

  // a dispatch holds at most kMaxBlocks workgroups: a longer call runs as
  // several launches of whole reflections (consecutive processing positions)
  const int64_t per_launch = std::max<int64_t>(1, kMaxBlocks / std::max<int64_t>(tiles, 1));
  auto launch_range = [&](int64_t lo, int64_t hi, bool guard) -> int {
    for (int64_t p0 = lo; p0 < hi; p0 += per_launch) {
      const int64_t Rc = std::min<int64_t>(per_launch, hi - p0);
      a.R = (int)Rc;
      a.pos0 = (int)p0;
      a.n_blocks = (int)(Rc * tiles);
      dispatch_trace(a, (int)((Rc * tiles + 7) / 8 * 8), guard, c->full != 0, layout, f64,
                     paths, !c->labels_over4, c->stream);
      HIP_TRY(hipGetLastError());
    }
    return ANACOR_OK;
  };
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  int rc = launch_range(0, RA, guard_of(0, RA));
  if (rc) return rc;
  if (RA < R) {
    rc = prepare_rest();
    if (rc) return rc;
    // the first part's trace reads positions and reflections below RA only,
    // so these uploads run beside it on the copy stream
    HIP_TRY(hipMemcpyAsync(B.rules + RA, h_rules + RA, (size_t)(R - RA) * sizeof(ReflRule),
                           hipMemcpyHostToDevice, c->copy_stream));
    HIP_TRY(hipMemcpyAsync(B.order + RA, h_order + RA, (size_t)(R - RA) * sizeof(int),
                           hipMemcpyHostToDevice, c->copy_stream));
    HIP_TRY(hipEventRecord(c->copied, c->copy_stream));
    HIP_TRY(hipStreamWaitEvent(c->stream, c->copied, 0));
    rc = launch_range(RA, R, guard_of(RA, R));
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  if (!paths) {
    hipLaunchKernelGGL(finish_kernel, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, c->stream,
                       (const float2*)B.partial, (const double*)B.partial, B.err, B.order,
                       (int)R, (int)tiles, c->N, f64 ? 1 : 0, B.out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, B.out, (size_t)R * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  } else {
    HIP_TRY(hipMemcpyAsync(paths_out, d_paths, (size_t)R * c->N * sizeof(PathRecord),
                           hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(&c->last_steps, B.steps, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipEventElapsedTime(&c->last_kernel_ms, c->ev0, c->ev1));
  c->last_blocks = nblk;
  if (paths) return ANACOR_OK;
  // report the first failing reflection (its result is already NaN); the
  // error words are read back only when some result is NaN
  int64_t bad = -1;
  for (int64_t i = 0; i < R && bad < 0; ++i)
    if (out[i] != out[i]) bad = i;
  if (bad < 0) return ANACOR_OK;
  std::vector<unsigned> tmp;
  if (!h_err) {
    tmp.resize((size_t)R);
    h_err = tmp.data();
  }
  HIP_TRY(hipMemcpy(h_err, B.err, (size_t)R * sizeof(unsigned), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < R; ++i)
    if (h_err[i]) {
      set_error(ANACOR_ERR_GEOMETRY,
                "reflection %lld: %s (the reference aborts here); result set to NaN",
                (long long)i,
                (h_err[i] & ANACOR_RAY_ERR_FACE) ? "no exit face (cube_face)"
                                                 : "exit face inconsistent with beam angles");
      return ANACOR_ERR_GEOMETRY;
    }
  return ANACOR_OK;
}

// With first in (0, R) and setup_rest set (no path records): only the rules
// of reflections [0, first) are ready; setup_rest() fills the rest into
// h_rules + first while the GPU traces the first part (trace_call), and each
// part is sorted by direction on its own.
static int run_rules(Context* c, const ReflRule* h_rules, int64_t R, int flags, double* out,
                     PathRecord* paths_out, int64_t first = -1,
                     const std::function<int()>& setup_rest = {}) {
  const bool f64 = (flags & ANACOR_FLAG_F64_MEAN) != 0;
  const bool paths = paths_out != nullptr;
  c->last_oob = 0;
  if (R == 0) return ANACOR_OK;
  if (c->N == 0) {
    for (int64_t i = 0; i < R; ++i) out ? (void)(out[i] = __builtin_nan("")) : (void)0;
    return ANACOR_OK;
  }
  int layout = raw_needs_wide(c) ? kLayoutWide : kLayoutRaw;   // path records: raw layout
  int rc = ANACOR_OK;
  if (!paths) {
    rc = ensure_layout(c);
    if (rc) return rc;
    layout = c->layout;
  }
  const int vpt = paths ? kBlock : choose_tile(c, R, layout);
  const int64_t tiles = (c->N + vpt - 1) / vpt;
  if (R > (int64_t)INT32_MAX / 2 || R * tiles > ((int64_t)1 << 40)) {
    set_error(ANACOR_ERR_ARG, "%lld reflections x %lld tiles in one call", (long long)R,
              (long long)tiles);
    return ANACOR_ERR_ARG;
  }
  rc = ensure_run_buffers(c, R, R * tiles, f64);
  if (rc) return rc;
  const int64_t RA = (!paths && setup_rest && first > 0 && first < R) ? first : R;
  order_reflections(c, h_rules, RA, layout);
  HIP_TRY(c->h_order_pin.reserve((size_t)R));
  for (int64_t i = 0; i < RA; ++i) c->h_order_pin[(size_t)i] = c->h_order[(size_t)i];
  // the second part: its setup, then its own direction order, as positions
  // [RA, R) of reflections [RA, R)
  auto prepare_rest = [&]() -> int {
    const int rc2 = setup_rest();
    if (rc2) return rc2;
    const int64_t n = R - RA;
    std::vector<int> ord((size_t)n);
    std::vector<uint64_t> key((size_t)n), tmp((size_t)n);
    direction_order(n, [&](int64_t i, int k) { return h_rules[RA + i].ray[k].dir; },
                    ord.data(), key.data(), tmp.data(), sort_major_for(c, layout));
    for (int64_t i = 0; i < n; ++i)
      c->h_order_pin[(size_t)(RA + i)] = (int)RA + ord[(size_t)i];
    return ANACOR_OK;
  };
  DeviceTemp paths_buf;   // path records, freed on every return below
  if (paths) HIP_TRY(hipMalloc(&paths_buf.p, (size_t)R * c->N * sizeof(PathRecord)));
  RunBuffers B;
  B.rules = c->d_rules;
  B.order = c->d_order;
  B.err = c->d_err;
  B.out = c->d_out;
  B.partial = c->d_partial;
  B.steps = c->d_steps;
  if (RA < R)
    return trace_call(c, B, h_rules, c->h_order_pin.data(), R, vpt, layout, flags, out,
                      (PathRecord*)paths_buf.p, paths_out, nullptr, RA, prepare_rest);
  return trace_call(c, B, h_rules, c->h_order_pin.data(), R, vpt, layout, flags, out,
                    (PathRecord*)paths_buf.p, paths_out, nullptr);
}

static void ctx_destroy(Context* c);

// Default grouping of crystal voxels into waves (voxel units, z y x): lanes
// run along x within a block, blocks in z, y, x order.  A wave's rays are
// parallel, so lanes that share z and y keep sharing them at every step and
// gather from one row of the [z][y][x] volume.
constexpr int kDefaultBlock[3] = {1, 2, 64};

// Order the voxel list for the kernels and upload it.  (bz, by, bx) > 0:
// blocks of that shape in z, y, x order, lexicographic (z, y, x) inside a
// block; (0, 0, 0): Morton order.
static int upload_voxel_order(Context* c, int bz, int by, int bx) {
  const size_t n = c->h_coords.size();
  if (n == 0) return ANACOR_OK;
  std::vector<uint64_t> key(n);
  const bool morton = bz <= 0 || by <= 0 || bx <= 0;
  for (size_t i = 0; i < n; ++i) {
    const int4 v = c->h_coords[i];
    const uint64_t z = (uint64_t)v.x, y = (uint64_t)v.y, x = (uint64_t)v.z;
    if (morton) {
      key[i] = (spread_bits(z) << 2) | (spread_bits(y) << 1) | spread_bits(x);
    } else {
      // block index in (z, y, x) order
      key[i] = ((z / bz) << 42) | ((y / by) << 21) | (x / bx);
    }
  }
  std::vector<int32_t> idx(n);
  for (size_t i = 0; i < n; ++i) idx[i] = (int32_t)i;
  if (morton) {
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int32_t a, int32_t b) { return key[(size_t)a] < key[(size_t)b]; });
  } else {
    // block index, then (z, y, x) inside the block
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
      if (key[(size_t)a] != key[(size_t)b]) return key[(size_t)a] < key[(size_t)b];
      const int4 u = c->h_coords[(size_t)a], w = c->h_coords[(size_t)b];
      if (u.x != w.x) return u.x < w.x;
      if (u.y != w.y) return u.y < w.y;
      return u.z < w.z;
    });
  }
  std::vector<int4> sorted(n);
  for (size_t i = 0; i < n; ++i) sorted[i] = c->h_coords[(size_t)idx[i]];
  if (!c->d_coords) HIP_TRY(hipMalloc(&c->d_coords, n * sizeof(int4)));
  HIP_TRY(hipMemcpy(c->d_coords, sorted.data(), n * sizeof(int4), hipMemcpyHostToDevice));
  c->block[0] = morton ? 0 : bz;
  c->block[1] = morton ? 0 : by;
  c->block[2] = morton ? 0 : bx;
  return ANACOR_OK;
}

static int ctx_create(const uint8_t* vol, const int64_t shape[3], const int64_t* coords,
                      const int32_t* coords32, int64_t n, const double vs[3], const double mu[4],
                      int full, int device, Context** out) {
  *out = nullptr;
  if (!vol || !shape) {
    set_error(ANACOR_ERR_ARG, "null volume or shape");
    return ANACOR_ERR_ARG;
  }
  for (int i = 0; i < 3; ++i)
    if (shape[i] <= 0 || shape[i] > (1 << 20)) {
      set_error(ANACOR_ERR_ARG, "bad volume shape %lld", (long long)shape[i]);
      return ANACOR_ERR_ARG;
    }
  const uint64_t nvox = (uint64_t)shape[0] * shape[1] * shape[2];
  const uint64_t npad = (uint64_t)(shape[0] + 2) * (shape[1] + 2) * (shape[2] + 2);
  if (npad >= (1ull << 31)) {
    set_error(ANACOR_ERR_ARG, "volume of %llu voxels exceeds 2^31 with its one-voxel shell",
              (unsigned long long)nvox);
    return ANACOR_ERR_ARG;
  }
  if (n < 0 || n >= (1ll << 31) || (n > 0 && !coords && !coords32)) {
    set_error(ANACOR_ERR_ARG, "bad voxel list (n=%lld)", (long long)n);
    return ANACOR_ERR_ARG;
  }
  // crystal voxels must lie inside the volume (generate_sampling only yields
  // such voxels; the kernels rely on it for their face-axis stepping)
  for (int64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      const int64_t v = coords ? coords[3 * i + a] : coords32[3 * i + a];
      if (v < 0 || v >= shape[a]) {
        set_error(ANACOR_ERR_ARG, "voxel %lld: coordinate %lld outside the volume", (long long)i,
                  (long long)v);
        return ANACOR_ERR_ARG;
      }
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error(ANACOR_ERR_NO_DEVICE, "no HIP device visible");
    return ANACOR_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= ndev) device = 0;
  Context* c = new Context();
  c->device = device;
  // everything that can fail from here on: on an error the half-built
  // context (and any temporary) is released before returning
  auto init = [&]() -> int {
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&c->ev0));
    HIP_TRY(hipEventCreate(&c->ev1));
    HIP_TRY(hipEventCreateWithFlags(&c->copied, hipEventDisableTiming));
    c->Z = (int)shape[0]; c->Y = (int)shape[1]; c->X = (int)shape[2];
    for (int i = 0; i < 3; ++i) c->vs[i] = vs ? vs[i] : 1.0;
    for (int i = 0; i < 4; ++i) c->mu[i] = mu ? mu[i] : 0.0;
    c->full = full;
    c->N = (int)n;
    {  // upload, scan which label ranges the volume holds (layout choice, walk
       // variant), and lay it out with its zero shell
      DeviceTemp in, flag;   // freed on every way out of this block
      unsigned h_flag = 0;
      HIP_TRY(hipMalloc(&in.p, nvox));
      HIP_TRY(hipMemcpy(in.p, vol, nvox, hipMemcpyHostToDevice));
      HIP_TRY(hipMalloc(&flag.p, sizeof(unsigned)));
      HIP_TRY(hipMemsetAsync(flag.p, 0, sizeof(unsigned), c->stream));
      hipLaunchKernelGGL(label_scan_kernel, dim3(2048), dim3(256), 0, c->stream,
                         (const uint8_t*)in.p, (long long)nvox, (unsigned*)flag.p);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMalloc(&c->d_vol, npad));
      c->vol_bytes = (long long)npad;
      hipLaunchKernelGGL(shell_volume_kernel, dim3(4096), dim3(256), 0, c->stream,
                         (const uint8_t*)in.p, c->Z, c->Y, c->X, (long long)npad, c->d_vol);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(&h_flag, flag.p, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      c->labels_over3 = (h_flag & 1u) != 0;
      c->labels_over4 = (h_flag & 2u) != 0;
    }
    {  // increments as doubles: a walk reaches at most extent + 1 plus one
       // unrolled group; 32 spare entries
      const int m = std::max(c->Z, std::max(c->Y, c->X)) + 32;
      std::vector<double> h((size_t)m);
      for (int i = 0; i < m; ++i) h[(size_t)i] = (double)i;
      HIP_TRY(hipMalloc(&c->d_incd, (size_t)m * sizeof(double)));
      HIP_TRY(hipMemcpy(c->d_incd, h.data(), (size_t)m * sizeof(double), hipMemcpyHostToDevice));
      c->incd_last = m - 1;
    }
    c->h_coords.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      const int64_t z = coords ? coords[3 * i] : coords32[3 * i];
      const int64_t y = coords ? coords[3 * i + 1] : coords32[3 * i + 1];
      const int64_t x = coords ? coords[3 * i + 2] : coords32[3 * i + 2];
      c->h_coords[(size_t)i] = make_int4((int)z, (int)y, (int)x, (int)i);
    }
    int blk[3] = {kDefaultBlock[0], kDefaultBlock[1], kDefaultBlock[2]};
    // volumes that take the bricked layout by default: Morton-ordered waves,
    // whose rays share bricks (measured on 1000x1000x800 with 100k sampled
    // voxels: 164 ms against 306 ms for row-shaped waves)
    if ((int64_t)nvox > (64ll << 20)) blk[0] = blk[1] = blk[2] = 0;
    if (const char* e = std::getenv("ANACOR_VOXEL_BLOCK"))
      if (std::sscanf(e, "%d,%d,%d", &blk[0], &blk[1], &blk[2]) != 3) blk[0] = blk[1] = blk[2] = 0;
    return upload_voxel_order(c, blk[0], blk[1], blk[2]);
  };
  const int rc = init();
  if (rc != ANACOR_OK) {
    ctx_destroy(c);
    return rc;
  }
  *out = c;
  return ANACOR_OK;
}

static void ctx_destroy(Context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->d_vol) (void)hipFree(c->d_vol);
  if (c->d_incd) (void)hipFree(c->d_incd);
  if (c->d_pack) (void)hipFree(c->d_pack);
  if (c->d_coords) (void)hipFree(c->d_coords);
  if (c->d_rules) (void)hipFree(c->d_rules);
  if (c->d_partial) (void)hipFree(c->d_partial);
  if (c->d_err) (void)hipFree(c->d_err);
  if (c->d_out) (void)hipFree(c->d_out);
  if (c->d_order) (void)hipFree(c->d_order);
  if (c->d_steps) (void)hipFree(c->d_steps);
  if (c->d_ib) (void)hipFree(c->d_ib);
  if (c->d_oob) (void)hipFree(c->d_oob);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->copied) (void)hipEventDestroy(c->copied);
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

static int host_threads(int requested) {
  if (requested > 0) return std::min(requested, 64);
  const char* e = std::getenv("ANACOR_HOST_THREADS");
  if (e && std::atoi(e) > 0) return std::min(std::atoi(e), 64);
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hw, 32u));
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static int max_dim(const Context* c) { return std::max(c->Z, std::max(c->Y, c->X)); }

constexpr int64_t kSplitMinReflections = 8192;

static int ctx_run(Context* c, const double* s1, const double* omega, int64_t R,
                   const double* raw_xray, const double* omega_axis, const double* F, int flags,
                   int n_threads, double* out, PathRecord* paths) {
  if (!c) {
    set_error(ANACOR_ERR_ARG, "null context");
    return ANACOR_ERR_ARG;
  }
  if (R < 0 || (R > 0 && (!s1 || !omega || !raw_xray || !omega_axis || !F))) {
    set_error(ANACOR_ERR_ARG, "bad reflection arguments");
    return ANACOR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(c->mu_run);
  HIP_TRY(hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(c->h_rules.reserve((size_t)std::max<int64_t>(R, 1)));
  const int threads = host_threads(n_threads);
  // large calls without path records in two parts: the host sets up and
  // sorts the last three quarters while the GPU traces the first quarter
  // (which holds thousands of workgroups: the call's tile is the whole call's)
  const int64_t first = (!paths && R >= kSplitMinReflections) ? R / 4 : R;
  host_setup(s1, omega, first, raw_xray, omega_axis, F, max_dim(c), c->h_rules.data(), nullptr,
             threads);
  c->last_host_ms = ms_since(t0);
  if (first == R) return run_rules(c, c->h_rules.data(), R, flags, out, paths);
  auto setup_rest = [&]() -> int {
    const auto t1 = std::chrono::steady_clock::now();
    host_setup(s1 + 3 * first, omega + first, R - first, raw_xray, omega_axis, F, max_dim(c),
               c->h_rules.data() + first, nullptr, threads);
    c->last_host_ms += ms_since(t1);
    return ANACOR_OK;
  };
  return run_rules(c, c->h_rules.data(), R, flags, out, paths, first, setup_rest);
}

// ------------------------------------------------- allocation-free trace
//
// anacor_ctx_run_noalloc traces on a resident context with every per-call
// buffer in one file-scope device array, for callers that may not allocate
// (graph capture, serving loops).  Each process holds one copy of the array
// per GPU for its lifetime; its address on a device comes from
// hipGetSymbolAddress on that device's first call (host code must never pass
// the array's own name to a kernel: that is a host-side stand-in).  The host
// staging of the rules and the processing order lives in static host arrays.
// One process-wide lock serialises the calls; two processes each have their
// own arrays.  The error words and the step counter are cleared by a kernel
// at the start of every call; everything else is written before it is read.
constexpr int64_t kNoAllocMaxR = ANACOR_NOALLOC_MAX_REFLECTIONS;
constexpr int64_t kNoAllocMaxPartials = ANACOR_NOALLOC_MAX_PARTIALS;
constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }
constexpr size_t kScratchRules = 0;
constexpr size_t kScratchPartial = kScratchRules + align256(kNoAllocMaxR * sizeof(ReflRule));
constexpr size_t kScratchOrder = kScratchPartial + align256(kNoAllocMaxPartials * 8);
constexpr size_t kScratchErr = kScratchOrder + align256(kNoAllocMaxR * sizeof(int));
constexpr size_t kScratchOut = kScratchErr + align256(kNoAllocMaxR * sizeof(unsigned));
constexpr size_t kScratchSteps = kScratchOut + align256(kNoAllocMaxR * sizeof(double));
constexpr size_t kScratchBytes = kScratchSteps + 256;
static_assert(kScratchBytes == ANACOR_NOALLOC_SCRATCH_BYTES, "scratch size in the header");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static unsigned char* g_scratch_dev[kMaxDevices];   // device address of `scratch` per device
static std::mutex g_noalloc_mu;
static ReflRule g_noalloc_rules[kNoAllocMaxR];
static int g_noalloc_order[kNoAllocMaxR];
static uint64_t g_noalloc_key[kNoAllocMaxR];
static uint64_t g_noalloc_key2[kNoAllocMaxR];   // sort work space
static unsigned g_noalloc_err[kNoAllocMaxR];

// Device address of `scratch` on the current device (set by the caller),
// looked up once per device.  Called with g_noalloc_mu held.
static int scratch_address(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices) {
    set_error(ANACOR_ERR_ARG, "device %d beyond the %d the scratch table covers", device,
              kMaxDevices);
    return ANACOR_ERR_ARG;
  }
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error(ANACOR_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
      return ANACOR_ERR_HIP;
    }
    g_scratch_dev[device] = (unsigned char*)p;
  }
  *out = g_scratch_dev[device];
  return ANACOR_OK;
}

static int ctx_run_noalloc(Context* c, const double* s1, const double* omega, int64_t R,
                           const double* raw_xray, const double* omega_axis, const double* F,
                           int flags, double* out) {
  if (!c) {
    set_error(ANACOR_ERR_ARG, "null context");
    return ANACOR_ERR_ARG;
  }
  if (R < 0 || (R > 0 && (!s1 || !omega || !raw_xray || !omega_axis || !F || !out))) {
    set_error(ANACOR_ERR_ARG, "bad reflection arguments");
    return ANACOR_ERR_ARG;
  }
  if (R > kNoAllocMaxR) {
    set_error(ANACOR_ERR_ARG, "%lld reflections exceed the %lld of one allocation-free call "
              "(ANACOR_NOALLOC_MAX_REFLECTIONS)", (long long)R, (long long)kNoAllocMaxR);
    return ANACOR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk_all(g_noalloc_mu);
  std::lock_guard<std::mutex> lk(c->mu_run);
  HIP_TRY(hipSetDevice(c->device));
  unsigned char* base = nullptr;
  int rc = scratch_address(c->device, &base);
  if (rc) return rc;
  c->last_oob = 0;
  if (R == 0) return ANACOR_OK;
  if (c->N == 0) {
    for (int64_t i = 0; i < R; ++i) out[i] = __builtin_nan("");
    return ANACOR_OK;
  }
  // the tile of anacor_ctx_run, enlarged until R * tiles fits the partials
  rc = ensure_layout(c, /*may_build=*/false);
  if (rc) return rc;
  int64_t vpt = choose_tile(c, R, c->layout);
  const int64_t max_tiles = kNoAllocMaxPartials / R;
  if ((c->N + vpt - 1) / vpt > max_tiles) vpt = (c->N + max_tiles - 1) / max_tiles;
  const auto t0 = std::chrono::steady_clock::now();
  // one host thread: the setup spawns none
  host_setup(s1, omega, R, raw_xray, omega_axis, F, max_dim(c), g_noalloc_rules, nullptr, 1);
  direction_order(R, [](int64_t i, int k) { return g_noalloc_rules[i].ray[k].dir; },
                  g_noalloc_order, g_noalloc_key, g_noalloc_key2, sort_major_for(c, c->layout));
  c->last_host_ms = ms_since(t0);
  RunBuffers B;
  B.rules = (ReflRule*)(base + kScratchRules);
  B.partial = base + kScratchPartial;
  B.order = (int*)(base + kScratchOrder);
  B.err = (unsigned*)(base + kScratchErr);
  B.out = (double*)(base + kScratchOut);
  B.steps = (unsigned long long*)(base + kScratchSteps);
  return trace_call(c, B, g_noalloc_rules, g_noalloc_order, R, (int)vpt, c->layout, flags, out,
                    nullptr, nullptr, g_noalloc_err);
}

// Build a contiguous copy of an int8*** row-pointer table (python_2_c_3d,
// utils/utils_os.py:45-68) unless its rows already are one block.
static const uint8_t* flatten_table(int8_t*** table, const int64_t shape[3],
                                    std::vector<uint8_t>& storage) {
  const int64_t Z = shape[0], Y = shape[1], X = shape[2];
  const int8_t* base = table[0][0];
  bool contiguous = true;
  for (int64_t z = 0; z < Z && contiguous; ++z)
    for (int64_t y = 0; y < Y; ++y)
      if (table[z][y] != base + (z * Y + y) * X) {
        contiguous = false;
        break;
      }
  if (contiguous) return reinterpret_cast<const uint8_t*>(base);
  storage.resize((size_t)(Z * Y * X));
  for (int64_t z = 0; z < Z; ++z)
    for (int64_t y = 0; y < Y; ++y)
      std::memcpy(&storage[(size_t)((z * Y + y) * X)], table[z][y], (size_t)X);
  return storage.data();
}


// ------------------------------------------------------------ batch (host)

struct Batch;

// What a sample's host node needs: the caller's arrays of the current run
// (rewritten before every launch of the graph) and what the node reports.
struct BatchJob {
  Batch* b = nullptr;
  int i = 0;                    // sample index
  const double* s1 = nullptr;
  const double* omega = nullptr;
  int64_t R = 0;
  bool guard = false;           // some ray needs the guarded kernel
  double host_ms = 0.0;
};

struct Batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<Context*> samples;
  std::vector<anacor_ctx> handles;
  int full = 0;
  BatchSample* d_desc = nullptr;
  size_t desc_cap = 0;
  unsigned long long* d_counters = nullptr;   // the step counter
  hipStream_t copy_stream = nullptr;          // host setup and rule uploads, ahead of the kernels
  std::vector<hipEvent_t> copied;             // per sample: its upload is done
  hipEvent_t fork = nullptr;                  // copy_stream joins the captured sequence
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.f;
  unsigned long long last_steps = 0;
  double last_host_ms = 0.0;
  long long blocks = 0;                       // workgroups of the last run's launches
  std::vector<BatchSample> h_desc;
  Pinned<BatchSample> h_desc_pin;             // its upload staging (a node of the graph)
  std::vector<int> seq;                       // launch order of the samples
  // the run as one graph: captured for one set of reflection counts, flags
  // and work shapes, relaunched while they stay the same
  std::vector<BatchJob> jobs;
  double geo[15] = {};                        // raw_xray, omega_axis, F of the current run
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<int64_t> graph_key;
  bool use_graph = false;                     // run as the batch graph (anacor_batch_set_graph)
  bool last_graph = false;                    // the last run launched the graph
  std::mutex mu;
};

static void batch_drop_graph(Batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key.clear();
}

static void batch_destroy(Batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->copy_stream) (void)hipStreamSynchronize(b->copy_stream);
  batch_drop_graph(b);
  for (hipEvent_t e : b->copied) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->copy_stream) (void)hipStreamDestroy(b->copy_stream);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_counters) (void)hipFree(b->d_counters);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  for (Context* c : b->samples) ctx_destroy(c);
  delete b;
}

static int batch_create(const anacor_sample_t* in, int n, int full, int device, Batch** out) {
  *out = nullptr;
  if (!in || n <= 0 || n > 4096) {
    set_error(ANACOR_ERR_ARG, "bad sample list (n=%d)", n);
    return ANACOR_ERR_ARG;
  }
  Batch* b = new Batch();
  b->full = full;
  for (int i = 0; i < n; ++i) {
    Context* c = nullptr;
    const int rc = ctx_create(in[i].volume, in[i].shape, in[i].coords, nullptr, in[i].n_coords,
                              in[i].voxel_size, in[i].coefficients, full, device, &c);
    if (rc != ANACOR_OK) {
      batch_destroy(b);
      return rc;
    }
    b->samples.push_back(c);
  }
  b->device = b->samples[0]->device;
  b->handles.resize((size_t)n);
  for (int i = 0; i < n; ++i) b->handles[(size_t)i].c = b->samples[(size_t)i];
#define BATCH_TRY(expr)                                                            \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error(ANACOR_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
      batch_destroy(b);                                                            \
      return ANACOR_ERR_HIP;                                                       \
    }                                                                              \
  } while (0)
  BATCH_TRY(hipSetDevice(b->device));
  BATCH_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  BATCH_TRY(hipStreamCreateWithFlags(&b->copy_stream, hipStreamNonBlocking));
  b->copied.assign((size_t)n, nullptr);
  for (int i = 0; i < n; ++i)
    BATCH_TRY(hipEventCreateWithFlags(&b->copied[(size_t)i], hipEventDisableTiming));
  BATCH_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  BATCH_TRY(hipEventCreate(&b->ev0));
  BATCH_TRY(hipEventCreate(&b->ev1));
  BATCH_TRY(hipMalloc(&b->d_counters, sizeof(unsigned long long)));
#undef BATCH_TRY
  *out = b;
  return ANACOR_OK;
}

// Host setup of one sample: rotation, angles, stepping rules and direction
// order into the sample's page-locked staging.  It makes no HIP call, so it
// can run as a host node of the batch graph as well as on the calling thread.
static void batch_setup_sample(void* p) {
  BatchJob* j = static_cast<BatchJob*>(p);
  Batch* b = j->b;
  Context* c = b->samples[(size_t)j->i];
  const auto t0 = std::chrono::steady_clock::now();
  host_setup(j->s1, j->omega, j->R, b->geo, b->geo + 3, b->geo + 6, max_dim(c),
             c->h_rules.data(), nullptr, host_threads(0));
  order_reflections(c, c->h_rules.data(), j->R, kLayoutRaw);
  bool guard = false;
  for (int64_t r = 0; r < j->R; ++r) {
    c->h_order_pin[(size_t)r] = c->h_order[(size_t)r];
    guard = guard || c->h_rules[(size_t)r].ray[0].guard || c->h_rules[(size_t)r].ray[1].guard;
  }
  j->guard = guard;
  j->host_ms = ms_since(t0);
}

// The batch's work on b->stream, the host setup and rule uploads of every
// sample on copy_stream ahead of it: clear, then per sample (in b->seq
// order) its setup, upload and trace launch behind that upload, then one
// finish kernel over every sample.  capture: the setups become host nodes
// (the stream is being captured into the batch graph); otherwise they run
// on this thread while the GPU traces the samples before them.
static int batch_enqueue(Batch* b, bool capture, bool f64, int is_exp, long long total_refl) {
  const int S = (int)b->samples.size();
  HIP_TRY(hipMemcpyAsync(b->d_desc, b->h_desc_pin.data(), (size_t)S * sizeof(BatchSample),
                         hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(batch_clear_kernel, dim3(64), dim3(256), 0, b->stream, b->d_desc, S,
                     b->d_counters);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(b->fork, b->stream));
  HIP_TRY(hipStreamWaitEvent(b->copy_stream, b->fork, 0));
  for (int j = 0; j < S; ++j) {
    const int i = b->seq[(size_t)j];
    Context* c = b->samples[(size_t)i];
    BatchJob& job = b->jobs[(size_t)i];
    if (capture) HIP_TRY(hipLaunchHostFunc(b->copy_stream, batch_setup_sample, &job));
    else batch_setup_sample(&job);
    if (job.R == 0) continue;
    HIP_TRY(hipMemcpyAsync(c->d_rules, c->h_rules.data(), (size_t)job.R * sizeof(ReflRule),
                           hipMemcpyHostToDevice, b->copy_stream));
    HIP_TRY(hipMemcpyAsync(c->d_order, c->h_order_pin.data(), (size_t)job.R * sizeof(int),
                           hipMemcpyHostToDevice, b->copy_stream));
    HIP_TRY(hipEventRecord(b->copied[(size_t)i], b->copy_stream));
    HIP_TRY(hipStreamWaitEvent(b->stream, b->copied[(size_t)i], 0));
    if (c->N == 0) continue;
    const BatchSample& d = b->h_desc[(size_t)i];
    TraceArgs t;
    t.vol = c->d_vol;
    t.vol_bytes = (int)c->vol_bytes;
    t.cz = t.cy = 0;
    t.Z = c->Z; t.Y = c->Y; t.X = c->X;
    t.coords = c->d_coords;
    t.N = c->N;
    t.rules = c->d_rules;
    t.R = d.R;
    t.tiles = d.tiles;
    t.vox_per_tile = d.vpt;
    t.n_blocks = d.R * d.tiles;
    for (int k = 0; k < 3; ++k) t.vs[k] = c->vs[k];
    for (int k = 0; k < 4; ++k) t.mu[k] = c->mu[k];
    t.is_exp = is_exp;
    t.partial32 = d.p32;
    t.partial64 = d.p64;
    t.err = d.err;
    t.steps = b->d_counters;
    t.paths = nullptr;
    t.order = d.order;
    t.incd = c->d_incd;
    t.incd_last = c->incd_last;
    t.group = launch_group(c, kLayoutRaw);
    t.pos0 = 0;
    // a graph is captured before any setup has run: it holds the unguarded
    // kernel, and a run whose rays need the guard is traced again without it
    const bool guard = !capture && job.guard;
    dispatch_trace(t, (t.n_blocks + 7) / 8 * 8, guard, b->full != 0,
                   raw_needs_wide(c) ? kLayoutWide : kLayoutRaw, f64, false, !c->labels_over4,
                   b->stream);
    HIP_TRY(hipGetLastError());
  }
  // rejoin copy_stream (a captured fork must end joined)
  HIP_TRY(hipEventRecord(b->fork, b->copy_stream));
  HIP_TRY(hipStreamWaitEvent(b->stream, b->fork, 0));
  if (total_refl > 0) {
    hipLaunchKernelGGL(batch_finish_kernel, dim3((unsigned)((total_refl + 255) / 256)),
                       dim3(256), 0, b->stream, b->d_desc, S, f64 ? 1 : 0, total_refl);
    HIP_TRY(hipGetLastError());
  }
  return ANACOR_OK;
}

// Capture the batch's sequence into a graph and instantiate it.
static int batch_capture(Batch* b, bool f64, int is_exp, long long total_refl) {
  batch_drop_graph(b);
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = batch_enqueue(b, true, f64, is_exp, total_refl);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc != ANACOR_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    set_error(ANACOR_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    return ANACOR_ERR_HIP;
  }
  b->graph = g;
  HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return ANACOR_OK;
}

static int batch_run(Batch* b, const double* const* s1, const double* const* omega,
                     const int64_t* R, const double* raw_xray, const double* omega_axis,
                     const double* F, int flags, double* const* out) {
  std::lock_guard<std::mutex> lk(b->mu);
  HIP_TRY(hipSetDevice(b->device));
  const bool f64 = (flags & ANACOR_FLAG_F64_MEAN) != 0;
  const int is_exp = (flags & ANACOR_FLAG_NO_EXP) ? 0 : 1;
  const int S = (int)b->samples.size();
  long long items = 0, total_refl = 0;
  b->h_desc.resize((size_t)S);
  b->jobs.resize((size_t)S);
  // Staging buffers, tiles and device buffers of every sample first: the
  // allocations stay on this thread and ahead of every launch (and outside
  // any capture).
  for (int i = 0; i < S; ++i) {
    Context* c = b->samples[(size_t)i];
    const int64_t Ri = R[i];
    if (Ri < 0 || (Ri > 0 && (!s1[i] || !omega[i] || !out[i]))) {
      set_error(ANACOR_ERR_ARG, "sample %d: bad reflection arguments", i);
      return ANACOR_ERR_ARG;
    }
    HIP_TRY(c->h_rules.reserve((size_t)std::max<int64_t>(Ri, 1)));
    HIP_TRY(c->h_order_pin.reserve((size_t)std::max<int64_t>(Ri, 1)));
    c->h_order.reserve((size_t)std::max<int64_t>(Ri, 1));
    int vpt = 1;
    if (c->N > 0) {
      vpt = fit_tile(c->N, Ri, choose_tile(c, std::max<int64_t>(Ri, 1)));
      if (vpt == 0) {
        set_error(ANACOR_ERR_ARG, "sample %d: %lld reflections exceed one dispatch", i,
                  (long long)Ri);
        return ANACOR_ERR_ARG;
      }
    }
    const int tiles = c->N > 0 ? (c->N + vpt - 1) / vpt : 0;
    const int rc = ensure_run_buffers(c, std::max<int64_t>(Ri, 1),
                                      std::max<int64_t>(Ri * tiles, 1), f64);
    if (rc) return rc;
    BatchSample& d = b->h_desc[(size_t)i];
    d.order = c->d_order;
    d.N = c->N;
    d.R = (int)Ri;
    d.tiles = std::max(tiles, 1);
    d.vpt = vpt;
    d.p32 = (float2*)c->d_partial;
    d.p64 = (double*)c->d_partial;
    d.err = c->d_err;
    d.out = c->d_out;
    BatchJob& job = b->jobs[(size_t)i];
    job.b = b;
    job.i = i;
    job.s1 = s1[i];
    job.omega = omega[i];
    job.R = Ri;
    job.guard = false;
    job.host_ms = 0.0;
    items += (c->N > 0) ? Ri * tiles : 0;
    total_refl += Ri;
  }
  if ((size_t)S > b->desc_cap) {
    if (b->d_desc) HIP_TRY(hipFree(b->d_desc));
    b->d_desc = nullptr;
    HIP_TRY(hipMalloc(&b->d_desc, (size_t)S * sizeof(BatchSample)));
    b->desc_cap = (size_t)S;
  }
  HIP_TRY(b->h_desc_pin.reserve((size_t)S));
  for (int i = 0; i < S; ++i) b->h_desc_pin[(size_t)i] = b->h_desc[(size_t)i];
  std::memcpy(b->geo, raw_xray, 3 * sizeof(double));
  std::memcpy(b->geo + 3, omega_axis, 3 * sizeof(double));
  std::memcpy(b->geo + 6, F, 9 * sizeof(double));
  b->blocks = items;
  // Sample order: a two-stage flow shop (host setup, then GPU trace) ordered
  // by Johnson's rule on estimated stage times (host ~0.085 us per
  // reflection on eight threads; GPU ~1.3 ps per voxel step of either ray, a
  // ray taking about a sixth of the extents' sum): samples whose trace
  // outlasts their setup first, by increasing setup, then the others by
  // decreasing trace.  Only the first setup is then exposed.  Results do not
  // depend on the order.
  b->seq.resize((size_t)S);
  {
    std::vector<double> setup((size_t)S), trace((size_t)S);
    for (int i = 0; i < S; ++i) {
      const Context* c = b->samples[(size_t)i];
      setup[(size_t)i] = 0.085e-6 * (double)R[i];
      trace[(size_t)i] = 1.3e-12 * (double)R[i] * (double)c->N * (c->Z + c->Y + c->X) / 6.0;
    }
    for (int i = 0; i < S; ++i) b->seq[(size_t)i] = i;
    auto first = [&](int i) { return setup[(size_t)i] < trace[(size_t)i]; };
    std::stable_sort(b->seq.begin(), b->seq.end(), [&](int x, int y) {
      if (first(x) != first(y)) return first(x);
      return first(x) ? setup[(size_t)x] < setup[(size_t)y] : trace[(size_t)x] > trace[(size_t)y];
    });
  }
  // The graph holds every buffer address and launch shape: it is captured
  // again whenever one of them changes.
  std::vector<int64_t> key = {S, f64 ? 1 : 0, is_exp, b->full};
  for (int i = 0; i < S; ++i) {
    const Context* c = b->samples[(size_t)i];
    const BatchSample& d = b->h_desc[(size_t)i];
    for (int64_t v : {(int64_t)R[i], (int64_t)d.vpt, (int64_t)launch_group(c, kLayoutRaw),
                      (int64_t)(intptr_t)c->d_rules, (int64_t)(intptr_t)c->d_order,
                      (int64_t)(intptr_t)c->d_partial, (int64_t)(intptr_t)c->d_err,
                      (int64_t)(intptr_t)c->d_out, (int64_t)(intptr_t)c->d_coords,
                      (int64_t)(intptr_t)c->h_rules.data(),
                      (int64_t)(intptr_t)c->h_order_pin.data()})
      key.push_back(v);
  }
  key.push_back((int64_t)(intptr_t)b->h_desc_pin.data());
  const bool use_graph = b->use_graph;
  if (use_graph && (!b->exec || key != b->graph_key)) {
    const int rc = batch_capture(b, f64, is_exp, total_refl);
    if (rc != ANACOR_OK) return rc;
    b->graph_key = key;
  }
  auto launch = [&](bool graph) -> int {
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    if (graph) {
      HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {
      const int rc = batch_enqueue(b, false, f64, is_exp, total_refl);
      if (rc != ANACOR_OK) return rc;
    }
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    for (int i = 0; i < S; ++i) {
      Context* c = b->samples[(size_t)i];
      if (R[i] > 0)
        HIP_TRY(hipMemcpyAsync(out[i], c->d_out, (size_t)R[i] * sizeof(double),
                               hipMemcpyDeviceToHost, b->stream));
    }
    HIP_TRY(hipMemcpyAsync(&b->last_steps, b->d_counters, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
    return ANACOR_OK;
  };
  int rc = launch(use_graph);
  if (rc != ANACOR_OK) return rc;
  bool guard = false;
  for (const BatchJob& j : b->jobs) guard = guard || j.guard;
  b->last_graph = use_graph;
  if (use_graph && guard) {   // rays beyond the int64 range: the guarded kernels, live
    rc = launch(false);
    if (rc != ANACOR_OK) return rc;
    b->last_graph = false;
  }
  double host_ms = 0.0;
  for (const BatchJob& j : b->jobs) host_ms += j.host_ms;
  b->last_host_ms = host_ms;
  for (int i = 0; i < S; ++i) {
    Context* c = b->samples[(size_t)i];
    if (c->N == 0)
      for (int64_t r = 0; r < R[i]; ++r) out[i][r] = __builtin_nan("");
  }
  // report the first reflection the reference would have aborted on
  for (int i = 0; i < S; ++i)
    for (int64_t r = 0; r < R[i]; ++r)
      if (out[i][r] != out[i][r] && b->samples[(size_t)i]->N > 0) {
        set_error(ANACOR_ERR_GEOMETRY, "sample %d reflection %lld: the reference aborts here; "
                  "result set to NaN", i, (long long)r);
        return ANACOR_ERR_GEOMETRY;
      }
  return ANACOR_OK;
}

}  // namespace anacor

namespace anacor {

// ------------------------------------------------------- gridding (host)

template <bool GUARD, bool FULL, int LAYOUT>
static void launch_grid_map(const GridArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((grid_map_kernel<kCap, unroll_for(LAYOUT), GUARD, FULL, LAYOUT>),
                     dim3(grid), dim3(kBlock), 0, s, a);
}

using LaunchGridFn = void (*)(const GridArgs&, int, hipStream_t);

template <int I>
static constexpr LaunchGridFn grid_launcher() {
  return &launch_grid_map<(I & 1) != 0, (I & 2) != 0, (I >> 2) & 3>;
}

template <int... I>
static constexpr std::array<LaunchGridFn, sizeof...(I)> grid_table(
    std::integer_sequence<int, I...>) {
  return {grid_launcher<I>()...};
}

static const std::array<LaunchGridFn, 16> kLaunchGrid =
    grid_table(std::make_integer_sequence<int, 16>{});

struct Grid {
  Context* c = nullptr;        // borrowed
  int n_theta = 0, n_phi = 0, D = 0;
  float* d_map = nullptr;      // [D][N]
  double* d_angles = nullptr;  // lookup buffers
  double* d_out = nullptr;
  unsigned* d_err = nullptr;
  int* d_order = nullptr;
  size_t cap = 0;
  double* d_lists = nullptr;
  size_t lists_cap = 0;
  float build_ms = 0.f, last_ms = 0.f;
  std::vector<ReflRule> h_rules;
  std::vector<ReflGeom> h_geom;
  std::vector<double> h_angles;
};

static void grid_free(Grid* g) {
  void* bufs[] = {g->d_map, g->d_angles, g->d_out, g->d_err, g->d_order, g->d_lists};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete g;
}

// Absorption maps of every crystal voxel of c for D directions.
static int grid_build(Context* c, int64_t n_theta, int64_t n_phi, const double* rotated_s1,
                      const double* theta_phi, int is_exp, Grid** out,
                      double* d_map64 = nullptr) {
  if (!c || n_theta < 1 || n_phi < 1 || !rotated_s1 || !theta_phi ||
      n_theta * n_phi > (1 << 24)) {
    set_error(ANACOR_ERR_ARG, "bad gridding arguments");
    return ANACOR_ERR_ARG;
  }
  Grid* g = new Grid();
  g->c = c;
  g->n_theta = (int)n_theta;
  g->n_phi = (int)n_phi;
  g->D = (int)(n_theta * n_phi);
  const int D = g->D, N = c->N;
  std::vector<RayRule> rays((size_t)D);
  host_setup_directions(rotated_s1, theta_phi, D, max_dim(c), rays.data());
  bool guard = false;
  for (const RayRule& r : rays) guard = guard || r.guard;
  int rc = ensure_layout(c);
  if (rc) { grid_free(g); return rc; }
  RayRule* d_rays = nullptr;
  unsigned* d_err = nullptr;
  auto fail = [&](hipError_t e, const char* what) {
    set_error(ANACOR_ERR_HIP, "gridding: %s failed: %s", what, hipGetErrorString(e));
    if (d_rays) (void)hipFree(d_rays);
    if (d_err) (void)hipFree(d_err);
    grid_free(g);
    return ANACOR_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipMalloc(&g->d_map, (size_t)D * std::max(N, 1) * sizeof(float))) != hipSuccess)
    return fail(e, "hipMalloc(map)");
  if ((e = hipMalloc(&d_rays, (size_t)D * sizeof(RayRule))) != hipSuccess)
    return fail(e, "hipMalloc");
  if ((e = hipMalloc(&d_err, (size_t)D * sizeof(unsigned))) != hipSuccess)
    return fail(e, "hipMalloc");
  if ((e = hipMemcpyAsync(d_rays, rays.data(), (size_t)D * sizeof(RayRule),
                          hipMemcpyHostToDevice, c->stream)) != hipSuccess)
    return fail(e, "upload");
  if ((e = hipMemsetAsync(d_err, 0, (size_t)D * sizeof(unsigned), c->stream)) != hipSuccess)
    return fail(e, "hipMemsetAsync");
  std::vector<unsigned> h_err((size_t)D, 0u);
  if (N > 0) {
    const int vpt = fit_tile(N, D, choose_tile(c, D));
    if (vpt == 0) {
      (void)hipFree(d_rays);
      (void)hipFree(d_err);
      grid_free(g);
      set_error(ANACOR_ERR_ARG, "gridding: %lld directions exceed one dispatch", (long long)D);
      return ANACOR_ERR_ARG;
    }
    const int64_t tiles = (N + vpt - 1) / vpt;
    const int64_t nblk = (int64_t)D * tiles;
    GridArgs a;
    const bool bricked = c->layout == kLayoutBrick2 || c->layout == kLayoutBrick4;
    a.vol = bricked ? c->d_pack : c->d_vol;
    a.vol_bytes = bricked ? (int)c->pack_bytes : (int)c->vol_bytes;
    a.cz = 512 * c->NBX * c->NBY - 512;
    a.cy = 512 * c->NBX - 64;
    a.Z = c->Z; a.Y = c->Y; a.X = c->X;
    a.coords = c->d_coords;
    a.N = N;
    a.rays = d_rays;
    a.D = D;
    a.tiles = (int)tiles;
    a.vox_per_tile = vpt;
    a.n_blocks = (int)nblk;
    for (int i = 0; i < 3; ++i) a.vs[i] = c->vs[i];
    for (int i = 0; i < 4; ++i) a.mu[i] = c->mu[i];
    a.is_exp = is_exp;
    a.map = g->d_map;
    a.map64 = d_map64;
    a.err = d_err;
    a.incd = c->d_incd;
    a.incd_last = c->incd_last;
    if ((e = hipEventRecord(c->ev0, c->stream)) != hipSuccess) return fail(e, "event");
    const int i = (guard ? 1 : 0) | (c->full ? 2 : 0) | (c->layout << 2);
    kLaunchGrid[i](a, (int)((nblk + 7) / 8 * 8), c->stream);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "grid_map_kernel");
    if ((e = hipEventRecord(c->ev1, c->stream)) != hipSuccess) return fail(e, "event");
  }
  if ((e = hipMemcpyAsync(h_err.data(), d_err, (size_t)D * sizeof(unsigned),
                          hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
    return fail(e, "download");
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return fail(e, "gridding");
  if (N > 0) (void)hipEventElapsedTime(&g->build_ms, c->ev0, c->ev1);
  (void)hipFree(d_rays);
  (void)hipFree(d_err);
  *out = g;
  for (int d = 0; d < D; ++d)
    if (h_err[(size_t)d]) {
      set_error(ANACOR_ERR_GEOMETRY, "gridding direction %d: the reference aborts here; its "
                "map entries are NaN", d);
      return ANACOR_ERR_GEOMETRY;
    }
  return ANACOR_OK;
}

struct GridLists {
  const double* theta_list;
  int64_t nx;
  const double* phi_list;
  int64_t ny;
  double theta_min, theta_max, phi_min, phi_max;
};

// Per-reflection angles (theta_1, phi_1, theta, phi) as
// nearest_neighbor_interpolate_overall computes them (gridding_interpolation.c:
// 120-146: rotation, then dials_2_thetaphi_22 of both rays).
static void reflection_angles(const double* s1, const double* omega, int64_t R,
                              const double* raw, const double* axis, const double* F,
                              int64_t max_dim_, std::vector<ReflRule>& rules,
                              std::vector<ReflGeom>& geom, std::vector<double>& ang, int threads) {
  rules.resize((size_t)R);
  geom.resize((size_t)R);
  ang.resize((size_t)R * 4);
  host_setup(s1, omega, R, raw, axis, F, max_dim_, rules.data(), geom.data(), threads);
  for (int64_t i = 0; i < R; ++i) {
    ang[4 * i + 0] = geom[(size_t)i].theta[0];
    ang[4 * i + 1] = geom[(size_t)i].phi[0];
    ang[4 * i + 2] = geom[(size_t)i].theta[1];
    ang[4 * i + 3] = geom[(size_t)i].phi[1];
  }
}

// Look-ups of R reflections with the given angles on the device; map is the
// resident map (float) or, when map64 is set, the reference's padded layout.
static int grid_lookup(hipStream_t stream, const float* map, const double* map64, int N,
                       int n_theta, int n_phi, const GridLists& L, const double* h_angles,
                       int64_t R, int method, double* out, double** d_bufs, size_t* cap,
                       double** d_lists, size_t* lists_cap, float* ms) {
  if (method != 1 && method != 2) {
    set_error(ANACOR_ERR_ARG, "interpolation method %d (1 nearest, 2 bilinear)", method);
    return ANACOR_ERR_ARG;
  }
  if (L.nx < 2 || L.ny < 2 || !L.theta_list || !L.phi_list) {
    set_error(ANACOR_ERR_ARG, "bad grid lists");
    return ANACOR_ERR_ARG;
  }
  if (R > kMaxBlocks) {   // one workgroup per reflection
    set_error(ANACOR_ERR_ARG, "%lld reflections in one call exceed one dispatch (at most %lld)",
              (long long)R, (long long)kMaxBlocks);
    return ANACOR_ERR_ARG;
  }
  const size_t need = (size_t)R * (4 * sizeof(double) + sizeof(double) + sizeof(unsigned)) + 64;
  if (need > *cap) {
    if (*d_bufs) HIP_TRY(hipFree(*d_bufs));
    *d_bufs = nullptr;
    HIP_TRY(hipMalloc(d_bufs, need));
    *cap = need;
  }
  const size_t nl = (size_t)(L.nx + L.ny);
  if (nl > *lists_cap) {
    if (*d_lists) HIP_TRY(hipFree(*d_lists));
    *d_lists = nullptr;
    HIP_TRY(hipMalloc(d_lists, nl * sizeof(double)));
    *lists_cap = nl;
  }
  double* d_ang = *d_bufs;
  double* d_out = d_ang + 4 * (size_t)R;
  unsigned* d_err = (unsigned*)(d_out + R);
  HIP_TRY(hipMemcpyAsync(d_ang, h_angles, (size_t)R * 4 * sizeof(double), hipMemcpyHostToDevice,
                         stream));
  HIP_TRY(hipMemcpyAsync(*d_lists, L.theta_list, (size_t)L.nx * sizeof(double),
                         hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(*d_lists + L.nx, L.phi_list, (size_t)L.ny * sizeof(double),
                         hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemsetAsync(d_err, 0, (size_t)R * sizeof(unsigned), stream));
  GridLookup g;
  g.map = map;
  g.map64 = map64;
  g.N = N;
  g.n_theta = n_theta;
  g.n_phi = n_phi;
  g.theta_extra = (int)((L.nx - n_theta) / 2);
  g.phi_extra = (int)((L.ny - n_phi) / 2);
  g.theta_list = *d_lists;
  g.phi_list = *d_lists + L.nx;
  g.nx = (int)L.nx;
  g.ny = (int)L.ny;
  g.theta_min = L.theta_min;
  g.theta_max = L.theta_max;
  g.phi_min = L.phi_min;
  g.phi_max = L.phi_max;
  g.angles = d_ang;
  g.order = nullptr;
  g.R = (int)R;
  g.out = d_out;
  g.err = d_err;
  struct Events {   // destroyed on every way out
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } ev;
  HIP_TRY(hipEventCreate(&ev.e0));
  HIP_TRY(hipEventCreate(&ev.e1));
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  HIP_TRY(hipEventRecord(e0, stream));
  const bool res = map != nullptr;
  if (method == 1 && res) hipLaunchKernelGGL((grid_lookup_kernel<1, true>), dim3((unsigned)R), dim3(kBlock), 0, stream, g);
  if (method == 2 && res) hipLaunchKernelGGL((grid_lookup_kernel<2, true>), dim3((unsigned)R), dim3(kBlock), 0, stream, g);
  if (method == 1 && !res) hipLaunchKernelGGL((grid_lookup_kernel<1, false>), dim3((unsigned)R), dim3(kBlock), 0, stream, g);
  if (method == 2 && !res) hipLaunchKernelGGL((grid_lookup_kernel<2, false>), dim3((unsigned)R), dim3(kBlock), 0, stream, g);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(e1, stream));
  std::vector<unsigned> h_err((size_t)R);
  HIP_TRY(hipMemcpyAsync(out, d_out, (size_t)R * sizeof(double), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipMemcpyAsync(h_err.data(), d_err, (size_t)R * sizeof(unsigned),
                         hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  if (ms) (void)hipEventElapsedTime(ms, e0, e1);
  for (int64_t i = 0; i < R; ++i)
    if (h_err[(size_t)i]) {
      set_error(ANACOR_ERR_GEOMETRY, "reflection %lld: direction outside the grid (result NaN)",
                (long long)i);
      return ANACOR_ERR_GEOMETRY;
    }
  return ANACOR_OK;
}

// ------------------------------------------------------ iterative bisection
//
// The reference's alternative traversal (ib_test, ray_tracing.c:25-160, with
// bisection.c:21-471): per ray, instead of visiting every voxel to the exit
// plane, bisect the distance along the exit face's axis for the outer air
// boundary, the crystal's outer boundary and, between them, the loop (class
// 2) and bubble (class 4) boundaries, each to a resolution of one step.
// One (reflection, voxel) pair per lane as in trace_kernel; a lane's label
// reads are a chain of dependent gathers (about 8 per search, 4 to 6 searches
// per ray), so the kernel is bound by gather latency, not by bytes.
//
// Positions are FP64 values of floor(): integers a double holds exactly, so
// differences and lengths equal the reference's int64 arithmetic.  A sample
// point outside the volume reads as vacuum (0); the reference reads its row
// table there unchecked.  Such reads are counted (anacor_stats_t).

struct IbArgs {
  const uint8_t* vol;          // raw volume or the packed one
  int vol_bytes;
  int cz, cy;                  // brick constants (bricked layouts)
  int Z, Y, X;
  const int4* coords;          // (z, y, x, caller index), wave-grouped order
  int N;
  const IbRule* rules;         // caller's order
  int R;
  int tiles, vox_per_tile, n_blocks;
  double vs[3];
  double mu[4];
  int is_exp;
  int num_cls;
  float2* partial32;
  double* partial64;
  unsigned* err;
  unsigned long long* reads;   // label reads
  unsigned long long* oob;     // of which outside the volume
  double* lengths;             // [R][N][2][4] in the caller's voxel order, or null
  const int* order;
  int pos0;                    // processing position of the launch's first block row
};

// Label at integer-valued FP64 position (z, y, x); 0 outside the volume.
// The positions are finite or infinite, never NaN (finite voxel, positive
// mid, finite or infinite ratio), so the saturating conversion keeps every
// outside position outside and the range test runs on integers.
// The addresses use full-rate 24-bit multiply-adds (v_mad_i32_i24) wherever
// the host has checked that every factor fits: the raw layout below 2^23
// voxels per z plane, the bricked layouts always (bricks_fit); the wide raw
// layout uses 32-bit multiplies.  yx: the z stride of the raw layouts (V.sz).
template <int LAYOUT>
__device__ __forceinline__ int ib_label(const Volume& V, int yx, double z, double y, double x,
                                        unsigned& oob) {
  const int iz = cvt_trunc_sat(z), iy = cvt_trunc_sat(y), ix = cvt_trunc_sat(x);
  if (!((unsigned)iz < (unsigned)V.Z && (unsigned)iy < (unsigned)V.Y &&
        (unsigned)ix < (unsigned)V.X)) {
    ++oob;
    return 0;
  }
  if (LAYOUT == kLayoutWide) return load_label(V, iz * yx + iy * V.sy + ix + V.base);
  if (LAYOUT == kLayoutRaw) return load_label(V, mad24(iz, yx, mad24(iy, V.sy, ix + V.base)));
  const int f = mad24(iz, 64, mad24(iz >> 3, V.cz, mad24(iy, 8, mad24(iy >> 3, V.cy,
                mad24(ix >> 3, 504, ix)))));
  return load_packed<LAYOUT>(V, f);
}

struct IbPoint {
  double p[3];                 // z, y, x
  double mid;
};

// bisection (bisection.c:156-235): halve [shortest, longest] to one step.
// outer = 0 pulls `longest` in while the sample holds `cls`; outer = 1 pushes
// `shortest` out while it does.  Returns the voxel plus floor(mid) steps.
template <int LAYOUT>
__device__ __forceinline__ IbPoint ib_search(const Volume& V, int yx, const double c[3],
                                             const double r[3], double longest, double shortest,
                                             bool outer, int cls, unsigned& reads,
                                             unsigned& oob) {
  double diff = longest - shortest;
  double mid = (longest + shortest) / 2.0;
  while (diff > 1.0) {
    mid = (longest + shortest) / 2.0;
    const int label = ib_label<LAYOUT>(V, yx, __builtin_floor(c[0] + mid * r[0]),
                                       __builtin_floor(c[1] - mid * r[1]),
                                       __builtin_floor(c[2] + mid * r[2]), oob);
    if ((label == cls) == outer) shortest = mid;
    else longest = mid;
    diff = longest - shortest;
    ++reads;
  }
  const double m = __builtin_floor(mid);
  IbPoint q;
  q.p[0] = __builtin_floor(c[0] + m * r[0]);
  q.p[1] = __builtin_floor(c[1] - m * r[1]);
  q.p[2] = __builtin_floor(c[2] + m * r[2]);
  q.mid = mid;
  return q;
}

__device__ __forceinline__ double ib_l1(const double a[3], const double b[3]) {
  return __builtin_fabs(a[0] - b[0]) + __builtin_fabs(a[1] - b[1]) + __builtin_fabs(a[2] - b[2]);
}

// difference_length (bisection.c:244-253), pow(v, 2) as v*v.
__device__ __forceinline__ double ib_dist(const double a[3], const double b[3],
                                          const double vs[3]) {
  const double dz = (a[0] - b[0]) * vs[0];
  const double dy = (a[1] - b[1]) * vs[1];
  const double dx = (a[2] - b[2]) * vs[2];
  return __builtin_sqrt(dz * dz + dy * dy + dx * dx);
}

// iterative_bisection (bisection.c:298-471) and cal_path_bisection (:255-296)
// of one ray from voxel c: L = [li, lo, cr, bu].  Each boundary pair is
// measured as soon as it is found, so only its length stays live.
template <int LAYOUT>
__device__ __forceinline__ void ib_ray(const Volume& V, int yx, const double c[3], const double r[3],
                                       double air_longest, int num_cls, const double vs[3],
                                       double L[4], unsigned& reads, unsigned& oob) {
  const IbPoint A = ib_search<LAYOUT>(V, yx, c, r, air_longest, 0.0, false, 0, reads, oob);
  const double total = ib_dist(c, A.p, vs);
  double cr;
  double cmid;
  {
    const IbPoint C = ib_search<LAYOUT>(V, yx, c, r, A.mid, 0.0, true, 3, reads, oob);
    cr = ib_dist(c, C.p, vs);
    cmid = C.mid;
  }
  // lengths of the pairs appended after (voxel, A, C), in order: at most two
  double pair0 = 0.0, pair1 = 0.0;
  int pairs = 0;
  do {
    const IbPoint Li = ib_search<LAYOUT>(V, yx, c, r, A.mid, cmid, false, 2, reads, oob);
    if (ib_l1(Li.p, A.p) / 3 > 1) {
      const IbPoint Lo = ib_search<LAYOUT>(V, yx, c, r, A.mid, Li.mid, true, 2, reads, oob);
      if (ib_l1(Li.p, Lo.p) / 3 <= 2) break;     // a segmentation artefact ends the search
      pair0 = ib_dist(Li.p, Lo.p, vs);
      pairs = 1;
    }
    if (num_cls > 3) {
      const IbPoint Bi = ib_search<LAYOUT>(V, yx, c, r, A.mid, cmid, false, 4, reads, oob);
      if (ib_l1(Bi.p, A.p) > 1) {
        const IbPoint Bo = ib_search<LAYOUT>(V, yx, c, r, A.mid, Bi.mid, true, 4, reads, oob);
        if (ib_l1(Bi.p, Bo.p) / 3 <= 2) break;
        const double d = ib_dist(Bi.p, Bo.p, vs);
        if (pairs == 0) pair0 = d;
        else pair1 = d;
        ++pairs;
      }
    }
  } while (false);
  // with one pair it is measured as the loop, with two only the second, as
  // the bubble (cal_path_bisection reads points 3-4 or 5-6)
  const double lo = pairs == 1 ? pair0 : 0.0;
  const double bu = pairs == 2 ? pair1 : 0.0;
  L[0] = total - lo - bu - cr - 0.0;
  L[1] = lo;
  L[2] = cr;
  L[3] = bu;
}

// Two lanes per (reflection, voxel): the even lane bisects the incident ray,
// the odd lane the diffracted one, so each lane's chain of dependent gathers
// is one ray long and a wave keeps twice as many in flight.  The pair swaps
// its four lengths (lane ^ 1) and the even lane applies cal_rate.
// Six waves per SIMD (at most 80 VGPRs, no spills): the kernel waits on
// gathers, so resident waves matter more than registers.
template <int LAYOUT, bool F64, bool LENGTHS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6)))
void ib_kernel(IbArgs a) {
  __shared__ ReduceLds red;
  const int lb = xcd_logical_block((int)blockIdx.x, (int)gridDim.x);
  if (lb >= a.n_blocks) return;
  const int row = lb / a.tiles;
  const int tile = lb - row * a.tiles;
  const int pos = a.pos0 + row;
  const int refl = a.order[pos];
  const Volume V = make_volume(a.vol, a.Z, a.Y, a.X, a.vol_bytes, a.cz, a.cy);
  const int k = (int)threadIdx.x & 1;
  const IbRay& ray = a.rules[refl].ray[k];
  TileAcc acc = {0.f, 0.f, 0.0, 0ull, 0u};
  unsigned oob = 0;
  const int v_begin = tile * a.vox_per_tile;
  const int v_end = min(a.N, v_begin + a.vox_per_tile);
  for (int v = v_begin + ((int)threadIdx.x >> 1); v < v_end; v += kBlock / 2) {
    const int4 co = a.coords[v];
    const double c[3] = {(double)co.x, (double)co.y, (double)co.z};
    double L[4] = {0.0, 0.0, 0.0, 0.0};
    unsigned verr = 0, reads = 0;
    const int face = exit_face<false>(co.x, co.y, co.z, ray.dir[0], ray.dir[1], ray.dir[2],
                                      0.0, 0.0, 0.0, a.Z, a.Y, a.X, ray.neg != 0);
    if (face == 0) {
      verr = ANACOR_RAY_ERR_FACE;
    } else if (!((ray.valid >> (face - 1)) & 1u)) {
      verr = ANACOR_RAY_ERR_ASSERT;
    } else {
      // AirLongest of increments: steps from the voxel to the face's last plane
      double air;
      switch (face) {
        case 4: air = (double)co.y; break;
        case 5: air = (double)(a.Y - 1 - co.y); break;
        case 6: air = (double)co.z; break;
        case 1: air = (double)(a.X - 1 - co.z); break;
        case 2: air = (double)co.x; break;
        default: air = (double)(a.Z - 1 - co.x); break;
      }
      const double r[3] = {ray.r[face - 1][0], ray.r[face - 1][1], ray.r[face - 1][2]};
      ib_ray<LAYOUT>(V, V.sz, c, r, air, a.num_cls, a.vs, L, reads, oob);
    }
    acc.err |= verr;
    acc.steps += reads;
    if (LENGTHS) {
      double* out = a.lengths + ((size_t)refl * a.N + co.w) * 8 + k * 4;
#pragma unroll
      for (int j = 0; j < 4; ++j) out[j] = L[j];
    }
    double M[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) M[j] = __shfl_xor(L[j], 1);
    // a pair with an error contributes nothing: its reflection ends as NaN
    verr |= __shfl_xor(verr, 1);
    if (k == 0 && verr == 0) {
      // cal_rate (ray_tracing_basic.c:1554-1583), incident + diffracted in
      // the reference's association order
      const double rsum = a.mu[0] * (L[0] + M[0]) + a.mu[1] * (L[1] + M[1]) +
                          a.mu[2] * (L[2] + M[2]) + a.mu[3] * (L[3] + M[3]);
      if (F64) {
        acc.acc64 += a.is_exp ? exp(-rsum) : rsum;
      } else {
        const float val = a.is_exp ? exp_neg(rsum) : (float)rsum;
        const float y = val - acc.kc;
        const float t = acc.ks + y;
        acc.kc = (t - acc.ks) - y;
        acc.ks = t;
      }
    }
  }
  if (oob) atomicAdd(a.oob, (unsigned long long)oob);
  const size_t slot = (size_t)pos * a.tiles + tile;   // finish_kernel reads [position][tile]
  reduce_tile<F64, false>(acc, red, &a.err[refl], a.reads, a.partial32 + slot,
                          a.partial64 + slot);
}

template <int LAYOUT, bool F64, bool LENGTHS>
static void launch_ib(const IbArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((ib_kernel<LAYOUT, F64, LENGTHS>), dim3((unsigned)grid), dim3(kBlock), 0, s,
                     a);
}

static void dispatch_ib(const IbArgs& a, int grid, int layout, bool f64, bool lengths,
                        hipStream_t s) {
  using Fn = void (*)(const IbArgs&, int, hipStream_t);
  static const Fn table[4][2][2] = {
      {{launch_ib<kLayoutRaw, false, false>, launch_ib<kLayoutRaw, false, true>},
       {launch_ib<kLayoutRaw, true, false>, launch_ib<kLayoutRaw, true, true>}},
      {{launch_ib<kLayoutWide, false, false>, launch_ib<kLayoutWide, false, true>},
       {launch_ib<kLayoutWide, true, false>, launch_ib<kLayoutWide, true, true>}},
      {{launch_ib<kLayoutBrick2, false, false>, launch_ib<kLayoutBrick2, false, true>},
       {launch_ib<kLayoutBrick2, true, false>, launch_ib<kLayoutBrick2, true, true>}},
      {{launch_ib<kLayoutBrick4, false, false>, launch_ib<kLayoutBrick4, false, true>},
       {launch_ib<kLayoutBrick4, true, false>, launch_ib<kLayoutBrick4, true, true>}}};
  const int li = layout == kLayoutWide ? 1 : layout == kLayoutBrick2 ? 2
                                                                  : layout == kLayoutBrick4 ? 3 : 0;
  table[li][f64 ? 1 : 0][lengths ? 1 : 0](a, grid, s);
}

// First failing reflection of a run (its result is already NaN) as the
// context's error; the error words are read back only when some result is NaN.
static int report_ray_errors(Context* c, const double* out, int64_t R) {
  int64_t bad = -1;
  for (int64_t i = 0; i < R && bad < 0; ++i)
    if (out[i] != out[i]) bad = i;
  if (bad < 0) return ANACOR_OK;
  std::vector<unsigned> herr((size_t)R);
  HIP_TRY(hipMemcpy(herr.data(), c->d_err, (size_t)R * sizeof(unsigned), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < R; ++i)
    if (herr[(size_t)i]) {
      set_error(ANACOR_ERR_GEOMETRY,
                "reflection %lld: %s (the reference aborts here); result set to NaN",
                (long long)i,
                (herr[(size_t)i] & ANACOR_RAY_ERR_FACE) ? "no exit face (cube_face)"
                                                        : "exit face inconsistent with beam angles");
      return ANACOR_ERR_GEOMETRY;
    }
  return ANACOR_OK;
}

// Bisection traversal of R reflections whose rules the caller has written
// into c->h_ib (pinned).
// With first in (0, R) and setup_rest set (no lengths): only the rules of
// reflections [0, first) are ready; setup_rest() fills the rest into
// h_ib + first while the GPU bisects the first part, and each part is sorted
// by direction on its own (as run_rules does for the walk).
static int ib_run(Context* c, int64_t R, int num_cls, int flags, double* out, double* lengths,
                  int64_t first = -1, const std::function<int()>& setup_rest = {}) {
  const bool f64 = (flags & ANACOR_FLAG_F64_MEAN) != 0;
  c->last_oob = 0;
  c->last_steps = 0;
  if (R == 0) return ANACOR_OK;
  if (c->N == 0) {
    for (int64_t i = 0; i < R; ++i) out[i] = __builtin_nan("");
    return ANACOR_OK;
  }
  const int vpt = fit_tile(c->N, R, choose_tile(c, R));
  if (vpt == 0) {
    set_error(ANACOR_ERR_ARG, "%lld reflections in one call exceed one dispatch (at most %lld)",
              (long long)R, (long long)kMaxBlocks);
    return ANACOR_ERR_ARG;
  }
  const int64_t tiles = (c->N + vpt - 1) / vpt;
  const int64_t nblk = R * tiles;
  int rc = ensure_run_buffers(c, R, nblk, f64);
  if (rc) return rc;
  if ((size_t)R > c->ib_cap) {
    if (c->d_ib) HIP_TRY(hipFree(c->d_ib));
    c->d_ib = nullptr;
    HIP_TRY(hipMalloc(&c->d_ib, (size_t)R * sizeof(IbRule)));
    c->ib_cap = (size_t)R;
  }
  if (!c->d_oob) HIP_TRY(hipMalloc(&c->d_oob, sizeof(unsigned long long)));
  const IbRule* rules = c->h_ib.data();
  const int64_t RA = (!lengths && setup_rest && first > 0 && first < R) ? first : R;
  order_by_direction(c, RA, [rules](int64_t i, int k) { return rules[i].ray[k].dir; }, 0);
  HIP_TRY(c->h_order_pin.reserve((size_t)R));
  for (int64_t i = 0; i < RA; ++i) c->h_order_pin[(size_t)i] = c->h_order[(size_t)i];
  // Automatic layout: bricks whenever they fit.  The bisection's gathers are
  // scattered along the ray, so a 4x or 2x smaller volume keeps more of it in
  // the L2s; the walk's choice (raw up to 64 MB) does not carry over.
  if (c->layout_req == ANACOR_LAYOUT_AUTO && bricks_fit(c)) {
    const char* env = std::getenv("ANACOR_VOLUME_LAYOUT");
    if (!(env && std::strcmp(env, "raw") == 0)) {
      c->layout_req = ANACOR_LAYOUT_BRICK;
      rc = ensure_layout(c);
      c->layout_req = ANACOR_LAYOUT_AUTO;
    } else {
      rc = ensure_layout(c);
    }
  } else {
    rc = ensure_layout(c);
  }
  if (rc) return rc;
  DeviceTemp len_buf;   // per-ray lengths, freed on every return below
  if (lengths) HIP_TRY(hipMalloc(&len_buf.p, (size_t)R * c->N * 8 * sizeof(double)));
  double* const d_len = (double*)len_buf.p;

  HIP_TRY(hipMemcpyAsync(c->d_ib, c->h_ib.data(), (size_t)RA * sizeof(IbRule),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->d_order, c->h_order_pin.data(), (size_t)RA * sizeof(int),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(c->d_err, 0, (size_t)R * sizeof(unsigned), c->stream));
  HIP_TRY(hipMemsetAsync(c->d_steps, 0, sizeof(unsigned long long), c->stream));
  HIP_TRY(hipMemsetAsync(c->d_oob, 0, sizeof(unsigned long long), c->stream));

  IbArgs a;
  const int layout = c->layout;
  const bool bricked = layout == kLayoutBrick2 || layout == kLayoutBrick4;
  a.vol = bricked ? c->d_pack : c->d_vol;
  a.vol_bytes = bricked ? (int)c->pack_bytes : (int)c->vol_bytes;
  a.cz = 512 * c->NBX * c->NBY - 512;
  a.cy = 512 * c->NBX - 64;
  a.Z = c->Z; a.Y = c->Y; a.X = c->X;
  a.coords = c->d_coords;
  a.N = c->N;
  a.rules = c->d_ib;
  a.R = (int)R;
  a.tiles = (int)tiles;
  a.vox_per_tile = vpt;
  a.n_blocks = (int)nblk;
  for (int i = 0; i < 3; ++i) a.vs[i] = c->vs[i];
  for (int i = 0; i < 4; ++i) a.mu[i] = c->mu[i];
  a.is_exp = (flags & ANACOR_FLAG_NO_EXP) ? 0 : 1;
  a.num_cls = num_cls;
  a.partial32 = (float2*)c->d_partial;
  a.partial64 = (double*)c->d_partial;
  a.err = c->d_err;
  a.reads = c->d_steps;
  a.oob = c->d_oob;
  a.lengths = d_len;
  a.order = c->d_order;
  auto launch_rows = [&](int64_t lo, int64_t hi) -> int {
    a.pos0 = (int)lo;
    a.n_blocks = (int)((hi - lo) * tiles);
    dispatch_ib(a, (int)((a.n_blocks + 7) / 8 * 8), layout, f64, lengths != nullptr, c->stream);
    HIP_TRY(hipGetLastError());
    return ANACOR_OK;
  };

  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  rc = launch_rows(0, RA);
  if (rc) return rc;
  if (RA < R) {
    // the second part: its setup and direction order on the host while the
    // GPU bisects the first; its uploads beside that launch, which reads
    // positions and reflections below RA only
    rc = setup_rest();
    if (rc) return rc;
    const int64_t n = R - RA;
    std::vector<int> ord((size_t)n);
    std::vector<uint64_t> key((size_t)n), tmp((size_t)n);
    direction_order(n, [&](int64_t i, int k) { return rules[RA + i].ray[k].dir; }, ord.data(),
                    key.data(), tmp.data(), 0);
    for (int64_t i = 0; i < n; ++i)
      c->h_order_pin[(size_t)(RA + i)] = (int)RA + ord[(size_t)i];
    HIP_TRY(hipMemcpyAsync(c->d_ib + RA, c->h_ib.data() + RA, (size_t)n * sizeof(IbRule),
                           hipMemcpyHostToDevice, c->copy_stream));
    HIP_TRY(hipMemcpyAsync(c->d_order + RA, c->h_order_pin.data() + RA, (size_t)n * sizeof(int),
                           hipMemcpyHostToDevice, c->copy_stream));
    HIP_TRY(hipEventRecord(c->copied, c->copy_stream));
    HIP_TRY(hipStreamWaitEvent(c->stream, c->copied, 0));
    rc = launch_rows(RA, R);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  hipLaunchKernelGGL(finish_kernel, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, c->stream,
                     (const float2*)c->d_partial, (const double*)c->d_partial, c->d_err,
                     c->d_order, (int)R, (int)tiles, c->N, f64 ? 1 : 0, c->d_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, c->d_out, (size_t)R * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  if (lengths)
    HIP_TRY(hipMemcpyAsync(lengths, d_len, (size_t)R * c->N * 8 * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&c->last_steps, c->d_steps, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&c->last_oob, c->d_oob, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipEventElapsedTime(&c->last_kernel_ms, c->ev0, c->ev1));
  c->last_blocks = nblk;
  return report_ray_errors(c, out, R);
}

}  // namespace anacor

// ===================================================================== C-ABI

using namespace anacor;

extern "C" {

int anacor_ctx_set_layout(anacor_ctx* ctx, int layout) {
  if (!ctx || !ctx->c || !(layout == ANACOR_LAYOUT_AUTO || layout == ANACOR_LAYOUT_RAW ||
                           layout == ANACOR_LAYOUT_BRICK || layout == ANACOR_LAYOUT_BRICK2 ||
                           layout == ANACOR_LAYOUT_BRICK4)) {
    set_error(ANACOR_ERR_ARG, "bad context or layout %d", layout);
    return ANACOR_ERR_ARG;
  }
  Context* c = ctx->c;
  std::lock_guard<std::mutex> lk(c->mu_run);
  HIP_TRY(hipSetDevice(c->device));
  c->layout_req = layout;
  return ensure_layout(c);
}

int anacor_ctx_get_layout(const anacor_ctx* ctx) {
  if (!ctx || !ctx->c) return ANACOR_LAYOUT_RAW;
  const int l = ctx->c->layout;
  return l == kLayoutBrick2 ? ANACOR_LAYOUT_BRICK2
                            : (l == kLayoutBrick4 ? ANACOR_LAYOUT_BRICK4 : ANACOR_LAYOUT_RAW);
}

int anacor_last_error(char* buf, size_t n) {
  if (buf && n) {
    std::strncpy(buf, g_err, n - 1);
    buf[n - 1] = 0;
  }
  return g_err_code.load();
}

int anacor_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

anacor_ctx* anacor_ctx_create(const uint8_t* volume, const int64_t shape[3],
                              const int64_t* coords, int64_t n_coords,
                              const double voxel_size[3], const double coefficients[4],
                              int full_iteration, int device) {
  clear_error();
  Context* c = nullptr;
  if (ctx_create(volume, shape, coords, nullptr, n_coords, voxel_size, coefficients,
                 full_iteration, device, &c) != ANACOR_OK)
    return nullptr;
  return new anacor_ctx{c};
}

void anacor_ctx_destroy(anacor_ctx* ctx) {
  if (!ctx) return;
  ctx_destroy(ctx->c);
  delete ctx;
}

int anacor_ctx_set_voxel_block(anacor_ctx* ctx, int bz, int by, int bx) {
  clear_error();
  if (!ctx) return ANACOR_ERR_ARG;
  std::lock_guard<std::mutex> lk(ctx->c->mu_run);
  HIP_TRY(hipSetDevice(ctx->c->device));
  return upload_voxel_order(ctx->c, bz, by, bx);
}

int anacor_ctx_set_order(anacor_ctx* ctx, int refl_group, int direction_major) {
  if (!ctx || !ctx->c || refl_group < 0 || refl_group > (1 << 16) ||
      direction_major < -1 || direction_major > 1) {
    set_error(ANACOR_ERR_ARG, "bad context or launch order (%d, %d)", refl_group,
              direction_major);
    return ANACOR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(ctx->c->mu_run);
  ctx->c->refl_group = refl_group;
  ctx->c->sort_major = direction_major;
  return ANACOR_OK;
}

int anacor_ctx_get_work_shape(const anacor_ctx* ctx, int block[3], int* tile, int* refl_group,
                              int* direction_major) {
  if (!ctx || !ctx->c) {
    set_error(ANACOR_ERR_ARG, "null context");
    return ANACOR_ERR_ARG;
  }
  const Context* c = ctx->c;
  if (block)
    for (int i = 0; i < 3; ++i) block[i] = c->block[i];
  if (tile) *tile = c->tile_voxels;
  if (refl_group) *refl_group = launch_group(c, c->layout);
  if (direction_major) *direction_major = sort_major_for(c, c->layout);
  return ANACOR_OK;
}

int anacor_ctx_set_tile(anacor_ctx* ctx, int voxels_per_tile) {
  if (!ctx) return ANACOR_ERR_ARG;
  ctx->c->tile_voxels = voxels_per_tile > 0 ? voxels_per_tile : 0;
  return ANACOR_OK;
}

int anacor_ctx_run(anacor_ctx* ctx, const double* s1, const double* omega, int64_t R,
                   const double raw_xray[3], const double omega_axis[3], const double F[9],
                   int flags, double* out) {
  clear_error();
  if (!ctx || (R > 0 && !out)) {
    set_error(ANACOR_ERR_ARG, "null context or output");
    return ANACOR_ERR_ARG;
  }
  return ctx_run(ctx->c, s1, omega, R, raw_xray, omega_axis, F, flags, 0, out, nullptr);
}

int anacor_ctx_run_noalloc(anacor_ctx* ctx, const double* s1, const double* omega, int64_t R,
                           const double* raw_xray, const double* omega_axis, const double* F,
                           int flags, double* out) {
  clear_error();
  if (!ctx) {
    set_error(ANACOR_ERR_ARG, "null context");
    return ANACOR_ERR_ARG;
  }
  return ctx_run_noalloc(ctx->c, s1, omega, R, raw_xray, omega_axis, F, flags, out);
}

int anacor_ctx_paths(anacor_ctx* ctx, const double* s1, const double* omega, int64_t R,
                     const double raw_xray[3], const double omega_axis[3], const double F[9],
                     anacor_path_t* out) {
  clear_error();
  if (!ctx || (R > 0 && !out)) {
    set_error(ANACOR_ERR_ARG, "null context or output");
    return ANACOR_ERR_ARG;
  }
  static_assert(sizeof(anacor_path_t) == sizeof(PathRecord), "path record ABI");
  return ctx_run(ctx->c, s1, omega, R, raw_xray, omega_axis, F, 0, 0, nullptr,
                 reinterpret_cast<PathRecord*>(out));
}

int anacor_ctx_stats(const anacor_ctx* ctx, anacor_stats_t* st) {
  if (!ctx || !st) return ANACOR_ERR_ARG;
  st->kernel_ms = ctx->c->last_kernel_ms;
  st->voxel_steps = ctx->c->last_steps;
  st->workgroups = ctx->c->last_blocks;
  st->n_voxels = ctx->c->N;
  st->host_ms = ctx->c->last_host_ms;
  st->outside_reads = ctx->c->last_oob;
  return ANACOR_OK;
}

anacor_batch* anacor_batch_create(const anacor_sample_t* samples, int n_samples,
                                  int full_iteration, int device) {
  clear_error();
  Batch* b = nullptr;
  if (batch_create(samples, n_samples, full_iteration, device, &b) != ANACOR_OK) return nullptr;
  return new anacor_batch{b};
}

void anacor_batch_destroy(anacor_batch* batch) {
  if (!batch) return;
  batch_destroy(batch->b);
  delete batch;
}

anacor_ctx* anacor_batch_sample(anacor_batch* batch, int i) {
  if (!batch || i < 0 || i >= (int)batch->b->handles.size()) return nullptr;
  return &batch->b->handles[(size_t)i];
}

int anacor_batch_run(anacor_batch* batch, const double* const* s1, const double* const* omega,
                     const int64_t* R, const double raw_xray[3], const double omega_axis[3],
                     const double F[9], int flags, double* const* out) {
  clear_error();
  if (!batch || !s1 || !omega || !R || !out || !raw_xray || !omega_axis || !F) {
    set_error(ANACOR_ERR_ARG, "null argument");
    return ANACOR_ERR_ARG;
  }
  return batch_run(batch->b, s1, omega, R, raw_xray, omega_axis, F, flags, out);
}

int anacor_batch_set_graph(anacor_batch* batch, int use_graph) {
  if (!batch || !batch->b) {
    set_error(ANACOR_ERR_ARG, "null batch");
    return ANACOR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(batch->b->mu);
  batch->b->use_graph = use_graph != 0;
  return ANACOR_OK;
}

int anacor_batch_last_graph(const anacor_batch* batch) {
  return (batch && batch->b && batch->b->last_graph) ? 1 : 0;
}

int anacor_batch_stats(const anacor_batch* batch, anacor_stats_t* st) {
  if (!batch || !st) return ANACOR_ERR_ARG;
  st->kernel_ms = batch->b->last_ms;
  st->voxel_steps = batch->b->last_steps;
  st->workgroups = batch->b->blocks;
  int64_t n = 0;
  for (Context* c : batch->b->samples) n += c->N;
  st->n_voxels = n;
  st->host_ms = batch->b->last_host_ms;
  st->outside_reads = 0;
  return ANACOR_OK;
}

int anacor_reflection_geometry(const double* s1, const double* omega, int64_t R,
                               const double raw_xray[3], const double omega_axis[3],
                               const double F[9], int64_t max_dim_, double* theta_phi,
                               double* d_table) {
  clear_error();
  if (R < 0) return ANACOR_ERR_ARG;
  std::vector<ReflRule> rules((size_t)std::max<int64_t>(R, 1));
  std::vector<ReflGeom> geom((size_t)std::max<int64_t>(R, 1));
  host_setup(s1, omega, R, raw_xray, omega_axis, F, max_dim_, rules.data(), geom.data(), 1);
  for (int64_t i = 0; i < R; ++i) {
    if (theta_phi) {
      theta_phi[i * 4 + 0] = geom[(size_t)i].theta[0];
      theta_phi[i * 4 + 1] = geom[(size_t)i].phi[0];
      theta_phi[i * 4 + 2] = geom[(size_t)i].theta[1];
      theta_phi[i * 4 + 3] = geom[(size_t)i].phi[1];
    }
    if (d_table)
      for (int k = 0; k < 2; ++k)
        for (int f = 0; f < 6; ++f)
          for (int a = 0; a < 3; ++a)
            d_table[((i * 2 + k) * 6 + f) * 3 + a] = rules[(size_t)i].ray[k].d[f][a];
  }
  return ANACOR_OK;
}

// ---- drop-in entry points with the reference's signatures

double* ray_tracing_overall(int64_t low, int64_t up, int64_t* coord_list, int64_t len_coord_list,
                            const double* scattering_vector_list, const double* omega_list,
                            const double* raw_xray, const double* omega_axis,
                            const double* kp_rotation_matrix, int64_t len_result,
                            double* voxel_size, double* coefficients, int8_t*** label_list,
                            int64_t* shape, int full_iteration, int store_paths, int num_workers,
                            int IsExp) {
  (void)low; (void)up; (void)store_paths;
  clear_error();
  const int64_t R = len_result > 0 ? len_result : 0;
  double* result = (double*)std::malloc((size_t)std::max<int64_t>(R, 1) * sizeof(double));
  if (!result) return nullptr;
  for (int64_t i = 0; i < R; ++i) result[i] = __builtin_nan("");
  if (!label_list || !shape) {
    set_error(ANACOR_ERR_ARG, "null label_list or shape");
    return result;
  }
  std::vector<uint8_t> storage;
  const uint8_t* vol = flatten_table(label_list, shape, storage);
  Context* c = nullptr;
  if (ctx_create(vol, shape, coord_list, nullptr, len_coord_list, voxel_size, coefficients,
                 full_iteration, 0, &c) != ANACOR_OK) {
    fprintf(stderr, "anacor: ray_tracing_overall: %s\n", g_err);
    return result;
  }
  const int flags = ((char)IsExp == 1) ? 0 : ANACOR_FLAG_NO_EXP;
  const int rc = ctx_run(c, scattering_vector_list, omega_list, R, raw_xray, omega_axis,
                         kp_rotation_matrix, flags, num_workers, result, nullptr);
  if (rc != ANACOR_OK && rc != ANACOR_ERR_GEOMETRY)
    fprintf(stderr, "anacor: ray_tracing_overall: %s\n", g_err);
  ctx_destroy(c);
  return result;
}

float* ray_tracing_gpu_overall(size_t low, size_t up, int* coord_list, size_t len_coord_list,
                               const float* scattering_vector_list, const float* omega_list,
                               const float* raw_xray, const float* omega_axis,
                               const float* kp_rotation_matrix, size_t len_result,
                               float* voxel_size, float* coefficients, int8_t*** label_list,
                               int8_t* label_list_1d, int* shape, int full_iteration,
                               int store_paths, int gpumethod) {
  (void)low; (void)up; (void)store_paths; (void)gpumethod;
  clear_error();
  const int64_t R = (int64_t)len_result;
  float* result = (float*)std::malloc((size_t)std::max<int64_t>(R, 1) * sizeof(float));
  if (!result) return nullptr;
  for (int64_t i = 0; i < R; ++i) result[i] = __builtin_nanf("");
  if (!shape || (!label_list_1d && !label_list)) {
    set_error(ANACOR_ERR_ARG, "null volume or shape");
    return result;
  }
  const int64_t shp[3] = {shape[0], shape[1], shape[2]};
  std::vector<uint8_t> storage;
  const uint8_t* vol = label_list_1d ? reinterpret_cast<const uint8_t*>(label_list_1d)
                                     : flatten_table(label_list, shp, storage);
  // float inputs are widened to double and traced with the FP64 rules
  std::vector<double> s1d((size_t)R * 3), omd((size_t)R);
  for (int64_t i = 0; i < R * 3; ++i) s1d[(size_t)i] = scattering_vector_list[i];
  for (int64_t i = 0; i < R; ++i) omd[(size_t)i] = omega_list[i];
  double xr[3], ax[3], Fm[9], vs[3], mu[4];
  for (int i = 0; i < 3; ++i) {
    xr[i] = raw_xray[i];
    ax[i] = omega_axis[i];
    vs[i] = voxel_size[i];
  }
  for (int i = 0; i < 9; ++i) Fm[i] = kp_rotation_matrix[i];
  for (int i = 0; i < 4; ++i) mu[i] = coefficients[i];
  Context* c = nullptr;
  if (ctx_create(vol, shp, nullptr, coord_list, (int64_t)len_coord_list, vs, mu, full_iteration,
                 0, &c) != ANACOR_OK) {
    fprintf(stderr, "anacor: ray_tracing_gpu_overall: %s\n", g_err);
    return result;
  }
  std::vector<double> tmp((size_t)std::max<int64_t>(R, 1));
  const int rc = ctx_run(c, s1d.data(), omd.data(), R, xr, ax, Fm, 0, 0, tmp.data(), nullptr);
  if (rc == ANACOR_OK || rc == ANACOR_ERR_GEOMETRY)
    for (int64_t i = 0; i < R; ++i) result[i] = (float)tmp[(size_t)i];
  else
    fprintf(stderr, "anacor: ray_tracing_gpu_overall: %s\n", g_err);
  ctx_destroy(c);
  return result;
}

static double single_impl(int64_t* coord_list, int64_t len_coord_list, const double* rotated_s1,
                          const double* xray, double* voxel_size, double* coefficients,
                          int8_t*** label_list, int64_t* shape, int full_iteration, int IsExp) {
  clear_error();
  if (!label_list || !shape || !rotated_s1 || !xray) {
    set_error(ANACOR_ERR_ARG, "null argument");
    return __builtin_nan("");
  }
  std::vector<uint8_t> storage;
  const uint8_t* vol = flatten_table(label_list, shape, storage);
  Context* c = nullptr;
  if (ctx_create(vol, shape, coord_list, nullptr, len_coord_list, voxel_size, coefficients,
                 full_iteration, 0, &c) != ANACOR_OK) {
    fprintf(stderr, "anacor: ray_tracing_single: %s\n", g_err);
    return __builtin_nan("");
  }
  ReflRule rule;
  host_setup_rotated(rotated_s1, xray, max_dim(c), &rule, nullptr);
  double out = __builtin_nan("");
  const int flags = ((char)IsExp == 1) ? 0 : ANACOR_FLAG_NO_EXP;
  {
    std::lock_guard<std::mutex> lk(c->mu_run);
    if (hipSetDevice(c->device) == hipSuccess) {
      const int rc = run_rules(c, &rule, 1, flags, &out, nullptr);
      if (rc != ANACOR_OK && rc != ANACOR_ERR_GEOMETRY)
        fprintf(stderr, "anacor: ray_tracing_single: %s\n", g_err);
    }
  }
  ctx_destroy(c);
  return out;
}

double ray_tracing_single(int64_t* coord_list, int64_t len_coord_list, const double* rotated_s1,
                          const double* xray, double* voxel_size, double* coefficients,
                          int8_t*** label_list, int64_t* shape, int full_iteration,
                          int64_t store_paths, int IsExp) {
  (void)store_paths;
  return single_impl(coord_list, len_coord_list, rotated_s1, xray, voxel_size, coefficients,
                     label_list, shape, full_iteration, IsExp);
}

double ray_tracing_single_mp(int64_t* coord_list, int64_t len_coord_list,
                             const double* rotated_s1, const double* xray, double* voxel_size,
                             double* coefficients, int8_t*** label_list, int64_t* shape,
                             int full_iteration, int64_t store_paths, int IsExp,
                             int num_workers) {
  (void)store_paths; (void)num_workers;
  return single_impl(coord_list, len_coord_list, rotated_s1, xray, voxel_size, coefficients,
                     label_list, shape, full_iteration, IsExp);
}


// ---- iterative bisection

static int ib_check_num_cls(int num_cls) {
  if (num_cls < 1) {
    set_error(ANACOR_ERR_ARG, "num_cls must be positive");
    return ANACOR_ERR_ARG;
  }
  return ANACOR_OK;
}

int anacor_ctx_bisection(anacor_ctx* ctx, const double* s1, const double* omega, int64_t R,
                         const double raw_xray[3], const double omega_axis[3], const double F[9],
                         int num_cls, int flags, double* out) {
  clear_error();
  if (!ctx || R < 0 ||
      (R > 0 && (!s1 || !omega || !raw_xray || !omega_axis || !F || !out))) {
    set_error(ANACOR_ERR_ARG, "bad arguments");
    return ANACOR_ERR_ARG;
  }
  if (ib_check_num_cls(num_cls)) return ANACOR_ERR_ARG;
  Context* c = ctx->c;
  std::lock_guard<std::mutex> lk(c->mu_run);
  HIP_TRY(hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(c->h_ib.reserve((size_t)std::max<int64_t>(R, 1)));
  const int threads = host_threads(0);
  // large calls in two parts, as anacor_ctx_run's
  const int64_t first = R >= kSplitMinReflections ? R / 4 : R;
  ib_setup(s1, omega, first, raw_xray, omega_axis, F, c->h_ib.data(), threads);
  c->last_host_ms = ms_since(t0);
  if (first == R) return ib_run(c, R, num_cls, flags, out, nullptr);
  auto setup_rest = [&]() -> int {
    const auto t1 = std::chrono::steady_clock::now();
    ib_setup(s1 + 3 * first, omega + first, R - first, raw_xray, omega_axis, F,
             c->h_ib.data() + first, threads);
    c->last_host_ms += ms_since(t1);
    return ANACOR_OK;
  };
  return ib_run(c, R, num_cls, flags, out, nullptr, first, setup_rest);
}

int anacor_ctx_bisection_rotated(anacor_ctx* ctx, const double* rotated_s1, const double* xray,
                                 int64_t R, int num_cls, int flags, double* out,
                                 double* lengths) {
  clear_error();
  if (!ctx || R < 0 || (R > 0 && (!rotated_s1 || !xray || !out))) {
    set_error(ANACOR_ERR_ARG, "bad arguments");
    return ANACOR_ERR_ARG;
  }
  if (ib_check_num_cls(num_cls)) return ANACOR_ERR_ARG;
  Context* c = ctx->c;
  std::lock_guard<std::mutex> lk(c->mu_run);
  HIP_TRY(hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(c->h_ib.reserve((size_t)std::max<int64_t>(R, 1)));
  ib_setup_rotated(rotated_s1, xray, R, c->h_ib.data(), host_threads(0));
  c->last_host_ms = ms_since(t0);
  return ib_run(c, R, num_cls, flags, out, lengths);
}

double ib_test(int64_t* coord_list, int64_t len_coord_list, double* rotated_s1, double* xray,
               double* voxel_size, double* coefficients, int8_t*** label_list, int64_t* shape,
               int full_iteration, int64_t store_paths, int num_cls) {
  (void)store_paths;
  clear_error();
  if (!label_list || !shape || !rotated_s1 || !xray) {
    set_error(ANACOR_ERR_ARG, "null argument");
    return __builtin_nan("");
  }
  std::vector<uint8_t> storage;
  const uint8_t* vol = flatten_table(label_list, shape, storage);
  Context* c = nullptr;
  if (ctx_create(vol, shape, coord_list, nullptr, len_coord_list, voxel_size, coefficients,
                 full_iteration, 0, &c) != ANACOR_OK) {
    fprintf(stderr, "anacor: ib_test: %s\n", g_err);
    return __builtin_nan("");
  }
  double out = __builtin_nan("");
  {
    std::lock_guard<std::mutex> lk(c->mu_run);
    if (hipSetDevice(c->device) == hipSuccess && c->h_ib.reserve(1) == hipSuccess) {
      ib_setup_rotated(rotated_s1, xray, 1, c->h_ib.data(), 1);
      const int rc = num_cls < 1 ? ib_check_num_cls(num_cls)
                                 : ib_run(c, 1, num_cls, 0, &out, nullptr);
      if (rc != ANACOR_OK && rc != ANACOR_ERR_GEOMETRY)
        fprintf(stderr, "anacor: ib_test: %s\n", g_err);
    }
  }
  ctx_destroy(c);
  return out;
}

// ------------------------------------------------------------- gridding

anacor_grid* anacor_grid_create(anacor_ctx* ctx, int64_t n_theta, int64_t n_phi,
                                const double* rotated_s1, const double* theta_phi, int is_exp) {
  clear_error();
  if (!ctx || !ctx->c) {
    set_error(ANACOR_ERR_ARG, "null context");
    return nullptr;
  }
  Context* c = ctx->c;
  std::lock_guard<std::mutex> lk(c->mu_run);
  if (hipSetDevice(c->device) != hipSuccess) {
    set_error(ANACOR_ERR_HIP, "hipSetDevice failed");
    return nullptr;
  }
  Grid* g = nullptr;
  const int rc = grid_build(c, n_theta, n_phi, rotated_s1, theta_phi, is_exp ? 1 : 0, &g);
  if (rc != ANACOR_OK && rc != ANACOR_ERR_GEOMETRY) return nullptr;
  return new anacor_grid{g};
}

void anacor_grid_destroy(anacor_grid* grid) {
  if (!grid) return;
  if (grid->g) {
    (void)hipSetDevice(grid->g->c->device);
    grid_free(grid->g);
  }
  delete grid;
}

int anacor_grid_map(const anacor_grid* grid, float* out) {
  if (!grid || !grid->g || !out) {
    set_error(ANACOR_ERR_ARG, "null grid or output");
    return ANACOR_ERR_ARG;
  }
  const Grid* g = grid->g;
  std::lock_guard<std::mutex> lk(g->c->mu_run);
  HIP_TRY(hipSetDevice(g->c->device));
  HIP_TRY(hipMemcpyAsync(out, g->d_map, (size_t)g->D * g->c->N * sizeof(float),
                         hipMemcpyDeviceToHost, g->c->stream));
  HIP_TRY(hipStreamSynchronize(g->c->stream));
  return ANACOR_OK;
}

int anacor_grid_run(anacor_grid* grid, const double* s1, const double* omega, int64_t R,
                    const double raw_xray[3], const double omega_axis[3], const double F[9],
                    const double* theta_list, int64_t nx, const double* phi_list, int64_t ny,
                    int method, double* out) {
  clear_error();
  if (!grid || !grid->g || R < 0 || (R > 0 && (!s1 || !omega || !out || !raw_xray ||
                                                !omega_axis || !F))) {
    set_error(ANACOR_ERR_ARG, "bad grid run arguments");
    return ANACOR_ERR_ARG;
  }
  Grid* g = grid->g;
  Context* c = g->c;
  if (nx < g->n_theta || ny < g->n_phi || ((nx - g->n_theta) & 1) || ((ny - g->n_phi) & 1)) {
    set_error(ANACOR_ERR_ARG, "padded lists (%lld x %lld) do not extend the %d x %d grid "
              "evenly", (long long)nx, (long long)ny, g->n_theta, g->n_phi);
    return ANACOR_ERR_ARG;
  }
  if (R == 0) return ANACOR_OK;
  if (c->N == 0) {
    for (int64_t i = 0; i < R; ++i) out[i] = __builtin_nan("");
    return ANACOR_OK;
  }
  std::lock_guard<std::mutex> lk(c->mu_run);
  HIP_TRY(hipSetDevice(c->device));
  reflection_angles(s1, omega, R, raw_xray, omega_axis, F, max_dim(c), g->h_rules, g->h_geom,
                    g->h_angles, host_threads(0));
  GridLists L{theta_list, nx, phi_list, ny, *std::min_element(theta_list, theta_list + nx),
              *std::max_element(theta_list, theta_list + nx),
              *std::min_element(phi_list, phi_list + ny), *std::max_element(phi_list, phi_list + ny)};
  return grid_lookup(c->stream, g->d_map, nullptr, c->N, g->n_theta, g->n_phi, L,
                     g->h_angles.data(), R, method, out, &g->d_angles, &g->cap, &g->d_lists,
                     &g->lists_cap, &g->last_ms);
}

int anacor_grid_stats(const anacor_grid* grid, double* build_ms, double* lookup_ms) {
  if (!grid || !grid->g) return ANACOR_ERR_ARG;
  if (build_ms) *build_ms = grid->g->build_ms;
  if (lookup_ms) *lookup_ms = grid->g->last_ms;
  return ANACOR_OK;
}

double* nearest_neighbor_interpolate_overall(
    int64_t low, int64_t up, int64_t* coord_list, int64_t len_coord_list,
    const double* scattering_vector_list, const double* omega_list, const double* raw_xray,
    const double* omega_axis, const double* kp_rotation_matrix, int64_t len_result,
    double* voxel_size, double* coefficients, int num_workers, int IsExp, double* theta_list,
    double* phi_list, double* gridding_data, size_t nx, size_t ny, double theta_min,
    double theta_max, double phi_min, double phi_max, int interpolation_method) {
  (void)low; (void)up; (void)coord_list; (void)voxel_size; (void)coefficients; (void)IsExp;
  clear_error();
  double* result = (double*)std::malloc((size_t)std::max<int64_t>(len_result, 1) * sizeof(double));
  if (!result) return nullptr;
  for (int64_t i = 0; i < len_result; ++i) result[i] = __builtin_nan("");
  if (len_result <= 0) return result;
  if (!scattering_vector_list || !omega_list || !raw_xray || !omega_axis ||
      !kp_rotation_matrix || !gridding_data || !theta_list || !phi_list || len_coord_list <= 0) {
    set_error(ANACOR_ERR_ARG, "null argument");
    return result;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error(ANACOR_ERR_NO_DEVICE, "no HIP device visible");
    fprintf(stderr, "anacor: %s\n", g_err);
    return result;
  }
  std::vector<ReflRule> rules;
  std::vector<ReflGeom> geom;
  std::vector<double> ang;
  reflection_angles(scattering_vector_list, omega_list, len_result, raw_xray, omega_axis,
                    kp_rotation_matrix, 1 << 20, rules, geom, ang, host_threads(num_workers));
  hipStream_t st = nullptr;
  double* d_map = nullptr;
  double* d_bufs = nullptr;
  double* d_lists = nullptr;
  size_t cap = 0, lcap = 0;
  const size_t n = (size_t)len_coord_list * ny * nx;
  auto cleanup = [&]() {
    if (d_map) (void)hipFree(d_map);
    if (d_bufs) (void)hipFree(d_bufs);
    if (d_lists) (void)hipFree(d_lists);
    if (st) (void)hipStreamDestroy(st);
  };
  if (hipSetDevice(0) != hipSuccess || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) !=
      hipSuccess || hipMalloc(&d_map, n * sizeof(double)) != hipSuccess ||
      hipMemcpyAsync(d_map, gridding_data, n * sizeof(double), hipMemcpyHostToDevice, st) !=
      hipSuccess) {
    set_error(ANACOR_ERR_HIP, "gridding map upload failed");
    fprintf(stderr, "anacor: %s\n", g_err);
    cleanup();
    return result;
  }
  GridLists L{theta_list, (int64_t)nx, phi_list, (int64_t)ny, theta_min, theta_max, phi_min,
              phi_max};
  const int rc = grid_lookup(st, nullptr, d_map, (int)len_coord_list, (int)nx, (int)ny, L,
                             ang.data(), len_result, interpolation_method, result, &d_bufs, &cap,
                             &d_lists, &lcap, nullptr);
  if (rc != ANACOR_OK && rc != ANACOR_ERR_GEOMETRY) fprintf(stderr, "anacor: %s\n", g_err);
  cleanup();
  return result;
}

double ray_tracing_single_gridding(int64_t* coord, const double* rotated_s1, double theta,
                                   double phi, double* voxel_size, double* coefficients,
                                   int8_t*** label_list, int64_t* shape, int full_iteration,
                                   int index, int IsExp) {
  clear_error();
  if (!coord || !rotated_s1 || !label_list || !shape) {
    set_error(ANACOR_ERR_ARG, "null argument");
    return __builtin_nan("");
  }
  // the reference asserts that the given angles are those of rotated_s1
  ReflRule rule;
  ReflGeom geo;
  host_setup_rotated(rotated_s1, rotated_s1, 1 << 20, &rule, &geo);
  if (!(std::fabs(geo.theta[1] - theta) < 1e-6) || !(std::fabs(geo.phi[1] - phi) < 1e-6)) {
    set_error(ANACOR_ERR_GEOMETRY, "direction %d: theta/phi do not match rotated_s1", index);
    return __builtin_nan("");
  }
  std::vector<uint8_t> storage;
  const uint8_t* vol = flatten_table(label_list, shape, storage);
  Context* c = nullptr;
  if (ctx_create(vol, shape, coord, nullptr, 1, voxel_size, coefficients, full_iteration, 0,
                 &c) != ANACOR_OK) {
    fprintf(stderr, "anacor: ray_tracing_single_gridding: %s\n", g_err);
    return __builtin_nan("");
  }
  const double tp[2] = {theta, phi};
  double out = __builtin_nan("");
  {
    std::lock_guard<std::mutex> lk(c->mu_run);
    Grid* g = nullptr;
    double* d64 = nullptr;
    if (hipSetDevice(c->device) == hipSuccess && hipMalloc(&d64, sizeof(double)) == hipSuccess &&
        grid_build(c, 1, 1, rotated_s1, tp, ((char)IsExp == 1) ? 1 : 0, &g, d64) == ANACOR_OK) {
      double v = 0.0;
      if (hipMemcpy(&v, d64, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess) out = v;
    }
    if (g) grid_free(g);
    if (d64) (void)hipFree(d64);
  }
  ctx_destroy(c);
  return out;
}

}  // extern "C"
