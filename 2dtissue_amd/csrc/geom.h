// Geometry of the step that decides discrete outcomes: the seam wrap and the UV -> 3-D point
// location. Written as __host__ __device__ so that the very same source is compiled for the host
// (tests/hostcheck) and checked decision for decision against the FP64 oracle.
//
// Both parts follow the reference's FP64 arithmetic as written (compiled with -ffp-contract=off)
// and only skip work whose outcome is already determined:
//  - a border test t = n_t / det in [0, 1] (and likewise s) holds exactly when n_t == 0 or n_t
//    and det have the same sign, and |n_t| <= |det|: for doubles, |n| > |d| implies
//    fl(|n| / |d|) > 1 and |n| <= |d| implies fl(|n| / |d|) <= 1, so the division is only done for
//    the border that is crossed;
//  - point location first ranks the candidate faces with FP32 distances and keeps that ranking
//    only when the best two differ by more than kLocateMargin, far above the FP32 rounding of a
//    distance between points of [0, 1]^2; otherwise the FP64 distances decide. The nearest-vertex
//    choice is screened the same way in 3-D.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tdev {

constexpr double kPi = 3.141592653589793;
constexpr float kLocateMargin = 1e-5f;   // UV units (coordinates in [0, 1])
constexpr float kVertexMargin = 1e-4f;   // 3-D units, relative to the mesh's extent (|X| <= ~10)
constexpr int kFaceChunk = 4;            // candidate faces per lane loaded together

struct MeshDev {
  int32_t nv, nf, grid, pad;
  const float2* uv;        // [nv]
  const float* xyz;        // [nv*3]
  const int32_t* faces;    // [nf*3]
  const int32_t* cell_start;
  const float4* cell_tri;  // [2 * n_cell_faces]: candidate faces of each cell with their UV corners
  const float4* face_pack; // [5 * nf]: each face's corner ids, UV and 3-D coordinates (pack_face)
  const int32_t* row_ptr;  // [nv+1]
  const int2* nbr;         // (vertex, float bits of the geodesic distance), rows ascending in distance
};

// ------------------------------------------------------------------ seam wrap (FP64)

// The predicates below combine their comparisons with & and | rather than && and ||: the
// comparisons have no side effects, and without short-circuit sequencing the compiler evaluates
// them side by side instead of branching on each one.
__host__ __device__ inline bool inside_unit(double x, double y) {
  return ((int)(x >= 0.0) & (int)(x <= 1.0) & (int)(y >= 0.0) & (int)(y <= 1.0)) != 0;
}

// ascending sort of 8 values in registers: Batcher's odd-even merge network (19 comparators,
// fixed indices, so the values stay in registers)
__host__ __device__ inline void sort8(int (&a)[8]) {
  constexpr int P[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                            {0, 4}, {1, 5}, {2, 6}, {3, 7}, {2, 4}, {3, 5}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
  for (int k = 0; k < 19; ++k) {
    const int lo = a[P[k][0]] < a[P[k][1]] ? a[P[k][0]] : a[P[k][1]];
    const int hi = a[P[k][0]] < a[P[k][1]] ? a[P[k][1]] : a[P[k][0]];
    a[P[k][0]] = lo;
    a[P[k][1]] = hi;
  }
}

__host__ __device__ inline uint64_t bits64(double x) {
  uint64_t b;
  __builtin_memcpy(&b, &x, sizeof b);
  return b;
}
__host__ __device__ inline double from_bits64(uint64_t b) {
  double x;
  __builtin_memcpy(&x, &b, sizeof x);
  return x;
}
constexpr uint64_t kSignBit = 0x8000000000000000ull;
// x with its sign bit xor-ed with m (an exact negation when m = kSignBit)
__host__ __device__ inline double flip(double x, uint64_t m) { return from_bits64(bits64(x) ^ m); }

// fl(n / d) in [0, 1] given nf = n with the sign of d folded in (n negated when d's sign bit is
// set) and ad = |d|: fl(n / d) >= 0 exactly when nf >= 0 (zero of either sign included), and for
// nf >= 0, fl(nf / ad) <= 1 exactly when nf <= ad (rounding is monotone and 1 is representable);
// NaN fails both ways, as the quotient would.
__host__ __device__ inline int folded_in_unit(double nf, double ad) { return (int)(nf >= 0.0) & (int)(nf <= ad); }

// fl(n / d) in [0, 1], decided without dividing
__host__ __device__ inline bool quotient_in_unit(double n, double d) {
  return folded_in_unit(flip(n, bits64(d) & kSignBit), fabs(d)) != 0;
}

// EuclideanTiling::intersection_point (:170-208) of the step segment A -> A + (dx, dy) with
// border k of the unit square (0 left, 1 right, 2 up, 3 down), the border's constant edge vector
// folded in: for the left border C = (0,0), D = (0,1), det = (bx-ax)*1 - (by-ay)*0 = bx-ax
// exactly, and so on (left / right: det = dx, up / down: det = -dy). The border is crossed when
// t = n / det and s = (a dy - b dx) / det both lie in [0, 1]: left n = -ax, right 1 - ax, up
// -(1 - ay), down ay; a = 1 - ax for the right border, -ax otherwise; b = 1 - ay for the upper
// border, -ay otherwise. Every rounding happens in the reference's operation order, so the tests
// decide as the general formula does (tests/test_device_math.py checks the loop against the
// oracle). is_point_on_segment (:156-168) reduces to "A lies on this edge", which rejects the
// border; so does |det| < 1e-9. omx = 1 - ax, omy = 1 - ay. With k a compile-time constant the
// selects fold away.
__host__ __device__ inline int border_test(int k, double ax, double ay, double dx, double dy, double omx, double omy) {
  const bool lr = k < 2;
  const double u = lr ? ax : ay, w = lr ? ay : ax;  // coordinate across / along the edge
  const double n = flip((k == 1 || k == 2) ? (lr ? omx : omy) : u, (k == 0 || k == 2) ? kSignBit : 0ull);
  const double det = lr ? dx : -dy;
  const double sv = (k == 1 ? omx : -ax) * dy - (k == 2 ? omy : -ay) * dx;
  const uint64_t f = bits64(det) & kSignBit;
  const double ad = fabs(det);
  const int on_edge = (int)(u == ((k == 0 || k == 3) ? 0.0 : 1.0)) & (int)(w >= 0.0) & (int)(w <= 1.0);
  return (on_edge ^ 1) & (int)!(ad < 1e-9) & folded_in_unit(flip(n, f), ad) & folded_in_unit(flip(sv, f), ad);
}

// 1 / d refined for quotient(): on the device the hardware reciprocal with two Newton-Raphson
// steps, started as soon as d is known (before the border decision); unused on the host
__host__ __device__ inline double quotient_reciprocal(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
#else
  return d;
#endif
}

// fl(n / d), IEEE round to nearest. Device: q = n r, then one remainder correction
// q + (n - d q) r (Markstein, IBM J. Res. Dev. 34(1), 1990), the same steps as the compiler's
// division without its exponent pre-scaling and special-value fix-up, neither of which applies
// to the operands here (|d| in [1e-9, 1e6], |n| <= 2, no zero or non-finite divisor).
// tests/test_gpu_wrap.py checks the GPU loop against the oracle bit for bit.
__host__ __device__ inline double quotient(double n, double d, double r) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double q = n * r;
  const double rem = fma(-d, q, n);
  return fma(rem, r, q);
#else
  (void)r;
  return n / d;
#endif
}

// One crossing once the border is known (border 4: none found): exit point by
// intersection_point with the exit snapped to 0 below 1e-3 (check_border_crossings :71-101), the
// end point mapped by processPoints (:112-154), the heading rotation, the quarter turns of the
// composed map, and the entry point for the next test (the swapped exit point). rx, ry: the
// quotient reciprocals of dx and -dy.
__host__ __device__ inline void cross_border(int border, double& ax, double& ay, double& px, double& py, double dx,
                                             double dy, double omx, double omy, double rx, double ry, double& dtheta,
                                             int& quarter) {
  const bool side = border <= 1;  // left / right: rotate by -90 degrees, else by +90
  const double nt = side ? (border == 0 ? -ax : omx) : (border == 2 ? -omy : ay);
  const double det = side ? dx : (border < 4 ? -dy : 1.0);
  const double t = quotient(nt, det, side ? rx : (border < 4 ? ry : 1.0));
  double ex = ax + t * dx, ey = ay + t * dy;
  if (fabs(ex) < 1e-3) ex = 0.0;
  if (fabs(ey) < 1e-3) ey = 0.0;
  if (border == 4) {
    ex = ax;
    ey = ay;
  }
  const double ox = px, oy = py;
  px = side ? oy : (border == 2 ? 2 - oy : -oy);
  py = border == 0 ? -ox : (border == 1 ? 2 - ox : ox);
  dtheta -= side ? 0.5 * kPi : 1.5 * kPi;
  quarter += side ? -1 : 1;
  ax = ey;  // entry point = swapped exit point
  ay = ex;
}

// Seam wrap of one particle: the reference's restart loop (EuclideanTiling.cpp:31-69) acts on
// each particle independently, so it runs here as a per-particle loop. (ax, ay) -> (px, py) is the
// step's segment; on return (px, py) is inside the unit square (or the crossing bound was hit)
// and dtheta holds the heading rotation (-pi/2 per left/right, -3 pi/2 per up/down crossing).
// The composed map is p -> R p + b with R a rotation by quarter_out * 90 degrees
// (counter-clockwise, 0..3); the adjoint needs only R. Returns false when more than
// max_crossings crossings would be needed. check_border_crossings takes the first of left,
// right, up, down that the segment crosses; the four tests are evaluated side by side.
__host__ __device__ inline bool seam_wrap(double ax, double ay, double& px, double& py, double& dtheta,
                                          int max_crossings, int* crossings_out = nullptr, int* quarter_out = nullptr) {
  int crossings = 0;
  int quarter = 0;
  dtheta = 0.0;
  bool ok = true;
  while (!inside_unit(px, py)) {
    if (crossings >= max_crossings) { ok = false; break; }
    ++crossings;
    const double dx = px - ax, dy = py - ay, omx = 1.0 - ax, omy = 1.0 - ay;
    const double rx = quotient_reciprocal(dx), ry = quotient_reciprocal(-dy);
    const int bits = border_test(0, ax, ay, dx, dy, omx, omy) | (border_test(1, ax, ay, dx, dy, omx, omy) << 1) |
                     (border_test(2, ax, ay, dx, dy, omx, omy) << 2) | (border_test(3, ax, ay, dx, dy, omx, omy) << 3);
    const int border = (bits & 1) ? 0 : ((bits & 2) ? 1 : ((bits & 4) ? 2 : ((bits & 8) ? 3 : 4)));
    cross_border(border, ax, ay, px, py, dx, dy, omx, omy, rx, ry, dtheta, quarter);
  }
  if (crossings_out) *crossings_out = crossings;
  if (quarter_out) *quarter_out = quarter & 3;
  return ok;
}

// The border tests and the crossing rewritten for the lane-parallel loop: the same roundings as
// border_test and cross_border, with the choices between borders made by exact negations and
// sign-bit flips instead of selects, so one crossing is 114 instructions instead of 136 (gfx950).
// The identities used (round to nearest is symmetric): fl(-0.0 - x) = -x bit for bit (zero
// signs included), fl(1 - x) = omx, fl(2 - x) as the reference forms it.
//
// Border k's test in the form one lane evaluates with its lane constants: lr = (k < 2),
// ca = (k == 1 ? 1 : -0), cb = (k == 2 ? 1 : -0). Then a = fl(ca - ax) and b = fl(cb - ay) are
// border_test's (k == 1 ? omx : -ax) and (k == 2 ? omy : -ay); its numerator is a (left /
// right) or -b (up / down), its determinant dx or -dy, and "the entry point lies on this edge"
// is a == 0 (left / right) or b == 0 (up / down), with the point's other coordinate in [0, 1].
// Folding the determinant's sign into both numerators leaves the numerator's flip independent
// of lr: nf = flip(lr ? a : b, sign(lr ? dx : dy)).
__host__ __device__ inline bool border_pass(bool lr, double ca, double cb, double ax, double ay, double dx,
                                            double dy) {
  const double a = ca - ax, b = cb - ay;
  const double N = lr ? a : b, D = lr ? dx : dy, W = lr ? ay : ax;
  const double sv = a * dy - b * dx;
  const uint64_t f = bits64(D) & kSignBit;
  const double ad = fabs(D);
  const double nf = flip(N, f), svf = flip(sv, f ^ (lr ? 0ull : kSignBit));
  const int off_edge = (int)!(N == 0.0) | (int)!(W >= 0.0) | (int)!(W <= 1.0);
  return (off_edge & (int)!(ad < 1e-9) & folded_in_unit(nf, ad) & folded_in_unit(svf, ad)) != 0;
}

// cross_border for border b (4: none passed) with the per-border choices written as sign flips:
// numerator flip(fl(c1 - X), sgn), X = ax (left / right) or ay, c1 = 1 for right / up else -0;
// determinant flip(lr ? dx : dy, sgn); new end point flip(fl(cx - oy), sgn ^ sign),
// flip(fl(cy - ox), sgn) with cx = 2 for up, cy = 2 for right, else -0; sgn = 0 for left / right,
// the sign bit otherwise.
__host__ __device__ inline void cross_border_flip(int b, double& ax, double& ay, double& px, double& py, double dx,
                                                  double dy, double& dtheta, int& quarter) {
  const bool side = b < 2;
  const uint64_t sgn = side ? 0ull : kSignBit;
  const double X = side ? ax : ay;
  const double c1 = ((b == 1) | (b == 2)) ? 1.0 : -0.0;
  const double nt = flip(c1 - X, sgn);
  const double det = flip(side ? dx : dy, sgn);
  const double t = quotient(nt, det, quotient_reciprocal(det));
  double ex = ax + t * dx, ey = ay + t * dy;
  if (fabs(ex) < 1e-3) ex = 0.0;
  if (fabs(ey) < 1e-3) ey = 0.0;
  if (b == 4) {
    ex = ax;
    ey = ay;
  }
  const double cx = b == 2 ? 2.0 : -0.0, cy = b == 1 ? 2.0 : -0.0;
  const double ox = px, oy = py;
  px = flip(cx - oy, sgn ^ kSignBit);
  py = flip(cy - ox, sgn);
  dtheta -= side ? 0.5 * kPi : 1.5 * kPi;
  quarter += side ? -1 : 1;
  ax = ey;
  ay = ex;
}

// One lane-group's choice on the host: border_pass for the four borders in order, the first
// passing one crossed (tests/hostcheck compares this with seam_wrap bit for bit)
__host__ __device__ inline bool seam_wrap_flip(double ax, double ay, double& px, double& py, double& dtheta,
                                               int max_crossings, int* crossings_out = nullptr,
                                               int* quarter_out = nullptr) {
  int crossings = 0, quarter = 0;
  dtheta = 0.0;
  bool ok = true;
  while (!inside_unit(px, py)) {
    if (crossings >= max_crossings) { ok = false; break; }
    ++crossings;
    const double dx = px - ax, dy = py - ay;
    int b = 4;
    for (int k = 3; k >= 0; --k)
      if (border_pass(k < 2, k == 1 ? 1.0 : -0.0, k == 2 ? 1.0 : -0.0, ax, ay, dx, dy)) b = k;
    cross_border_flip(b, ax, ay, px, py, dx, dy, dtheta, quarter);
  }
  if (crossings_out) *crossings_out = crossings;
  if (quarter_out) *quarter_out = quarter & 3;
  return ok;
}

// ---- the last flight of a wave, three crossings per pass
//
// Once a single flight of a wave still needs crossings, all 64 lanes advance it together, three
// crossings per pass: lane l follows the path that crosses border (l >> 4) & 3 first, (l >> 2) & 3
// second and l & 3 third, and before each of its crossings tests border l & 3 (border_pass) on
// its path's current segment. The first passing border of each level is read from the ballots of
// those tests (level 1: lanes 0..3; level 2: lanes 16 c1 + 0..3; level 3: lanes 16 c1 + 4 c2 +
// 0..3), and the flight's state after the crossings taken is read from the lane whose path they
// are. Every lane performs exactly the sequential loop's operations along its own path, so the
// outcome is the sequential loop's bit for bit; a pass stops early where the flight lands inside
// the square, reaches its crossing bound, or meets a level whose four tests all fail (the
// reference's fall-through to the down mapping with the exit at the start), which every lane then
// crosses alike.

struct WrapState {
  double ax, ay, px, py, dth;
  int q;
};

// one crossing along border b (0..4) of the state, by cross_border_flip
__host__ __device__ inline void wrap_cross(int b, WrapState& s) {
  const double dx = s.px - s.ax, dy = s.py - s.ay;
  cross_border_flip(b, s.ax, s.ay, s.px, s.py, dx, dy, s.dth, s.q);
}

// lane l's path from s: its three border tests and its states after one, two and three crossings
struct WrapPath {
  bool t1, t2, t3, in1, in2;
  WrapState s1, s2, s3;
};

__host__ __device__ inline void wrap_path(int l, const WrapState& s, WrapPath& w) {
  const int k = l & 3;
  const bool lr = k < 2;
  const double ca = k == 1 ? 1.0 : -0.0, cb = k == 2 ? 1.0 : -0.0;
  w.t1 = border_pass(lr, ca, cb, s.ax, s.ay, s.px - s.ax, s.py - s.ay);
  w.s1 = s;
  wrap_cross((l >> 4) & 3, w.s1);
  w.in1 = inside_unit(w.s1.px, w.s1.py);
  w.t2 = border_pass(lr, ca, cb, w.s1.ax, w.s1.ay, w.s1.px - w.s1.ax, w.s1.py - w.s1.ay);
  w.s2 = w.s1;
  wrap_cross((l >> 2) & 3, w.s2);
  w.in2 = inside_unit(w.s2.px, w.s2.py);
  w.t3 = border_pass(lr, ca, cb, w.s2.ax, w.s2.ay, w.s2.px - w.s2.ax, w.s2.py - w.s2.ay);
  w.s3 = w.s2;
  wrap_cross(k, w.s3);
}

__host__ __device__ inline int first_of4(uint64_t bits) { return (bits & 0xFull) ? __builtin_ctzll(bits & 0xFull) : 4; }

// The pass loop. Lanes supplies run(s) (every lane's wrap_path from the uniform state s, and the
// 64-bit masks of t1, t2, t3, in1, in2 over the lanes) and take(level, lane) (lane's state after
// level crossings). crossings counts as in seam_wrap; returns false at the crossing bound.
struct WrapMasks {
  uint64_t t1, t2, t3, in1, in2;
};

template <class Lanes>
__host__ __device__ inline bool seam_wrap_passes(Lanes& L, WrapState& s, int& crossings, int max_crossings) {
  for (;;) {
    if (inside_unit(s.px, s.py)) return true;
    if (crossings >= max_crossings) return false;
    const WrapMasks m = L.run(s);
    const int c1 = first_of4(m.t1);
    ++crossings;
    if (c1 == 4) {
      wrap_cross(4, s);
      continue;
    }
    int lane = c1 << 4;
    if (((m.in1 >> lane) & 1) || crossings >= max_crossings) {
      s = L.take(1, lane);
      continue;
    }
    const int c2 = first_of4(m.t2 >> lane);
    ++crossings;
    if (c2 == 4) {
      s = L.take(1, lane);
      wrap_cross(4, s);
      continue;
    }
    lane |= c2 << 2;
    if (((m.in2 >> lane) & 1) || crossings >= max_crossings) {
      s = L.take(2, lane);
      continue;
    }
    const int c3 = first_of4(m.t3 >> lane);
    ++crossings;
    if (c3 == 4) {
      s = L.take(2, lane);
      wrap_cross(4, s);
      continue;
    }
    s = L.take(3, lane | c3);
  }
}

__device__ inline double readlane_f64(double x, int l) {
  const uint64_t b = bits64(x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return from_bits64(((uint64_t)hi << 32) | lo);
}

__device__ inline WrapState readlane_state(const WrapState& s, int l) {
  return WrapState{readlane_f64(s.ax, l), readlane_f64(s.ay, l), readlane_f64(s.px, l), readlane_f64(s.py, l),
                   readlane_f64(s.dth, l), __builtin_amdgcn_readlane(s.q, l)};
}

// the 64 lanes of a wave (all active) as seam_wrap_passes' Lanes
struct WaveLanes {
  WrapPath w;
  __device__ WrapMasks run(const WrapState& s) {
    wrap_path((int)__lane_id(), s, w);
    return WrapMasks{__builtin_amdgcn_ballot_w64(w.t1), __builtin_amdgcn_ballot_w64(w.t2),
                     __builtin_amdgcn_ballot_w64(w.t3), __builtin_amdgcn_ballot_w64(w.in1),
                     __builtin_amdgcn_ballot_w64(w.in2)};
  }
  __device__ WrapState take(int level, int l) const {
    return readlane_state(level == 1 ? w.s1 : (level == 2 ? w.s2 : w.s3), l);
  }
};

// One crossing of the flight of a lane group of LANES lanes (aligned within the wave, all of its
// lanes active): the four border tests are spread over the lanes (lane % 4 tests border lane % 4;
// with two lanes, lane k tests border k and border k + 2; one lane tests all four) and the first
// passing border is read from a wave ballot. k, ca, cb: the lane's border constants
// (group_constants); base: the group's first lane.
template <int LANES>
__device__ inline void group_constants(int lane, int& k, double& ca, double& cb) {
  k = LANES >= 4 ? (lane & 3) : (lane & 1);
  ca = k == 1 ? 1.0 : -0.0;
  cb = (LANES >= 4 ? k == 2 : k == 0) ? 1.0 : -0.0;
}

template <int LANES>
__device__ inline void group_crossing(int k, double ca, double cb, unsigned base, double& ax, double& ay, double& px,
                                      double& py, double& dtheta, int& quarter) {
  const double dx = px - ax, dy = py - ay;
  unsigned bits;
  if constexpr (LANES >= 4) {
    bits = (unsigned)(__builtin_amdgcn_ballot_w64(border_pass(k < 2, ca, cb, ax, ay, dx, dy)) >> base);
  } else if constexpr (LANES == 2) {
    const unsigned lo = (unsigned)(__builtin_amdgcn_ballot_w64(border_pass(true, ca, -0.0, ax, ay, dx, dy)) >> base) & 0x3u;
    const unsigned hi = (unsigned)(__builtin_amdgcn_ballot_w64(border_pass(false, -0.0, cb, ax, ay, dx, dy)) >> base) & 0x3u;
    bits = lo | (hi << 2);
  } else {  // one lane: seam_wrap's crossing (the four tests side by side, reciprocals started early)
    (void)k; (void)ca; (void)cb; (void)base;
    const double omx = 1.0 - ax, omy = 1.0 - ay;
    const double rx = quotient_reciprocal(dx), ry = quotient_reciprocal(-dy);
    const int b4 = border_test(0, ax, ay, dx, dy, omx, omy) | (border_test(1, ax, ay, dx, dy, omx, omy) << 1) |
                   (border_test(2, ax, ay, dx, dy, omx, omy) << 2) | (border_test(3, ax, ay, dx, dy, omx, omy) << 3);
    const int border = (b4 & 1) ? 0 : ((b4 & 2) ? 1 : ((b4 & 4) ? 2 : ((b4 & 8) ? 3 : 4)));
    cross_border(border, ax, ay, px, py, dx, dy, omx, omy, rx, ry, dtheta, quarter);
    return;
  }
  // first passing border of the group (bits above the group's four belong to the next group: a
  // lowest set bit at 4 or above, or none, means none of ours passed)
  const unsigned low = (unsigned)__builtin_ctz(bits | 0x10u);
  cross_border_flip((int)low, ax, ay, px, py, dx, dy, dtheta, quarter);
}

// The seam wrap run by the LANES lanes of one particle (group_crossing per crossing). Same
// decisions and the same FP64 values as seam_wrap.
// WAVE = false: each group loops on its own (finished groups wait masked).
// WAVE = true (for crowded systems, whose longest flights run hundreds of crossings while the
// other waves of the SIMD compete for issue): the groups of the wave stay in the loop together,
// and when a full wave has a single flight left, its 64 lanes take it over three crossings per
// pass (seam_wrap_passes); the flight's own lanes keep the result.
template <int LANES, bool WAVE = false>
__device__ inline bool seam_wrap_group(double ax, double ay, double& px, double& py, double& dtheta, int max_crossings,
                                       int lane, int* quarter_out) {
  const unsigned base = __lane_id() & ~(unsigned)(LANES - 1);
  int k;
  double ca, cb;
  group_constants<LANES>(lane, k, ca, cb);
  int crossings = 0;
  int quarter = 0;
  dtheta = 0.0;
  bool ok = true;
  if constexpr (!WAVE && LANES == 1) {
    (void)lane;
    return seam_wrap(ax, ay, px, py, dtheta, max_crossings, nullptr, quarter_out);
  } else if constexpr (!WAVE) {
    while (!inside_unit(px, py)) {
      if (crossings >= max_crossings) {
        ok = false;
        break;
      }
      ++crossings;
      group_crossing<LANES>(k, ca, cb, base, ax, ay, px, py, dtheta, quarter);
    }
  } else {
    const bool full = __builtin_amdgcn_ballot_w64(true) == ~0ull;  // every lane of the wave is here
    constexpr uint64_t kGroup = LANES == 64 ? ~0ull : ((1ull << LANES) - 1);
    bool run = !inside_unit(px, py);  // the flight needs another crossing (bound not yet checked)
    for (;;) {
      const uint64_t act = __builtin_amdgcn_ballot_w64(run);
      if (act == 0) break;
      if (full) {
        const int src = __builtin_ctzll(act);
        if ((act & ~(kGroup << (src & ~(LANES - 1)))) == 0) {  // a single flight left
          WrapState s = readlane_state(WrapState{ax, ay, px, py, dtheta, quarter}, src);
          int c = __builtin_amdgcn_readlane(crossings, src);
          WaveLanes wl;
          const bool good = seam_wrap_passes(wl, s, c, max_crossings);
          if (run) {
            px = s.px;
            py = s.py;
            dtheta = s.dth;
            quarter = s.q;
            ok = good;
          }
          break;
        }
      }
      if (run) {
        if (crossings >= max_crossings) {
          ok = false;
          run = false;
        } else {
          ++crossings;
          group_crossing<LANES>(k, ca, cb, base, ax, ay, px, py, dtheta, quarter);
          run = !inside_unit(px, py);
        }
      }
    }
  }
  if (quarter_out) *quarter_out = quarter & 3;
  return ok;
}

// heading kept in [-pi, pi)
__host__ __device__ inline double wrap_heading(double t) { return t - 2.0 * kPi * floor((t + kPi) / (2.0 * kPi)); }

// ------------------------------------------------------------------ point location

template <class T>
__host__ __device__ inline T sqr(T a) { return a * a; }

// CellHelper::pointSegmentDistance (CellHelper.cpp:216-222), z = 0
template <class T>
__host__ __device__ inline T point_segment_distance(T px, T py, T ax, T ay, T bx, T by) {
  const T abx = bx - ax, aby = by - ay;
  T t = (abx * (px - ax) + aby * (py - ay)) / (abx * abx + aby * aby);
  t = t < T(0) ? T(0) : (T(1) < t ? T(1) : t);
  const T qx = ax + abx * t - px, qy = ay + aby * t - py;
  return sqrt(qx * qx + qy * qy);
}

// CellHelper::pointTriangleDistance (CellHelper.cpp:162-214), z = 0
template <class T>
__host__ __device__ inline T point_triangle_distance(T px, T py, T ax, T ay, T bx, T by, T cx, T cy) {
  const T abx = bx - ax, aby = by - ay;
  const T acx = cx - ax, acy = cy - ay;
  const T apx = px - ax, apy = py - ay;
  const T bpx = px - bx, bpy = py - by;
  const T cpx = px - cx, cpy = py - cy;
  const T d_ab_ap = abx * apx + aby * apy;
  const T d_ac_ap = acx * apx + acy * apy;
  const T d_ab_bp = abx * bpx + aby * bpy;
  const T d_ac_bp = acx * bpx + acy * bpy;
  const T d_ab_cp = abx * cpx + aby * cpy;
  const T d_ac_cp = acx * cpx + acy * cpy;
  if (d_ab_ap <= T(0) && d_ac_ap <= T(0)) return sqrt(apx * apx + apy * apy);
  if (d_ab_bp >= T(0) && d_ac_bp <= d_ab_bp) return sqrt(bpx * bpx + bpy * bpy);
  if (d_ac_cp >= T(0) && d_ab_cp <= d_ac_cp) return sqrt(cpx * cpx + cpy * cpy);
  const T vc = d_ab_ap * d_ac_bp - d_ab_bp * d_ac_ap;
  if (vc <= T(0) && d_ab_ap >= T(0) && d_ab_bp <= T(0)) return point_segment_distance(px, py, ax, ay, bx, by);
  const T vb = d_ab_cp * d_ac_ap - d_ab_ap * d_ac_cp;
  if (vb <= T(0) && d_ac_ap >= T(0) && d_ac_cp <= T(0)) return point_segment_distance(px, py, ax, ay, cx, cy);
  const T va = d_ab_bp * d_ac_cp - d_ab_cp * d_ac_bp;
  if (va <= T(0) && (d_ac_bp - d_ab_bp) >= T(0) && (d_ab_cp - d_ac_cp) >= T(0))
    return point_segment_distance(px, py, bx, by, cx, cy);
  const T denom = T(1) / (va + vb + vc);
  const T v = vb * denom;
  const T w = vc * denom;
  const T qx = ax + abx * v + acx * w - px;
  const T qy = ay + aby * v + acy * w - py;
  return sqrt(qx * qx + qy * qy);
}

// Candidate record of a face bucket: entry q of the cell's face list as two float4,
// (ax, ay, bx, by) and (cx, cy, face id bits, 0), so one load level yields a candidate's corners.
inline void pack_cell_triangle(const float* uv, const int32_t* faces, int32_t f, float4* out) {
  const int32_t a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
  float fbits;
  static_assert(sizeof(fbits) == sizeof(f), "face id bits");
  __builtin_memcpy(&fbits, &f, sizeof f);
  out[0] = make_float4(uv[2 * a], uv[2 * a + 1], uv[2 * b], uv[2 * b + 1]);
  out[1] = make_float4(uv[2 * c], uv[2 * c + 1], fbits, 0.0f);
}

// Face record read once the face is chosen: corner ids a, b, c (faces[3f ..]) with their UV and
// 3-D coordinates, (ua, ub), (uc, a, b), (A, c), (B, 0), (C, 0), ids as float bits; one load level
// instead of the face's ids and then its corners.
inline void pack_face(const float* uv, const float* xyz, const int32_t* faces, int32_t f, float4* out) {
  const int32_t a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
  float ab, bb, cb;
  __builtin_memcpy(&ab, &a, sizeof a);
  __builtin_memcpy(&bb, &b, sizeof b);
  __builtin_memcpy(&cb, &c, sizeof c);
  out[0] = make_float4(uv[2 * a], uv[2 * a + 1], uv[2 * b], uv[2 * b + 1]);
  out[1] = make_float4(uv[2 * c], uv[2 * c + 1], ab, bb);
  out[2] = make_float4(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2], cb);
  out[3] = make_float4(xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2], 0.0f);
  out[4] = make_float4(xyz[3 * c], xyz[3 * c + 1], xyz[3 * c + 2], 0.0f);
}

__host__ __device__ inline int float_bits(float x) {
  int b;
  __builtin_memcpy(&b, &x, sizeof b);
  return b;
}

__host__ __device__ inline float approx_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);  // v_rcp_f32, about 1 ulp: enough for the FP32 ranking
#else
  return 1.0f / x;
#endif
}

// Squared distance from P to triangle ABC (z = 0) for the FP32 ranking: the closest point is
// chosen by the tests of point_triangle_distance in the same order (vertex A, B, C, edge AB, AC,
// BC, interior), evaluated side by side and combined with selects, so lanes whose candidates fall
// in different regions do not diverge. A vertex region is the segment from that vertex with
// t = 0. Same mathematics as point_triangle_distance; reciprocals are approximate (the ranking
// is only kept when the best two distances differ by kLocateMargin).
__host__ __device__ inline float point_triangle_dist2(float px, float py, float ax, float ay, float bx, float by,
                                                      float cx, float cy) {
  const float abx = bx - ax, aby = by - ay, acx = cx - ax, acy = cy - ay;
  const float apx = px - ax, apy = py - ay, bpx = px - bx, bpy = py - by, cpx = px - cx, cpy = py - cy;
  const float d_ab_ap = abx * apx + aby * apy, d_ac_ap = acx * apx + acy * apy;
  const float d_ab_bp = abx * bpx + aby * bpy, d_ac_bp = acx * bpx + acy * bpy;
  const float d_ab_cp = abx * cpx + aby * cpy, d_ac_cp = acx * cpx + acy * cpy;
  const float vc = d_ab_ap * d_ac_bp - d_ab_bp * d_ac_ap;
  const float vb = d_ab_cp * d_ac_ap - d_ab_ap * d_ac_cp;
  const float va = d_ab_bp * d_ac_cp - d_ab_cp * d_ac_bp;
  const int in_a = (int)(d_ab_ap <= 0.0f) & (int)(d_ac_ap <= 0.0f);
  const int in_b = (int)(d_ab_bp >= 0.0f) & (int)(d_ac_bp <= d_ab_bp);
  const int in_c = (int)(d_ac_cp >= 0.0f) & (int)(d_ab_cp <= d_ac_cp);
  const int on_ab = (int)(vc <= 0.0f) & (int)(d_ab_ap >= 0.0f) & (int)(d_ab_bp <= 0.0f);
  const int on_ac = (int)(vb <= 0.0f) & (int)(d_ac_ap >= 0.0f) & (int)(d_ac_cp <= 0.0f);
  const int on_bc = (int)(va <= 0.0f) & (int)((d_ac_bp - d_ab_bp) >= 0.0f) & (int)((d_ab_cp - d_ac_cp) >= 0.0f);
  const int region = in_a ? 0 : (in_b ? 1 : (in_c ? 2 : (on_ab ? 3 : (on_ac ? 4 : (on_bc ? 5 : 6)))));
  // segment S0 -> S1 of the vertex / edge regions
  const bool from_a = region == 0 || region == 3 || region == 4, from_b = region == 1 || region == 5;
  const float s0x = from_a ? ax : (from_b ? bx : cx), s0y = from_a ? ay : (from_b ? by : cy);
  const float s1x = region == 3 ? bx : cx, s1y = region == 3 ? by : cy;
  const bool vertex = region <= 2;
  const float ex = s1x - s0x, ey = s1y - s0y;
  const float num = vertex ? 0.0f : ex * (px - s0x) + ey * (py - s0y);
  const float den = vertex ? 1.0f : ex * ex + ey * ey;
  float t = num * approx_rcp(den);
  t = t < 0.0f ? 0.0f : (1.0f < t ? 1.0f : t);
  float qx = s0x + ex * t - px, qy = s0y + ey * t - py;
  // interior: barycentric closest point
  const float denom = approx_rcp(va + vb + vc);
  const float v = vb * denom, w = vc * denom;
  if (region == 6) {
    qx = ax + abx * v + acx * w - px;
    qy = ay + aby * v + acy * w - py;
  }
  return qx * qx + qy * qy;
}

struct Located {
  int face, vid;
  double X, Y, Z;
  int fp64;  // 1: the face ranking needed FP64, 2: the vertex choice needed FP64 (bit mask)
};

__host__ __device__ inline bool lex_less(double d, int f, double bd, int bf) { return d < bd || (d == bd && f < bf); }

// combine the (best, runner-up) distances of two lanes
__host__ __device__ inline void merge_best(float& b1, int& f1, float& b2, float o1, int of1, float o2) {
  if (o1 < b1 || (o1 == b1 && of1 < f1)) {
    b2 = fminf(b1, o2);
    b1 = o1;
    f1 = of1;
  } else {
    b2 = fminf(b2, o1);
  }
}

template <int LANES, class T>
__host__ __device__ inline T lane_xor(T v, int off) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __shfl_xor(v, off, LANES);
#else
  (void)off;
  return v;
#endif
}

// CellHelper::calculate_barycentric_3D_coord (CellHelper.cpp:104-160) with the candidate faces
// taken from the face bucket of the point's grid cell. Faces containing the point (distance 0
// up to rounding) always share its cell, so the lexicographic (distance, face) minimum over the
// bucket is the minimum over all faces. LANES lanes split the bucket and reduce the minimum.
template <int LANES>
__host__ __device__ inline Located locate(const MeshDev& m, double px, double py, int lane) {
  const int G = m.grid;
  const double gx = px * G, gy = py * G;
  const int cx = gx >= (double)G ? G - 1 : (gx >= 1.0 ? (int)gx : 0);
  const int cy = gy >= (double)G ? G - 1 : (gy >= 1.0 ? (int)gy : 0);
  const int cell = cy * G + cx;
  const int cs = m.cell_start[cell], ce = m.cell_start[cell + 1];
  Located r;
  r.fp64 = 0;
  // FP32 ranking: best and runner-up distance
  const float qx = (float)px, qy = (float)py;
  float b1 = __builtin_inff(), b2 = __builtin_inff();  // squared distances
  int f1 = 0x7fffffff;
  // candidates in chunks of kFaceChunk per lane, their records loaded together (same visiting
  // order as one at a time)
  for (int q0 = cs + lane; q0 < ce; q0 += kFaceChunk * LANES) {
    float4 e0[kFaceChunk], e1[kFaceChunk];
#pragma unroll
    for (int k = 0; k < kFaceChunk; ++k) {
      const int q = q0 + k * LANES;
      if (q < ce) { e0[k] = m.cell_tri[2 * q]; e1[k] = m.cell_tri[2 * q + 1]; }
    }
#pragma unroll
    for (int k = 0; k < kFaceChunk; ++k)
      if (q0 + k * LANES < ce) {
        const float d2 = point_triangle_dist2(qx, qy, e0[k].x, e0[k].y, e0[k].z, e0[k].w, e1[k].x, e1[k].y);
        merge_best(b1, f1, b2, d2, float_bits(e1[k].z), __builtin_inff());
      }
  }
  if constexpr (LANES > 1) {
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) {
      const float o1 = lane_xor<LANES>(b1, off), o2 = lane_xor<LANES>(b2, off);
      const int of1 = lane_xor<LANES>(f1, off);
      merge_best(b1, f1, b2, o1, of1, o2);
    }
  }
  int best_f = f1;
  if (!(sqrtf(b2) - sqrtf(b1) > kLocateMargin)) {  // close call (or non-finite input): rank in FP64
    r.fp64 |= 1;
    double bd = __builtin_inf();
    int bf = 0x7fffffff;
    for (int q = cs + lane; q < ce; q += LANES) {
      const float4 e0 = m.cell_tri[2 * q], e1 = m.cell_tri[2 * q + 1];
      const int f = float_bits(e1.z);
      const double d = point_triangle_distance<double>(px, py, e0.x, e0.y, e0.z, e0.w, e1.x, e1.y);
      if (lex_less(d, f, bd, bf)) { bd = d; bf = f; }
    }
    if constexpr (LANES > 1) {
#pragma unroll
      for (int off = LANES / 2; off > 0; off >>= 1) {
        const double od = lane_xor<LANES>(bd, off);
        const int of = lane_xor<LANES>(bf, off);
        if (lex_less(od, of, bd, bf)) { bd = od; bf = of; }
      }
    }
    best_f = bf;
  }
  if (best_f == 0x7fffffff) best_f = 0;  // non-finite point: any valid face keeps indices in range
  const float4* fp = m.face_pack + 5 * (size_t)best_f;
  const float4 p0 = fp[0], p1 = fp[1], p2 = fp[2], p3 = fp[3], p4 = fp[4];
  const int ia = float_bits(p1.z), ib = float_bits(p1.w), ic = float_bits(p2.w);
  const float2 ua = make_float2(p0.x, p0.y), ub = make_float2(p0.z, p0.w), uc = make_float2(p1.x, p1.y);
  const float Ax = p2.x, Ay = p2.y, Az = p2.z;
  const float Bx = p3.x, By = p3.y, Bz = p3.z;
  const float Cx = p4.x, Cy = p4.y, Cz = p4.z;
  // FP32 blend and nearest-vertex choice, kept when the choice is clear
  {
    float wa = sqrtf(sqr(qx - ua.x) + sqr(qy - ua.y));
    float wb = sqrtf(sqr(qx - ub.x) + sqr(qy - ub.y));
    float wc = sqrtf(sqr(qx - uc.x) + sqr(qy - uc.y));
    const float sw = wa + wb + wc;
    wa /= sw; wb /= sw; wc /= sw;
    const float X = wa * Ax + wb * Bx + wc * Cx, Y = wa * Ay + wb * By + wc * Cy, Z = wa * Az + wb * Bz + wc * Cz;
    const float da = sqrtf(sqr(X - Ax) + sqr(Y - Ay) + sqr(Z - Az));
    const float db = sqrtf(sqr(X - Bx) + sqr(Y - By) + sqr(Z - Bz));
    const float dc = sqrtf(sqr(X - Cx) + sqr(Y - Cy) + sqr(Z - Cz));
    // gap between the smallest distance and the other two
    const float mn = fminf(da, fminf(db, dc));
    const float gap = (da == mn ? fminf(db, dc) : (db == mn ? fminf(da, dc) : fminf(da, db))) - mn;
    if (gap > kVertexMargin && (da == mn) + (db == mn) + (dc == mn) == 1) {
      r.X = X; r.Y = Y; r.Z = Z;
      r.vid = da == mn ? ia : (db == mn ? ib : ic);
      r.face = best_f;
      return r;
    }
  }
  r.fp64 |= 2;
  const double Ax64 = Ax, Ay64 = Ay, Az64 = Az, Bx64 = Bx, By64 = By, Bz64 = Bz, Cx64 = Cx, Cy64 = Cy, Cz64 = Cz;
  double w_a = sqrt(sqr(px - (double)ua.x) + sqr(py - (double)ua.y));
  double w_b = sqrt(sqr(px - (double)ub.x) + sqr(py - (double)ub.y));
  double w_c = sqrt(sqr(px - (double)uc.x) + sqr(py - (double)uc.y));
  const double sum_weights = w_a + w_b + w_c;  // CellHelper::normalize_weights (:224-230)
  w_a /= sum_weights;
  w_b /= sum_weights;
  w_c /= sum_weights;
  r.X = w_a * Ax64 + w_b * Bx64 + w_c * Cx64;
  r.Y = w_a * Ay64 + w_b * By64 + w_c * Cy64;
  r.Z = w_a * Az64 + w_b * Bz64 + w_c * Cz64;
  const double da = sqrt(sqr(r.X - Ax64) + sqr(r.Y - Ay64) + sqr(r.Z - Az64));
  const double db = sqrt(sqr(r.X - Bx64) + sqr(r.Y - By64) + sqr(r.Z - Bz64));
  const double dc = sqrt(sqr(r.X - Cx64) + sqr(r.Y - Cy64) + sqr(r.Z - Cz64));
  double mn = da;
  if (db < mn) mn = db;
  if (dc < mn) mn = dc;
  r.vid = (mn == da) ? ia : ((mn == db) ? ib : ic);
  r.face = best_f;
  return r;
}

}  // namespace tdev
