// Device-side building blocks of the per-step update (gfx950, wave64). Shared by the per-step
// kernel, the batched persistent kernel and the adjoint.
//
// Precision plan: state and pair arithmetic in FP32; the decisive or ill-conditioned parts in
// FP64 (seam-wrap geometry, point location, nearest vertex, order-parameter sign, and the
// alignment mean when the neighbour sum nearly cancels). All code is compiled with
// -ffp-contract=off, so every expression rounds as written; the FP64 point-location code uses
// the same operation order as the host oracle and therefore makes the same decisions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tdev {

constexpr int kSlots = 8;                       // particles held per vertex bucket before overflow
constexpr double kPi = 3.141592653589793;
constexpr float kPiF = 3.14159274f;             // pi rounded to float
constexpr float kTwoPiF = 6.28318548f;

struct MeshDev {
  int32_t nv, nf, grid, pad;
  const float2* uv;        // [nv]
  const float* xyz;        // [nv*3]
  const int32_t* faces;    // [nf*3]
  const int32_t* cell_start;
  const int32_t* cell_faces;
  const int32_t* row_ptr;  // [nv+1]
  const int2* nbr;         // (vertex, float bits of the geodesic distance), rows ascending in distance
};

// per-simulation constants (FP32 values of the FP64 parameters; thresholds reproduce the FP64
// comparisons d < 2 sigma and d <= 2.4 sigma exactly for FP32 d)
struct StepParams {
  float t_near;    // largest float f with (double)f < 2 sigma
  float t_count;   // largest float f with (double)f <= 2.4 sigma
  float two_sigma, k, v0, mu, dt, rate;  // rate = dt / tau
  int32_t max_crossings, pad;
  double two_sigma64, k64, v064, mu64, dt64, rate64;  // the same parameters for the FP64 paths
};

// displacement (UV units) above which the force, speed and Euler update are redone in FP64:
// an FP32 speed carries ~1e-7 relative error, which a displacement D turns into ~1e-7 D absolute
constexpr float kBigStep = 0.25f;

// vertex buckets of one step: count[v], the first kSlots ids in slot[v*kSlots..], the rest as
// (vertex, id) pairs in the overflow list
struct BucketsIn { const int* count; const int* slot; const int* ovf_n; const int2* ovf; };
struct BucketsOut { int* count; int* slot; int* ovf_n; int2* ovf; };

// ------------------------------------------------------------------ neighbour traversal

// visit the particles of bucket w in ascending id order (fixed order => reproducible sums)
template <class B, class F>
__device__ __forceinline__ void for_bucket(const B& b, int w, F&& f) {
  const int c = b.count[w];
  if (c == 0) return;
  const int* s = b.slot + (size_t)w * kSlots;
  if (c == 1) { f(s[0]); return; }
  if (c == 2) {
    const int a = s[0], d = s[1];
    if (a < d) { f(a); f(d); } else { f(d); f(a); }
    return;
  }
  const int ns = c < kSlots ? c : kSlots;
  const int novf = c > kSlots ? *b.ovf_n : 0;
  int prev = -1;
  for (int r = 0; r < c; ++r) {
    int best = 0x7fffffff;
    for (int q = 0; q < ns; ++q) {
      const int id = s[q];
      if (id > prev && id < best) best = id;
    }
    for (int q = 0; q < novf; ++q) {
      const int2 e = b.ovf[q];
      if (e.x == w && e.y > prev && e.y < best) best = e.y;
    }
    f(best);
    prev = best;
  }
}

// force coefficient Fij / d for a pair at table distance d < 2 sigma
// (ForceHelper.cpp:55-70, 84-104: d == 0 becomes 0.001; the adhesion branch is unreachable)
__device__ __forceinline__ float pair_coef(float d, const StepParams& p) {
  const float dd = (d == 0.0f) ? 0.001f : d;
  const float fij = (-p.k * (p.two_sigma - dd)) / p.two_sigma;
  return fij / dd;
}

template <int LANES, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, LANES);
  return v;
}

// ------------------------------------------------------------------ seam wrap (FP64)

__device__ __forceinline__ bool inside_unit(double x, double y) { return x >= 0.0 && x <= 1.0 && y >= 0.0 && y <= 1.0; }

// EuclideanTiling::is_point_on_segment (Locomotion/EuclideanTiling.cpp:156-168)
__device__ __forceinline__ bool on_segment(double px, double py, double ax, double ay, double bx, double by) {
  if (px < fmin(ax, bx) || px > fmax(ax, bx) || py < fmin(ay, by) || py > fmax(ay, by)) return false;
  const double cr = (px - ax) * (by - ay) - (py - ay) * (bx - ax);
  return fabs(cr) < 1e-9;
}

// EuclideanTiling::intersection_point for one straight border segment C-D (:170-208)
__device__ __forceinline__ bool seg_intersect(double ax, double ay, double bx, double by, double cx, double cy,
                                              double dx, double dy, double& ox, double& oy) {
  if (on_segment(ax, ay, cx, cy, dx, dy)) return false;
  const double det = (bx - ax) * (dy - cy) - (by - ay) * (dx - cx);
  if (fabs(det) < 1e-9) return false;
  const double t = ((cx - ax) * (dy - cy) - (cy - ay) * (dx - cx)) / det;
  const double s = ((cx - ax) * (by - ay) - (cy - ay) * (bx - ax)) / det;
  if (t >= 0 && t <= 1 && s >= 0 && s <= 1) {
    ox = ax + t * (bx - ax);
    oy = ay + t * (by - ay);
    return true;
  }
  return false;
}

// Seam wrap of one particle: the reference's restart loop (EuclideanTiling.cpp:31-69) acts on
// each particle independently, so it runs here as a per-particle loop. (ax, ay) -> (px, py) is the
// step's segment; on return (px, py) is inside the unit square (or the crossing bound was hit)
// and dtheta holds the heading rotation (-pi/2 per left/right, -3 pi/2 per up/down crossing).
// Returns false when more than max_crossings crossings would be needed.
__device__ inline bool seam_wrap(double ax, double ay, double& px, double& py, double& dtheta, int max_crossings) {
  int crossings = 0;
  dtheta = 0.0;
  while (!inside_unit(px, py)) {
    if (crossings >= max_crossings) return false;
    ++crossings;
    // check_border_crossings (:71-101): left, right, up, down; exit snaps to 0 below 1e-3
    double ex = ax, ey = ay;
    int border = 4;
    if (seg_intersect(ax, ay, px, py, 0, 0, 0, 1, ex, ey)) border = 0;
    else if (seg_intersect(ax, ay, px, py, 1, 0, 1, 1, ex, ey)) border = 1;
    else if (seg_intersect(ax, ay, px, py, 0, 1, 1, 1, ex, ey)) border = 2;
    else if (seg_intersect(ax, ay, px, py, 0, 0, 1, 0, ex, ey)) border = 3;
    if (border < 4) {
      if (fabs(ex) < 1e-3) ex = 0.0;
      if (fabs(ey) < 1e-3) ey = 0.0;
    }
    // processPoints (:112-154)
    const double ox = px, oy = py;
    if (border == 0) { px = oy; py = -ox; dtheta -= 0.5 * kPi; }
    else if (border == 1) { px = oy; py = 2 - ox; dtheta -= 0.5 * kPi; }
    else if (border == 2) { px = 2 - oy; py = ox; dtheta -= 1.5 * kPi; }
    else { px = -oy; py = ox; dtheta -= 1.5 * kPi; }
    ax = ey;  // entry point = swapped exit point
    ay = ex;
  }
  return true;
}

// ------------------------------------------------------------------ point location (FP64)

__device__ __forceinline__ double sqd(double a) { return a * a; }

// CellHelper::pointSegmentDistance (CellHelper.cpp:216-222), z = 0
__device__ __forceinline__ double point_segment_distance(double px, double py, double ax, double ay, double bx,
                                                         double by) {
  const double abx = bx - ax, aby = by - ay;
  double t = (abx * (px - ax) + aby * (py - ay)) / (abx * abx + aby * aby);
  t = t < 0.0 ? 0.0 : (1.0 < t ? 1.0 : t);
  const double qx = ax + abx * t - px, qy = ay + aby * t - py;
  return sqrt(qx * qx + qy * qy);
}

// CellHelper::pointTriangleDistance (CellHelper.cpp:162-214), z = 0
__device__ inline double point_triangle_distance(double px, double py, double ax, double ay, double bx, double by,
                                                 double cx, double cy) {
  const double abx = bx - ax, aby = by - ay;
  const double acx = cx - ax, acy = cy - ay;
  const double apx = px - ax, apy = py - ay;
  const double bpx = px - bx, bpy = py - by;
  const double cpx = px - cx, cpy = py - cy;
  const double d_ab_ap = abx * apx + aby * apy;
  const double d_ac_ap = acx * apx + acy * apy;
  const double d_ab_bp = abx * bpx + aby * bpy;
  const double d_ac_bp = acx * bpx + acy * bpy;
  const double d_ab_cp = abx * cpx + aby * cpy;
  const double d_ac_cp = acx * cpx + acy * cpy;
  if (d_ab_ap <= 0.0 && d_ac_ap <= 0.0) return sqrt(apx * apx + apy * apy);
  if (d_ab_bp >= 0.0 && d_ac_bp <= d_ab_bp) return sqrt(bpx * bpx + bpy * bpy);
  if (d_ac_cp >= 0.0 && d_ab_cp <= d_ac_cp) return sqrt(cpx * cpx + cpy * cpy);
  const double vc = d_ab_ap * d_ac_bp - d_ab_bp * d_ac_ap;
  if (vc <= 0.0 && d_ab_ap >= 0.0 && d_ab_bp <= 0.0) return point_segment_distance(px, py, ax, ay, bx, by);
  const double vb = d_ab_cp * d_ac_ap - d_ab_ap * d_ac_cp;
  if (vb <= 0.0 && d_ac_ap >= 0.0 && d_ac_cp <= 0.0) return point_segment_distance(px, py, ax, ay, cx, cy);
  const double va = d_ab_bp * d_ac_cp - d_ab_cp * d_ac_bp;
  if (va <= 0.0 && (d_ac_bp - d_ab_bp) >= 0.0 && (d_ab_cp - d_ac_cp) >= 0.0)
    return point_segment_distance(px, py, bx, by, cx, cy);
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom;
  const double w = vc * denom;
  const double qx = ax + abx * v + acx * w - px;
  const double qy = ay + aby * v + acy * w - py;
  return sqrt(qx * qx + qy * qy);
}

struct Located {
  int face, vid;
  double X, Y, Z;
};

// CellHelper::calculate_barycentric_3D_coord (CellHelper.cpp:104-160) with the candidate faces
// taken from the face bucket of the point's grid cell. Faces containing the point (distance 0
// up to rounding) always share its cell, so the lexicographic (distance, face) minimum over the
// bucket is the minimum over all faces. LANES lanes split the bucket and reduce the minimum.
template <int LANES>
__device__ inline Located locate(const MeshDev& m, double px, double py, int lane) {
  const int G = m.grid;
  const double fx = px * G, fy = py * G;
  const int cx = fx >= (double)G ? G - 1 : (fx >= 1.0 ? (int)fx : 0);
  const int cy = fy >= (double)G ? G - 1 : (fy >= 1.0 ? (int)fy : 0);
  const int cell = cy * G + cx;
  const int cs = m.cell_start[cell], ce = m.cell_start[cell + 1];
  double best_d = __builtin_inf();
  int best_f = 0x7fffffff;
  for (int q = cs + lane; q < ce; q += LANES) {
    const int f = m.cell_faces[q];
    const int ia = m.faces[3 * f], ib = m.faces[3 * f + 1], ic = m.faces[3 * f + 2];
    const float2 a = m.uv[ia], b = m.uv[ib], c = m.uv[ic];
    const double d = point_triangle_distance(px, py, a.x, a.y, b.x, b.y, c.x, c.y);
    if (d < best_d || (d == best_d && f < best_f)) { best_d = d; best_f = f; }
  }
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) {
    const double od = __shfl_xor(best_d, off, LANES);
    const int of = __shfl_xor(best_f, off, LANES);
    if (od < best_d || (od == best_d && of < best_f)) { best_d = od; best_f = of; }
  }
  if (best_f == 0x7fffffff) best_f = 0;  // non-finite point: any valid face keeps indices in range
  const int ia = m.faces[3 * best_f], ib = m.faces[3 * best_f + 1], ic = m.faces[3 * best_f + 2];
  const float2 ua = m.uv[ia], ub = m.uv[ib], uc = m.uv[ic];
  const double Ax = m.xyz[3 * ia], Ay = m.xyz[3 * ia + 1], Az = m.xyz[3 * ia + 2];
  const double Bx = m.xyz[3 * ib], By = m.xyz[3 * ib + 1], Bz = m.xyz[3 * ib + 2];
  const double Cx = m.xyz[3 * ic], Cy = m.xyz[3 * ic + 1], Cz = m.xyz[3 * ic + 2];
  double w_a = sqrt(sqd(px - (double)ua.x) + sqd(py - (double)ua.y));
  double w_b = sqrt(sqd(px - (double)ub.x) + sqd(py - (double)ub.y));
  double w_c = sqrt(sqd(px - (double)uc.x) + sqd(py - (double)uc.y));
  const double sum_weights = w_a + w_b + w_c;  // CellHelper::normalize_weights (:224-230)
  w_a /= sum_weights;
  w_b /= sum_weights;
  w_c /= sum_weights;
  Located r;
  r.X = w_a * Ax + w_b * Bx + w_c * Cx;
  r.Y = w_a * Ay + w_b * By + w_c * Cy;
  r.Z = w_a * Az + w_b * Bz + w_c * Cz;
  const double da = sqrt(sqd(r.X - Ax) + sqd(r.Y - Ay) + sqd(r.Z - Az));
  const double db = sqrt(sqd(r.X - Bx) + sqd(r.Y - By) + sqd(r.Z - Bz));
  const double dc = sqrt(sqd(r.X - Cx) + sqd(r.Y - Cy) + sqd(r.Z - Cz));
  double mn = da;
  if (db < mn) mn = db;
  if (dc < mn) mn = dc;
  r.vid = (mn == da) ? ia : ((mn == db) ? ib : ic);
  r.face = best_f;
  return r;
}

// heading kept in [-pi, pi)
__device__ __forceinline__ double wrap_heading(double t) {
  return t - 2.0 * kPi * floor((t + kPi) / (2.0 * kPi));
}

constexpr unsigned kFlagLost = 1u, kFlagNonfinite = 2u, kFlagTiling = 4u;

// result of one particle's step
struct StepOut {
  float xo, yo, to, co, so;  // new position, heading and its cos / sin
  float vx, vy;              // r_dot (unrotated, as the reference outputs it)
  Located L;                 // remap of (xo, yo)
  int nb;                    // neighbour count
  unsigned fl;               // validation flags
  int sgn;                   // sign of x vy - y vx (order parameter)
};

// Global-memory neighbour source: the previous step's SoA state and vertex buckets.
struct GlobalSource {
  BucketsIn b;
  const float4* xycs;
  const float* th;
  __device__ __forceinline__ int count(int w) const { return b.count[w]; }
  template <class F>
  __device__ __forceinline__ void each(int w, F&& f) const { for_bucket(b, w, f); }
  __device__ __forceinline__ float4 pos(int j) const { return xycs[j]; }
  __device__ __forceinline__ float theta(int j) const { return th[j]; }
};

// One particle's update (U1-U6 of the reference step) by a group of LANES lanes.
//   me = (x, y, cos theta, sin theta) of particle i, th its heading, v its vertex id.
// Neighbours: particles bucketed on the vertices w of row v of the geodesic table; pair
// distance d = D[v, w]. Force and alignment for d < 2 sigma (ForceHelper.cpp:34-73,
// OrientationHelper.cpp:29-74, self included in the mean), count for 0 < d <= 2.4 sigma
// (2DTissue.cpp:254-268). FP32 sums; FP64 re-evaluation when the displacement is large or the
// alignment sum nearly cancels.
template <int LANES, class Src>
__device__ inline StepOut particle_step(const MeshDev& mesh, const StepParams& prm, const Src& src, int i, float4 me,
                                        float th, int v, float eta, int lane) {
  float Fx = 0.f, Fy = 0.f, Sx = 0.f, Sy = 0.f;
  int na = 0, nb = 0;
  const int rs = mesh.row_ptr[v], re = mesh.row_ptr[v + 1];
  for (int e = rs + lane; e < re; e += LANES) {
    const int2 ent = mesh.nbr[e];
    const float d = __int_as_float(ent.y);
    if (d > prm.t_count) break;  // rows ascend in distance
    if (d > prm.t_near) {        // counted only (2 sigma <= d <= 2.4 sigma)
      nb += src.count(ent.x);
      continue;
    }
    const float g = pair_coef(d, prm);
    const bool counted = d != 0.0f;
    src.each(ent.x, [&](int j) {
      const float4 q = src.pos(j);
      if (j != i) {
        Fx += g * (me.x - q.x);
        Fy += g * (me.y - q.y);
      }
      Sx += q.z;
      Sy += q.w;
      ++na;
      if (counted) ++nb;
    });
  }
  Fx = group_sum<LANES>(Fx);
  Fy = group_sum<LANES>(Fy);
  Sx = group_sum<LANES>(Sx);
  Sy = group_sum<LANES>(Sy);
  na = group_sum<LANES>(na);
  nb = group_sum<LANES>(nb);

  // speed, velocity, Euler step (Locomotion.cpp:71-84)
  const float Fm = sqrtf(Fx * Fx + Fy * Fy);
  const float spd = prm.v0 + prm.mu * Fm;
  float vx = me.z * spd, vy = me.w * spd;
  double xn = (double)(me.x + vx * prm.dt), yn = (double)(me.y + vy * prm.dt);

  // The FP32 mean heading has error ~ na * 6e-8 / |S|, scaled by dt/tau in the update; the FP32
  // displacement has error ~ 1e-7 * |D|. Either beyond ~1e-7 => redo that part in FP64.
  const float Sm = sqrtf(Sx * Sx + Sy * Sy);
  const bool weak = !(Sm >= 0.6f * prm.rate * (float)na);
  const bool big = !(spd * prm.dt <= kBigStep);
  double thn;
  if (weak || big) {
    double Fx64 = 0.0, Fy64 = 0.0, Sx64 = 0.0, Sy64 = 0.0;
    for (int e = rs + lane; e < re; e += LANES) {
      const int2 ent = mesh.nbr[e];
      const float d = __int_as_float(ent.y);
      if (d > prm.t_near) break;
      const double dd = d == 0.0f ? 0.001 : (double)d;
      const double fij = (-prm.k64 * (prm.two_sigma64 - dd)) / prm.two_sigma64;
      src.each(ent.x, [&](int j) {
        if (big && j != i) {
          const float4 q = src.pos(j);
          Fx64 += fij * (((double)me.x - (double)q.x) / dd);
          Fy64 += fij * (((double)me.y - (double)q.y) / dd);
        }
        if (weak) {
          double sj, cj;
          sincos((double)src.theta(j), &sj, &cj);
          Sx64 += cj;
          Sy64 += sj;
        }
      });
    }
    if (big) {
      Fx64 = group_sum<LANES>(Fx64);
      Fy64 = group_sum<LANES>(Fy64);
      double st, ct;
      sincos((double)th, &st, &ct);
      const double spd64 = prm.v064 + prm.mu64 * sqrt(Fx64 * Fx64 + Fy64 * Fy64);
      const double vx64 = ct * spd64, vy64 = st * spd64;
      xn = (double)me.x + vx64 * prm.dt64;
      yn = (double)me.y + vy64 * prm.dt64;
      vx = (float)vx64;
      vy = (float)vy64;
    }
    if (weak) {
      Sx64 = group_sum<LANES>(Sx64);
      Sy64 = group_sum<LANES>(Sy64);
      const double mean = atan2(Sy64, Sx64);
      double delta = mean - (double)th;
      if (delta > kPi) delta -= 2.0 * kPi;
      else if (delta <= -kPi) delta += 2.0 * kPi;
      thn = (double)th + prm.rate64 * delta + (double)eta;
    }
  }
  if (!weak) {
    const float mean = atan2f(Sy, Sx);
    float delta = mean - th;
    if (delta > kPiF) delta -= kTwoPiF;
    else if (delta <= -kPiF) delta += kTwoPiF;
    thn = (double)(th + prm.rate * delta + eta);
  }

  // seam wrap (EuclideanTiling.cpp:31-69) and validation (Validation.cpp:40-72)
  StepOut o;
  o.fl = 0;
  double px = xn, py = yn, dth = 0.0;
  const bool finite = isfinite(xn) && isfinite(yn) && isfinite(thn);
  if (!finite) o.fl |= kFlagNonfinite;
  else if (!seam_wrap(me.x, me.y, px, py, dth, prm.max_crossings)) o.fl |= kFlagTiling;
  if (finite && !inside_unit(px, py)) o.fl |= kFlagLost;
  if (finite) thn = wrap_heading(thn + dth);
  o.xo = (float)px;
  o.yo = (float)py;
  o.to = (float)thn;
  o.vx = vx;
  o.vy = vy;
  o.nb = nb;
  // UV -> 3D remap of the stored FP32 position (CellHelper.cpp:71-85)
  o.L = locate<LANES>(mesh, (double)o.xo, (double)o.yo, lane);
  sincosf(o.to, &o.so, &o.co);
  // order parameter (LinearAlgebra.cpp:90-105, z = 0): sign of the cross product's z component
  const double cz = (double)o.xo * (double)vy - (double)o.yo * (double)vx;
  o.sgn = cz > 0.0 ? 1 : (cz < 0.0 ? -1 : 0);
  return o;
}

}  // namespace tdev
