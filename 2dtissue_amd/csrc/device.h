// Device-side building blocks of the per-step update (gfx950, wave64). Shared by the per-step
// kernel, the batched persistent kernel and the adjoint.
//
// Precision plan: state and pair arithmetic in FP32; the decisive or ill-conditioned parts in
// FP64 (seam-wrap geometry, point location, nearest vertex, order-parameter sign, and the
// alignment mean when the neighbour sum nearly cancels). All code is compiled with
// -ffp-contract=off, so every expression rounds as written. Point location ranks candidate faces
// in FP32 and falls back to FP64 (the host oracle's operation order, so the same decisions)
// whenever the FP32 ranking or the nearest-vertex choice is within its error margin.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "geom.h"

namespace tdev {

constexpr int kSlots = 8;                       // particles held per vertex bucket before overflow
constexpr float kPiF = 3.14159274f;             // pi rounded to float
constexpr float kTwoPiF = 6.28318548f;

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
constexpr float kBigStep = 2.0f;

// Vertex buckets of one step: count[v] particles. Sorted (after k_bucket_finalize): up to kSlots
// of them in slot[v*kSlots ..] in ascending id order; a bucket with more holds its ids, ascending,
// in big[slot[v*kSlots] ..] instead. Raw (no finalize pass): the first kSlots arrivals in
// slot[v*kSlots ..] in arrival order, later ones in the (vertex, id) overflow list ovf[0 .. *ovf_n);
// readers order them themselves. Either way a bucket is visited in ascending id order.
struct BucketsIn {
  const int* count;
  const int* slot;
  const int* big;
  const int2* ovf;
  const int* ovf_n;
  int sorted;
};
// The same buckets while the step fills them: the first kSlots arrivals of a vertex take its
// slots, later ones go to the (vertex, id) overflow list; the second arrival at a vertex also
// lists the vertex in multi (the buckets the finalize pass must sort; null when no finalize pass
// follows); the first arrival sets the vertex's bit in the occupancy mask occ (bit v % 32 of word
// v / 32).
struct BucketsOut { int* count; int* slot; int* ovf_n; int2* ovf; int* multi_n; int* multi; unsigned* occ; };

// words of a vertex-occupancy mask over nv vertices
__host__ __device__ constexpr int occ_words(int nv) { return (nv + 31) >> 5; }

// ------------------------------------------------------------------ neighbour traversal

// visit the particles of bucket w in ascending id order (fixed order => reproducible sums)
template <class B, class F>
__device__ __forceinline__ void for_bucket(const B& b, int w, F&& f) {
  const int c = b.count[w];
  if (c == 0) return;
  const int* s = b.slot + (size_t)w * kSlots;
  if (b.sorted) {
    if (c <= kSlots) {
      for (int q = 0; q < c; ++q) f(s[q]);
      return;
    }
    const int* g = b.big + s[0];
    for (int q = 0; q < c; ++q) f(g[q]);
    return;
  }
  if (c <= kSlots) {  // raw: the slots in registers, sorted by the network (8 slots are in bounds)
    const int4 s0 = reinterpret_cast<const int4*>(s)[0], s1 = reinterpret_cast<const int4*>(s)[1];
    int a[kSlots] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
    for (int q = 0; q < kSlots; ++q) a[q] = q < c ? a[q] : 0x7fffffff;
    sort8(a);
    for (int q = 0; q < c; ++q) f(a[q]);
    return;
  }
  // raw and crowded (rare for the systems that skip the finalize pass): the next id each time is
  // the least one above the previous among the slots and the bucket's overflow entries
  const int n_ovf = *b.ovf_n;
  int prev = -1;
  for (int k = 0; k < c; ++k) {
    int best = 0x7fffffff;
    for (int q = 0; q < kSlots; ++q) {
      const int id = s[q];
      best = (id > prev && id < best) ? id : best;
    }
    for (int o = 0; o < n_ovf; ++o) {
      const int2 e = b.ovf[o];
      best = (e.x == w && e.y > prev && e.y < best) ? e.y : best;
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

constexpr unsigned kFlagLost = 1u, kFlagNonfinite = 2u, kFlagTiling = 4u;

// result of one particle's step
struct StepOut {
  float xo, yo, to, co, so;  // new position, heading and its cos / sin
  float vx, vy;              // r_dot (unrotated, as the reference outputs it)
  Located L;                 // remap of (xo, yo)
  int nb;                    // neighbour count
  unsigned fl;               // validation flags
  int sgn;                   // sign of x vy - y vx (order parameter)
  // what the adjoint replays and the visit check records (dead, and removed by the compiler, in
  // kernels that do not record them)
  int na;                    // particles visited within 2 sigma (the neighbour set, self included)
  unsigned vh;               // sum of visit_mix(j) over them (mod 2^32): the set, order-free
  float Fx, Fy;              // force sum
  float Sx, Sy;              // alignment sums (cos, sin) over the neighbour set incl. self
  int q;                     // seam wrap: counter-clockwise quarter turns of the composed map
};

// table entries per lane loaded together in the neighbour walk (scan phase, and the FP64 redo)
constexpr int kRowChunk = 4;
constexpr int kScanChunk = 8;  // measured on MI355X: 4 and 16 are slower
// lanes per particle from which the walk loads bucket heads in its scan (measured on MI355X: 1.7 %
// faster at 16 lanes, 1 % slower at 4 lanes with chunks of 4), and its chunk of entries per lane
constexpr int kHeadScanLanes = 8;

// 32-bit mixing of a particle index (the MurmurHash3 finaliser): summed over a neighbour set it
// names the set independently of the visiting order
__host__ __device__ inline unsigned visit_mix(unsigned h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
constexpr int kHeadChunk = 2;
static_assert(32 % kScanChunk == 0, "scan chunks must tile the 32-entry window");

// Global-memory neighbour source: the previous step's SoA state and vertex buckets. With OCC the
// buckets' occupancy mask is staged in LDS by the workgroup: the table walk tests occupancy there
// and touches a bucket in global memory only when its vertex holds particles; without it the walk
// reads each entry's bucket count. Both visit the same buckets in the same order.
template <bool OCC>
struct GlobalSource {
  static constexpr bool kOcc = OCC;
  BucketsIn b;
  const float4* xycs;
  const float* th;
  const unsigned* occ;  // LDS copy of the replica's occupancy mask of b (OCC only)
  __device__ __forceinline__ bool occupied(int w) const { return (occ[w >> 5] >> (w & 31)) & 1u; }
  struct Head { int c = 0, s0 = 0, s1 = 0; };
  // bucket head: count and the first two slots, loaded together (the slots are in bounds
  // whatever the count)
  __device__ __forceinline__ Head head(int w) const {
    Head h;
    h.c = b.count[w];
    h.s0 = b.slot[(size_t)w * kSlots];
    h.s1 = b.slot[(size_t)w * kSlots + 1];
    return h;
  }
  __device__ __forceinline__ int count(int w) const { return b.count[w]; }
  template <class F>
  __device__ __forceinline__ void each(int w, const Head& h, F&& f) const {
    if (h.c == 0) return;
    if (h.c == 1) { f(h.s0); return; }
    if (h.c == 2) {  // in order when sorted by the finalize pass; else ordered here
      const int lo = b.sorted ? h.s0 : min(h.s0, h.s1), hi = b.sorted ? h.s1 : max(h.s0, h.s1);
      f(lo);
      f(hi);
      return;
    }
    for_bucket(b, w, f);
  }
  __device__ __forceinline__ float4 pos(int j) const { return xycs[j]; }
  __device__ __forceinline__ float theta(int j) const { return th[j]; }
};

// LDS-resident neighbour source of the persistent kernel: vertex buckets as a CSR over the
// particles counting-sorted by vertex (begin[v] .. begin[v+1]), visited in ascending id order like
// the global buckets, so both paths add neighbours in the same order.
struct LdsSource {
  const uint32_t* begin;   // [V+1]
  const uint16_t* sorted;  // [n]
  const float4* xycs;
  const float* th;
  struct Head { int b = 0, c = 0; };
  __device__ __forceinline__ Head head(int w) const {
    Head h;
    h.b = (int)begin[w];
    h.c = (int)begin[w + 1] - h.b;
    return h;
  }
  static constexpr bool kOcc = false;  // counts are LDS reads already
  __device__ __forceinline__ int count(int w) const { return (int)begin[w + 1] - (int)begin[w]; }
  __device__ __forceinline__ bool occupied(int w) const { return begin[w + 1] != begin[w]; }
  template <class F>
  __device__ __forceinline__ void each(int, const Head& h, F&& f) const {
    const int b = h.b, c = h.c;
    if (c == 0) return;
    if (c == 1) { f((int)sorted[b]); return; }
    if (c == 2) {
      const int a = sorted[b], d = sorted[b + 1];
      if (a < d) { f(a); f(d); } else { f(d); f(a); }
      return;
    }
    int prev = -1;
    for (int r = 0; r < c; ++r) {
      int best = 0x7fffffff;
      for (int q = 0; q < c; ++q) {
        const int id = sorted[b + q];
        if (id > prev && id < best) best = id;
      }
      f(best);
      prev = best;
    }
  }
  __device__ __forceinline__ float4 pos(int j) const { return xycs[j]; }
  __device__ __forceinline__ float theta(int j) const { return th[j]; }
};

// One particle's update (U1-U6 of the reference step) by a group of LANES lanes.
//   me = (x, y, cos theta, sin theta) of particle i, th its heading, v its vertex id.
// Neighbours: particles bucketed on the vertices w of row v of the geodesic table; pair
// distance d = D[v, w]. Force and alignment for d < 2 sigma (ForceHelper.cpp:34-73,
// OrientationHelper.cpp:29-74, self included in the mean), count for 0 < d <= 2.4 sigma
// (2DTissue.cpp:254-268). The force is summed in FP64 (its FP32 rounding feeds the FP32 speed
// and Euler step; the FP64 value the large-displacement path); the alignment sums are FP32, redone
// in FP64 when they nearly cancel.
// WAVE selects the seam-wrap loop of crowded systems (seam_wrap_group). Always inlined: as an
// out-of-line call (which the inliner picks once two kernels share an instantiation) its result
// and live state go through scratch memory.
template <int LANES, bool WAVE = false, class Src>
__device__ __forceinline__ StepOut particle_step(const MeshDev& mesh, const StepParams& prm, const Src& src, int i, float4 me,
                                        float th, int v, float eta, int lane) {
  float Sx = 0.f, Sy = 0.f;
  double Fx64 = 0.0, Fy64 = 0.0;  // the force, summed in FP64 (its FP32 rounding drives the FP32 path)
  const double kc64 = prm.k64 / prm.two_sigma64;  // Fij / d = kc (d - 2 sigma) / d, to a few ulps
  int na = 0, nb = 0;
  unsigned vh = 0u;
  const int rs = mesh.row_ptr[v], re = mesh.row_ptr[v + 1];
  // pair set-up of a near entry (vertex w at distance d, bucket head h) and its particles
  auto visit = [&](int w, float d, const typename Src::Head& h) __attribute__((always_inline)) {
    const double dd = d == 0.0f ? 0.001 : (double)d;
    const double g64 = (kc64 * (dd - prm.two_sigma64)) / dd;
    const bool counted = d != 0.0f;
    src.each(w, h, [&](int j) {
      const float4 q = src.pos(j);
      if (j != i) {
        Fx64 += g64 * ((double)me.x - (double)q.x);
        Fy64 += g64 * ((double)me.y - (double)q.y);
      }
      Sx += q.z;
      Sy += q.w;
      ++na;
      vh += visit_mix((unsigned)j);
      if (counted) ++nb;
    });
  };
  // Each lane walks its entries rs + lane + m * LANES, m = 0, 1, .., in order.
  if constexpr (LANES >= kHeadScanLanes) {
    // At most a few entries per lane: the scan loads each near entry's bucket head (count and
    // first slots) together with the entries' counts, and the lane visits its occupied entries
    // from registers: table entry -> bucket head -> particles, no reload of marked entries.
    for (int base = rs + lane; base < re; base += kHeadChunk * LANES) {
      int2 ent[kHeadChunk];
      typename Src::Head hd[kHeadChunk];
#pragma unroll
      for (int k = 0; k < kHeadChunk; ++k) {
        const int e = base + k * LANES;
        ent[k] = e < re ? mesh.nbr[e] : make_int2(0, 0x7f800000);  // +inf ends the walk
      }
      bool oc[kHeadChunk];
#pragma unroll
      for (int k = 0; k < kHeadChunk; ++k)
        oc[k] = __int_as_float(ent[k].y) <= prm.t_count && (!Src::kOcc || src.occupied(ent[k].x));
#pragma unroll
      for (int k = 0; k < kHeadChunk; ++k) {
        const float d = __int_as_float(ent[k].y);
        if (oc[k] && d <= prm.t_near) hd[k] = src.head(ent[k].x);
        else if (oc[k]) hd[k].c = src.count(ent[k].x);
      }
      bool stop = false;
#pragma unroll
      for (int k = 0; k < kHeadChunk; ++k) {
        const float d = __int_as_float(ent[k].y);
        if (d > prm.t_count) {  // rows ascend in distance
          stop = true;
          break;
        }
        if (d > prm.t_near) nb += hd[k].c;  // counted only
        else if (hd[k].c > 0) visit(ent[k].x, d, hd[k]);
      }
      if (stop) break;
    }
  } else {
    // Windows of 32 entries per lane:
    //  (1) scan: the window's table entries are loaded kScanChunk at a time (one load latency per
    //      chunk) and their occupancy found (from the LDS mask, or from the bucket counts);
    //      occupied counted-only entries (2 sigma <= d <= 2.4 sigma) add their bucket count to nb,
    //      and occupied near entries are marked;
    //  (2) visit: only the marked entries, in order, set up their pair coefficient and add their
    //      particles. Most table entries have empty buckets, so the lanes of a wave loop over a
    //      few occupied entries instead of paying the pair set-up for every entry; the order, and
    //      so every sum, is that of a walk over all entries.
    bool done = false;
    for (int wbase = rs + lane; !done && wbase < re; wbase += 32 * LANES) {
      unsigned mark = 0u;
      for (int m0 = 0; m0 < 32 && !done; m0 += kScanChunk) {
        int2 ent[kScanChunk];
        int cnt[kScanChunk];
#pragma unroll
        for (int k = 0; k < kScanChunk; ++k) {
          const int e = wbase + (m0 + k) * LANES;
          ent[k] = e < re ? mesh.nbr[e] : make_int2(0, 0x7f800000);  // +inf ends the walk
        }
        bool oc[kScanChunk];
        if constexpr (Src::kOcc) {
          // occupancy from LDS; bucket counts only for the occupied entries counted without a
          // pair (2 sigma <= d <= 2.4 sigma)
#pragma unroll
          for (int k = 0; k < kScanChunk; ++k) oc[k] = __int_as_float(ent[k].y) <= prm.t_count && src.occupied(ent[k].x);
#pragma unroll
          for (int k = 0; k < kScanChunk; ++k) cnt[k] = oc[k] && __int_as_float(ent[k].y) > prm.t_near ? src.count(ent[k].x) : 0;
        } else {
#pragma unroll
          for (int k = 0; k < kScanChunk; ++k) cnt[k] = __int_as_float(ent[k].y) <= prm.t_count ? src.count(ent[k].x) : 0;
#pragma unroll
          for (int k = 0; k < kScanChunk; ++k) oc[k] = cnt[k] > 0;
        }
#pragma unroll
        for (int k = 0; k < kScanChunk; ++k) {
          const float d = __int_as_float(ent[k].y);
          if (d > prm.t_count) {  // rows ascend in distance
            done = true;
            break;
          }
          if (d > prm.t_near) nb += cnt[k];  // counted only
          else if (oc[k]) mark |= 1u << (m0 + k);
        }
      }
      while (mark) {
        const int m = __builtin_ctz(mark);
        mark &= mark - 1u;
        const int2 en = mesh.nbr[wbase + m * LANES];
        visit(en.x, __int_as_float(en.y), src.head(en.x));
      }
    }
  }
  Fx64 = group_sum<LANES>(Fx64);
  Fy64 = group_sum<LANES>(Fy64);
  float Fx = (float)Fx64, Fy = (float)Fy64;
  Sx = group_sum<LANES>(Sx);
  Sy = group_sum<LANES>(Sy);
  na = group_sum<LANES>(na);
  vh = group_sum<LANES>(vh);
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
  if (big) {  // uniform over the particle's lanes: the sums were reduced over the group
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
    // the alignment sum again, over the neighbours' headings in FP64
    double Sx64 = 0.0, Sy64 = 0.0;
    for (int base = rs + lane; base < re; base += kRowChunk * LANES) {
      int2 ent[kRowChunk];
      typename Src::Head head[kRowChunk];
#pragma unroll
      for (int k = 0; k < kRowChunk; ++k) {
        const int e = base + k * LANES;
        ent[k] = e < re ? mesh.nbr[e] : make_int2(0, 0x7f800000);
      }
#pragma unroll
      for (int k = 0; k < kRowChunk; ++k)
        if (__int_as_float(ent[k].y) <= prm.t_near && (!Src::kOcc || src.occupied(ent[k].x))) head[k] = src.head(ent[k].x);
      bool stop = false;
#pragma unroll
      for (int k = 0; k < kRowChunk; ++k) {
        const float d = __int_as_float(ent[k].y);
        if (d > prm.t_near) { stop = true; break; }
        src.each(ent[k].x, head[k], [&](int j) {
          double sj, cj;
          sincos((double)src.theta(j), &sj, &cj);
          Sx64 += cj;
          Sy64 += sj;
        });
      }
      if (stop) break;
    }
    {
      Sx64 = group_sum<LANES>(Sx64);
      Sy64 = group_sum<LANES>(Sy64);
      const double mean = atan2(Sy64, Sx64);
      double delta = mean - (double)th;
      if (delta > kPi) delta -= 2.0 * kPi;
      else if (delta <= -kPi) delta += 2.0 * kPi;
      thn = (double)th + prm.rate64 * delta + (double)eta;
      Sx = (float)Sx64;
      Sy = (float)Sy64;
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
  o.q = 0;
  o.Fx = Fx;
  o.Fy = Fy;
  o.Sx = Sx;
  o.Sy = Sy;
  double px = xn, py = yn, dth = 0.0;
  const bool finite = isfinite(xn) && isfinite(yn) && isfinite(thn);
  if (!finite) o.fl |= kFlagNonfinite;
  else if (!seam_wrap_group<LANES, WAVE>(me.x, me.y, px, py, dth, prm.max_crossings, lane, &o.q)) o.fl |= kFlagTiling;
  if (finite && !inside_unit(px, py)) o.fl |= kFlagLost;
  if (finite) thn = wrap_heading(thn + dth);
  o.xo = (float)px;
  o.yo = (float)py;
  o.to = (float)thn;
  o.vx = vx;
  o.vy = vy;
  o.nb = nb;
  o.na = na;
  o.vh = vh;
  // UV -> 3D remap of the stored FP32 position (CellHelper.cpp:71-85)
  o.L = locate<LANES>(mesh, (double)o.xo, (double)o.yo, lane);
  sincosf(o.to, &o.so, &o.co);
  // order parameter (LinearAlgebra.cpp:90-105, z = 0): sign of the cross product's z component
  const double cz = (double)o.xo * (double)vy - (double)o.yo * (double)vx;
  o.sgn = cz > 0.0 ? 1 : (cz < 0.0 ? -1 : 0);
  return o;
}

}  // namespace tdev
