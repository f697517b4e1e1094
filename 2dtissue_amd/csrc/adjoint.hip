// Reverse pass of the step (gfx950), through the forward's frozen decisions.
//
// Step t maps the state (r_i, theta_i) to
//   p_i      = r_i + dt s_i e_i,  e_i = (cos theta_i, sin theta_i),  s_i = v0 + mu |F_i|,
//   F_i      = sum_{j in A_i, j != i} c_ij (r_i - r_j),    c_ij = f(d~_ij) / d~_ij  (frozen d_ij),
//   r_i'     = R_i p_i + b_i                                (seam wrap: rotation by q_i quarter turns),
//   theta_i' = theta_i + rho wrap(atan2(S_i, C_i) - theta_i) + eta_i + const,  rho = dt / tau,
//   S_i, C_i = sum_{j in A_i} (sin, cos) theta_j            (A_i includes i),
// and adds w_v |rdot_i - u(r_i)|^2 to the loss (w_v = w_velocity / (T n)).
// Its adjoint, for incoming (lr_i, lt_i) of the new state:
//   lp = R_i^T lr;  lrdot = dt lp + 2 w_v (rdot - u);  ls = lrdot . e;  lF = mu ls F / |F|;
//   lr_i  = lp - 2 w_v J_u^T (rdot - u) + sum_{j in A_i, j != i} c_ij (lF_i - lF_j);
//   lS_i  = rho lt_i C_i / Q_i,  lC_i = -rho lt_i S_i / Q_i,  Q_i = C_i^2 + S_i^2;
//   lt_i  = (1 - rho) lt_i + s_i lrdot . (-sin, cos) + cos theta_i sum_{j in A_i} lS_j - sin theta_i sum_{j in A_i} lC_j.
// The neighbour relation is symmetric (the geodesic table is min-symmetrised), so the transposed
// sums are gathers over the particle's own neighbour set: k_adj_local computes the per-particle
// factors, k_adj_gather walks the same geodesic rows and vertex buckets as the forward step.
// The adjoint state of replica r is held as lambda * 2^-E_r (FP32, E_r >= 0 in sexp[r]): the map
// is linear in lambda, so only the loss sources are scaled by 2^-E_r and the FP64 parameter
// accumulators by 2^E_r. E_r is raised whenever |lambda| outgrows 2^20: k_adj_gather records each
// replica's max |lambda| (one atomic per wave), and the next k_adj_local scales its inputs by the
// matching power of two and carries E_r forward (E and the max alternate between two slots per
// replica by step parity, so no kernel reads a slot another thread of it writes).
#include <hip/hip_runtime.h>

#include "adjoint.h"

using namespace tdev;

namespace tissue {

namespace {

__device__ __forceinline__ void field_eval(const LossDev& L, float x, float y, float& ux, float& uy, float& jxx,
                                           float& jxy, float& jyx, float& jyy) {
  ux = uy = jxx = jxy = jyx = jyy = 0.f;
#pragma unroll
  for (int m = 0; m < kMaxModes; ++m) {
    if (m < L.n_modes) {
      float s, c;
      sincosf(L.kx[m] * x + L.ky[m] * y + L.phase[m], &s, &c);
      const float v = L.amp[m] * s, d = L.amp[m] * c;
      if (L.comp[m] == 0) { ux += v; jxx += d * L.kx[m]; jxy += d * L.ky[m]; }
      else { uy += v; jyx += d * L.kx[m]; jyy += d * L.ky[m]; }
    }
  }
}

// R_q^T v for a rotation by q counter-clockwise quarter turns
__device__ __forceinline__ float2 rot_t(int q, float2 v) {
  switch (q & 3) {
    case 1: return make_float2(v.y, -v.x);
    case 2: return make_float2(-v.x, -v.y);
    case 3: return make_float2(-v.y, v.x);
    default: return v;
  }
}

// deterministic FP64 sum over a block (wave butterflies, waves added in order)
__device__ double block_sum(double v, double* tmp) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) tmp[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s += tmp[w];
  return s;
}

// one block per replica: polar order of the final state and the seeds of the heading adjoint
__global__ void k_adj_seed(const ReplicaDesc* reps, const float4* xycs, LossDev L, const unsigned* halt, float2* lr,
                           float* lt, double* out) {
  __shared__ double tmp[kAdjBlock / 64];
  const int r = blockIdx.x;
  const int off = reps[r].off, n = reps[r].n;
  double cx = 0.0, cy = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float4 q = xycs[off + i];
    cx += (double)q.z;
    cy += (double)q.w;
  }
  const double Px = block_sum(cx, tmp);
  const double Py = block_sum(cy, tmp);
  const double norm = sqrt(Px * Px + Py * Py);
  const double phi = norm / n;
  const bool ok = halt[r] == 0u;
  // d phi / d theta_i = (-Px sin theta_i + Py cos theta_i) / (n |P|)
  const double coef = (ok && norm > 0.0) ? L.w_order * 2.0 * (phi - L.target) / (n * norm) : 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float4 q = xycs[off + i];
    lr[off + i] = make_float2(0.f, 0.f);
    lt[off + i] = (float)(coef * (-Px * (double)q.w + Py * (double)q.z));
  }
  if (threadIdx.x == 0) {
    out[2 * r] = phi;
    out[2 * r + 1] = L.w_order * (phi - L.target) * (phi - L.target);
  }
}

// power-of-two exponent that brings a replica's max |lambda| (float bits of a non-negative value)
// below 2^21; 0 when no rescale is needed or the max is not finite
__device__ __forceinline__ int rescale_exponent(unsigned mx_bits) {
  const float mx = __uint_as_float(mx_bits);
  if (!(mx > 0.f) || !isfinite(mx)) return 0;
  const int s = ilogbf(mx) - 20;
  return s > 0 ? s : 0;
}

// particle j's factors of step t's reverse, from its tape entry and its incoming adjoint (scaled
// by 2^-E, then by down = 2^-s_up): the factors its neighbours gather, its own terms of the new
// adjoint, and what the accumulators take.
struct AdjLocal {
  float2 lF, lsc;          // d loss / d F, (d loss / d sum sin, d loss / d sum cos)
  float lrx, lry, lth;     // the particle's own terms of the new adjoint
  float ls, m, ltn, wv, dvx, dvy;
};

__device__ __forceinline__ AdjLocal adj_local(const StepParams& prm, const LossDev& L, int n_r, float down, float src,
                                              float4 me, float4 f, int quarter, float2 lr_s, float lt_s) {
  AdjLocal o;
  const float2 lp = rot_t(quarter, make_float2(lr_s.x * down, lr_s.y * down));
  o.ltn = lt_s * down;
  o.m = sqrtf(f.x * f.x + f.y * f.y);
  const float s = prm.v0 + prm.mu * o.m;
  const float rdx = me.z * s, rdy = me.w * s;
  float ux, uy, jxx, jxy, jyx, jyy;
  field_eval(L, me.x, me.y, ux, uy, jxx, jxy, jyx, jyy);
  o.wv = L.w_velocity / ((float)L.n_steps * (float)n_r);
  o.dvx = rdx - ux;
  o.dvy = rdy - uy;
  const float sw = 2.f * o.wv * src;
  const float lrdx = prm.dt * lp.x + sw * o.dvx, lrdy = prm.dt * lp.y + sw * o.dvy;
  o.lrx = lp.x - sw * (jxx * o.dvx + jyx * o.dvy);
  o.lry = lp.y - sw * (jxy * o.dvx + jyy * o.dvy);
  o.ls = lrdx * me.z + lrdy * me.w;
  const float lm = prm.mu * o.ls;
  o.lF = o.m > 0.f ? make_float2(lm * f.x / o.m, lm * f.y / o.m) : make_float2(0.f, 0.f);
  const float Q = f.z * f.z + f.w * f.w;
  const float lmean = prm.rate * o.ltn;
  o.lsc = Q > 0.f ? make_float2(lmean * f.z / Q, -lmean * f.w / Q) : make_float2(0.f, 0.f);
  o.lth = (1.f - prm.rate) * o.ltn + s * (-lrdx * me.w + lrdy * me.z);
  return o;
}

// parameter gradients and the velocity loss of particle g, accumulated in FP64 (true units)
__device__ __forceinline__ void adj_accumulate(const StepParams& prm, const AdjLocal& o, float4 f, float theta,
                                               double unscale, double* acc, size_t ntot, int g) {
  double delta = atan2((double)f.w, (double)f.z) - (double)theta;
  if (delta > kPi) delta -= 2.0 * kPi;
  else if (delta <= -kPi) delta += 2.0 * kPi;
  const double tau = prm.dt64 / prm.rate64;
  acc[g] += unscale * (double)o.ls;                                            // d v0
  acc[ntot + g] += unscale * (double)o.ls * (double)o.m;                       // d mu
  if (prm.k64 != 0.0)                                                          // d k
    acc[2 * ntot + g] += unscale * ((double)o.lF.x * f.x + (double)o.lF.y * f.y) / prm.k64;
  acc[3 * ntot + g] += unscale * (double)o.ltn * delta * (-prm.dt64 / (tau * tau));  // d tau
  acc[4 * ntot + g] += (double)o.wv * ((double)o.dvx * o.dvx + (double)o.dvy * o.dvy);  // loss
}

__global__ void k_adj_local(const ReplicaDesc* reps, const int* rid, int ntot, LossDev L, const unsigned* halt,
                            const int* sexp_in, int* sexp_out, const unsigned* mx_in, unsigned* mx_next,
                            const float4* xycs, const float* th, const float4* fs, const int* quarter,
                            const float2* lr_in, const float* lt_in, float2* gF, float2* gSC, float4* part,
                            double* acc) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  const int r = rid[g];
  // rescale of the incoming adjoint (exact: a power of two); the replica's first particle carries
  // the exponent forward and clears the max slot the following gather fills
  const int s_up = rescale_exponent(mx_in[r]);
  const int E = sexp_in[r] + s_up;
  if (g == reps[r].off) {
    sexp_out[r] = E;
    mx_next[r] = 0u;
  }
  if (halt[r]) {
    gF[g] = make_float2(0.f, 0.f);
    gSC[g] = make_float2(0.f, 0.f);
    part[g] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const StepParams& prm = reps[r].prm;
  const float4 f = fs[g];  // (Fx, Fy, sum cos, sum sin)
  const AdjLocal o = adj_local(prm, L, reps[r].n, ldexpf(1.f, -s_up), ldexpf(1.f, -E), xycs[g], f, quarter[g],
                               lr_in[g], lt_in[g]);
  gF[g] = o.lF;
  gSC[g] = o.lsc;
  part[g] = make_float4(o.lrx, o.lry, o.lth, 0.f);
  adj_accumulate(prm, o, f, th[g], ldexp(1.0, E), acc, (size_t)ntot, g);
}

// each replica's max |lambda| over the wave: the wave's particles of one replica occupy a
// contiguous range of lanes, so an inclusive max-scan restricted to that range leaves the max on
// its last lane, which posts it (float bits of non-negative values order like the values;
// replicas are concatenated, so r never decreases along the wave). Every lane of the wave calls it.
__device__ __forceinline__ void post_replica_max(int r, float mx, unsigned* mx_out) {
  const int me_lane = (int)__lane_id();
  const int r_prev = __shfl_up(r, 1, 64), r_next = __shfl_down(r, 1, 64);
  const unsigned long long heads = __ballot(me_lane == 0 || r_prev != r);
  const unsigned long long upto = me_lane == 63 ? ~0ull : (2ull << me_lane) - 1ull;
  const int first = 63 - __clzll(heads & upto);
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const float up = __shfl_up(mx, o, 64);
    if (me_lane - o >= first) mx = fmaxf(mx, up);
  }
  if ((me_lane == 63 || r_next != r) && r >= 0 && mx > 0.f) atomicMax(&mx_out[r], __float_as_uint(mx));
}

template <int LANES>
__global__ void __launch_bounds__(256) k_adj_gather(MeshDev mesh, const ReplicaDesc* reps, const int* rid, int ntot,
                                                   BucketSet bs, const unsigned* halt, const float4* xycs,
                                                   const int* vid, const float2* gF, const float2* gSC,
                                                   const float4* part, float2* lr_out, float* lt_out,
                                                   unsigned* mx_out) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = tid / LANES, lane = tid & (LANES - 1);
  const int r = g < ntot ? rid[g] : -1;  // -1: past the end (the lanes stay for the wave reduction)
  float mx = 0.f;                        // |lambda| of this particle (on lane 0 of its group)
  if (r >= 0 && halt[r]) {
    if (lane == 0) { lr_out[g] = make_float2(0.f, 0.f); lt_out[g] = 0.f; }
  } else if (r >= 0) {
    const ReplicaDesc& rep = reps[r];
    const StepParams& prm = rep.prm;
    const int nv = mesh.nv;
    const BucketsIn b = replica_in(bs, r, nv, rep.off);
    const float2 mine = gF[g];
    float ax = 0.f, ay = 0.f, aS = 0.f, aC = 0.f;
    const int v = vid[g];
    const int rs = mesh.row_ptr[v], re = mesh.row_ptr[v + 1];
    for (int e = rs + lane; e < re; e += LANES) {
      const int2 ent = mesh.nbr[e];
      const float d = __int_as_float(ent.y);
      if (d > prm.t_near) break;  // rows ascend in distance
      const float c = pair_coef(d, prm);
      for_bucket(b, ent.x, [&](int j) {
        const float2 oj = gF[j];
        const float2 sc = gSC[j];
        if (j != g) {
          ax += c * (mine.x - oj.x);
          ay += c * (mine.y - oj.y);
        }
        aS += sc.x;
        aC += sc.y;
      });
    }
    ax = group_sum<LANES>(ax);
    ay = group_sum<LANES>(ay);
    aS = group_sum<LANES>(aS);
    aC = group_sum<LANES>(aC);
    if (lane == 0) {
      const float4 pt = part[g];
      const float4 me = xycs[g];
      const float2 o = make_float2(pt.x + ax, pt.y + ay);
      const float ot = pt.z + me.z * aS - me.w * aC;
      lr_out[g] = o;
      lt_out[g] = ot;
      mx = fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(ot)));  // fmaxf skips NaN
    }
  }
  post_replica_max(r, mx, mx_out);
}

// one block per replica: sums of the per-particle accumulators
__global__ void k_adj_reduce(const ReplicaDesc* reps, int ntot, const double* acc, double* out) {
  __shared__ double tmp[kAdjBlock / 64];
  const int r = blockIdx.x;
  const int off = reps[r].off, n = reps[r].n;
  for (int k = 0; k < kAdjAcc; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += acc[(size_t)k * ntot + off + i];
    s = block_sum(s, tmp);
    if (threadIdx.x == 0) out[kAdjAcc * r + k] = s;
  }
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace

hipError_t launch_adj_seed(int R, const ReplicaDesc* reps, const float4* xycs, const LossDev& L, const unsigned* halt,
                           float2* lr, float* lt, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_seed, dim3(R), dim3(kAdjBlock), 0, st, reps, xycs, L, halt, lr, lt, out);
  return hipGetLastError();
}

hipError_t launch_adj_local(const ReplicaDesc* reps, const int* rid, int ntot, const LossDev& L, const unsigned* halt,
                            const int* sexp_in, int* sexp_out, const unsigned* mx_in, unsigned* mx_next,
                            const float4* xycs, const float* th, const float4* fs, const int* quarter,
                            const float2* lr_in, const float* lt_in, float2* gF, float2* gSC, float4* part,
                            double* acc, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_local, dim3(cdiv(ntot, 256)), dim3(256), 0, st, reps, rid, ntot, L, halt, sexp_in, sexp_out,
                     mx_in, mx_next, xycs, th, fs, quarter, lr_in, lt_in, gF, gSC, part, acc);
  return hipGetLastError();
}

hipError_t launch_adj_gather(int lanes, const MeshDev& mesh, const ReplicaDesc* reps, const int* rid, int ntot,
                             const BucketSet& bs, const unsigned* halt, const float4* xycs, const int* vid,
                             const float2* gF, const float2* gSC, const float4* part, float2* lr_out, float* lt_out,
                             unsigned* mx_out, hipStream_t st) {
  const int blocks = cdiv((long)ntot * lanes, 256);
#define TISSUE_ADJ_GATHER(L)                                                                                     \
  hipLaunchKernelGGL(k_adj_gather<L>, dim3(blocks), dim3(256), 0, st, mesh, reps, rid, ntot, bs, halt, xycs, vid, \
                     gF, gSC, part, lr_out, lt_out, mx_out)
  switch (lanes) {
    case 1: TISSUE_ADJ_GATHER(1); break;
    case 2: TISSUE_ADJ_GATHER(2); break;
    case 4: TISSUE_ADJ_GATHER(4); break;
    case 8: TISSUE_ADJ_GATHER(8); break;
    case 16: TISSUE_ADJ_GATHER(16); break;
    case 32: TISSUE_ADJ_GATHER(32); break;
    case 64: TISSUE_ADJ_GATHER(64); break;
    default: return hipErrorInvalidValue;
  }
#undef TISSUE_ADJ_GATHER
  return hipGetLastError();
}

hipError_t launch_adj_reduce(int R, const ReplicaDesc* reps, int ntot, const double* acc, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_reduce, dim3(R), dim3(kAdjBlock), 0, st, reps, ntot, acc, out);
  return hipGetLastError();
}

}  // namespace tissue
