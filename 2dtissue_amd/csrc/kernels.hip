// gfx950 kernels of the per-step update (one kernel launch per step).
//
// k_step<LANES>: LANES lanes of a wave64 cooperate on one particle. They stride over the
// particle's row of the truncated geodesic table (vertices within 2.4 sigma of its vertex,
// ascending in distance), visit the particles bucketed on those vertices, and reduce force,
// alignment sum and neighbour count with butterfly shuffles. Every lane then runs the
// per-particle epilogue (speed, Euler, heading, seam wrap) and the lanes split the point
// location candidates of the face bucket. Lane 0 writes the SoA outputs, inserts the particle
// into next step's vertex buckets and contributes to the block's order-parameter partials.
#include <hip/hip_runtime.h>

#include "launch.h"

using namespace tdev;

namespace tissue {

namespace {

__device__ __forceinline__ void bucket_insert(const BucketsOut& b, int v, int i) {
  const int old = atomicAdd(&b.count[v], 1);
  if (old < kSlots) {
    b.slot[(size_t)v * kSlots + old] = i;
  } else {
    const int o = atomicAdd(b.ovf_n, 1);
    b.ovf[o] = make_int2(v, i);
  }
}

// block-wide deterministic sum of (sgn, cs, sn): butterfly within each wave64, then wave
// partials added in wave order by thread 0
__device__ void block_order_partial(int sgn, double cs, double sn, OrderPartial* out) {
  __shared__ double s_cs[16], s_sn[16];
  __shared__ int s_sg[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sgn += __shfl_xor(sgn, off, 64);
    cs += __shfl_xor(cs, off, 64);
    sn += __shfl_xor(sn, off, 64);
  }
  const int wave = threadIdx.x / 64, nw = (blockDim.x + 63) / 64;
  if ((threadIdx.x & 63) == 0) { s_cs[wave] = cs; s_sn[wave] = sn; s_sg[wave] = sgn; }
  __syncthreads();
  if (threadIdx.x == 0) {
    OrderPartial p{0.0, 0.0, 0, 0};
    for (int w = 0; w < nw; ++w) { p.cs += s_cs[w]; p.sn += s_sn[w]; p.sgn += s_sg[w]; }
    *out = p;
  }
}

template <int LANES>
__global__ void __launch_bounds__(kStepMaxBlock) k_step(MeshDev mesh, StepParams prm, int n, BucketsIn bin, BucketsOut bout,
                                               int* clear_count, int* clear_ovf_n, StepIO io) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (LANES - 1);
  const int i = tid / LANES;
  // a step after a validation failure passes the state through unchanged (the reference stops
  // at the first failure)
  const bool halted = *(volatile const unsigned*)io.flags != 0u;
  // clear the bucket set consumed two steps ago (it is next step's output)
  for (int v = tid; v < mesh.nv; v += gridDim.x * blockDim.x) clear_count[v] = 0;
  if (tid == 0) *clear_ovf_n = 0;

  int sgn = 0;
  double pcs = 0.0, psn = 0.0;
  if (i < n) {
    const float4 me = io.xycs_in[i];
    const float th = io.th_in[i];
    const int v = io.vid_in[i];
    if (halted) {
      if (lane == 0) {
        io.xycs_out[i] = me;
        io.th_out[i] = th;
        io.vid_out[i] = v;
        bucket_insert(bout, v, i);
      }
    } else {
      const float eta = io.noise ? io.noise[i] : 0.0f;
      const GlobalSource src{bin, io.xycs_in, io.th_in};
      const StepOut o = particle_step<LANES>(mesh, prm, src, i, me, th, v, eta, lane);
      if (lane == 0) {
        io.xycs_out[i] = make_float4(o.xo, o.yo, o.co, o.so);
        io.th_out[i] = o.to;
        io.vid_out[i] = o.L.vid;
        io.vx[i] = o.vx;
        io.vy[i] = o.vy;
        io.r3x[i] = (float)o.L.X;
        io.r3y[i] = (float)o.L.Y;
        io.r3z[i] = (float)o.L.Z;
        io.face[i] = o.L.face;
        io.nbr[i] = o.nb;  // count_particle_neighbors (2DTissue.cpp:254-268)
        bucket_insert(bout, o.L.vid, i);
        if (o.fl) {
          atomicOr(io.flags, o.fl);
          atomicMin(io.first_bad, io.step);
        }
        sgn = o.sgn;
        pcs = (double)o.co;
        psn = (double)o.so;
      }
    }
  }
  block_order_partial(sgn, pcs, psn, io.partial + blockIdx.x);
}

__global__ void k_prepare(int n, const float* x, const float* y, const float* th, float4* xycs, float* th_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s, c;
  sincosf(th[i], &s, &c);
  xycs[i] = make_float4(x[i], y[i], c, s);
  th_out[i] = th[i];
}

__global__ void k_locate(MeshDev mesh, int n, const float4* xycs, int* vid, int* face, float* r3x, float* r3y,
                         float* r3z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = xycs[i];
  const Located L = locate<1>(mesh, (double)p.x, (double)p.y, 0);
  vid[i] = L.vid;
  if (face) face[i] = L.face;
  if (r3x) { r3x[i] = (float)L.X; r3y[i] = (float)L.Y; r3z[i] = (float)L.Z; }
}

__global__ void k_clear(int nv, int* count, int* ovf_n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) count[i] = 0;
  if (i == 0) *ovf_n = 0;
}

__global__ void k_insert(int n, const int* vid, BucketsOut bout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) bucket_insert(bout, vid[i], i);
}

__global__ void k_order_finalize(const OrderPartial* partial, int steps, int blocks, int n, double* phi) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= steps) return;
  double cs = 0.0, sn = 0.0;
  long sg = 0;
  for (int b = 0; b < blocks; ++b) {
    const OrderPartial p = partial[(size_t)t * blocks + b];
    cs += p.cs;
    sn += p.sn;
    sg += p.sgn;
  }
  phi[2 * t] = (1.0 / n) * fabs((double)sg);
  phi[2 * t + 1] = sqrt(cs * cs + sn * sn) / n;
}

__global__ void k_neighbor_lists(MeshDev mesh, StepParams prm, int n, BucketsIn bin, const int* vid, int cap,
                                 int* counts, int* cols) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = vid[i];
  int k = 0;
  for (int e = mesh.row_ptr[v]; e < mesh.row_ptr[v + 1]; ++e) {
    const int2 ent = mesh.nbr[e];
    if (__int_as_float(ent.y) > prm.t_near) break;
    for_bucket(bin, ent.x, [&](int j) {
      if (k < cap) cols[(size_t)i * cap + k] = j;
      ++k;
    });
  }
  counts[i] = k;
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace

int step_blocks(int n, int lanes, int block) { return cdiv((long)n * lanes, block); }

hipError_t launch_step(int lanes, int block, const MeshDev& mesh, const StepParams& prm, int n, const BucketsIn& bin,
                       const BucketsOut& bout, int* clear_count, int* clear_ovf_n, const StepIO& io, hipStream_t st) {
  const int blocks = step_blocks(n, lanes, block);
  switch (lanes) {
    case 1: hipLaunchKernelGGL(k_step<1>, dim3(blocks), dim3(block), 0, st, mesh, prm, n, bin, bout, clear_count, clear_ovf_n, io); break;
    case 2: hipLaunchKernelGGL(k_step<2>, dim3(blocks), dim3(block), 0, st, mesh, prm, n, bin, bout, clear_count, clear_ovf_n, io); break;
    case 4: hipLaunchKernelGGL(k_step<4>, dim3(blocks), dim3(block), 0, st, mesh, prm, n, bin, bout, clear_count, clear_ovf_n, io); break;
    case 8: hipLaunchKernelGGL(k_step<8>, dim3(blocks), dim3(block), 0, st, mesh, prm, n, bin, bout, clear_count, clear_ovf_n, io); break;
    case 16: hipLaunchKernelGGL(k_step<16>, dim3(blocks), dim3(block), 0, st, mesh, prm, n, bin, bout, clear_count, clear_ovf_n, io); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_prepare(int n, const float* x, const float* y, const float* th, float4* xycs, float* th_out,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_prepare, dim3(cdiv(n, 256)), dim3(256), 0, st, n, x, y, th, xycs, th_out);
  return hipGetLastError();
}

hipError_t launch_locate(const MeshDev& mesh, int n, const float4* xycs, int* vid, int* face, float* r3x, float* r3y,
                         float* r3z, hipStream_t st) {
  hipLaunchKernelGGL(k_locate, dim3(cdiv(n, 256)), dim3(256), 0, st, mesh, n, xycs, vid, face, r3x, r3y, r3z);
  return hipGetLastError();
}

hipError_t launch_clear(int nv, int* count, int* ovf_n, hipStream_t st) {
  hipLaunchKernelGGL(k_clear, dim3(cdiv(nv > 0 ? nv : 1, 256)), dim3(256), 0, st, nv, count, ovf_n);
  return hipGetLastError();
}

hipError_t launch_insert(int n, const int* vid, const BucketsOut& bout, hipStream_t st) {
  hipLaunchKernelGGL(k_insert, dim3(cdiv(n, 256)), dim3(256), 0, st, n, vid, bout);
  return hipGetLastError();
}

hipError_t launch_order_finalize(const OrderPartial* partial, int steps, int blocks, int n, double* phi,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_order_finalize, dim3(cdiv(steps, 64)), dim3(64), 0, st, partial, steps, blocks, n, phi);
  return hipGetLastError();
}

hipError_t launch_neighbor_lists(const MeshDev& mesh, const StepParams& prm, int n, const BucketsIn& bin,
                                 const int* vid, int cap, int* counts, int* cols, hipStream_t st) {
  hipLaunchKernelGGL(k_neighbor_lists, dim3(cdiv(n, 256)), dim3(256), 0, st, mesh, prm, n, bin, vid, cap, counts, cols);
  return hipGetLastError();
}

}  // namespace tissue
