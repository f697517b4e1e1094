// Persistent, LDS-resident kernel: one workgroup runs one replica for many steps.
//
// Per workgroup the replica's state (x, y, cos, sin, theta, vertex id) and the step's vertex
// buckets live in LDS. Each step:
//   1. counting sort of the particles by vertex id into a CSR (LDS atomics + block scan);
//   2. every LANES-lane group updates its particles with the shared particle_step (the
//      geodesic-table rows and the mesh stay in L2) and stages the new state in a global slice;
//   3. barrier, then the staged state is copied back to LDS (the update is Jacobi: every
//      particle reads the old state) and the order parameters are reduced and written.
// Per step only the noise row is read from HBM and the order parameters written; the staged
// state stays in L2; the state goes back to its HBM home once per launch.
#include <hip/hip_runtime.h>

#include "launch.h"

using namespace tdev;

namespace tissue {

namespace {

constexpr int kMaxBlock = kResidentMaxBlock;

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

struct LdsLayout {
  size_t begin, xycs, th, vid, sorted, scan, total;
  __host__ __device__ LdsLayout(int nv, int nmax) {
    size_t o = 0;
    begin = o; o = align16(o + 4 * (size_t)(nv + 1));
    xycs = o; o = align16(o + 16 * (size_t)nmax);
    th = o; o = align16(o + 4 * (size_t)nmax);
    vid = o; o = align16(o + 4 * (size_t)nmax);
    sorted = o; o = align16(o + 2 * (size_t)nmax);
    scan = o; o = align16(o + 4 * (kMaxBlock / 64));
    total = o;
  }
};

__device__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* tmp) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  uint32_t inc = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) tmp[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int w = 0; w < nw; ++w) {
      const uint32_t t = tmp[w];
      tmp[w] = acc;
      acc += t;
    }
  }
  __syncthreads();
  return tmp[wave] + inc - x;
}

// counting sort of particles by vertex: afterwards begin[v] .. begin[v+1] lists the particles on
// vertex v (begin[nv] = n)
__device__ void build_buckets(uint32_t* begin, uint16_t* sorted, const int* vid, int n, int nv, uint32_t* tmp) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int v = tid; v <= nv; v += nthr) begin[v] = 0u;
  __syncthreads();
  for (int i = tid; i < n; i += nthr) atomicAdd(&begin[vid[i]], 1u);
  __syncthreads();
  const int chunk = (nv + nthr - 1) / nthr;
  const int lo = min(tid * chunk, nv), hi = min(lo + chunk, nv);
  uint32_t sum = 0;
  for (int v = lo; v < hi; ++v) {
    sum += begin[v];
    begin[v] = sum;
  }
  const uint32_t excl = block_exclusive_scan(sum, tmp);
  for (int v = lo; v < hi; ++v) begin[v] += excl;
  __syncthreads();
  for (int i = tid; i < n; i += nthr) {
    const uint32_t p = atomicSub(&begin[vid[i]], 1u) - 1u;
    sorted[p] = (uint16_t)i;
  }
  if (tid == 0) begin[nv] = (uint32_t)n;
  __syncthreads();
}

template <int LANES>
__global__ void __launch_bounds__(kMaxBlock) k_resident(MeshDev mesh, const ReplicaDesc* reps, int nmax, int steps,
                                                        ResidentIO io) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red_cs[kMaxBlock / 64], red_sn[kMaxBlock / 64];
  __shared__ int red_sg[kMaxBlock / 64];
  __shared__ unsigned s_flags;
  const LdsLayout L(mesh.nv, nmax);
  uint32_t* begin = (uint32_t*)(smem + L.begin);
  float4* xycs = (float4*)(smem + L.xycs);
  float* th = (float*)(smem + L.th);
  int* vid = (int*)(smem + L.vid);
  uint16_t* sorted = (uint16_t*)(smem + L.sorted);
  uint32_t* scan_tmp = (uint32_t*)(smem + L.scan);

  const int r = blockIdx.x;
  const ReplicaDesc rep = reps[r];
  const int n = rep.n, off = rep.off;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = tid & (LANES - 1), group = tid / LANES, ngroups = nthr / LANES;
  float4* nxt_xycs = io.stage_xycs + off;
  float* nxt_th = io.stage_th + off;
  int* nxt_vid = io.stage_vid + off;

  for (int i = tid; i < n; i += nthr) {
    xycs[i] = io.xycs[off + i];
    th[i] = io.th[off + i];
    vid[i] = io.vid[off + i];
  }
  if (tid == 0) s_flags = io.flags_in[r];
  __syncthreads();

  const LdsSource src{begin, sorted, xycs, th};
  for (int s = 0; s < steps; ++s) {
    const int t = io.t_base + s;
    double* phi = io.phi + 2 * ((size_t)r * io.phi_stride + t);
    if (s_flags != 0u) {  // a failed step stops the replica (the reference stops at the first failure)
      if (tid == 0) { phi[0] = 0.0; phi[1] = 0.0; }
      continue;
    }
    build_buckets(begin, sorted, vid, n, mesh.nv, scan_tmp);
    const size_t row = (size_t)s + (io.noise_row0 ? (size_t)io.noise_row0[0] : 0);
    const float* eta_row = io.noise ? io.noise + row * io.noise_stride + off : nullptr;
    int sgn = 0;
    double cs = 0.0, sn = 0.0;
    unsigned fl = 0;
    // new state staged in global memory (this workgroup's slice, L2-resident): the update is
    // Jacobi, every particle must read the old state of its neighbours
    for (int i = group; i < n; i += ngroups) {
      const float eta = eta_row ? eta_row[i] : 0.0f;
      const StepOut o = particle_step<LANES>(mesh, rep.prm, src, i, xycs[i], th[i], vid[i], eta, lane);
      if (lane == 0) {
        nxt_xycs[i] = make_float4(o.xo, o.yo, o.co, o.so);
        nxt_th[i] = o.to;
        nxt_vid[i] = o.L.vid;
        sgn += o.sgn;
        cs += (double)o.co;
        sn += (double)o.so;
        fl |= o.fl;
        if (s == steps - 1) {
          const int g = off + i;
          io.vx[g] = o.vx; io.vy[g] = o.vy;
          io.r3x[g] = (float)o.L.X; io.r3y[g] = (float)o.L.Y; io.r3z[g] = (float)o.L.Z;
          io.face[g] = o.L.face;
          io.nbr[g] = o.nb;
        }
      }
    }
    if (fl) {
      atomicOr(&s_flags, fl);
      atomicMin(&io.first_bad[r], t);
    }
    // order parameters of the step: butterfly per wave, waves summed in order by thread 0
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      sgn += __shfl_xor(sgn, d, 64);
      cs += __shfl_xor(cs, d, 64);
      sn += __shfl_xor(sn, d, 64);
    }
    const int wave = tid >> 6;
    if ((tid & 63) == 0) { red_cs[wave] = cs; red_sn[wave] = sn; red_sg[wave] = sgn; }
    __syncthreads();  // all reads of the old state done; the staged state is visible to the workgroup
    for (int i = tid; i < n; i += nthr) {
      xycs[i] = nxt_xycs[i];
      th[i] = nxt_th[i];
      vid[i] = nxt_vid[i];
    }
    if (tid == 0) {
      double C = 0.0, S = 0.0;
      long G = 0;
      for (int w = 0; w < (nthr + 63) / 64; ++w) { C += red_cs[w]; S += red_sn[w]; G += red_sg[w]; }
      phi[0] = (1.0 / n) * fabs((double)G);
      phi[1] = sqrt(C * C + S * S) / n;
    }
    __syncthreads();
  }

  for (int i = tid; i < n; i += nthr) {
    io.xycs[off + i] = xycs[i];
    io.th[off + i] = th[i];
    io.vid[off + i] = vid[i];
  }
  if (tid == 0) io.flags_out[r] = s_flags;
}

template <int LANES>
hipError_t launch_one(int block, const MeshDev& mesh, const ReplicaDesc* reps, int R, int nmax, int steps,
                      const ResidentIO& io, hipStream_t st) {
  const size_t lds = LdsLayout(mesh.nv, nmax).total;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(st, &cap);
  if (e != hipSuccess) return e;
  if (cap == hipStreamCaptureStatusNone) {  // under capture the attribute was set by resident_prepare
    e = hipFuncSetAttribute((const void*)k_resident<LANES>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_resident<LANES>), dim3(R), dim3(block), lds, st, mesh, reps, nmax, steps, io);
  return hipGetLastError();
}

}  // namespace

size_t resident_lds_bytes(int nv, int nmax) {
  return LdsLayout(nv, nmax).total + sizeof(double) * 2 * (kMaxBlock / 64) + sizeof(int) * (kMaxBlock / 64) + 16;
}

int resident_max_particles(int nv) {
  const size_t cap = 160 * 1024;
  int lo = 0, hi = 65535;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (resident_lds_bytes(nv, mid) <= cap) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

hipError_t resident_prepare(int nv, int nmax) {
  const int lds = (int)LdsLayout(nv, nmax).total;
  const void* fns[] = {(const void*)k_resident<1>, (const void*)k_resident<2>, (const void*)k_resident<4>,
                       (const void*)k_resident<8>};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_resident(int lanes, int block, const MeshDev& mesh, const ReplicaDesc* reps, int R, int nmax,
                           int steps, const ResidentIO& io, hipStream_t st) {
  if (block < 64 || block > kMaxBlock || block % 64 || nmax > resident_max_particles(mesh.nv)) return hipErrorInvalidValue;
  switch (lanes) {
    case 1: return launch_one<1>(block, mesh, reps, R, nmax, steps, io, st);
    case 2: return launch_one<2>(block, mesh, reps, R, nmax, steps, io, st);
    case 4: return launch_one<4>(block, mesh, reps, R, nmax, steps, io, st);
    case 8: return launch_one<8>(block, mesh, reps, R, nmax, steps, io, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace tissue
