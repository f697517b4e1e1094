// Launchers and device-side constants of the reverse pass (adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "launch.h"

namespace tissue {

constexpr int kMaxModes = 16;  // modes of the target velocity field
constexpr int kAdjBlock = 256; // threads of the per-replica reduction kernels
constexpr int kAdjAcc = 5;     // per-particle FP64 accumulators: d v0, d mu, d k, d tau, velocity loss

// loss of a trajectory: w_order (phi_T - target)^2 + w_velocity mean_{t<T,i} |rdot - u(r)|^2 with
// u_c(x, y) = sum over modes of component c of amp sin(kx x + ky y + phase)
struct LossDev {
  float amp[kMaxModes], kx[kMaxModes], ky[kMaxModes], phase[kMaxModes];
  int comp[kMaxModes];
  int n_modes;
  int n_steps;  // T
  float target, w_order, w_velocity;
};

// per replica: out[2r] = phi_T, out[2r+1] = order loss; seeds lr = 0, lt = d loss / d theta_T
hipError_t launch_adj_seed(int R, const ReplicaDesc* reps, const float4* xycs, const LossDev& L, const unsigned* halt,
                           float2* lr, float* lt, double* out, hipStream_t st);
// per particle factors of step t's reverse (see adjoint.hip), accumulators acc[kAdjAcc][ntot];
// the adjoint state of replica r is stored scaled by 2^-E_r: the kernel reads E_r from sexp_in and
// the max |lambda| of its input from mx_in (the previous gather's), scales its input by a power of
// two when that max exceeds 2^21, writes the new E_r to sexp_out and clears mx_next
hipError_t launch_adj_local(const ReplicaDesc* reps, const int* rid, int ntot, const LossDev& L, const unsigned* halt,
                            const int* sexp_in, int* sexp_out, const unsigned* mx_in, unsigned* mx_next,
                            const float4* xycs, const float* th, const float4* fs, const int* quarter,
                            const float2* lr_in, const float* lt_in, float2* gF, float2* gSC, float4* part,
                            double* acc, hipStream_t st);
// the transposed neighbour sums over step t's vertex buckets bs; mx_out[r] = max |lambda| of the
// output of replica r (float bits, combined with atomicMax: clear it before)
hipError_t launch_adj_gather(int lanes, const tdev::MeshDev& mesh, const ReplicaDesc* reps, const int* rid, int ntot,
                             const BucketSet& bs, const unsigned* halt, const float4* xycs, const int* vid,
                             const float2* gF, const float2* gSC, const float4* part, float2* lr_out, float* lt_out,
                             unsigned* mx_out, hipStream_t st);
// out[kAdjAcc r + k] = sum over replica r of acc[k]
hipError_t launch_adj_reduce(int R, const ReplicaDesc* reps, int ntot, const double* acc, double* out, hipStream_t st);

}  // namespace tissue
