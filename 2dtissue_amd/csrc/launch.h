// Host-side launchers for the kernels in kernels.hip (plain C++ signatures, callable from the
// C-ABI layer).
#pragma once
#include <hip/hip_runtime.h>

#include "device.h"

namespace tissue {

struct OrderPartial {
  double cs, sn;   // sums of cos / sin of the new headings
  int sgn, pad;    // sum of sign(x vy - y vx)
};

// everything one step reads and writes besides the mesh and the buckets
struct StepIO {
  const float4* xycs_in;  // (x, y, cos theta, sin theta)
  const float* th_in;
  const int* vid_in;
  float4* xycs_out;
  float* th_out;
  int* vid_out;
  float* vx;
  float* vy;
  float* r3x;
  float* r3y;
  float* r3z;
  int* face;
  int* nbr;
  const float* noise;      // eta of this step, [n] (may be null)
  OrderPartial* partial;   // [blocks] for this step
  unsigned* flags;
  int* first_bad;
  int step;
};

int step_blocks(int n, int lanes, int block);
hipError_t launch_step(int lanes, int block, const tdev::MeshDev& mesh, const tdev::StepParams& prm, int n,
                       const tdev::BucketsIn& bin, const tdev::BucketsOut& bout, int* clear_count, int* clear_ovf_n,
                       const StepIO& io, hipStream_t st);
hipError_t launch_prepare(int n, const float* x, const float* y, const float* th, float4* xycs, float* th_out,
                          hipStream_t st);
hipError_t launch_locate(const tdev::MeshDev& mesh, int n, const float4* xycs, int* vid, int* face, float* r3x,
                         float* r3y, float* r3z, hipStream_t st);
hipError_t launch_clear(int nv, int* count, int* ovf_n, hipStream_t st);
hipError_t launch_insert(int n, const int* vid, const tdev::BucketsOut& bout, hipStream_t st);
hipError_t launch_order_finalize(const OrderPartial* partial, int steps, int blocks, int n, double* phi,
                                 hipStream_t st);
hipError_t launch_neighbor_lists(const tdev::MeshDev& mesh, const tdev::StepParams& prm, int n,
                                 const tdev::BucketsIn& bin, const int* vid, int cap, int* counts, int* cols,
                                 hipStream_t st);

}  // namespace tissue
