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

// one replica of the persistent kernel: particles [off, off + n) of the batch arrays
struct ReplicaDesc {
  int off, n;
  tdev::StepParams prm;
};

struct ResidentIO {
  float4* xycs;           // batch state, read at launch start and written back at the end
  float* th;
  int* vid;
  float4* stage_xycs;     // per-step staging of the new state (same layout as the state)
  float* stage_th;
  int* stage_vid;
  float* vx;              // outputs of the launch's last step
  float* vy;
  float* r3x;
  float* r3y;
  float* r3z;
  int* face;
  int* nbr;
  const float* noise;     // eta of the launch's first step; step s at noise + s * noise_stride
  size_t noise_stride;
  double* phi;            // (phi_ref, phi_polar) of replica r, global step t at phi[2 * (r * phi_stride + t)]
  int phi_stride;
  int t_base;             // global step index of the launch's first step
  unsigned* flags;        // [R]
  int* first_bad;         // [R]
};

// largest workgroup of the persistent kernel (its register budget is set for 8 waves per CU)
constexpr int kResidentMaxBlock = 512;

// lanes per particle for the persistent kernel when not tuned: wider groups for smaller replicas
inline int resident_default_lanes(int nmax) { return nmax <= 256 ? 8 : (nmax <= 1024 ? 4 : (nmax <= 2048 ? 2 : 1)); }

// LDS bytes of one workgroup of the persistent kernel for a mesh with nv vertices and replicas of
// at most nmax particles; resident_max_particles is the largest nmax that fits one CU
size_t resident_lds_bytes(int nv, int nmax);
int resident_max_particles(int nv);
// allow the persistent kernels the LDS of replicas up to nmax particles (call outside capture)
hipError_t resident_prepare(int nv, int nmax);
hipError_t launch_resident(int lanes, int block, const tdev::MeshDev& mesh, const ReplicaDesc* reps, int R, int nmax,
                           int steps, const ResidentIO& io, hipStream_t st);

// largest workgroup of the one-launch-per-step kernel (register budget of 2 waves per SIMD)
constexpr int kStepMaxBlock = 512;

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
