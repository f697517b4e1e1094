// Host-side launchers for the kernels in kernels.hip and resident.hip (plain C++ signatures,
// callable from the C-ABI layer).
#pragma once
#include <hip/hip_runtime.h>

#include "device.h"

namespace tissue {

// fixed-point scale of the order-parameter sums: cos and sin of the headings are summed as
// integers llrint(value * 2^40), which is exact and associative, so the sums do not depend on the
// order of the particles or on how they are grouped; each term is rounded by at most 2^-41, and
// up to 2^22 particles per replica stay inside int64
constexpr double kPhiScale = 1099511627776.0;  // 2^40

struct OrderPartial {
  long long cs, sn;  // sums of cos / sin of the new headings, fixed point (kPhiScale)
  int sgn, pad;      // sum of sign(x vy - y vx)
};

// one replica: particles [off, off + n) of the concatenated arrays
struct ReplicaDesc {
  int off, n;
  tdev::StepParams prm;
};

// one workgroup of the tiled step kernel: particles [first, first + block / lanes) of replica r
struct TileDesc {
  int r;        // replica; -1 marks an idle workgroup (padding of the XCD-balanced order)
  int first;    // first particle of the tile, counted within the replica
  int q, nq;    // tile index within the replica, number of tiles of the replica
  int logical;  // the tile's order-parameter partials start at logical * partials_per_tile(block)
                // (tiles of a replica are consecutive)
  int pad;
};

// one vertex-bucket set of all replicas: replica r owns count[r*nv ..], slot[r*nv*kSlots ..],
// the occupancy mask occ[r*occ_words(nv) ..], the counters ovf_n[r], multi_n[r], big_n[r] and
// the entries ovf, multi and big at [off_r .. off_r + n_r). As the set a step fills, multi may be
// null: no finalize pass follows, so the step lists no bucket for sorting.
struct BucketSet {
  unsigned* occ;
  int* count;
  int* slot;
  int* ovf_n;
  int2* ovf;
  int* multi_n;
  int* multi;
  int* big_n;
  int* big;
  int sorted;  // 1: readers see it after k_bucket_finalize; 0: raw (readers order each bucket)
};

// buckets of replica r (particles [off, ...)) within a set, as the step's readers see them
__device__ __forceinline__ tdev::BucketsIn replica_in(const BucketSet& s, int r, int nv, int off) {
  return {s.count + (size_t)r * nv, s.slot + (size_t)r * nv * tdev::kSlots, s.big + off, s.ovf + off, s.ovf_n + r,
          s.sorted};
}

// everything one step of the tiled kernel reads and writes besides the mesh and the buckets
struct TiledIO {
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
  const float* noise;       // eta of this step, [ntot] (may be null: zero noise), shifted by
  const int* noise_row0;    //   noise_row0[0] rows of noise_stride (null: no shift); lets one
  size_t noise_stride;      //   captured graph serve any first noise row
  OrderPartial* partial;    // [logical tiles][partials_per_tile] of this step
  const unsigned* halt_in;  // [R] flags raised before this step (a flagged replica passes through)
  unsigned* halt_out;       // [R] flags raised up to and including this step
  unsigned* halt_clr;       // [R] the third flag row, cleared for the next step
  int* first_bad;           // [R] first step of the run that raised a flag
  int step;                 // step index within the run
  float4* tape_fs;          // adjoint tape of this step, [ntot] (Fx, Fy, Sx, Sy); used by the TAPE variant
  int* tape_q;              // [ntot] quarter turns of the seam wrap
  int2* visit;              // [ntot] (na, vh) of StepOut: the neighbour set the step's own walk
                            //   visited (-1, 0 for a replica passed through); null: not recorded
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
  const float* noise;     // eta of the launch's first step; step s at noise + s * noise_stride,
  size_t noise_stride;    // shifted by noise_row0[0] rows (null: no shift)
  const int* noise_row0;
  double* phi;            // (phi_ref, phi_polar) of replica r, run step t at phi[2 * (r * phi_stride + t)]
  int phi_stride;
  int t_base;             // run step index of the launch's first step
  const unsigned* flags_in;  // [R] flags before the launch
  unsigned* flags_out;       // [R] flags after the launch
  int* first_bad;            // [R]
};

// largest workgroup of the tiled step kernel and of the persistent kernel (register budget of
// two waves per SIMD)
constexpr int kStepMaxBlock = 512;

// Whether the tiled step kernel stages the replica's vertex-occupancy mask in LDS: each workgroup
// loads the whole mask (one word per 32 vertices) before its first particle, which pays when the
// workgroup holds enough particles per word (measured on MI355X: 256 replicas, 64 particles per
// workgroup and 324 words, 7 % faster; 2,000 particles at 8 per workgroup, 6 % slower; 100,000
// particles at 64 per workgroup and 3,134 words, 1 % slower). Results are the same either way.
constexpr int kOccWordsPerParticle = 8;
inline bool stage_occupancy(int nv, int block, int lanes) {
  return tdev::occ_words(nv) <= kOccWordsPerParticle * (block / lanes);
}
constexpr int kResidentMaxBlock = 512;

// lanes per particle for the persistent kernel when not tuned: wider groups for smaller replicas
inline int resident_default_lanes(int nmax) { return nmax <= 256 ? 8 : (nmax <= 1024 ? 4 : (nmax <= 2048 ? 2 : 1)); }

// Default launch shape of the tiled step kernel (lanes per particle, threads per workgroup) for
// R replicas of ntot particles in all. Measured on MI355X: a single system of a few thousand
// particles fills few CUs, so more lanes per particle shorten its critical path, and one-wave
// workgroups let every wave start its walk without waiting for another at the staging barrier
// (2,000 particles: 32 lanes in 64 threads 18.97 us, 16 in 128 19.30 us, 64 in 64 19.19 us, 8 in
// 64 20.7 us); a single 100,000-particle system is fastest with 4 (2 lanes 10 % slower), a large
// batch with 2 (autotuned).
inline void tiled_default_shape(int R, int ntot, int* lanes, int* block) {
  const int big_lanes = R == 1 ? 4 : 2;
  *lanes = ntot <= 4096 ? 32 : (ntot <= 8192 ? 8 : (ntot <= 65536 ? 4 : big_lanes));
  *block = ntot <= 4096 ? 64 : 256;
}

// Replicas of a few thousand particles, at most one per two vertices, leave their vertex buckets
// raw: the finalize pass would cost a kernel boundary per step, and the readers order the few
// particles of each bucket themselves (the same visiting order, so the same results). Denser or
// larger systems pile particles onto vertices (the exit snapping), and a crowded raw bucket costs
// each reader a scan of the overflow list.
constexpr int kRawBucketsMax = 8192;
inline bool raw_buckets_for(int nmax, int nv) { return nmax <= kRawBucketsMax && 2L * nmax <= nv; }

// LDS bytes of one workgroup of the persistent kernel for a mesh with nv vertices and replicas of
// at most nmax particles; resident_max_particles is the largest nmax that fits one CU
size_t resident_lds_bytes(int nv, int nmax);
int resident_max_particles(int nv);
// allow the persistent kernels the LDS of replicas up to nmax particles (call outside capture)
hipError_t resident_prepare(int nv, int nmax);
hipError_t launch_resident(int lanes, int block, const tdev::MeshDev& mesh, const ReplicaDesc* reps, int R, int nmax,
                           int steps, const ResidentIO& io, hipStream_t st);

// one step of every replica: n_tiles workgroups of `block` threads, `lanes` lanes per particle;
// reads bucket set bin, writes bout, clears bclr; tape selects the variant that records the
// adjoint tape; wave the seam-wrap loop of crowded systems (same results, see seam_wrap_group)
hipError_t launch_step_tiled(int lanes, int block, bool tape, bool wave, const tdev::MeshDev& mesh, const ReplicaDesc* reps,
                             const TileDesc* tiles, int n_tiles, const BucketSet& bin, const BucketSet& bout,
                             const BucketSet& bclr, const TiledIO& io, hipStream_t st);
// sort the buckets of set bs that hold two or more particles (over the same tiles as the step);
// crowded buckets (more than kSlots) move, sorted, into the big area. single_n: the particle
// count when there is one replica (its tiles are then not read), else 0
hipError_t launch_bucket_finalize(int block, const ReplicaDesc* reps, const TileDesc* tiles, int n_tiles, int nv,
                                  int single_n, const BucketSet& bs, hipStream_t st);
// order parameters of `steps` steps of every replica from the step kernel's partials, one per
// wave: per_tile = waves per workgroup, partial[(s * n_logical + tile) * per_tile + wave];
// replica r's tiles are tile_begin[r] .. tile_begin[r+1]
hipError_t launch_phi_tiles(int R, int steps, const int* tile_begin, int n_logical, int per_tile,
                            const OrderPartial* partial, const ReplicaDesc* reps, double* phi, int phi_stride,
                            int t_base, hipStream_t st);
// order-parameter partials the tiled step kernel writes per tile (one per wave)
inline int partials_per_tile(int block) { return block / 64; }
// run start: rows[0] = halt, rows[1] = 0, first_bad = INT_MAX; run end: halt = last row
hipError_t launch_flags_begin(int R, const unsigned* halt, unsigned* rows, int* first_bad, hipStream_t st);
// halt = the run's last flag row; with tail, also tail[0 .. R) = the flags and tail[R .. 2R) =
// first_bad (the run's read-back words); then, for the next run, next_out_row (the output row of
// its first step) = 0 and first_bad = INT_MAX
hipError_t launch_flags_end(int R, const unsigned* last_row, unsigned* halt, int* first_bad, unsigned* tail,
                            unsigned* next_out_row, hipStream_t st);

hipError_t launch_prepare(int n, const float* x, const float* y, const float* th, float4* xycs, float* th_out,
                          hipStream_t st);
hipError_t launch_locate(const tdev::MeshDev& mesh, int n, const float4* xycs, int* vid, int* face, float* r3x,
                         float* r3y, float* r3z, hipStream_t st);
// copy step t's outputs into the per-step output ring (when its global step number, ctl[1] + t + 1,
// is a multiple of every): ring [depth][5][ntot] = x, y, X, Y, Z; ring_nbr [depth][ntot]
hipError_t launch_record(int ntot, const int* ctl, int t, int every, int depth, const float4* xycs, const float* r3x,
                         const float* r3y, const float* r3z, const int* nbr, float* ring, int* ring_nbr, hipStream_t st);
// area-uniform initial positions and headings of every replica (replica r from seeds[r])
hipError_t launch_init_random(int ntot, const int* rid, const ReplicaDesc* reps, const uint64_t* seeds,
                              const double* cdf, const tdev::MeshDev& mesh, float* x, float* y, float* th, hipStream_t st);
// seam wrap of n moves start -> end (end overwritten), lanes = 1, 2, 4, 8, 16, 32 or 64 per move;
// wave: a wave's last flight is finished by all of its lanes (seam_wrap_group)
hipError_t launch_seam_wrap(int lanes, int n, const double2* start, double2* end, double* dtheta, int* quarter, int* ok,
                            int max_crossings, int wave, hipStream_t st);
// bucket every particle of every replica (rid[g] = replica of particle g) into set `bs`
hipError_t launch_insert(int ntot, const int* vid, const int* rid, const ReplicaDesc* reps, int nv,
                         const BucketSet& bs, hipStream_t st);
// neighbour sets {j : d_ij < 2 sigma} of the current state (ids counted within the replica)
hipError_t launch_neighbor_lists(const tdev::MeshDev& mesh, const ReplicaDesc* reps, const int* rid, int ntot,
                                 const BucketSet& bs, const int* vid, int cap, int* counts, int* cols,
                                 hipStream_t st);

}  // namespace tissue
