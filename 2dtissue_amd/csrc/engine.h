// Host engine shared by tissue_sim (one replica) and tissue_batch (R replicas): device state,
// vertex buckets, launch-shape resolution, hipGraph capture of runs, autotuning, and the
// bookkeeping of order parameters and validation flags.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "adjoint.h"
#include "capi_common.h"

namespace tissue {

// launch shape of a run; 0 in a field of a request means "automatic"
struct Shape {
  int mode = 0;   // TISSUE_MODE_STEP_KERNEL (tiled, one launch per step) or TISSUE_MODE_RESIDENT
  int lanes = 0;  // lanes per particle
  int block = 0;  // threads per workgroup
  int spl = 0;    // steps per launch (resident) / per order-parameter reduction (tiled)
  int xcd = -1;   // tiled: 1 = each replica's tiles on one XCD, 0 = linear order, -1 = automatic
  std::tuple<int, int, int, int, int> key() const { return {mode, lanes, block, spl, xcd}; }
};

struct Engine {
  Engine(tissue_mesh* m, int R, const int* n, const tissue_params* p);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // state of all replicas, concatenated (replica r at off[r])
  void adopt_state(const float* dx, const float* dy, const float* dt, const int32_t* vid_in);
  void set_state(const float* x, const float* y, const float* theta, const int32_t* vid);
  void init_random(const uint64_t* seeds);
  void get_state(float* x, float* y, float* theta, int32_t* vid);
  void get_outputs(float* vx, float* vy, float* r3d, int32_t* face, int32_t* nbr);
  // neighbour sets {j : d_ij < 2 sigma} of the current state, ids counted within each replica
  void neighbor_lists(int32_t* counts, int32_t* cols, int32_t cap);
  // visit check: while on, every step of the tiled kernel records, per particle, the size of the
  // neighbour set its own walk visited and the sum of visit_mix over the visited (global) ids
  void set_visit_check(bool on);
  void get_visits(int32_t* na, uint32_t* hash);

  // noise: rows [rows][ntot] kept on the device for later runs
  void stage_noise(const float* h, int rows);
  void gen_noise(const uint64_t* seeds, const double* stds, int rows);
  const float* staged_rows(int first_row, int n_steps) const;
  // one call's rows in a separate buffer (leaves the staged rows alone)
  const float* upload_call_noise(const float* h, int n_steps);
  // first_row, checked against the staged rows (throws when rows first_row .. + n_steps are absent)
  int checked_first_row(int first_row, int n_steps) const;

  // n_steps steps; row s of the noise at noise + s * stride (null: no noise). Returns the status.
  // row0: first row of the noise used (row s of the run at noise + (row0 + s) * stride); the
  // captured step graphs do not depend on it (it is read from device memory)
  int32_t run(int n_steps, const float* noise, size_t stride, int row0 = 0);
  int32_t run_shape(const Shape& s, int n_steps, const float* noise, size_t stride, int row0 = 0);
  int32_t run_once(const Shape& s, int n_steps, const float* noise, size_t stride, int row0);
  void record_collect(long base, int n_steps);
  Shape resolve() const;
  // the request resolved for the tiled step kernel (the adjoint records its tape there)
  Shape resolve_tiled() const;
  Shape autotune(int n_steps);
  void set_request(const Shape& s);
  void synchronize();
  // gradients of a trajectory loss over n_steps from the current state (engine_adjoint.cpp);
  // g[R]; d_theta0, d_x0, d_y0 [ntot] may be null
  int32_t adjoint(int n_steps, const float* noise, size_t stride, int ckpt_every, const tissue_loss& loss,
                  tissue_grads* g, double* d_theta0, double* d_x0, double* d_y0);

  Shape req;
  std::vector<std::vector<double>> hist_ref, hist_polar;
  std::vector<unsigned> flags_acc;
  std::vector<int> first_bad_acc;
  long steps_done = 0;
  double last_ms = 0.0;

  tissue_mesh* mesh = nullptr;
  tdev::MeshDev md{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // control words of the current run, read by its kernels: [0] first noise row, [1] steps done
  // since the output stream's origin (0 when not recording); uploaded from pinned host memory
  // only when they differ from what the device holds (h_row0 mirrors it once row0_known)
  int* d_row0 = nullptr;
  int* h_row0 = nullptr;
  bool row0_known = false;
  // per-step output stream (engine_record.cpp); null when not recording
  std::unique_ptr<struct Recorder> rec;
  void record_start(const char* dir, int every);
  int32_t record_stop();
  int R = 0, ntot = 0, nmax = 0, resident_cap = 0;
  std::vector<int> n, off;
  std::vector<tissue_params> p;
  std::vector<ReplicaDesc> reps;
  bool has_state = false;
  // vertex buckets left raw (no finalize pass; readers order each bucket): raw_buckets_for
  bool raw_buckets = false;

  ReplicaDesc* d_reps = nullptr;
  int* d_rid = nullptr;  // replica of each particle
  int2* d_visit = nullptr;  // visit check records (set_visit_check), null while off
  float4* xycs[2] = {nullptr, nullptr};
  float* th[2] = {nullptr, nullptr};
  int* vid[2] = {nullptr, nullptr};
  int sphase = 0;  // state buffer holding the current state
  float *vx = nullptr, *vy = nullptr, *r3x = nullptr, *r3y = nullptr, *r3z = nullptr;
  int *face = nullptr, *nbr = nullptr;
  unsigned* d_halt = nullptr;   // [R] flags of replicas that have stopped (kept across runs)
  unsigned* d_rows = nullptr;   // [3][R] rotating flag rows of a run
  int* d_first_bad = nullptr;   // [R]

  // tiled strategy
  BucketSet bs[3] = {};
  int bphase = 0;  // bucket set holding the current state's buckets
  bool buckets_ready = false, buckets_stale = true;
  TileDesc* d_tiles = nullptr;
  int* d_tile_begin = nullptr;
  int n_phys = 0, n_logical = 0;
  std::tuple<int, int, int> tiles_key{-1, -1, -1};
  OrderPartial* partial = nullptr;
  size_t partial_cap = 0;

  // a run's order history (2 doubles per replica and step), then its flag words and first
  // failing steps (R each): coherent pinned host memory that the run's kernels write directly,
  // read by the host once the run has synchronised (no read-back copy); phi_dev is its device
  // address
  double* phi = nullptr;
  double* phi_dev = nullptr;
  size_t phi_cap = 0;
  float* noise = nullptr;  // staged rows
  int noise_rows = 0;      // rows of the last staging (the valid ones)
  int noise_cap = 0;       // rows the buffer holds
  float* call_noise = nullptr;
  size_t call_noise_cap = 0;

  std::map<std::tuple<int, const float*, size_t, std::tuple<int, int, int, int, int>, int, int>, hipGraphExec_t> graphs;

  // one segment's tape: states [K+1][ntot], force / alignment sums and seam-wrap turns [K][ntot],
  // the bucket sets of its states when they fit (else empty: each reverse step re-buckets), and
  // the scratch its recompute writes (flag rows, order partials, per-step outputs)
  struct Tape {
    float4* xycs = nullptr; float* th = nullptr; int* vid = nullptr;
    float4* fs = nullptr; int* q = nullptr;
    std::vector<BucketSet> sbs;
    float *vx = nullptr, *vy = nullptr, *r3x = nullptr, *r3y = nullptr, *r3z = nullptr;
    int *face = nullptr, *nbr = nullptr;
    unsigned* rows = nullptr; int* first_bad = nullptr;
    OrderPartial* partial = nullptr; int partial_n = 0;
  };
  static constexpr int kMaxTapes = 3;
  // adjoint buffers (grown on demand): checkpoints [C][ntot], tapes, adjoint state, factors.
  // With ntape > 1 tapes, segment c is recomputed into tape c % ntape on recompute stream
  // c % (ntape - 1) while the reverse passes of later segments run on the engine's stream;
  // events hand the tapes over.
  struct AdjointBuffers {
    int K = 0, C = 0;
    float4* ck_xycs = nullptr; float* ck_th = nullptr; int* ck_vid = nullptr;
    Tape tp[kMaxTapes];
    int ntape = 0;
    hipStream_t rec_stream[kMaxTapes - 1] = {};
    hipEvent_t ev_fwd = nullptr, ev_ck[kMaxTapes] = {}, ev_rec[kMaxTapes] = {}, ev_rev[kMaxTapes] = {};
    float2* lr[2] = {nullptr, nullptr}; float* lt[2] = {nullptr, nullptr};
    float2* gF = nullptr; float2* gSC = nullptr; float4* part = nullptr;
    double* acc = nullptr; double* out_seed = nullptr; double* out_red = nullptr;
    int* sexp = nullptr;  // [2][R]: the adjoint state is stored scaled by 2^-sexp (slots alternate by step)
    unsigned* mx = nullptr;  // [2][R]: max |lambda| of the last gather's output (float bits)
  } adj;
  void free_tape(Tape& t);
  void adjoint_buffers(int K, int C);
  void adjoint_release();

  void drop_graphs();
  void release();
  void ensure_buckets();
  void rebuild_buckets();
  void bucket_state(const int* vid, int set);
  void bucket_state_into(const int* vid, BucketSet& b, hipStream_t st = nullptr);
  void alloc_bucket_set(BucketSet& b);
  void free_bucket_set(BucketSet& b);
  void clear_bucket_set(BucketSet& b, hipStream_t st = nullptr);
  size_t bucket_set_bytes() const;
  void ensure_tiles(const Shape& s);
  void enqueue(const Shape& s, int n_steps, const float* noise, size_t stride);
  Shape resolve_for(Shape s) const;
};

}  // namespace tissue

// the C-ABI handles: a single simulation is the one-replica case of the engine
struct tissue_sim {
  std::unique_ptr<tissue::Engine> e;
};
struct tissue_batch {
  std::unique_ptr<tissue::Engine> e;
};
