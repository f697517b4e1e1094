// C-ABI of the batched persistent path (include/tissue.h, tissue_batch_*): many independent
// replicas, one workgroup each, launches captured in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "capi_common.h"
#include "rng.h"

namespace {

constexpr int kDefaultStepsPerLaunch = 100;

// autotuned launch shapes, keyed by (largest replica, replica count, mesh vertices)
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, std::pair<int, int>> g_tuned;

unsigned host_threads() {
  unsigned t = std::thread::hardware_concurrency();
  return std::max(1u, std::min(t ? t : 1u, 16u));
}

}  // namespace

struct tissue_batch {
  tissue_mesh* mesh = nullptr;
  tdev::MeshDev md{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int R = 0, ntot = 0, nmax = 0;
  std::vector<int> n, off;
  std::vector<tissue_params> p;
  tissue::ReplicaDesc* d_reps = nullptr;
  float4* xycs = nullptr;
  float* th = nullptr;
  int* vid = nullptr;
  float4* sxycs = nullptr;  // per-step staging of the persistent kernel
  float* sth = nullptr;
  int* svid = nullptr;
  float *vx = nullptr, *vy = nullptr, *r3x = nullptr, *r3y = nullptr, *r3z = nullptr;
  int *face = nullptr, *nbr = nullptr;
  unsigned* flags = nullptr;
  int* first_bad = nullptr;
  float* noise = nullptr;
  int noise_rows = 0;
  double* phi = nullptr;
  size_t phi_cap = 0;
  std::vector<std::vector<double>> hist_ref, hist_polar;
  std::vector<unsigned> flags_acc;
  std::vector<int> first_bad_acc;
  long steps_done = 0;
  int lanes = 0, block = 0, spl = 0;
  bool has_state = false;
  double last_ms = 0.0;
  hipGraphExec_t gexec = nullptr;
  std::tuple<int, int, int, int, int> gkey{-1, -1, -1, -1, -1};

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gkey = {-1, -1, -1, -1, -1};
  }

  void release() {
    (void)hipSetDevice(device);
    drop_graph();
    dfree(d_reps); dfree(xycs); dfree(th); dfree(vid); dfree(sxycs); dfree(sth); dfree(svid);
    dfree(vx); dfree(vy); dfree(r3x); dfree(r3y); dfree(r3z); dfree(face); dfree(nbr);
    dfree(flags); dfree(first_bad); dfree(noise); dfree(phi);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr; ev0 = ev1 = nullptr;
  }

  int eff_lanes() const {
    if (lanes > 0) return lanes;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find({nmax, R, md.nv});
    if (it != g_tuned.end()) return it->second.first;
    return tissue::resident_default_lanes(nmax);
  }
  int eff_block() const {
    if (block > 0) return block;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find({nmax, R, md.nv});
    if (it != g_tuned.end()) return it->second.second;
    return tissue::kResidentMaxBlock;
  }
  int eff_spl() const { return spl > 0 ? spl : kDefaultStepsPerLaunch; }
};

namespace {

void check_batch(tissue_batch* b) {
  if (!b) throw ApiError(TISSUE_E_INVALID_ARG, "null batch handle");
  HIP_OK(hipSetDevice(b->device));
}

// enqueue the launches of one run (no host synchronisation inside, so it can be captured)
void enqueue_run(tissue_batch* b, int n_steps, const float* noise_dev, int lanes, int block, int spl) {
  tissue::ResidentIO io;
  io.xycs = b->xycs; io.th = b->th; io.vid = b->vid;
  io.stage_xycs = b->sxycs; io.stage_th = b->sth; io.stage_vid = b->svid;
  io.vx = b->vx; io.vy = b->vy; io.r3x = b->r3x; io.r3y = b->r3y; io.r3z = b->r3z;
  io.face = b->face; io.nbr = b->nbr;
  io.noise_stride = (size_t)b->ntot;
  io.phi = b->phi;
  io.phi_stride = n_steps;
  io.flags = b->flags;
  io.first_bad = b->first_bad;
  for (int t0 = 0; t0 < n_steps; t0 += spl) {
    const int k = std::min(spl, n_steps - t0);
    io.noise = noise_dev ? noise_dev + (size_t)t0 * b->ntot : nullptr;
    io.t_base = t0;
    HIP_OK(tissue::launch_resident(lanes, block, b->md, b->d_reps, b->R, b->nmax, k, io, b->stream));
  }
}

void reset_flags(tissue_batch* b) {
  std::vector<unsigned> z(b->R, 0u);
  std::vector<int> big(b->R, INT_MAX);
  HIP_OK(hipMemcpyAsync(b->flags, z.data(), b->R * sizeof(unsigned), hipMemcpyHostToDevice, b->stream));
  HIP_OK(hipMemcpyAsync(b->first_bad, big.data(), b->R * sizeof(int), hipMemcpyHostToDevice, b->stream));
}

void ensure_phi(tissue_batch* b, int n_steps) {
  const size_t need = 2 * (size_t)b->R * n_steps;
  if (b->phi_cap < need) {
    HIP_OK(hipStreamSynchronize(b->stream));
    b->drop_graph();
    dfree(b->phi);
    b->phi = dalloc<double>(need);
    b->phi_cap = need;
  }
}

int32_t run(tissue_batch* b, int n_steps, int first_row) {
  if (!b->has_state) throw ApiError(TISSUE_E_INVALID_ARG, "no state: call tissue_batch_init_random or _set_state");
  if (n_steps < 0) throw ApiError(TISSUE_E_INVALID_ARG, "negative step count");
  if (first_row >= 0 && (long)first_row + n_steps > b->noise_rows)
    throw ApiError(TISSUE_E_INVALID_ARG, "staged noise rows out of range");
  if (n_steps == 0) return TISSUE_OK;
  ensure_phi(b, n_steps);
  const int lanes = b->eff_lanes(), block = b->eff_block(), spl = b->eff_spl();
  const float* nz = first_row >= 0 ? b->noise + (size_t)first_row * b->ntot : nullptr;
  reset_flags(b);
  const auto key = std::make_tuple(n_steps, first_row, lanes, block, spl);
  if (!b->gexec || b->gkey != key) {
    b->drop_graph();
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_run(b, n_steps, nz, lanes, block, spl);
    } catch (...) {
      (void)hipStreamEndCapture(b->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_OK(hipStreamEndCapture(b->stream, &g));
    hipError_t e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_OK(e);
    b->gkey = key;
  }
  HIP_OK(hipEventRecord(b->ev0, b->stream));
  HIP_OK(hipGraphLaunch(b->gexec, b->stream));
  HIP_OK(hipEventRecord(b->ev1, b->stream));
  std::vector<double> phi(2 * (size_t)b->R * n_steps);
  std::vector<unsigned> fl(b->R);
  std::vector<int> fb(b->R);
  HIP_OK(hipMemcpyAsync(phi.data(), b->phi, phi.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIP_OK(hipMemcpyAsync(fl.data(), b->flags, b->R * sizeof(unsigned), hipMemcpyDeviceToHost, b->stream));
  HIP_OK(hipMemcpyAsync(fb.data(), b->first_bad, b->R * sizeof(int), hipMemcpyDeviceToHost, b->stream));
  HIP_OK(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  unsigned all = 0;
  for (int r = 0; r < b->R; ++r) {
    for (int t = 0; t < n_steps; ++t) {
      b->hist_ref[r].push_back(phi[2 * ((size_t)r * n_steps + t)]);
      b->hist_polar[r].push_back(phi[2 * ((size_t)r * n_steps + t) + 1]);
    }
    if (fl[r]) {
      b->flags_acc[r] |= fl[r];
      const int abs_step = (int)std::min<long>(b->steps_done + fb[r], INT_MAX - 1);
      b->first_bad_acc[r] = std::min(b->first_bad_acc[r], abs_step);
      all |= fl[r];
    }
  }
  b->steps_done += n_steps;
  if (all & TISSUE_FLAG_NONFINITE) return TISSUE_E_NONFINITE;
  if (all & TISSUE_FLAG_TILING) return TISSUE_E_TILING_NO_CONVERGE;
  if (all & TISSUE_FLAG_LOST_PARTICLE) return TISSUE_E_LOST_PARTICLE;
  return TISSUE_OK;
}

void fill_noise(tissue_batch* b, const uint64_t* seeds, const double* stds, int n_rows, float* out) {
  const unsigned nt = host_threads();
  std::vector<std::thread> th;
  for (unsigned w = 0; w < nt; ++w)
    th.emplace_back([&, w] {
      for (int t = (int)w; t < n_rows; t += (int)nt)
        for (int r = 0; r < b->R; ++r)
          for (int i = 0; i < b->n[r]; ++i)
            out[(size_t)t * b->ntot + b->off[r] + i] =
                (float)(stds[r] * tissue::normal_draw(seeds[r], (uint32_t)i, (uint32_t)t));
    });
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

int32_t tissue_batch_create(const tissue_mesh* mc, int32_t n_replicas, const int32_t* n, const tissue_params* params,
                            tissue_batch** out) {
  return guarded([&] {
    if (!mc || !n || !params || !out || n_replicas < 1) throw ApiError(TISSUE_E_INVALID_ARG, "bad batch arguments");
    auto* m = const_cast<tissue_mesh*>(mc);
    auto* b = new tissue_batch;
    try {
      HIP_OK(hipGetDevice(&b->device));
      b->mesh = m;
      b->md = m->device_view();
      b->R = n_replicas;
      const int cap = tissue::resident_max_particles(b->md.nv);
      std::vector<tissue::ReplicaDesc> reps(n_replicas);
      long tot = 0;
      for (int r = 0; r < n_replicas; ++r) {
        if (n[r] < 1 || n[r] > cap) throw ApiError(TISSUE_E_INVALID_ARG, "replica size outside [1, resident capacity]");
        check_params(&params[r], m->sigma_max);
        b->n.push_back(n[r]);
        b->off.push_back((int)tot);
        b->p.push_back(params[r]);
        reps[r] = tissue::ReplicaDesc{(int)tot, n[r], make_step_params(params[r])};
        tot += n[r];
        b->nmax = std::max(b->nmax, n[r]);
      }
      if (tot > INT_MAX / 2) throw ApiError(TISSUE_E_INVALID_ARG, "too many particles");
      b->ntot = (int)tot;
      HIP_OK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
      HIP_OK(hipEventCreate(&b->ev0));
      HIP_OK(hipEventCreate(&b->ev1));
      HIP_OK(tissue::resident_prepare(b->md.nv, b->nmax));
      b->d_reps = dalloc<tissue::ReplicaDesc>(n_replicas);
      upload(b->d_reps, reps.data(), reps.size());
      b->xycs = dalloc<float4>(b->ntot); b->th = dalloc<float>(b->ntot); b->vid = dalloc<int>(b->ntot);
      b->sxycs = dalloc<float4>(b->ntot); b->sth = dalloc<float>(b->ntot); b->svid = dalloc<int>(b->ntot);
      b->vx = dalloc<float>(b->ntot); b->vy = dalloc<float>(b->ntot);
      b->r3x = dalloc<float>(b->ntot); b->r3y = dalloc<float>(b->ntot); b->r3z = dalloc<float>(b->ntot);
      b->face = dalloc<int>(b->ntot); b->nbr = dalloc<int>(b->ntot);
      b->flags = dalloc<unsigned>(n_replicas);
      b->first_bad = dalloc<int>(n_replicas);
      b->hist_ref.resize(n_replicas);
      b->hist_polar.resize(n_replicas);
      b->flags_acc.assign(n_replicas, 0u);
      b->first_bad_acc.assign(n_replicas, INT_MAX);
    } catch (...) {
      b->release();
      delete b;
      throw;
    }
    *out = b;
    return TISSUE_OK;
  });
}

void tissue_batch_destroy(tissue_batch* b) {
  if (!b) return;
  b->release();
  delete b;
}

int32_t tissue_batch_set_state(tissue_batch* b, const float* x, const float* y, const float* theta, const int32_t* vid) {
  return guarded([&] {
    check_batch(b);
    if (!x || !y || !theta) throw ApiError(TISSUE_E_INVALID_ARG, "null state array");
    const int n = b->ntot;
    for (int i = 0; i < n; ++i) {
      if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(theta[i]))
        throw ApiError(TISSUE_E_INVALID_ARG, "non-finite state value");
      if (x[i] < 0.f || x[i] > 1.f || y[i] < 0.f || y[i] > 1.f)
        throw ApiError(TISSUE_E_INVALID_ARG, "position outside the UV square");
      if (vid && (vid[i] < 0 || vid[i] >= b->md.nv)) throw ApiError(TISSUE_E_INVALID_ARG, "vertex id out of range");
    }
    float *dx = dalloc<float>(n), *dy = dalloc<float>(n), *dt = dalloc<float>(n);
    try {
      HIP_OK(hipMemcpyAsync(dx, x, n * sizeof(float), hipMemcpyHostToDevice, b->stream));
      HIP_OK(hipMemcpyAsync(dy, y, n * sizeof(float), hipMemcpyHostToDevice, b->stream));
      HIP_OK(hipMemcpyAsync(dt, theta, n * sizeof(float), hipMemcpyHostToDevice, b->stream));
      HIP_OK(tissue::launch_prepare(n, dx, dy, dt, b->xycs, b->th, b->stream));
      HIP_OK(tissue::launch_locate(b->md, n, b->xycs, b->vid, b->face, b->r3x, b->r3y, b->r3z, b->stream));
      if (vid) HIP_OK(hipMemcpyAsync(b->vid, vid, n * sizeof(int), hipMemcpyHostToDevice, b->stream));
      HIP_OK(hipMemsetAsync(b->vx, 0, n * sizeof(float), b->stream));
      HIP_OK(hipMemsetAsync(b->vy, 0, n * sizeof(float), b->stream));
      HIP_OK(hipMemsetAsync(b->nbr, 0, n * sizeof(int), b->stream));
      HIP_OK(hipStreamSynchronize(b->stream));
    } catch (...) {
      dfree(dx); dfree(dy); dfree(dt);
      throw;
    }
    dfree(dx); dfree(dy); dfree(dt);
    b->has_state = true;
    return TISSUE_OK;
  });
}

int32_t tissue_batch_init_random(tissue_batch* b, const uint64_t* seeds) {
  return guarded([&] {
    check_batch(b);
    if (!seeds) throw ApiError(TISSUE_E_INVALID_ARG, "null seeds");
    std::vector<float> x(b->ntot), y(b->ntot), th(b->ntot);
    for (int r = 0; r < b->R; ++r)
      random_init(b->mesh->m, seeds[r], b->n[r], x.data() + b->off[r], y.data() + b->off[r], th.data() + b->off[r]);
    return tissue_batch_set_state(b, x.data(), y.data(), th.data(), nullptr);
  });
}

int32_t tissue_batch_get_state(tissue_batch* b, float* x, float* y, float* theta, int32_t* vid) {
  return guarded([&] {
    check_batch(b);
    const int n = b->ntot;
    std::vector<float4> p(n);
    HIP_OK(hipMemcpyAsync(p.data(), b->xycs, n * sizeof(float4), hipMemcpyDeviceToHost, b->stream));
    if (theta) HIP_OK(hipMemcpyAsync(theta, b->th, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    if (vid) HIP_OK(hipMemcpyAsync(vid, b->vid, n * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    for (int i = 0; i < n; ++i) {
      if (x) x[i] = p[i].x;
      if (y) y[i] = p[i].y;
    }
    return TISSUE_OK;
  });
}

int32_t tissue_batch_get_outputs(tissue_batch* b, float* vx, float* vy, float* r3d, int32_t* face, int32_t* nbr) {
  return guarded([&] {
    check_batch(b);
    const int n = b->ntot;
    std::vector<float> a(n), c(n), d(n);
    if (vx) HIP_OK(hipMemcpyAsync(vx, b->vx, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    if (vy) HIP_OK(hipMemcpyAsync(vy, b->vy, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    if (face) HIP_OK(hipMemcpyAsync(face, b->face, n * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    if (nbr) HIP_OK(hipMemcpyAsync(nbr, b->nbr, n * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    if (r3d) {
      HIP_OK(hipMemcpyAsync(a.data(), b->r3x, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
      HIP_OK(hipMemcpyAsync(c.data(), b->r3y, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
      HIP_OK(hipMemcpyAsync(d.data(), b->r3z, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    }
    HIP_OK(hipStreamSynchronize(b->stream));
    if (r3d)
      for (int i = 0; i < n; ++i) { r3d[3 * i] = a[i]; r3d[3 * i + 1] = c[i]; r3d[3 * i + 2] = d[i]; }
    return TISSUE_OK;
  });
}

int32_t tissue_batch_stage_noise(tissue_batch* b, const float* noise, int32_t n_rows) {
  return guarded([&] {
    check_batch(b);
    if (!noise || n_rows < 1) throw ApiError(TISSUE_E_INVALID_ARG, "bad noise rows");
    if (n_rows > b->noise_rows) {
      HIP_OK(hipStreamSynchronize(b->stream));
      b->drop_graph();
      dfree(b->noise);
      b->noise = dalloc<float>((size_t)n_rows * b->ntot);
      b->noise_rows = n_rows;
    }
    HIP_OK(hipMemcpyAsync(b->noise, noise, (size_t)n_rows * b->ntot * sizeof(float), hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    return TISSUE_OK;
  });
}

int32_t tissue_batch_gen_noise(tissue_batch* b, const uint64_t* seeds, const double* stds, int32_t n_rows) {
  return guarded([&] {
    check_batch(b);
    if (!seeds || !stds || n_rows < 1) throw ApiError(TISSUE_E_INVALID_ARG, "bad noise arguments");
    std::vector<float> h((size_t)n_rows * b->ntot);
    fill_noise(b, seeds, stds, n_rows, h.data());
    return tissue_batch_stage_noise(b, h.data(), n_rows);
  });
}

int32_t tissue_batch_run(tissue_batch* b, int32_t n_steps, int32_t first_row) {
  return guarded([&] {
    check_batch(b);
    return run(b, n_steps, first_row);
  });
}

int32_t tissue_batch_get_order(tissue_batch* b, int32_t replica, double* phi_ref, double* phi_polar, int32_t cap,
                               int32_t* count) {
  return guarded([&] {
    if (!b || replica < 0 || replica >= b->R) throw ApiError(TISSUE_E_INVALID_ARG, "bad replica index");
    const auto& a = b->hist_ref[replica];
    const auto& c = b->hist_polar[replica];
    const int k = (int)std::min<size_t>(a.size(), (size_t)std::max(cap, 0));
    if (phi_ref) std::copy(a.begin(), a.begin() + k, phi_ref);
    if (phi_polar) std::copy(c.begin(), c.begin() + k, phi_polar);
    if (count) *count = (int32_t)a.size();
    return TISSUE_OK;
  });
}

int32_t tissue_batch_get_flags(tissue_batch* b, uint32_t* flags, int32_t* first_step) {
  return guarded([&] {
    if (!b) throw ApiError(TISSUE_E_INVALID_ARG, "null batch handle");
    for (int r = 0; r < b->R; ++r) {
      if (flags) flags[r] = b->flags_acc[r];
      if (first_step) first_step[r] = b->flags_acc[r] ? b->first_bad_acc[r] : -1;
      b->flags_acc[r] = 0;
      b->first_bad_acc[r] = INT_MAX;
    }
    return TISSUE_OK;
  });
}

int32_t tissue_batch_set_launch(tissue_batch* b, int32_t lanes, int32_t block, int32_t steps_per_launch) {
  return guarded([&] {
    if (!b) throw ApiError(TISSUE_E_INVALID_ARG, "null batch handle");
    if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8)
      throw ApiError(TISSUE_E_INVALID_ARG, "lanes must be 0, 1, 2, 4 or 8");
    if (block != 0 && (block < 64 || block > tissue::kResidentMaxBlock || block % 64))
      throw ApiError(TISSUE_E_INVALID_ARG, "block must be a multiple of 64 in [64, 512]");
    if (steps_per_launch < 0) throw ApiError(TISSUE_E_INVALID_ARG, "negative steps per launch");
    b->lanes = lanes;
    b->block = block;
    b->spl = steps_per_launch;
    return TISSUE_OK;
  });
}

int32_t tissue_batch_autotune(tissue_batch* b, int32_t n_steps, int32_t* lanes_out, int32_t* block_out) {
  return guarded([&] {
    check_batch(b);
    if (!b->has_state) throw ApiError(TISSUE_E_INVALID_ARG, "no state to tune on");
    if (n_steps < 1) throw ApiError(TISSUE_E_INVALID_ARG, "autotune needs at least one step");
    const int n = b->ntot;
    float4* sx = dalloc<float4>(n);
    float* st = dalloc<float>(n);
    int* sv = dalloc<int>(n);
    const int keep_l = b->lanes, keep_b = b->block;
    const long keep_steps = b->steps_done;
    std::vector<size_t> keep_hist(b->R);
    for (int r = 0; r < b->R; ++r) keep_hist[r] = b->hist_ref[r].size();
    double best = 1e300;
    int bl = 0, bb = 0;
    try {
      HIP_OK(hipMemcpyAsync(sx, b->xycs, n * sizeof(float4), hipMemcpyDeviceToDevice, b->stream));
      HIP_OK(hipMemcpyAsync(st, b->th, n * sizeof(float), hipMemcpyDeviceToDevice, b->stream));
      HIP_OK(hipMemcpyAsync(sv, b->vid, n * sizeof(int), hipMemcpyDeviceToDevice, b->stream));
      const int row = b->noise_rows >= n_steps ? 0 : -1;
      for (int l : {1, 2, 4, 8})
        for (int blk : {256, 512}) {
          b->lanes = l;
          b->block = blk;
          try {
            run(b, n_steps, row);
          } catch (const ApiError&) {
            continue;  // this shape does not launch here (registers or LDS); skip it
          }
          if (b->last_ms < best) { best = b->last_ms; bl = l; bb = blk; }
          HIP_OK(hipMemcpyAsync(b->xycs, sx, n * sizeof(float4), hipMemcpyDeviceToDevice, b->stream));
          HIP_OK(hipMemcpyAsync(b->th, st, n * sizeof(float), hipMemcpyDeviceToDevice, b->stream));
          HIP_OK(hipMemcpyAsync(b->vid, sv, n * sizeof(int), hipMemcpyDeviceToDevice, b->stream));
        }
      HIP_OK(hipStreamSynchronize(b->stream));
    } catch (...) {
      dfree(sx); dfree(st); dfree(sv);
      b->lanes = keep_l; b->block = keep_b;
      throw;
    }
    dfree(sx); dfree(st); dfree(sv);
    if (bl == 0) throw ApiError(TISSUE_E_HIP, "no launch shape ran");
    for (int r = 0; r < b->R; ++r) {
      b->hist_ref[r].resize(keep_hist[r]);
      b->hist_polar[r].resize(keep_hist[r]);
      b->flags_acc[r] = 0;
      b->first_bad_acc[r] = INT_MAX;
    }
    b->steps_done = keep_steps;
    b->lanes = keep_l;
    b->block = keep_b;
    b->drop_graph();
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tuned[{b->nmax, b->R, b->md.nv}] = {bl, bb};
    }
    if (lanes_out) *lanes_out = bl;
    if (block_out) *block_out = bb;
    return TISSUE_OK;
  });
}

int32_t tissue_batch_last_run_ms(tissue_batch* b, double* ms) {
  return guarded([&] {
    if (!b || !ms) throw ApiError(TISSUE_E_INVALID_ARG, "null argument");
    *ms = b->last_ms;
    return TISSUE_OK;
  });
}

}  // extern "C"
