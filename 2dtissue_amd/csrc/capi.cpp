// C-ABI (include/tissue.h): mesh objects, simulation handles, the per-step launch sequence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_common.h"
#include "rng.h"

using tissue::OrderPartial;
using tissue::StepIO;
using namespace tcapi;

thread_local std::string tcapi::g_err;

namespace {

}  // namespace

void tcapi::random_init(const tissue::UVMesh& M, uint64_t seed, int n, float* x, float* y, float* th) {
  std::vector<double> cdf(M.nf);
  double tot = 0;
  for (int f = 0; f < M.nf; ++f) {
    const double* A = &M.xyz[3 * M.faces[3 * f]];
    const double* B = &M.xyz[3 * M.faces[3 * f + 1]];
    const double* C = &M.xyz[3 * M.faces[3 * f + 2]];
    double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
    tot += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    cdf[f] = tot;
  }
  for (int i = 0; i < n; ++i) {
    const tissue::Draw4 d = tissue::philox_draw(seed, (uint32_t)i, 0u, 0u);
    const double u0 = tissue::u53(d.w[0], d.w[1]);
    int f = (int)(std::upper_bound(cdf.begin(), cdf.end(), u0 * tot) - cdf.begin());
    f = std::min(f, M.nf - 1);
    const tissue::Draw4 e = tissue::philox_draw(seed, (uint32_t)i, 0u, 2u);
    double r1 = tissue::u53(d.w[2], d.w[3]), r2 = tissue::u53(e.w[0], e.w[1]);
    if (r1 + r2 > 1.0) { r1 = 1.0 - r1; r2 = 1.0 - r2; }
    const double* a = &M.uv[2 * M.faces[3 * f]];
    const double* b = &M.uv[2 * M.faces[3 * f + 1]];
    const double* c = &M.uv[2 * M.faces[3 * f + 2]];
    x[i] = (float)std::clamp(a[0] + r1 * (b[0] - a[0]) + r2 * (c[0] - a[0]), 0.0, 1.0);
    y[i] = (float)std::clamp(a[1] + r1 * (b[1] - a[1]) + r2 * (c[1] - a[1]), 0.0, 1.0);
    double t = 2.0 * M_PI * tissue::u53(e.w[2], e.w[3]);
    t = t - 2.0 * M_PI * std::floor((t + M_PI) / (2.0 * M_PI));
    th[i] = (float)t;
  }
}

namespace {

int32_t finish_mesh(tissue::SurfaceMesh& s, tissue_mesh** out) {
  if (!out) throw ApiError(TISSUE_E_INVALID_ARG, "null output pointer");
  auto* tm = new tissue_mesh;
  try {
    tm->m = tissue::cut_and_parameterize(s);
  } catch (...) {
    delete tm;
    throw;
  }
  *out = tm;
  return TISSUE_OK;
}

}  // namespace

extern "C" {

const char* tissue_error_string(void) { return g_err.c_str(); }

int32_t tissue_mesh_ellipsoid(double a, double b, double c, int32_t freq, tissue_mesh** out) {
  return guarded([&] {
    if (!(a > 0 && b > 0 && c > 0) || freq < 1 || freq > 2000) throw ApiError(TISSUE_E_INVALID_ARG, "bad ellipsoid arguments");
    tissue::SurfaceMesh s = tissue::icosphere(freq);
    tissue::scale_axes(s, a, b, c);
    return finish_mesh(s, out);
  });
}

int32_t tissue_mesh_perturbed_sphere(double amp, int32_t freq, tissue_mesh** out) {
  return guarded([&] {
    if (!(std::fabs(amp) < 0.5) || freq < 1 || freq > 2000) throw ApiError(TISSUE_E_INVALID_ARG, "bad sphere arguments");
    tissue::SurfaceMesh s = tissue::icosphere(freq);
    tissue::perturb_y32(s, amp);
    return finish_mesh(s, out);
  });
}

int32_t tissue_mesh_load_off(const char* path, tissue_mesh** out) {
  return guarded([&] {
    if (!path) throw ApiError(TISSUE_E_INVALID_ARG, "null path");
    tissue::SurfaceMesh s;
    std::string err;
    if (!tissue::load_off(path, s, err)) throw ApiError(TISSUE_E_IO, err);
    return finish_mesh(s, out);
  });
}

int32_t tissue_mesh_build_tables(tissue_mesh* m, double sigma_max, int32_t grid) {
  return guarded([&] {
    if (!m || !(sigma_max > 0) || !std::isfinite(sigma_max)) throw ApiError(TISSUE_E_INVALID_ARG, "bad table arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    if (m->dev_ready) throw ApiError(TISSUE_E_INVALID_ARG, "mesh tables are in use on a device");
    tissue::build_face_buckets(m->m, grid);
    tissue::build_distance_table(m->m, 2.4 * sigma_max);
    m->sigma_max = sigma_max;
    m->tables = true;
    return TISSUE_OK;
  });
}

void tissue_mesh_destroy(tissue_mesh* m) {
  if (!m) return;
  m->free_device();
  delete m;
}

int32_t tissue_mesh_info(const tissue_mesh* m, int32_t* nv, int32_t* nf, int32_t* nv3d, int64_t* nnz, int32_t* grid,
                         int64_t* n_cell_faces) {
  return guarded([&] {
    if (!m) throw ApiError(TISSUE_E_INVALID_ARG, "null mesh");
    if (nv) *nv = m->m.nv;
    if (nf) *nf = m->m.nf;
    if (nv3d) *nv3d = m->m.nv3d;
    if (nnz) *nnz = (int64_t)m->m.col.size();
    if (grid) *grid = m->m.grid;
    if (n_cell_faces) *n_cell_faces = (int64_t)m->m.cell_faces.size();
    return TISSUE_OK;
  });
}

int32_t tissue_mesh_export(const tissue_mesh* m, double* uv, double* xyz, int32_t* faces, int32_t* row_ptr,
                           int32_t* col, float* dist, int32_t* cell_start, int32_t* cell_faces) {
  return guarded([&] {
    if (!m) throw ApiError(TISSUE_E_INVALID_ARG, "null mesh");
    const auto& M = m->m;
    auto cp = [](auto* dst, const auto& v) {
      if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(uv, M.uv); cp(xyz, M.xyz); cp(faces, M.faces); cp(row_ptr, M.row_ptr); cp(col, M.col); cp(dist, M.dist);
    cp(cell_start, M.cell_start); cp(cell_faces, M.cell_faces);
    return TISSUE_OK;
  });
}

int32_t tissue_mesh_save_off(const tissue_mesh* m, const char* path_3d, const char* path_uv) {
  return guarded([&] {
    if (!m) throw ApiError(TISSUE_E_INVALID_ARG, "null mesh");
    std::string err;
    const auto& M = m->m;
    if (path_3d && !tissue::save_off(path_3d, M.xyz.data(), M.nv, M.faces.data(), M.nf, 3, err)) throw ApiError(TISSUE_E_IO, err);
    if (path_uv && !tissue::save_off(path_uv, M.uv.data(), M.nv, M.faces.data(), M.nf, 2, err)) throw ApiError(TISSUE_E_IO, err);
    return TISSUE_OK;
  });
}

}  // extern "C"

// ------------------------------------------------------------------ simulation

struct tissue_sim {
  tissue_mesh* mesh = nullptr;
  tdev::MeshDev md{};
  tissue_params p{};
  tdev::StepParams sp{};
  int n = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float4* xycs[2] = {nullptr, nullptr};
  float* th[2] = {nullptr, nullptr};
  int* vid[2] = {nullptr, nullptr};
  float *vx = nullptr, *vy = nullptr, *r3x = nullptr, *r3y = nullptr, *r3z = nullptr;
  int *face = nullptr, *nbr = nullptr;
  int* bcount[3] = {nullptr, nullptr, nullptr};
  int* bslot[3] = {nullptr, nullptr, nullptr};
  int* bovf_n[3] = {nullptr, nullptr, nullptr};
  int2* bovf[3] = {nullptr, nullptr, nullptr};
  unsigned* flags = nullptr;
  int* first_bad = nullptr;
  float* noise = nullptr;
  size_t noise_cap = 0;
  int staged_rows = 0;
  OrderPartial* partial = nullptr;
  size_t partial_cap = 0;
  double* phi = nullptr;
  size_t phi_cap = 0;
  std::vector<double> hist_ref, hist_polar;
  unsigned flags_acc = 0;
  int first_bad_acc = INT_MAX;
  long steps_done = 0;
  int sphase = 0;  // state buffer holding the current state
  int bphase = 0;  // bucket set holding the current state's buckets
  int lanes = 0, block = 0;
  int mode = TISSUE_MODE_AUTO;
  int resident_cap = 0;  // largest n the persistent kernel holds for this mesh
  tissue::ReplicaDesc* d_rep = nullptr;
  bool buckets_stale = false;
  bool has_state = false;
  double last_ms = 0.0;

  // AUTO takes the persistent kernel only for systems that one pass of a workgroup covers: beyond
  // that, each extra pass pays the slowest lane's seam-wrap / point-location chain again, and
  // the one-launch-per-step kernel (one wave per CU) is faster (N = 2,000: 41 us vs 320 us)
  static constexpr int kAutoResidentMax = 256;
  bool use_resident() const {
    if (mode == TISSUE_MODE_STEP_KERNEL) return false;
    if (mode == TISSUE_MODE_AUTO && n > kAutoResidentMax) return false;
    return n <= resident_cap;
  }
  int res_lanes() const { return lanes > 0 ? std::min(lanes, 8) : tissue::resident_default_lanes(n); }
  int res_block() const { return block > 0 ? std::min(block, tissue::kResidentMaxBlock) : tissue::kResidentMaxBlock; }

  tdev::BucketsIn bin(int k) const { return {bcount[k], bslot[k], bovf_n[k], bovf[k]}; }
  tdev::BucketsOut bout(int k) const { return {bcount[k], bslot[k], bovf_n[k], bovf[k]}; }

  void release() {
    (void)hipSetDevice(device);
    for (int k = 0; k < 2; ++k) { dfree(xycs[k]); dfree(th[k]); dfree(vid[k]); }
    for (int k = 0; k < 3; ++k) { dfree(bcount[k]); dfree(bslot[k]); dfree(bovf_n[k]); dfree(bovf[k]); }
    dfree(vx); dfree(vy); dfree(r3x); dfree(r3y); dfree(r3z); dfree(face); dfree(nbr);
    dfree(flags); dfree(first_bad); dfree(noise); dfree(partial); dfree(phi); dfree(d_rep);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr; ev0 = ev1 = nullptr;
  }

  int eff_lanes() const {
    if (lanes > 0) return lanes;
    if (n <= 8192) return 8;
    if (n <= 65536) return 4;
    return 2;
  }
  int eff_block() const { return block > 0 ? block : 256; }
};

namespace {

void check_sim(tissue_sim* s) {
  if (!s) throw ApiError(TISSUE_E_INVALID_ARG, "null simulation handle");
  HIP_OK(hipSetDevice(s->device));
}

void rebuild_buckets(tissue_sim* s) {
  for (int k = 0; k < 3; ++k) HIP_OK(tissue::launch_clear(s->md.nv, s->bcount[k], s->bovf_n[k], s->stream));
  HIP_OK(tissue::launch_insert(s->n, s->vid[s->sphase], s->bout(s->bphase), s->stream));
}

int32_t status_from_flags(unsigned f) {
  if (f & TISSUE_FLAG_NONFINITE) return TISSUE_E_NONFINITE;
  if (f & TISSUE_FLAG_TILING) return TISSUE_E_TILING_NO_CONVERGE;
  if (f & TISSUE_FLAG_LOST_PARTICLE) return TISSUE_E_LOST_PARTICLE;
  return TISSUE_OK;
}

// steps per launch of the persistent kernel (bounds one kernel's duration)
constexpr int kResidentStepsPerLaunch = 256;

// collect the order parameters and flags of a run_steps call; first_bad is relative to the call
// for the resident kernel and absolute for the per-step kernel
int32_t finish_steps(tissue_sim* s, int n_steps, bool relative_first_bad) {
  std::vector<double> phi(2 * (size_t)n_steps);
  unsigned fl = 0;
  int fb = INT_MAX;
  HIP_OK(hipMemcpyAsync(phi.data(), s->phi, phi.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipMemcpyAsync(&fl, s->flags, sizeof fl, hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipMemcpyAsync(&fb, s->first_bad, sizeof fb, hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  s->last_ms = ms;
  for (int t = 0; t < n_steps; ++t) {
    s->hist_ref.push_back(phi[2 * t]);
    s->hist_polar.push_back(phi[2 * t + 1]);
  }
  if (fl && relative_first_bad) fb = (int)std::min<long>(s->steps_done + fb, INT_MAX - 1);
  s->steps_done += n_steps;
  s->flags_acc |= fl;
  if (fl && fb < s->first_bad_acc) s->first_bad_acc = fb;
  const int32_t st = status_from_flags(fl);
  if (st != TISSUE_OK) {
    char buf[128];
    std::snprintf(buf, sizeof buf, "step %d raised validation flags 0x%x", fb, fl);
    g_err = buf;
  }
  return st;
}

int32_t run_steps(tissue_sim* s, int n_steps, const float* noise_dev, size_t noise_stride) {
  if (!s->has_state) throw ApiError(TISSUE_E_INVALID_ARG, "no state: call tissue_init_random or tissue_set_state");
  if (n_steps < 0) throw ApiError(TISSUE_E_INVALID_ARG, "negative step count");
  if (n_steps == 0) return TISSUE_OK;
  if (s->phi_cap < 2 * (size_t)n_steps) {
    dfree(s->phi);
    s->phi = dalloc<double>(2 * (size_t)n_steps);
    s->phi_cap = 2 * (size_t)n_steps;
  }
  const unsigned zero = 0;
  const int big = INT_MAX;
  HIP_OK(hipMemcpyAsync(s->flags, &zero, sizeof zero, hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipMemcpyAsync(s->first_bad, &big, sizeof big, hipMemcpyHostToDevice, s->stream));
  if (s->use_resident()) {
    // persistent LDS-resident kernel, in place on the current state buffers
    const int lanes = s->res_lanes(), block = s->res_block();
    tissue::ResidentIO io;
    const int cur = s->sphase;
    io.xycs = s->xycs[cur]; io.th = s->th[cur]; io.vid = s->vid[cur];
    io.stage_xycs = s->xycs[1 - cur]; io.stage_th = s->th[1 - cur]; io.stage_vid = s->vid[1 - cur];
    io.vx = s->vx; io.vy = s->vy; io.r3x = s->r3x; io.r3y = s->r3y; io.r3z = s->r3z;
    io.face = s->face; io.nbr = s->nbr;
    io.noise_stride = noise_stride;
    io.phi = s->phi;
    io.phi_stride = n_steps;
    io.flags = s->flags;
    io.first_bad = s->first_bad;
    HIP_OK(hipEventRecord(s->ev0, s->stream));
    for (int t0 = 0; t0 < n_steps; t0 += kResidentStepsPerLaunch) {
      const int k = std::min(kResidentStepsPerLaunch, n_steps - t0);
      io.noise = noise_dev ? noise_dev + (size_t)t0 * noise_stride : nullptr;
      io.t_base = t0;
      HIP_OK(tissue::launch_resident(lanes, block, s->md, s->d_rep, 1, s->n, k, io, s->stream));
    }
    HIP_OK(hipEventRecord(s->ev1, s->stream));
    s->buckets_stale = true;
    return finish_steps(s, n_steps, true);
  }
  if (s->buckets_stale) {
    rebuild_buckets(s);
    s->buckets_stale = false;
  }
  const int lanes = s->eff_lanes(), block = s->eff_block();
  const int blocks = tissue::step_blocks(s->n, lanes, block);
  const size_t need_partial = (size_t)n_steps * blocks;
  if (s->partial_cap < need_partial) {
    dfree(s->partial);
    s->partial = dalloc<OrderPartial>(need_partial);
    s->partial_cap = need_partial;
  }
  HIP_OK(hipEventRecord(s->ev0, s->stream));
  for (int t = 0; t < n_steps; ++t) {
    const int si = s->sphase, so = 1 - s->sphase;
    const int bi = s->bphase, bo = (s->bphase + 1) % 3, bc = (s->bphase + 2) % 3;
    StepIO io;
    io.xycs_in = s->xycs[si]; io.th_in = s->th[si]; io.vid_in = s->vid[si];
    io.xycs_out = s->xycs[so]; io.th_out = s->th[so]; io.vid_out = s->vid[so];
    io.vx = s->vx; io.vy = s->vy; io.r3x = s->r3x; io.r3y = s->r3y; io.r3z = s->r3z;
    io.face = s->face; io.nbr = s->nbr;
    io.noise = noise_dev ? noise_dev + (size_t)t * noise_stride : nullptr;
    io.partial = s->partial + (size_t)t * blocks;
    io.flags = s->flags;
    io.first_bad = s->first_bad;
    io.step = (int)std::min<long>(s->steps_done + t, INT_MAX - 1);
    HIP_OK(tissue::launch_step(lanes, block, s->md, s->sp, s->n, s->bin(bi), s->bout(bo), s->bcount[bc], s->bovf_n[bc],
                               io, s->stream));
    s->sphase = so;
    s->bphase = bo;
  }
  HIP_OK(hipEventRecord(s->ev1, s->stream));
  HIP_OK(tissue::launch_order_finalize(s->partial, n_steps, blocks, s->n, s->phi, s->stream));
  return finish_steps(s, n_steps, false);
}

}  // namespace

extern "C" {

int32_t tissue_create(const tissue_mesh* mc, const tissue_params* p, int32_t n, tissue_sim** out) {
  return guarded([&] {
    if (!mc || !out) throw ApiError(TISSUE_E_INVALID_ARG, "null argument");
    if (n < 1) throw ApiError(TISSUE_E_INVALID_ARG, "particle count must be positive");
    auto* m = const_cast<tissue_mesh*>(mc);
    check_params(p, m->sigma_max);
    auto* s = new tissue_sim;
    try {
      HIP_OK(hipGetDevice(&s->device));
      s->mesh = m;
      s->md = m->device_view();
      s->p = *p;
      s->sp = make_step_params(*p);
      s->n = n;
      HIP_OK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
      HIP_OK(hipEventCreate(&s->ev0));
      HIP_OK(hipEventCreate(&s->ev1));
      for (int k = 0; k < 2; ++k) {
        s->xycs[k] = dalloc<float4>(n);
        s->th[k] = dalloc<float>(n);
        s->vid[k] = dalloc<int>(n);
      }
      const int nv = s->md.nv;
      for (int k = 0; k < 3; ++k) {
        s->bcount[k] = dalloc<int>(nv);
        s->bslot[k] = dalloc<int>((size_t)nv * tdev::kSlots);
        s->bovf_n[k] = dalloc<int>(1);
        s->bovf[k] = dalloc<int2>(n);
      }
      s->vx = dalloc<float>(n); s->vy = dalloc<float>(n);
      s->r3x = dalloc<float>(n); s->r3y = dalloc<float>(n); s->r3z = dalloc<float>(n);
      s->face = dalloc<int>(n); s->nbr = dalloc<int>(n);
      s->flags = dalloc<unsigned>(1);
      s->first_bad = dalloc<int>(1);
      s->resident_cap = tissue::resident_max_particles(s->md.nv);
      if (n <= s->resident_cap) HIP_OK(tissue::resident_prepare(s->md.nv, n));
      tissue::ReplicaDesc rep{0, n, s->sp};
      s->d_rep = dalloc<tissue::ReplicaDesc>(1);
      upload(s->d_rep, &rep, 1);
    } catch (...) {
      s->release();
      delete s;
      throw;
    }
    *out = s;
    return TISSUE_OK;
  });
}

void tissue_destroy(tissue_sim* s) {
  if (!s) return;
  s->release();
  delete s;
}

int32_t tissue_set_state(tissue_sim* s, const float* x, const float* y, const float* theta, const int32_t* vid) {
  return guarded([&] {
    check_sim(s);
    if (!x || !y || !theta) throw ApiError(TISSUE_E_INVALID_ARG, "null state array");
    const int n = s->n;
    for (int i = 0; i < n; ++i) {
      if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(theta[i]))
        throw ApiError(TISSUE_E_INVALID_ARG, "non-finite state value");
      if (x[i] < 0.f || x[i] > 1.f || y[i] < 0.f || y[i] > 1.f)
        throw ApiError(TISSUE_E_INVALID_ARG, "position outside the UV square");
      if (vid && (vid[i] < 0 || vid[i] >= s->md.nv)) throw ApiError(TISSUE_E_INVALID_ARG, "vertex id out of range");
    }
    float *dx = dalloc<float>(n), *dy = dalloc<float>(n), *dt = dalloc<float>(n);
    try {
      HIP_OK(hipMemcpyAsync(dx, x, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
      HIP_OK(hipMemcpyAsync(dy, y, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
      HIP_OK(hipMemcpyAsync(dt, theta, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
      const int cur = s->sphase;
      HIP_OK(tissue::launch_prepare(n, dx, dy, dt, s->xycs[cur], s->th[cur], s->stream));
      HIP_OK(tissue::launch_locate(s->md, n, s->xycs[cur], s->vid[cur], s->face, s->r3x, s->r3y, s->r3z, s->stream));
      if (vid) HIP_OK(hipMemcpyAsync(s->vid[cur], vid, n * sizeof(int), hipMemcpyHostToDevice, s->stream));
      HIP_OK(hipMemsetAsync(s->vx, 0, n * sizeof(float), s->stream));
      HIP_OK(hipMemsetAsync(s->vy, 0, n * sizeof(float), s->stream));
      HIP_OK(hipMemsetAsync(s->nbr, 0, n * sizeof(int), s->stream));
      rebuild_buckets(s);
      s->buckets_stale = false;
      HIP_OK(hipStreamSynchronize(s->stream));
    } catch (...) {
      dfree(dx); dfree(dy); dfree(dt);
      throw;
    }
    dfree(dx); dfree(dy); dfree(dt);
    s->has_state = true;
    return TISSUE_OK;
  });
}

int32_t tissue_init_random(tissue_sim* s, uint64_t seed) {
  return guarded([&] {
    check_sim(s);
    const int n = s->n;
    std::vector<float> x(n), y(n), th(n);
    random_init(s->mesh->m, seed, n, x.data(), y.data(), th.data());
    return tissue_set_state(s, x.data(), y.data(), th.data(), nullptr);
  });
}

int32_t tissue_get_state(tissue_sim* s, float* x, float* y, float* theta, int32_t* vid) {
  return guarded([&] {
    check_sim(s);
    const int n = s->n, cur = s->sphase;
    std::vector<float4> p(n);
    HIP_OK(hipMemcpyAsync(p.data(), s->xycs[cur], n * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
    if (theta) HIP_OK(hipMemcpyAsync(theta, s->th[cur], n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if (vid) HIP_OK(hipMemcpyAsync(vid, s->vid[cur], n * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    for (int i = 0; i < n; ++i) {
      if (x) x[i] = p[i].x;
      if (y) y[i] = p[i].y;
    }
    return TISSUE_OK;
  });
}

int32_t tissue_get_outputs(tissue_sim* s, float* vx, float* vy, float* r3d, int32_t* face, int32_t* nbr) {
  return guarded([&] {
    check_sim(s);
    const int n = s->n;
    std::vector<float> a(n), b(n), c(n);
    if (vx) HIP_OK(hipMemcpyAsync(vx, s->vx, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if (vy) HIP_OK(hipMemcpyAsync(vy, s->vy, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if (face) HIP_OK(hipMemcpyAsync(face, s->face, n * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    if (nbr) HIP_OK(hipMemcpyAsync(nbr, s->nbr, n * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    if (r3d) {
      HIP_OK(hipMemcpyAsync(a.data(), s->r3x, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
      HIP_OK(hipMemcpyAsync(b.data(), s->r3y, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
      HIP_OK(hipMemcpyAsync(c.data(), s->r3z, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    }
    HIP_OK(hipStreamSynchronize(s->stream));
    if (r3d)
      for (int i = 0; i < n; ++i) { r3d[3 * i] = a[i]; r3d[3 * i + 1] = b[i]; r3d[3 * i + 2] = c[i]; }
    return TISSUE_OK;
  });
}

int32_t tissue_step(tissue_sim* s, int32_t n_steps, const float* noise) {
  return guarded([&] {
    check_sim(s);
    if (!noise) return run_steps(s, n_steps, nullptr, 0);
    const size_t need = (size_t)std::max(n_steps, 0) * s->n;
    if (s->noise_cap < need) {
      dfree(s->noise);
      s->noise = dalloc<float>(need);
      s->noise_cap = need;
    }
    HIP_OK(hipMemcpyAsync(s->noise, noise, need * sizeof(float), hipMemcpyHostToDevice, s->stream));
    s->staged_rows = 0;  // the staging buffer now holds this call's rows
    return run_steps(s, n_steps, s->noise, (size_t)s->n);
  });
}

int32_t tissue_step_device(tissue_sim* s, int32_t n_steps, const float* noise_device) {
  return guarded([&] {
    check_sim(s);
    return run_steps(s, n_steps, noise_device, (size_t)s->n);
  });
}

int32_t tissue_stage_noise(tissue_sim* s, const float* noise, int32_t n_rows) {
  return guarded([&] {
    check_sim(s);
    if (!noise || n_rows < 1) throw ApiError(TISSUE_E_INVALID_ARG, "bad noise rows");
    const size_t need = (size_t)n_rows * s->n;
    if (s->noise_cap < need) {
      HIP_OK(hipStreamSynchronize(s->stream));
      dfree(s->noise);
      s->noise = dalloc<float>(need);
      s->noise_cap = need;
    }
    HIP_OK(hipMemcpyAsync(s->noise, noise, need * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    s->staged_rows = n_rows;
    return TISSUE_OK;
  });
}

int32_t tissue_step_staged(tissue_sim* s, int32_t n_steps, int32_t first_row) {
  return guarded([&] {
    check_sim(s);
    if (first_row < 0 || n_steps < 0 || (long)first_row + n_steps > s->staged_rows)
      throw ApiError(TISSUE_E_INVALID_ARG, "staged noise rows out of range");
    return run_steps(s, n_steps, s->noise + (size_t)first_row * s->n, (size_t)s->n);
  });
}

int32_t tissue_get_order(tissue_sim* s, double* phi_ref, double* phi_polar, int32_t cap, int32_t* count) {
  return guarded([&] {
    if (!s) throw ApiError(TISSUE_E_INVALID_ARG, "null simulation handle");
    const int k = (int)std::min<size_t>(s->hist_ref.size(), (size_t)std::max(cap, 0));
    if (phi_ref) std::copy(s->hist_ref.begin(), s->hist_ref.begin() + k, phi_ref);
    if (phi_polar) std::copy(s->hist_polar.begin(), s->hist_polar.begin() + k, phi_polar);
    if (count) *count = (int32_t)s->hist_ref.size();
    return TISSUE_OK;
  });
}

int32_t tissue_reset_order(tissue_sim* s) {
  return guarded([&] {
    if (!s) throw ApiError(TISSUE_E_INVALID_ARG, "null simulation handle");
    s->hist_ref.clear();
    s->hist_polar.clear();
    return TISSUE_OK;
  });
}

int32_t tissue_get_flags(tissue_sim* s, uint32_t* flags, int32_t* first_step) {
  return guarded([&] {
    if (!s) throw ApiError(TISSUE_E_INVALID_ARG, "null simulation handle");
    if (flags) *flags = s->flags_acc;
    if (first_step) *first_step = s->flags_acc ? s->first_bad_acc : -1;
    s->flags_acc = 0;
    s->first_bad_acc = INT_MAX;
    return TISSUE_OK;
  });
}

int32_t tissue_neighbor_lists(tissue_sim* s, int32_t* counts, int32_t* cols, int32_t cap) {
  return guarded([&] {
    check_sim(s);
    if (!counts || !cols || cap < 1) throw ApiError(TISSUE_E_INVALID_ARG, "bad neighbour-list buffers");
    const int n = s->n;
    int* dc = dalloc<int>(n);
    int* dl = dalloc<int>((size_t)n * cap);
    try {
      if (s->buckets_stale) {
        rebuild_buckets(s);
        s->buckets_stale = false;
      }
      HIP_OK(tissue::launch_neighbor_lists(s->md, s->sp, n, s->bin(s->bphase), s->vid[s->sphase], cap, dc, dl, s->stream));
      HIP_OK(hipMemcpyAsync(counts, dc, n * sizeof(int), hipMemcpyDeviceToHost, s->stream));
      HIP_OK(hipMemcpyAsync(cols, dl, (size_t)n * cap * sizeof(int), hipMemcpyDeviceToHost, s->stream));
      HIP_OK(hipStreamSynchronize(s->stream));
    } catch (...) {
      dfree(dc); dfree(dl);
      throw;
    }
    dfree(dc); dfree(dl);
    for (int i = 0; i < n; ++i) std::sort(cols + (size_t)i * cap, cols + (size_t)i * cap + std::min(counts[i], cap));
    return TISSUE_OK;
  });
}

int32_t tissue_set_mode(tissue_sim* s, int32_t mode) {
  return guarded([&] {
    if (!s) throw ApiError(TISSUE_E_INVALID_ARG, "null simulation handle");
    if (mode != TISSUE_MODE_AUTO && mode != TISSUE_MODE_STEP_KERNEL && mode != TISSUE_MODE_RESIDENT)
      throw ApiError(TISSUE_E_INVALID_ARG, "unknown mode");
    if (mode == TISSUE_MODE_RESIDENT && s->n > s->resident_cap)
      throw ApiError(TISSUE_E_INVALID_ARG, "too many particles for the resident kernel on this mesh");
    s->mode = mode;
    return TISSUE_OK;
  });
}

int32_t tissue_resident_capacity(const tissue_mesh* m, int32_t* max_particles) {
  return guarded([&] {
    if (!m || !max_particles) throw ApiError(TISSUE_E_INVALID_ARG, "null argument");
    *max_particles = tissue::resident_max_particles(m->m.nv);
    return TISSUE_OK;
  });
}

int32_t tissue_set_launch(tissue_sim* s, int32_t lanes, int32_t block) {
  return guarded([&] {
    if (!s) throw ApiError(TISSUE_E_INVALID_ARG, "null simulation handle");
    if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16)
      throw ApiError(TISSUE_E_INVALID_ARG, "lanes must be 0, 1, 2, 4, 8 or 16");
    if (block != 0 && (block < 64 || block > tissue::kStepMaxBlock || block % 64))
      throw ApiError(TISSUE_E_INVALID_ARG, "block must be a multiple of 64 in [64, 512]");
    s->lanes = lanes;
    s->block = block;
    return TISSUE_OK;
  });
}

int32_t tissue_synchronize(tissue_sim* s) {
  return guarded([&] {
    check_sim(s);
    HIP_OK(hipStreamSynchronize(s->stream));
    return TISSUE_OK;
  });
}

int32_t tissue_last_step_ms(tissue_sim* s, double* ms) {
  return guarded([&] {
    if (!s || !ms) throw ApiError(TISSUE_E_INVALID_ARG, "null argument");
    *ms = s->last_ms;
    return TISSUE_OK;
  });
}

int32_t tissue_gen_noise(uint64_t seed, double std_dev, int32_t n_steps, int32_t n, float* out) {
  return guarded([&] {
    if (!out || n_steps < 0 || n < 0 || !std::isfinite(std_dev)) throw ApiError(TISSUE_E_INVALID_ARG, "bad noise arguments");
    for (int t = 0; t < n_steps; ++t)
      for (int i = 0; i < n; ++i)
        out[(size_t)t * n + i] = (float)(std_dev * tissue::normal_draw(seed, (uint32_t)i, (uint32_t)t));
    return TISSUE_OK;
  });
}

void tissue_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  tissue::philox4x32_10(ctr, key, out);
}

}  // extern "C"
