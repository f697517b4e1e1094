// C-ABI (include/tissue.h): error strings, mesh objects, particle initialisation, host helpers.
// The simulation handles are in capi_sim.cpp, the batched replicas in capi_batch.cpp.
#include <atomic>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_common.h"
#include "rng.h"
#include "uvsolve.h"

using namespace tcapi;

thread_local std::string tcapi::g_err;

std::vector<double> tcapi::face_area_cdf(const tissue::UVMesh& M) {
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
  return cdf;
}

void tcapi::random_init(const tissue::UVMesh& M, uint64_t seed, int n, float* x, float* y, float* th) {
  const std::vector<double> cdf = face_area_cdf(M);
  for (int i = 0; i < n; ++i) {
    const tissue::InitDraw d = tissue::area_uniform_draw(seed, i, cdf.data(), M.nf);
    const double* a = &M.uv[2 * M.faces[3 * d.face]];
    const double* b = &M.uv[2 * M.faces[3 * d.face + 1]];
    const double* c = &M.uv[2 * M.faces[3 * d.face + 2]];
    tissue::barycentric_uv(a[0], a[1], b[0], b[1], c[0], c[1], d.r1, d.r2, &x[i], &y[i]);
    th[i] = d.th;
  }
}

namespace {

// where the mesh constructors solve the UV map (tissue_set_uv_solve)
std::atomic<int> g_uv_where{TISSUE_BUILD_AUTO};

int32_t finish_mesh(tissue::SurfaceMesh& s, tissue_mesh** out) {
  if (!out) throw ApiError(TISSUE_E_INVALID_ARG, "null output pointer");
  // AUTO solves large systems on the GPU: on MI355X its iterations take the same time as 16 host
  // threads' at 100,000 free vertices and less above (200,000: mesh build 1.1 against 1.4 s;
  // 900,000: 3.0 against 4.2 s); the result is the same either way
  constexpr int kGpuMinRows = 100000;
  const int where = g_uv_where.load();
  tissue::UvSolver solver = nullptr;
  if (where == TISSUE_BUILD_GPU || (where == TISSUE_BUILD_AUTO && tissue::uv_solve_gpu_available())) {
    if (!tissue::uv_solve_gpu_available()) throw ApiError(TISSUE_E_HIP, "no GPU for the UV solve");
    const bool always = where == TISSUE_BUILD_GPU;
    solver = [always](const tissue::UvSystem& sys, std::vector<double>* sol) {
      return (always || sys.ni >= kGpuMinRows) && tissue::uv_solve_gpu(sys, sol);
    };
  }
  auto* tm = new tissue_mesh;
  try {
    tm->m = tissue::cut_and_parameterize(s, -1, -1, -1, solver);
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

int32_t tissue_mesh_build_tables_on(tissue_mesh* m, double sigma_max, int32_t grid, int32_t where) {
  return guarded([&] {
    if (!m || !(sigma_max > 0) || !std::isfinite(sigma_max)) throw ApiError(TISSUE_E_INVALID_ARG, "bad table arguments");
    if (where < TISSUE_BUILD_AUTO || where > TISSUE_BUILD_GPU) throw ApiError(TISSUE_E_INVALID_ARG, "bad build place");
    if (where == TISSUE_BUILD_GPU)
      throw ApiError(TISSUE_E_INVALID_ARG,
                     "the geodesic table is built on the host's threads (TISSUE_BUILD_HOST or TISSUE_BUILD_AUTO)");
    std::lock_guard<std::mutex> lk(m->mu);
    if (m->dev_ready) throw ApiError(TISSUE_E_INVALID_ARG, "mesh tables are in use on a device");
    tissue::build_face_buckets(m->m, grid);
    tissue::build_distance_table(m->m, 2.4 * sigma_max);
    m->sigma_max = sigma_max;
    m->tables = true;
    return TISSUE_OK;
  });
}

int32_t tissue_set_uv_solve(int32_t where) {
  return guarded([&] {
    if (where < TISSUE_BUILD_AUTO || where > TISSUE_BUILD_GPU) throw ApiError(TISSUE_E_INVALID_ARG, "bad build place");
    g_uv_where.store(where);
    return TISSUE_OK;
  });
}

int32_t tissue_get_uv_solve(int32_t* where) {
  return guarded([&] {
    if (!where) throw ApiError(TISSUE_E_INVALID_ARG, "null output pointer");
    *where = g_uv_where.load();
    return TISSUE_OK;
  });
}

int32_t tissue_mesh_build_tables(tissue_mesh* m, double sigma_max, int32_t grid) {
  return tissue_mesh_build_tables_on(m, sigma_max, grid, TISSUE_BUILD_AUTO);
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

int32_t tissue_resident_capacity(const tissue_mesh* m, int32_t* max_particles) {
  return guarded([&] {
    if (!m || !max_particles) throw ApiError(TISSUE_E_INVALID_ARG, "null argument");
    *max_particles = tissue::resident_max_particles(m->m.nv);
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

int32_t tissue_random_init_host(const tissue_mesh* m, uint64_t seed, int32_t n, float* x, float* y, float* theta) {
  return guarded([&] {
    if (!m || n < 0 || (n > 0 && (!x || !y || !theta))) throw ApiError(TISSUE_E_INVALID_ARG, "bad init arguments");
    random_init(m->m, seed, n, x, y, theta);
    return TISSUE_OK;
  });
}

int32_t tissue_seam_wrap(int32_t n, const double* start_xy, double* end_xy, double* dtheta, int32_t* quarter,
                         int32_t max_crossings, int32_t lanes, int32_t wave) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!start_xy || !end_xy || !dtheta))) throw ApiError(TISSUE_E_INVALID_ARG, "bad seam-wrap arguments");
    if (!valid_lanes(lanes)) throw ApiError(TISSUE_E_INVALID_ARG, "lanes must be 1, 2, 4, 8, 16, 32 or 64");
    if ((int64_t)n * lanes > INT_MAX - 256) throw ApiError(TISSUE_E_INVALID_ARG, "too many moves for one call");
    for (int64_t q = 0; q < 2 * (int64_t)n; ++q)
      if (!std::isfinite(start_xy[q]) || !std::isfinite(end_xy[q])) throw ApiError(TISSUE_E_INVALID_ARG, "non-finite position");
    if (n == 0) return TISSUE_OK;
    struct Buffers {
      double2 *s = nullptr, *e = nullptr;
      double* d = nullptr;
      int *q = nullptr, *ok = nullptr;
      ~Buffers() { dfree(s); dfree(e); dfree(d); dfree(q); dfree(ok); }
    } b;
    b.s = dalloc<double2>(n);
    b.e = dalloc<double2>(n);
    b.d = dalloc<double>(n);
    b.q = dalloc<int>(n);
    b.ok = dalloc<int>(n);
    HIP_OK(hipMemcpy(b.s, start_xy, 2 * sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b.e, end_xy, 2 * sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
    const int mc = max_crossings > 0 ? max_crossings : kDefaultMaxCrossings;
    HIP_OK(tissue::launch_seam_wrap(lanes, n, b.s, b.e, b.d, b.q, b.ok, mc, wave != 0 ? 1 : 0, nullptr));
    HIP_OK(hipStreamSynchronize(nullptr));
    std::vector<int> ok((size_t)n), qv((size_t)n);
    HIP_OK(hipMemcpy(end_xy, b.e, 2 * sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(dtheta, b.d, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(qv.data(), b.q, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(ok.data(), b.ok, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
    if (quarter) std::copy(qv.begin(), qv.end(), quarter);
    for (int v : ok)
      if (!v) return TISSUE_E_TILING_NO_CONVERGE;
    return TISSUE_OK;
  });
}

void tissue_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  tissue::philox4x32_10(ctr, key, out);
}

}  // extern "C"
