// Shared pieces of the C-ABI implementation: error handling, device allocation helpers, the mesh
// object (host tables + lazily uploaded device copy) and the per-simulation step constants.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tissue.h"
#include "launch.h"
#include "mesh.h"

namespace tcapi {

extern thread_local std::string g_err;


struct ApiError : std::runtime_error {
  int32_t code;
  ApiError(int32_t c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_OK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) throw ApiError(TISSUE_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class F>
int32_t guarded(F&& f) {
  try {
    return f();
  } catch (const ApiError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host out of memory";
    return TISSUE_E_INVALID_ARG;
  } catch (const std::exception& e) {
    g_err = e.what();
    return TISSUE_E_MESH;
  }
}

template <class T>
T* dalloc(size_t n) {
  void* p = nullptr;
  HIP_OK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return (T*)p;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

template <class T>
void upload(T* dst, const T* src, size_t n) {
  if (n) HIP_OK(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
}

// largest float f with (double)f < T (strict) or <= T: FP32 thresholds that reproduce the
// FP64 comparisons of the reference exactly for FP32 table distances
inline float float_below(double T, bool strict) {
  float f = (float)T;
  auto bad = [&](float x) { return strict ? !((double)x < T) : !((double)x <= T); };
  while (bad(f)) f = std::nextafter(f, -INFINITY);
  for (;;) {
    float g = std::nextafter(f, INFINITY);
    if (bad(g)) break;
    f = g;
  }
  return f;
}

// lanes per particle the kernels are built for
inline bool valid_lanes(int l) { return l == 1 || l == 2 || l == 4 || l == 8 || l == 16 || l == 32 || l == 64; }

// the reference's tiling loop has no bound; this one only stops pathological inputs
constexpr int kDefaultMaxCrossings = 16384;

inline void check_params(const tissue_params* p, double sigma_max) {
  if (!p) throw ApiError(TISSUE_E_INVALID_ARG, "null params");
  const double v[] = {p->sigma, p->v0, p->k, p->mu, p->tau, p->dt};
  for (double x : v)
    if (!std::isfinite(x)) throw ApiError(TISSUE_E_INVALID_ARG, "non-finite parameter");
  if (!(p->sigma > 0) || !(p->tau > 0) || !(p->dt > 0))
    throw ApiError(TISSUE_E_INVALID_ARG, "sigma, tau and dt must be positive");
  if (p->sigma > sigma_max * (1 + 1e-12))
    throw ApiError(TISSUE_E_INVALID_ARG, "sigma exceeds the sigma_max the mesh tables were built for");
}

inline tdev::StepParams make_step_params(const tissue_params& p) {
  tdev::StepParams s{};
  s.t_near = float_below(2 * p.sigma, true);
  s.t_count = float_below(2.4 * p.sigma, false);
  s.two_sigma = (float)(2 * p.sigma);
  s.k = (float)p.k;
  s.v0 = (float)p.v0;
  s.mu = (float)p.mu;
  s.dt = (float)p.dt;
  s.rate = (float)(p.dt / p.tau);
  s.max_crossings = p.max_crossings > 0 ? p.max_crossings : kDefaultMaxCrossings;
  s.two_sigma64 = 2 * p.sigma;
  s.k64 = p.k;
  s.v064 = p.v0;
  s.mu64 = p.mu;
  s.dt64 = p.dt;
  s.rate64 = p.dt / p.tau;
  return s;
}

// particles placed uniformly by area: a face with probability proportional to its 3-D area,
// then a uniform point of it (the UV image of a triangle is affine, so uniform in UV too);
// headings uniform on [0, 2 pi), stored in [-pi, pi). Counter-based: particle i uses the
// Philox4x32-10 blocks (i, 0, 0) and (i, 0, 2) of the seed.
// cumulative FP64 3-D face areas, the face-choice table of the area-uniform initialisation
std::vector<double> face_area_cdf(const tissue::UVMesh& M);
// the area-uniform initialisation on the host (the GPU kernel draws the same values)
void random_init(const tissue::UVMesh& M, uint64_t seed, int n, float* x, float* y, float* th);

}  // namespace tcapi

using namespace tcapi;

// ------------------------------------------------------------------ mesh

struct tissue_mesh {
  tissue::UVMesh m;
  bool tables = false;
  double sigma_max = 0.0;
  std::mutex mu;
  bool dev_ready = false;
  int device = -1;
  float2* d_uv = nullptr;
  float* d_xyz = nullptr;
  int32_t* d_faces = nullptr;
  int32_t* d_cell_start = nullptr;
  float4* d_cell_tri = nullptr;
  float4* d_face_pack = nullptr;
  double* d_cdf = nullptr;  // cumulative face areas (area-uniform initialisation)
  int32_t* d_row_ptr = nullptr;
  int2* d_nbr = nullptr;

  void free_device() {
    if (!dev_ready) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    dfree(d_uv); dfree(d_xyz); dfree(d_faces); dfree(d_cell_start); dfree(d_cell_tri); dfree(d_face_pack); dfree(d_cdf); dfree(d_row_ptr); dfree(d_nbr);
    (void)hipSetDevice(cur);
    dev_ready = false;
  }

  tdev::MeshDev device_view() {
    std::lock_guard<std::mutex> lk(mu);
    if (!tables) throw ApiError(TISSUE_E_INVALID_ARG, "mesh tables not built (tissue_mesh_build_tables)");
    int cur = 0;
    HIP_OK(hipGetDevice(&cur));
    if (dev_ready && device != cur) throw ApiError(TISSUE_E_INVALID_ARG, "mesh already resident on another device");
    if (!dev_ready) {
      device = cur;
      std::vector<float2> uv(m.nv);
      std::vector<float> xyz(3 * (size_t)m.nv);
      for (int v = 0; v < m.nv; ++v) uv[v] = make_float2((float)m.uv[2 * v], (float)m.uv[2 * v + 1]);
      for (size_t q = 0; q < xyz.size(); ++q) xyz[q] = (float)m.xyz[q];
      std::vector<int2> nbr(m.col.size());
      for (size_t e = 0; e < nbr.size(); ++e) {
        int bits;
        std::memcpy(&bits, &m.dist[e], 4);
        nbr[e] = make_int2(m.col[e], bits);
      }
      d_uv = dalloc<float2>(uv.size()); upload(d_uv, uv.data(), uv.size());
      d_xyz = dalloc<float>(xyz.size()); upload(d_xyz, xyz.data(), xyz.size());
      d_faces = dalloc<int32_t>(m.faces.size()); upload(d_faces, m.faces.data(), m.faces.size());
      d_cell_start = dalloc<int32_t>(m.cell_start.size()); upload(d_cell_start, m.cell_start.data(), m.cell_start.size());
      std::vector<float4> tri(2 * m.cell_faces.size());
      for (size_t q = 0; q < m.cell_faces.size(); ++q)
        tdev::pack_cell_triangle(reinterpret_cast<const float*>(uv.data()), m.faces.data(), m.cell_faces[q], &tri[2 * q]);
      d_cell_tri = dalloc<float4>(tri.size()); upload(d_cell_tri, tri.data(), tri.size());
      std::vector<float4> fpk(5 * (size_t)m.nf);
      for (int f = 0; f < m.nf; ++f)
        tdev::pack_face(reinterpret_cast<const float*>(uv.data()), xyz.data(), m.faces.data(), f, &fpk[5 * (size_t)f]);
      d_face_pack = dalloc<float4>(fpk.size()); upload(d_face_pack, fpk.data(), fpk.size());
      const std::vector<double> cdf = face_area_cdf(m);
      d_cdf = dalloc<double>(cdf.size()); upload(d_cdf, cdf.data(), cdf.size());
      d_row_ptr = dalloc<int32_t>(m.row_ptr.size()); upload(d_row_ptr, m.row_ptr.data(), m.row_ptr.size());
      d_nbr = dalloc<int2>(nbr.size()); upload(d_nbr, nbr.data(), nbr.size());
      dev_ready = true;
    }
    return view();
  }

  // the device tables when they are already on GPU `dev` (no upload, no allocation)
  tdev::MeshDev resident_view(int dev) {
    std::lock_guard<std::mutex> lk(mu);
    if (!tables) throw ApiError(TISSUE_E_INVALID_ARG, "mesh tables not built (tissue_mesh_build_tables)");
    if (!dev_ready || device != dev)
      throw ApiError(TISSUE_E_INVALID_ARG, "mesh tables are not on this GPU (a tissue_create on the mesh uploads them)");
    return view();
  }

 private:
  tdev::MeshDev view() const {
    tdev::MeshDev v;
    v.nv = m.nv; v.nf = m.nf; v.grid = m.grid; v.pad = 0;
    v.uv = d_uv; v.xyz = d_xyz; v.faces = d_faces; v.cell_start = d_cell_start; v.cell_tri = d_cell_tri;
    v.face_pack = d_face_pack;
    v.row_ptr = d_row_ptr; v.nbr = d_nbr;
    return v;
  }
};
