// Mesh toolkit: closed genus-0 surfaces, the seam cut and unit-square UV map with the
// diagonal gluing the seam wrap expects, face buckets for point location, and the
// truncated geodesic distance table on the cut-open mesh.
//
// It replaces the init-time tables that 2DTissue takes from MeshCartographyLib:
//   create_uv_surface      -> vertice_UV / vertice_3D / face_UV   (src/simulation/2DTissue.cpp:89-91,107)
//   Tessellation::get_borders -> the four unit-square borders     (src/simulation/Locomotion/EuclideanTiling.cpp:34)
//   CachedGeodesicDistanceHelper(<stem>_open.off) -> distance_matrix (src/simulation/2DTissue.cpp:102-105)
// All coordinates handed to the step are FP32-representable doubles, so an FP32 device
// copy and an FP64 host copy hold identical values.
#pragma once
#include <cstdint>
#include <functional>
#include <string>
#include <utility>
#include <vector>

namespace tissue {

struct SurfaceMesh {              // closed, consistently oriented (outward normals) triangle mesh
  std::vector<double> xyz;        // V*3
  std::vector<int32_t> faces;     // F*3
  int32_t nv() const { return (int32_t)(xyz.size() / 3); }
  int32_t nf() const { return (int32_t)(faces.size() / 3); }
};

struct UVMesh {
  int32_t nv = 0;                 // UV vertices (closed-surface vertices + seam duplicates)
  int32_t nf = 0;
  int32_t nv3d = 0;               // vertices of the closed surface
  std::vector<double> uv;         // nv*2, inside [0,1]^2, FP32-representable
  std::vector<double> xyz;        // nv*3, FP32-representable
  std::vector<int32_t> faces;     // nf*3 into uv/xyz, counter-clockwise in UV
  std::vector<int32_t> orig;      // nv: UV vertex -> closed-surface vertex
  std::vector<int32_t> cut_path;  // closed-surface ids Pa ... Pm ... Pb
  int32_t pa = -1, pm = -1, pb = -1;
  // face buckets: grid x grid cells over [0,1]^2, CSR lists of the faces whose UV triangle meets the cell (grown by 1e-9)
  int32_t grid = 0;
  std::vector<int32_t> cell_start, cell_faces;
  // truncated geodesic table on the open mesh: CSR rows sorted by (dist, col); dist <= cutoff;
  // symmetric (min of both directions); the diagonal entry (v, v, 0) is present in every row
  double cutoff = 0.0;
  std::vector<int32_t> row_ptr, col;
  std::vector<float> dist;
};

// ---- closed surfaces ----
SurfaceMesh icosphere(int freq);                          // V = 10 f^2 + 2, F = 20 f^2, unit sphere
void scale_axes(SurfaceMesh& m, double a, double b, double c);
void perturb_y32(SurfaceMesh& m, double amp);             // r = 1 + amp * Y_3^2(theta, phi) (unit sphere input)
bool load_off(const std::string& path, SurfaceMesh& m, std::string& err);
bool save_off(const std::string& path, const double* xyz, int nv, const int32_t* faces, int nf, int dim_uv,
              std::string& err);

// ---- cut + UV map (throws std::runtime_error on failure) ----
// The interior of the UV map solves A x = b for both coordinates (A: the symmetric mean-value
// Laplacian of the free vertices, CSR off-diagonals rp/ci/val plus diag) by Jacobi-preconditioned
// CG with fixed reduction blocks: every dot product is the sum, in block order, of per-block sums
// of kUvBlock consecutive rows taken in row order, so the arithmetic is the same on any number of
// threads and on the GPU (uv_solve_gpu). A coordinate stops at |r| <= kUvTol |b|.
constexpr int kUvBlock = 2048;
constexpr double kUvTol = 1e-13;
struct UvSystem {
  int ni = 0, nb = 0, max_it = 0;  // rows, reduction blocks, iteration limit
  const int* rp = nullptr;
  const int* ci = nullptr;
  const double* val = nullptr;
  const double* diag = nullptr;
  const double* b[2] = {nullptr, nullptr};   // right-hand sides: the starting residuals
  const double* z0[2] = {nullptr, nullptr};  // starting preconditioned residuals = directions
  double rz[2] = {0, 0}, bn[2] = {0, 0};     // r.z and |b| at the start
  bool live[2] = {false, false};             // coordinates with a nonzero right-hand side
};
// An alternative solver of the same iteration: returns true with both solutions in sol[0..1], or
// false to leave the system to the host; throws when the iteration does not converge.
using UvSolver = std::function<bool(const UvSystem&, std::vector<double>* sol)>;

// pa/pm/pb: closed-surface vertex ids of the cone points; pass -1 to choose automatically.
UVMesh cut_and_parameterize(const SurfaceMesh& m, int pa = -1, int pm = -1, int pb = -1,
                            const UvSolver& solver = nullptr);

// ---- acceleration structures ----
void build_face_buckets(UVMesh& uv, int grid = 0);        // grid <= 0: automatic
// geodesic distance on the open mesh, truncated at cutoff: exact polyhedral distances by window
// propagation (geodesic.h), one source per task on the host's threads
void build_distance_table(UVMesh& uv, double cutoff);

}  // namespace tissue
