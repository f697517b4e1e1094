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
// pa/pm/pb: closed-surface vertex ids of the cone points; pass -1 to choose automatically.
UVMesh cut_and_parameterize(const SurfaceMesh& m, int pa = -1, int pm = -1, int pb = -1);

// ---- acceleration structures ----
void build_face_buckets(UVMesh& uv, int grid = 0);        // grid <= 0: automatic
// geodesic distance on the open mesh, truncated at cutoff: exact polyhedral distances by window
// propagation (geodesic.h), one source per task on the host's threads
void build_distance_table(UVMesh& uv, double cutoff);

}  // namespace tissue
