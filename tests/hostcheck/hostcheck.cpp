// Host build of the device geometry (2dtissue_amd/csrc/geom.h) for exactness tests against the
// FP64 oracle, and of the geodesic table's window propagation (geodesic.cpp) for closed-form
// tests on meshes the library does not build (open sheets). Test infrastructure: compiled by
// tests/hostcheck/Makefile, loaded by tests/test_device_math.py and tests/test_geodesic.py;
// nothing in the library uses it.
#include <vector>

#include "../../2dtissue_amd/csrc/geodesic.h"
#include "../../2dtissue_amd/csrc/geom.h"

extern "C" {

int hc_quotient_in_unit(double n, double d) { return tdev::quotient_in_unit(n, d) ? 1 : 0; }

void hc_sort8_many(int n, int* v /* n*8, sorted in place */) {
  for (int i = 0; i < n; ++i) {
    int a[8];
    for (int k = 0; k < 8; ++k) a[k] = v[8 * i + k];
    tdev::sort8(a);
    for (int k = 0; k < 8; ++k) v[8 * i + k] = a[k];
  }
}

void hc_seam_wrap_many(int n, const double* a, const double* p, int max_crossings, double* out_p, double* dth,
                       int* ok, int* crossings) {
  for (int i = 0; i < n; ++i) {
    double px = p[2 * i], py = p[2 * i + 1], dt = 0.0;
    int c = 0;
    ok[i] = tdev::seam_wrap(a[2 * i], a[2 * i + 1], px, py, dt, max_crossings, &c) ? 1 : 0;
    out_p[2 * i] = px;
    out_p[2 * i + 1] = py;
    dth[i] = dt;
    crossings[i] = c;
  }
}

// the lane-parallel loop's formulation (border_pass, cross_border_flip) run on the host
void hc_seam_wrap_flip_many(int n, const double* a, const double* p, int max_crossings, double* out_p, double* dth,
                            int* ok, int* crossings, int* quarter) {
  for (int i = 0; i < n; ++i) {
    double px = p[2 * i], py = p[2 * i + 1], dt = 0.0;
    int c = 0, q = 0;
    ok[i] = tdev::seam_wrap_flip(a[2 * i], a[2 * i + 1], px, py, dt, max_crossings, &c, &q) ? 1 : 0;
    out_p[2 * i] = px;
    out_p[2 * i + 1] = py;
    dth[i] = dt;
    crossings[i] = c;
    quarter[i] = q;
  }
}

void hc_seam_wrap_quarter_many(int n, const double* a, const double* p, int max_crossings, int* quarter) {
  for (int i = 0; i < n; ++i) {
    double px = p[2 * i], py = p[2 * i + 1], dt = 0.0;
    int q = 0;
    tdev::seam_wrap(a[2 * i], a[2 * i + 1], px, py, dt, max_crossings, nullptr, &q);
    quarter[i] = q;
  }
}

// FP32 branch-free squared distance and the FP64 region-by-region distance, for n (point,
// triangle) pairs: tri = (ax, ay, bx, by, cx, cy) per pair
void hc_tri_dist_many(int n, const float* pts, const float* tri, float* d2_f32, double* d_f64) {
  for (int i = 0; i < n; ++i) {
    const float* t = tri + 6 * i;
    d2_f32[i] = tdev::point_triangle_dist2(pts[2 * i], pts[2 * i + 1], t[0], t[1], t[2], t[3], t[4], t[5]);
    d_f64[i] = tdev::point_triangle_distance<double>(pts[2 * i], pts[2 * i + 1], t[0], t[1], t[2], t[3], t[4], t[5]);
  }
}

void hc_locate_many(int n, const double* pts, int nv, int nf, const float* uv, const float* xyz, const int* faces,
                    int grid, const int* cell_start, const int* cell_faces, int* face, int* vid, double* r3d,
                    int* fp64) {
  tdev::MeshDev m{};
  m.nv = nv;
  m.nf = nf;
  m.grid = grid;
  m.uv = (const float2*)uv;
  m.xyz = xyz;
  m.faces = faces;
  m.cell_start = cell_start;
  std::vector<float4> tri(2 * (size_t)cell_start[grid * grid]);
  for (int q = 0; q < cell_start[grid * grid]; ++q) tdev::pack_cell_triangle(uv, faces, cell_faces[q], &tri[2 * (size_t)q]);
  m.cell_tri = tri.data();
  std::vector<float4> fpk(5 * (size_t)nf);
  for (int f = 0; f < nf; ++f) tdev::pack_face(uv, xyz, faces, f, &fpk[5 * (size_t)f]);
  m.face_pack = fpk.data();
  for (int i = 0; i < n; ++i) {
    const tdev::Located L = tdev::locate<1>(m, pts[2 * i], pts[2 * i + 1], 0);
    face[i] = L.face;
    vid[i] = L.vid;
    r3d[3 * i] = L.X;
    r3d[3 * i + 1] = L.Y;
    r3d[3 * i + 2] = L.Z;
    fp64[i] = L.fp64;
  }
}


// row of source s of the exact truncated geodesic distances on a triangle mesh (any manifold
// mesh, open or closed): writes up to cap (vertex, distance) pairs, returns the row length
int hc_geodesic_row(int nv, int nf, const double* xyz, const int* faces, double lim, int s, int cap, int* out_v,
                    double* out_d) {
  const tissue::geo::WindowMesh m = tissue::geo::window_mesh(nv, nf, xyz, faces);
  tissue::geo::WindowPropagation wp(m, lim);
  std::vector<std::pair<int, double>> row;
  wp.run(s, row);
  for (int k = 0; k < (int)row.size() && k < cap; ++k) {
    out_v[k] = row[k].first;
    out_d[k] = row[k].second;
  }
  return (int)row.size();
}

}  // extern "C"
