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

// the wave's three-crossings-per-pass loop (seam_wrap_passes) with its 64 lanes emulated on the
// host, entered after `before` crossings of the lane-group loop (the lane-parallel form above);
// nones (may be null) counts, per move, the crossings of that first part whose four tests failed
struct HostLanes {
  tdev::WrapPath w[64];
  tdev::WrapMasks run(const tdev::WrapState& s) {
    tdev::WrapMasks m{0, 0, 0, 0, 0};
    for (int l = 0; l < 64; ++l) {
      tdev::wrap_path(l, s, w[l]);
      m.t1 |= (uint64_t)w[l].t1 << l;
      m.t2 |= (uint64_t)w[l].t2 << l;
      m.t3 |= (uint64_t)w[l].t3 << l;
      m.in1 |= (uint64_t)w[l].in1 << l;
      m.in2 |= (uint64_t)w[l].in2 << l;
    }
    return m;
  }
  tdev::WrapState take(int level, int l) const { return level == 1 ? w[l].s1 : (level == 2 ? w[l].s2 : w[l].s3); }
};

void hc_seam_wrap_passes_many(int n, const double* a, const double* p, int max_crossings, int before, double* out_p,
                              double* dth, int* ok, int* crossings, int* quarter, int* nones) {
  HostLanes L;
  for (int i = 0; i < n; ++i) {
    tdev::WrapState s{a[2 * i], a[2 * i + 1], p[2 * i], p[2 * i + 1], 0.0, 0};
    int c = 0;
    bool good = true;
    // the lane-group loop's crossings first (cross_border_flip after border_pass, as on the GPU)
    while (c < before && !tdev::inside_unit(s.px, s.py)) {
      if (c >= max_crossings) {
        good = false;
        break;
      }
      ++c;
      const double dx = s.px - s.ax, dy = s.py - s.ay;
      int b = 4;
      for (int k = 3; k >= 0; --k)
        if (tdev::border_pass(k < 2, k == 1 ? 1.0 : -0.0, k == 2 ? 1.0 : -0.0, s.ax, s.ay, dx, dy)) b = k;
      if (nones && b == 4) ++nones[i];
      tdev::cross_border_flip(b, s.ax, s.ay, s.px, s.py, dx, dy, s.dth, s.q);
    }
    if (good) good = tdev::seam_wrap_passes(L, s, c, max_crossings);
    out_p[2 * i] = s.px;
    out_p[2 * i + 1] = s.py;
    dth[i] = s.dth;
    ok[i] = good ? 1 : 0;
    crossings[i] = c;
    quarter[i] = s.q & 3;
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
