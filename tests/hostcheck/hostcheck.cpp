// Host build of the device geometry (2dtissue_amd/csrc/geom.h) for exactness tests against the
// FP64 oracle. Test infrastructure: compiled by tests/hostcheck/Makefile, loaded by
// tests/test_device_math.py; nothing in the library uses it.
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
  m.cell_faces = cell_faces;
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

}  // extern "C"
