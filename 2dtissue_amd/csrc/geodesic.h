// Exact geodesic distances on a triangle mesh, truncated at a radius: the distance table's rows.
//
// The distance between two vertices is the length of the shortest path on the piecewise-flat
// surface (the polyhedral geodesic distance). It is computed per source vertex by window
// propagation (Chen & Han 1990; windows as in Surazhsky et al., "Fast exact and approximate
// geodesics on meshes", SIGGRAPH 2005; processed in order of their least distance, as in Xin &
// Wang, "Improving Chen and Han's algorithm on the discrete geodesic problem", 2009):
//   - a window is an interval [b0, b1] of a mesh edge together with the unfolded image of the
//     source (or of a pseudo-source) in the plane of the face it enters: every point t of the
//     interval is reached by the straight segment from the image, at distance sigma + |image - t|;
//   - propagating a window across its face unfolds the face's third vertex c into the edge's
//     plane; the rays from the image through the interval hit c (which then gets its distance)
//     and split into the parts that leave through the two other edges (the child windows);
//   - shortest paths bend only at saddle vertices (angle sum above 2 pi) and at boundary
//     vertices (the mesh is cut open along the seam); those become pseudo-sources that start
//     windows of their own, with sigma = their distance;
//   - a window is dropped when one of its edge's endpoints reaches every point of it by a
//     shorter path (the endpoint's distance plus the straight path along the edge): then no
//     shortest path runs through the window. The distance of the window over its interval minus
//     that alternative is monotone along the edge, so the test at one end of the interval covers
//     all of it;
//   - events (windows and pseudo-sources) are processed in order of their least distance, so a
//     pseudo-source emits once, with its final distance, and everything beyond the radius stops.
// Only vertex-to-vertex distances are kept. Everything is FP64.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <queue>
#include <utility>
#include <vector>

namespace tissue {
namespace geo {

inline double dist3(const double* a, const double* b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// the mesh as the propagation reads it
struct WindowMesh {
  int nv = 0, nf = 0;
  const double* xyz = nullptr;     // nv * 3
  const int32_t* faces = nullptr;  // nf * 3
  std::vector<int32_t> across;     // 3 * nf: face across the edge opposite corner q of face f, -1 on the boundary
  std::vector<int32_t> vf_ptr, vf_idx;  // faces around each vertex (CSR)
  std::vector<char> pseudo;        // saddle or boundary vertex
  bool manifold = true;            // no edge with more than two faces
};

// adjacency, faces around each vertex, and the pseudo-source vertices
WindowMesh window_mesh(int nv, int nf, const double* xyz, const int32_t* faces);

// one source's propagation (one per thread: it keeps its scratch between sources)
class WindowPropagation {
 public:
  WindowPropagation(const WindowMesh& m, double lim);
  // (vertex, distance) of every vertex within lim of source s, ascending in (distance, vertex)
  void run(int s, std::vector<std::pair<int, double>>& row);

 private:
  struct Window {
    int face, corner;  // the face the window leaves and its corner opposite the window's edge
    double b0, b1;     // the interval, along the edge from its first vertex
    double sx, sy;     // the source image in the edge's frame (sy < 0: behind the edge)
    double sigma;      // distance of the image from the source
  };
  struct Event {
    double key;  // least distance of any point the event reaches
    int id;      // window index, or -1 - v for the pseudo-source vertex v
    bool operator>(const Event& o) const { return key > o.key; }
  };
  void relax(int v, double t);
  void emit(int v, double sigma);
  void push_window(int f, int corner, double b0, double b1, double sx, double sy, double sigma);
  void propagate(const Window& w);
  int corner_vertex(int f, int q) const { return m_.faces[3 * f + q % 3]; }

  const WindowMesh& m_;
  double lim_;
  std::vector<double> d_;
  std::vector<int> touched_;
  std::vector<Window> windows_;
  std::priority_queue<Event, std::vector<Event>, std::greater<Event>> queue_;
};

}  // namespace geo
}  // namespace tissue
