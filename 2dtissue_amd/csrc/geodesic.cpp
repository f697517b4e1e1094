// Exact truncated geodesic distances by window propagation (geodesic.h).
#include "geodesic.h"

namespace tissue {
namespace geo {

WindowMesh window_mesh(int nv, int nf, const double* xyz, const int32_t* faces) {
  WindowMesh m;
  m.nv = nv;
  m.nf = nf;
  m.xyz = xyz;
  m.faces = faces;
  // faces across each edge: (edge key, face corner) pairs sorted and grouped by edge
  auto key = [](int a, int b) { return a < b ? ((int64_t)a << 32) | b : ((int64_t)b << 32) | a; };
  std::vector<std::pair<int64_t, int>> ek(3 * (size_t)nf);
  for (int f = 0; f < nf; ++f)
    for (int q = 0; q < 3; ++q) ek[3 * (size_t)f + q] = {key(faces[3 * f + (q + 1) % 3], faces[3 * f + (q + 2) % 3]), 3 * f + q};
  std::sort(ek.begin(), ek.end());
  m.across.assign(3 * (size_t)nf, -1);
  std::vector<char> boundary(nv, 0);
  for (size_t k = 0; k < ek.size();) {
    size_t e = k;
    while (e < ek.size() && ek[e].first == ek[k].first) ++e;
    if (e - k == 2) {
      m.across[ek[k].second] = ek[k + 1].second / 3;
      m.across[ek[k + 1].second] = ek[k].second / 3;
    } else {
      if (e - k > 2) m.manifold = false;
      for (size_t j = k; j < e; ++j) {  // boundary (or non-manifold) edge: its vertices bend paths
        const int f = ek[j].second / 3, q = ek[j].second % 3;
        boundary[faces[3 * f + (q + 1) % 3]] = 1;
        boundary[faces[3 * f + (q + 2) % 3]] = 1;
      }
    }
    k = e;
  }
  // faces around each vertex and the angle sum at it
  m.vf_ptr.assign(nv + 1, 0);
  for (int f = 0; f < nf; ++f)
    for (int q = 0; q < 3; ++q) ++m.vf_ptr[faces[3 * f + q] + 1];
  for (int v = 0; v < nv; ++v) m.vf_ptr[v + 1] += m.vf_ptr[v];
  m.vf_idx.assign(m.vf_ptr[nv], 0);
  std::vector<int32_t> fill(m.vf_ptr.begin(), m.vf_ptr.end() - 1);
  std::vector<double> angle(nv, 0.0);
  for (int f = 0; f < nf; ++f)
    for (int q = 0; q < 3; ++q) {
      const int v = faces[3 * f + q];
      m.vf_idx[fill[v]++] = f;
      const double* P = xyz + 3 * v;
      const double* A = xyz + 3 * faces[3 * f + (q + 1) % 3];
      const double* B = xyz + 3 * faces[3 * f + (q + 2) % 3];
      const double u[3] = {A[0] - P[0], A[1] - P[1], A[2] - P[2]};
      const double w[3] = {B[0] - P[0], B[1] - P[1], B[2] - P[2]};
      const double c = (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / (dist3(A, P) * dist3(B, P));
      angle[v] += std::acos(std::clamp(c, -1.0, 1.0));
    }
  // a vertex whose angle sum exceeds 2 pi (beyond rounding) is a saddle; a flat or convex one
  // never lies inside a shortest path
  m.pseudo.assign(nv, 0);
  for (int v = 0; v < nv; ++v) m.pseudo[v] = boundary[v] || angle[v] > 2 * M_PI + 1e-12;
  return m;
}

WindowPropagation::WindowPropagation(const WindowMesh& m, double lim)
    : m_(m), lim_(lim), d_(m.nv, std::numeric_limits<double>::infinity()) {}

void WindowPropagation::relax(int v, double t) {
  if (!(t < d_[v])) return;
  if (d_[v] == std::numeric_limits<double>::infinity()) touched_.push_back(v);
  d_[v] = t;
  if (m_.pseudo[v] && t <= lim_) queue_.push({t, -1 - v});
}

// The edge opposite corner `corner` of face f, in its frame: first vertex p = corner + 1 at the
// origin, second vertex e = corner + 2 at (L, 0), the face f below (y < 0), the face across
// above. A window on it whose interval is empty, beyond the radius, or dominated by an endpoint
// is dropped.
void WindowPropagation::push_window(int f, int corner, double b0, double b1, double sx, double sy, double sigma) {
  if (m_.across[3 * (size_t)f + corner] < 0) return;  // boundary: nothing beyond
  const int p = corner_vertex(f, corner + 1), e = corner_vertex(f, corner + 2);
  const double L = dist3(m_.xyz + 3 * p, m_.xyz + 3 * e);
  b0 = std::max(0.0, b0);
  b1 = std::min(L, b1);
  if (!(b1 - b0 > 1e-12 * L)) return;
  const double near = sx < b0 ? std::hypot(b0 - sx, sy) : (sx > b1 ? std::hypot(sx - b1, sy) : std::fabs(sy));
  const double key = sigma + near;
  if (key > lim_) return;
  // sigma + |image - t| - (d[p] + t) does not increase along the edge, sigma + |image - t| -
  // (d[e] + L - t) does not decrease: each is least at one end of the interval
  const double tol = 1e-12 * (1.0 + lim_);
  if (sigma + std::hypot(b1 - sx, sy) > d_[p] + b1 + tol) return;
  if (sigma + std::hypot(b0 - sx, sy) > d_[e] + (L - b0) + tol) return;
  windows_.push_back({f, corner, b0, b1, sx, sy, sigma});
  queue_.push({key, (int)windows_.size() - 1});
}

// a pseudo-source v at distance sigma: straight paths along its edges, and a window over the
// whole edge opposite v in every face around it
void WindowPropagation::emit(int v, double sigma) {
  const double* V = m_.xyz + 3 * v;
  for (int k = m_.vf_ptr[v]; k < m_.vf_ptr[v + 1]; ++k) {
    const int f = m_.vf_idx[k];
    const int qv = m_.faces[3 * f] == v ? 0 : (m_.faces[3 * f + 1] == v ? 1 : 2);
    const int p = corner_vertex(f, qv + 1), e = corner_vertex(f, qv + 2);
    const double* P = m_.xyz + 3 * p;
    const double* E = m_.xyz + 3 * e;
    const double lp = dist3(V, P), le = dist3(V, E), L = dist3(P, E);
    relax(p, sigma + lp);
    relax(e, sigma + le);
    const double sx = (lp * lp - le * le + L * L) / (2 * L);
    const double sy = -std::sqrt(std::max(lp * lp - sx * sx, 0.0));
    push_window(f, qv, 0.0, L, sx, sy, sigma);
  }
}

void WindowPropagation::propagate(const Window& w) {
  const int f = w.face;
  const int g = m_.across[3 * (size_t)f + w.corner];
  const int p = corner_vertex(f, w.corner + 1), e = corner_vertex(f, w.corner + 2);
  int qc = 0;
  while (m_.faces[3 * g + qc] == p || m_.faces[3 * g + qc] == e) ++qc;
  const int c = m_.faces[3 * g + qc];
  // g's vertices in the window's frame: p = (0, 0), e = (L, 0), c = (cx, cy) above the edge
  const double* P = m_.xyz + 3 * p;
  const double* E = m_.xyz + 3 * e;
  const double* Cv = m_.xyz + 3 * c;
  const double L = dist3(P, E), lpc = dist3(P, Cv), lec = dist3(E, Cv);
  const double cx = (lpc * lpc - lec * lec + L * L) / (2 * L);
  const double cy = std::sqrt(std::max(lpc * lpc - cx * cx, 0.0));
  const double sx = w.sx, sy = w.sy;
  // the ray from the image through c crosses the edge at tc: c is seen when tc is in the interval
  const double tc = sx + (cx - sx) * (0.0 - sy) / (cy - sy);
  const double eps = 1e-12 * L;
  if (tc >= w.b0 - eps && tc <= w.b1 + eps) relax(c, w.sigma + std::hypot(cx - sx, cy - sy));
  // the rays through [t0, t1] that leave g through its edge opposite vertex `opp`
  auto child = [&](int opp, double t0, double t1) {
    int qo = 0;
    while (m_.faces[3 * g + qo] != opp) ++qo;
    const int a = corner_vertex(g, qo + 1), b = corner_vertex(g, qo + 2);
    auto at = [&](int v, double* o) {
      o[0] = v == p ? 0.0 : (v == e ? L : cx);
      o[1] = v == c ? cy : 0.0;
    };
    double A[2], B[2], O[2];
    at(a, A);
    at(b, B);
    at(opp, O);
    // the child's frame: a at the origin, x toward b, y away from g (away from opp)
    const double lab = std::hypot(B[0] - A[0], B[1] - A[1]);
    const double ex = (B[0] - A[0]) / lab, ey = (B[1] - A[1]) / lab;
    double nx = -ey, ny = ex;
    if (nx * (O[0] - A[0]) + ny * (O[1] - A[1]) > 0) { nx = -nx; ny = -ny; }
    auto frame = [&](double x, double y, double* u, double* v) {
      *u = (x - A[0]) * ex + (y - A[1]) * ey;
      *v = (x - A[0]) * nx + (y - A[1]) * ny;
    };
    double Sx, Sy;
    frame(sx, sy, &Sx, &Sy);
    // where the ray from the image through (t, 0) meets the child's edge line
    auto meet = [&](double t, double* u) {
      double Qx, Qy;
      frame(t, 0.0, &Qx, &Qy);
      const double den = Qy - Sy;
      if (!(den > 0)) return false;
      *u = Sx + (Qx - Sx) * (0.0 - Sy) / den;
      return true;
    };
    double u0, u1;
    if (!meet(t0, &u0) || !meet(t1, &u1)) return;
    if (u0 > u1) std::swap(u0, u1);
    push_window(g, qo, u0, u1, Sx, Sy, w.sigma);
  };
  if (tc > w.b0) child(e, w.b0, std::min(w.b1, tc));  // out through edge (p, c)
  if (tc < w.b1) child(p, std::max(w.b0, tc), w.b1);  // out through edge (c, e)
}

void WindowPropagation::run(int s, std::vector<std::pair<int, double>>& row) {
  row.clear();
  relax(s, 0.0);
  if (!m_.pseudo[s]) queue_.push({0.0, -1 - s});
  while (!queue_.empty()) {
    const Event ev = queue_.top();
    queue_.pop();
    if (ev.key > lim_) break;
    if (ev.id < 0) {
      const int v = -1 - ev.id;
      if (ev.key > d_[v]) continue;  // stale: v was reached more closely since
      emit(v, d_[v]);
    } else {
      const Window w = windows_[ev.id];
      propagate(w);
    }
  }
  for (int v : touched_)
    if (d_[v] <= lim_) row.push_back({v, d_[v]});
  std::sort(row.begin(), row.end(),
            [](const auto& x, const auto& y) { return x.second < y.second || (x.second == y.second && x.first < y.first); });
  for (int v : touched_) d_[v] = std::numeric_limits<double>::infinity();
  touched_.clear();
  windows_.clear();
  while (!queue_.empty()) queue_.pop();
}

}  // namespace geo
}  // namespace tissue
