// Mesh toolkit (host, FP64). See mesh.h for what each table replaces in the reference.
#include "mesh.h"

#include "geodesic.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <numeric>
#include <queue>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <unordered_map>

namespace tissue {

int build_threads();

namespace {

inline double f32(double x) { return (double)(float)x; }

// Threads that run one function in lockstep: wait() returns once every member has reached it.
// Members spin briefly, then yield (more members than free cores stays correct, only slower).
class SpinBarrier {
 public:
  explicit SpinBarrier(int n) : n_(n) {}
  void wait() {
    const unsigned g = gen_.load(std::memory_order_acquire);
    if (count_.fetch_add(1, std::memory_order_acq_rel) == n_ - 1) {
      count_.store(0, std::memory_order_relaxed);
      gen_.fetch_add(1, std::memory_order_acq_rel);
      return;
    }
    for (int spin = 0; gen_.load(std::memory_order_acquire) == g; ++spin)
      if (spin > 256) std::this_thread::yield();
  }

 private:
  const int n_;
  std::atomic<int> count_{0};
  std::atomic<unsigned> gen_{0};
};

// run f(member, team size, barrier) on `threads` threads (the caller is member 0)
template <class F>
void run_team(int threads, F&& f) {
  SpinBarrier bar(threads);
  std::vector<std::thread> pool;
  for (int k = 1; k < threads; ++k) pool.emplace_back([&, k] { f(k, threads, bar); });
  f(0, threads, bar);
  for (auto& t : pool) t.join();
}

inline void sub3(const double* a, const double* b, double* o) {
  o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double norm3(const double* a) { return std::sqrt(dot3(a, a)); }
inline void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

// ------------------------------------------------------------------ closed surfaces

SurfaceMesh icosphere(int freq) {
  if (freq < 1) throw std::runtime_error("icosphere: frequency must be >= 1");
  const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<std::array<double, 3>> ico;
  for (int s1 = -1; s1 <= 1; s1 += 2)
    for (int s2 = -1; s2 <= 1; s2 += 2) {
      ico.push_back({0.0, (double)s1, s2 * phi});
      ico.push_back({(double)s1, s2 * phi, 0.0});
      ico.push_back({s2 * phi, 0.0, (double)s1});
    }
  // faces of the icosahedron: triples of mutually adjacent corners (edge length 2)
  std::vector<std::array<int, 3>> ifaces;
  auto adj = [&](int i, int j) {
    double d[3];
    sub3(ico[i].data(), ico[j].data(), d);
    return std::fabs(norm3(d) - 2.0) < 1e-9;
  };
  for (int i = 0; i < 12; ++i)
    for (int j = i + 1; j < 12; ++j)
      for (int k = j + 1; k < 12; ++k)
        if (adj(i, j) && adj(j, k) && adj(i, k)) {
          double e1[3], e2[3], n[3], c[3];
          sub3(ico[j].data(), ico[i].data(), e1);
          sub3(ico[k].data(), ico[i].data(), e2);
          cross3(e1, e2, n);
          for (int q = 0; q < 3; ++q) c[q] = ico[i][q] + ico[j][q] + ico[k][q];
          if (dot3(n, c) > 0) ifaces.push_back({i, j, k});
          else ifaces.push_back({i, k, j});
        }
  if (ifaces.size() != 20) throw std::runtime_error("icosphere: bad icosahedron");

  SurfaceMesh m;
  std::map<std::array<int, 6>, int> key2v;
  auto vid = [&](const int* corner, const int* w) -> int {
    // canonical key: (corner, weight) pairs with weight > 0, sorted by corner id
    std::array<std::pair<int, int>, 3> cw;
    int n = 0;
    for (int q = 0; q < 3; ++q)
      if (w[q] > 0) cw[n++] = {corner[q], w[q]};
    std::sort(cw.begin(), cw.begin() + n);
    std::array<int, 6> key{-1, -1, -1, -1, -1, -1};
    for (int q = 0; q < n; ++q) { key[2 * q] = cw[q].first; key[2 * q + 1] = cw[q].second; }
    auto it = key2v.find(key);
    if (it != key2v.end()) return it->second;
    double p[3] = {0, 0, 0};
    for (int q = 0; q < 3; ++q)
      for (int c = 0; c < 3; ++c) p[c] += (double)w[q] * ico[corner[q]][c];
    double r = norm3(p);
    int id = (int)(m.xyz.size() / 3);
    for (int c = 0; c < 3; ++c) m.xyz.push_back(p[c] / r);
    key2v.emplace(key, id);
    return id;
  };
  for (const auto& f : ifaces) {
    const int corner[3] = {f[0], f[1], f[2]};
    auto P = [&](int i, int j) {
      int w[3] = {freq - i - j, i, j};
      return vid(corner, w);
    };
    for (int i = 0; i < freq; ++i)
      for (int j = 0; i + j < freq; ++j) {
        int a = P(i, j), b = P(i + 1, j), c = P(i, j + 1);
        m.faces.insert(m.faces.end(), {a, b, c});
        if (i + j <= freq - 2) {
          int d = P(i + 1, j + 1);
          m.faces.insert(m.faces.end(), {b, d, c});
        }
      }
  }
  if (m.nv() != 10 * freq * freq + 2 || m.nf() != 20 * freq * freq)
    throw std::runtime_error("icosphere: unexpected vertex/face count");
  return m;
}

void scale_axes(SurfaceMesh& m, double a, double b, double c) {
  for (int v = 0; v < m.nv(); ++v) {
    m.xyz[3 * v] *= a;
    m.xyz[3 * v + 1] *= b;
    m.xyz[3 * v + 2] *= c;
  }
}

void perturb_y32(SurfaceMesh& m, double amp) {
  // real orthonormal Y_3^2 = (1/4) sqrt(105/pi) sin^2(theta) cos(theta) cos(2 phi)
  const double kY = 0.25 * std::sqrt(105.0 / M_PI);
  for (int v = 0; v < m.nv(); ++v) {
    double* p = &m.xyz[3 * v];
    double r = norm3(p);
    double ct = p[2] / r;
    double st2 = std::max(0.0, 1.0 - ct * ct);
    double ph = std::atan2(p[1], p[0]);
    double rr = 1.0 + amp * kY * st2 * ct * std::cos(2.0 * ph);
    for (int q = 0; q < 3; ++q) p[q] = p[q] / r * rr;
  }
}

bool load_off(const std::string& path, SurfaceMesh& m, std::string& err) {
  std::ifstream in(path);
  if (!in) { err = "cannot open " + path; return false; }
  std::vector<std::string> tok;
  std::string line;
  while (std::getline(in, line)) {
    auto h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    std::istringstream ls(line);
    std::string t;
    while (ls >> t) tok.push_back(t);
  }
  size_t p = 0;
  if (tok.empty()) { err = "empty OFF file"; return false; }
  if (tok[0] == "OFF") p = 1;
  else if (tok[0].size() > 3 && tok[0].compare(0, 3, "OFF") == 0) { err = "unsupported OFF variant " + tok[0]; return false; }
  if (tok.size() < p + 3) { err = "truncated OFF header"; return false; }
  long nv = std::stol(tok[p]), nf = std::stol(tok[p + 1]);
  p += 3;
  if (nv <= 0 || nf <= 0) { err = "OFF file without vertices or faces"; return false; }
  m.xyz.assign(3 * nv, 0.0);
  m.faces.clear();
  for (long v = 0; v < nv; ++v)
    for (int c = 0; c < 3; ++c) {
      if (p >= tok.size()) { err = "truncated OFF vertex list"; return false; }
      m.xyz[3 * v + c] = std::stod(tok[p++]);
    }
  for (long f = 0; f < nf; ++f) {
    if (p >= tok.size()) { err = "truncated OFF face list"; return false; }
    long k = std::stol(tok[p++]);
    if (k != 3) { err = "only triangle faces are supported"; return false; }
    for (int c = 0; c < 3; ++c) {
      long id = std::stol(tok[p++]);
      if (id < 0 || id >= nv) { err = "face index out of range"; return false; }
      m.faces.push_back((int32_t)id);
    }
  }
  return true;
}

bool save_off(const std::string& path, const double* xyz, int nv, const int32_t* faces, int nf, int dim, std::string& err) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) { err = "cannot write " + path; return false; }
  std::fprintf(f, "OFF\n%d %d 0\n", nv, nf);
  for (int v = 0; v < nv; ++v) {
    if (dim == 2) std::fprintf(f, "%.17g %.17g 0\n", xyz[2 * v], xyz[2 * v + 1]);
    else std::fprintf(f, "%.17g %.17g %.17g\n", xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]);
  }
  for (int i = 0; i < nf; ++i) std::fprintf(f, "3 %d %d %d\n", faces[3 * i], faces[3 * i + 1], faces[3 * i + 2]);
  std::fclose(f);
  return true;
}

// ------------------------------------------------------------------ cut + UV map

namespace {

struct Graph {
  std::vector<std::vector<std::pair<int, double>>> nbr;
};

Graph edge_graph(const std::vector<double>& xyz, const std::vector<int32_t>& faces, int nv) {
  Graph g;
  g.nbr.resize(nv);
  auto add = [&](int a, int b) {
    for (auto& e : g.nbr[a])
      if (e.first == b) return;
    double d[3];
    sub3(&xyz[3 * a], &xyz[3 * b], d);
    double L = norm3(d);
    g.nbr[a].push_back({b, L});
    g.nbr[b].push_back({a, L});
  };
  for (size_t f = 0; f < faces.size() / 3; ++f)
    for (int q = 0; q < 3; ++q) add(faces[3 * f + q], faces[3 * f + (q + 1) % 3]);
  return g;
}

std::vector<int> dijkstra_path(const Graph& g, int src, int dst, const std::vector<char>& blocked) {
  int nv = (int)g.nbr.size();
  std::vector<double> d(nv, std::numeric_limits<double>::infinity());
  std::vector<int> prev(nv, -1);
  using QE = std::pair<double, int>;
  std::priority_queue<QE, std::vector<QE>, std::greater<QE>> pq;
  d[src] = 0;
  pq.push({0, src});
  while (!pq.empty()) {
    auto [dv, v] = pq.top();
    pq.pop();
    if (dv > d[v]) continue;
    if (v == dst) break;
    for (auto [w, L] : g.nbr[v]) {
      if (blocked[w] && w != dst) continue;
      if (dv + L < d[w]) { d[w] = dv + L; prev[w] = v; pq.push({d[w], w}); }
    }
  }
  if (prev[dst] < 0 && src != dst) throw std::runtime_error("cut: no path between cone points");
  std::vector<int> path;
  for (int v = dst; v != -1; v = prev[v]) path.push_back(v);
  std::reverse(path.begin(), path.end());
  return path;
}

bool has_halfedge(const int32_t* f, int u, int w) {
  for (int q = 0; q < 3; ++q)
    if (f[q] == u && f[(q + 1) % 3] == w) return true;
  return false;
}

// symmetric positive edge weights: mean-value weights (Floater 2003) averaged over both
// directions, computed from 3D geometry
std::vector<std::unordered_map<int, double>> mean_value_weights(const std::vector<double>& xyz,
                                                                 const std::vector<int32_t>& faces, int nv) {
  std::vector<std::unordered_map<int, double>> w(nv);
  for (size_t f = 0; f < faces.size() / 3; ++f) {
    for (int q = 0; q < 3; ++q) {
      int i = faces[3 * f + q], j = faces[3 * f + (q + 1) % 3], l = faces[3 * f + (q + 2) % 3];
      double eij[3], eil[3];
      sub3(&xyz[3 * j], &xyz[3 * i], eij);
      sub3(&xyz[3 * l], &xyz[3 * i], eil);
      double lij = norm3(eij), lil = norm3(eil);
      double c = std::clamp(dot3(eij, eil) / (lij * lil), -1.0, 1.0);
      double t = std::tan(0.5 * std::acos(c));
      w[i][j] += 0.5 * t / lij;   // half: averaged with the j->i weight below
      w[j][i] += 0.5 * t / lij;
      w[i][l] += 0.5 * t / lil;
      w[l][i] += 0.5 * t / lil;
    }
  }
  return w;
}

}  // namespace

UVMesh cut_and_parameterize(const SurfaceMesh& m, int pa, int pm, int pb, const UvSolver& solver) {
  const int V = m.nv(), F = m.nf();
  if (V < 4 || F < 4) throw std::runtime_error("cut: mesh too small");
  // Euler characteristic check (closed genus 0: V - E + F = 2)
  Graph g = edge_graph(m.xyz, m.faces, V);
  long E = 0;
  for (auto& n : g.nbr) E += (long)n.size();
  E /= 2;
  if (V - E + F != 2) throw std::runtime_error("cut: mesh is not a closed genus-0 surface");

  if (pa < 0 || pb < 0 || pm < 0) {
    // cone points: the two extremes along x and the top in z
    int ia = 0, ib = 0, im = 0;
    for (int v = 1; v < V; ++v) {
      if (m.xyz[3 * v] < m.xyz[3 * ia]) ia = v;
      if (m.xyz[3 * v] > m.xyz[3 * ib]) ib = v;
      if (m.xyz[3 * v + 2] > m.xyz[3 * im + 2]) im = v;
    }
    if (pa < 0) pa = ia;
    if (pb < 0) pb = ib;
    if (pm < 0) pm = im;
  }
  if (pa == pb || pa == pm || pb == pm || pa >= V || pb >= V || pm >= V)
    throw std::runtime_error("cut: cone points must be three distinct vertices");

  std::vector<char> blocked(V, 0);
  std::vector<int> p1 = dijkstra_path(g, pa, pm, blocked);
  for (int v : p1) blocked[v] = 1;
  blocked[pm] = 0;
  std::vector<int> p2 = dijkstra_path(g, pm, pb, blocked);
  std::vector<int> path = p1;
  path.insert(path.end(), p2.begin() + 1, p2.end());
  const int K = (int)path.size() - 1;  // path = P0..PK, P0 = pa, PK = pb
  const int m_idx = (int)p1.size() - 1;

  // split the face fan of every interior path vertex into the two sides of the cut
  std::vector<int32_t> faces = m.faces;
  std::vector<std::vector<int>> vf(V);
  for (int f = 0; f < F; ++f)
    for (int q = 0; q < 3; ++q) vf[faces[3 * f + q]].push_back(f);
  std::vector<int> dup(V, -1);
  int nv = V;
  std::vector<int32_t> orig(V);
  std::iota(orig.begin(), orig.end(), 0);
  for (int i = 1; i < K; ++i) {
    const int v = path[i], prev = path[i - 1], next = path[i + 1];
    const auto& fan = vf[v];
    const int nfan = (int)fan.size();
    std::vector<int> parent(nfan);
    std::iota(parent.begin(), parent.end(), 0);
    std::function<int(int)> find = [&](int x) { return parent[x] == x ? x : parent[x] = find(parent[x]); };
    std::unordered_map<int, std::vector<int>> by_other;
    for (int a = 0; a < nfan; ++a)
      for (int q = 0; q < 3; ++q) {
        int x = faces[3 * fan[a] + q];
        if (x != v) by_other[x].push_back(a);
      }
    for (auto& [x, lst] : by_other) {
      if (x == prev || x == next || (dup[prev] >= 0 && x == dup[prev])) continue;
      if (lst.size() == 2) parent[find(lst[0])] = find(lst[1]);
    }
    int side_a = -1;
    for (int a = 0; a < nfan; ++a)
      if (has_halfedge(&faces[3 * fan[a]], prev, v)) side_a = find(a);
    if (side_a < 0) throw std::runtime_error("cut: path edge without a forward face");
    const int d = nv++;
    dup[v] = d;
    orig.push_back(v);
    int nb = 0;
    for (int a = 0; a < nfan; ++a)
      if (find(a) != side_a) {
        ++nb;
        for (int q = 0; q < 3; ++q)
          if (faces[3 * fan[a] + q] == v) faces[3 * fan[a] + q] = d;
      }
    if (nb == 0 || nb == nfan) throw std::runtime_error("cut: fan did not split in two");
  }

  UVMesh out;
  out.nv = nv;
  out.nf = F;
  out.nv3d = V;
  out.faces = faces;
  out.orig = orig;
  out.cut_path = path;
  out.pa = pa;
  out.pm = pm;
  out.pb = pb;
  out.xyz.resize(3 * nv);
  for (int v = 0; v < nv; ++v)
    for (int c = 0; c < 3; ++c) out.xyz[3 * v + c] = f32(m.xyz[3 * orig[v] + c]);

  // boundary: arc-length parameter along Pa->Pm and Pm->Pb (3D edge lengths)
  std::vector<double> uv(2 * nv, 0.0);
  std::vector<char> fixed(nv, 0);
  auto seglen = [&](int a, int b) {
    double d[3];
    sub3(&m.xyz[3 * a], &m.xyz[3 * b], d);
    return norm3(d);
  };
  std::vector<double> s(K + 1, 0.0);
  for (int i = 1; i <= K; ++i) s[i] = s[i - 1] + seglen(path[i - 1], path[i]);
  const double L1 = s[m_idx], L2 = s[K] - s[m_idx];
  for (int i = 0; i <= K; ++i) {
    const int v = path[i];
    double t;
    if (i <= m_idx) {
      t = f32(s[i] / L1);
      if (i == 0) t = 0.0;
      if (i == m_idx) t = 1.0;
      uv[2 * v] = t;      // side A: down edge (t, 0)
      uv[2 * v + 1] = 0.0;
      if (dup[v] >= 0) { uv[2 * dup[v]] = 0.0; uv[2 * dup[v] + 1] = t; }  // side B: left edge (0, t)
    } else {
      t = f32((s[i] - s[m_idx]) / L2);
      if (i == K) t = 1.0;
      uv[2 * v] = 1.0;    // side A: right edge (1, t)
      uv[2 * v + 1] = t;
      if (dup[v] >= 0) { uv[2 * dup[v]] = t; uv[2 * dup[v] + 1] = 1.0; }  // side B: up edge (t, 1)
    }
    fixed[v] = 1;
    if (dup[v] >= 0) fixed[dup[v]] = 1;
  }

  // interior: harmonic map with symmetric mean-value weights, solved by Jacobi-preconditioned CG
  auto w = mean_value_weights(out.xyz, faces, nv);
  std::vector<int> idx(nv, -1);
  int ni = 0;
  for (int v = 0; v < nv; ++v)
    if (!fixed[v]) idx[v] = ni++;
  std::vector<int> rp(ni + 1, 0), ci;
  std::vector<double> val, diag(ni, 0.0), bx(ni, 0.0), by(ni, 0.0);
  for (int v = 0; v < nv; ++v) {
    if (fixed[v]) continue;
    const int r = idx[v];
    std::vector<std::pair<int, double>> row(w[v].begin(), w[v].end());
    std::sort(row.begin(), row.end());
    for (auto [u, wt] : row) {
      diag[r] += wt;
      if (fixed[u]) { bx[r] += wt * uv[2 * u]; by[r] += wt * uv[2 * u + 1]; }
      else { ci.push_back(idx[u]); val.push_back(-wt); }
    }
    rp[r + 1] = (int)ci.size();
  }
  // Jacobi-preconditioned CG for both coordinates at once (the two systems share the matrix),
  // rows split into fixed blocks of kUvBlock: every dot product is the sum of per-block partials
  // (rows in order within a block) added in block order, so the solution does not depend on the
  // number of threads. Threads take blocks round-robin and meet at a barrier after each of the
  // three phases of an iteration; every thread forms the same scalars from the partials and so
  // takes the same convergence decisions. A coordinate stops updating once converged (each
  // coordinate's arithmetic is that of its own solve).
  constexpr int kBlock = kUvBlock;
  const int nb = (ni + kBlock - 1) / kBlock;
  const std::vector<double>* bs[2] = {&bx, &by};
  std::vector<double> sol[2], res[2], zz[2], pp[2], Ap[2];
  std::vector<double> part(6 * (size_t)nb);  // per block: pAp, r.r, r.z for x and y
  for (int k = 0; k < 2; ++k) {
    sol[k].assign(ni, 0.0);
    res[k] = *bs[k];
    zz[k].assign(ni, 0.0);
    pp[k].assign(ni, 0.0);
    Ap[k].assign(ni, 0.0);
  }
  double rz[2] = {0, 0}, bn[2] = {0, 0};
  bool live[2] = {false, false};
  for (int k = 0; k < 2; ++k) {
    for (int i = 0; i < ni; ++i) {
      zz[k][i] = res[k][i] / diag[i];
      pp[k][i] = zz[k][i];
    }
    for (int i = 0; i < ni; ++i) {
      rz[k] += res[k][i] * zz[k][i];
      bn[k] += (*bs[k])[i] * (*bs[k])[i];
    }
    bn[k] = std::sqrt(bn[k]);
    live[k] = bn[k] != 0;
  }
  bool converged = !live[0] && !live[1];
  const int max_it = 20 * ni + 100;
  if (!converged && solver) {  // the same iteration elsewhere (the GPU), if it takes the system
    UvSystem sys;
    sys.ni = ni;
    sys.nb = nb;
    sys.max_it = max_it;
    sys.rp = rp.data();
    sys.ci = ci.data();
    sys.val = val.data();
    sys.diag = diag.data();
    for (int k = 0; k < 2; ++k) {
      sys.b[k] = bs[k]->data();
      sys.z0[k] = zz[k].data();
      sys.rz[k] = rz[k];
      sys.bn[k] = bn[k];
      sys.live[k] = live[k];
    }
    converged = solver(sys, sol);
  }
  if (!converged) {
    const int threads = std::min(build_threads(), nb);
    run_team(threads, [&](int me, int nt, SpinBarrier& bar) {
      double rzl[2] = {rz[0], rz[1]};
      bool on[2] = {live[0], live[1]};
      auto total = [&](int slot) {
        double t = 0;
        for (int q = 0; q < nb; ++q) t += part[6 * (size_t)q + slot];
        return t;
      };
      for (int it = 0; it < max_it; ++it) {
        // (1) Ap = A p and the partials of p.Ap (one pass over the matrix for both coordinates)
        for (int q = me; q < nb; q += nt) {
          const int i0 = q * kBlock, i1 = std::min(ni, i0 + kBlock);
          if (on[0] && on[1]) {
            const double *p0 = pp[0].data(), *p1 = pp[1].data();
            double acc0 = 0, acc1 = 0;
            for (int i = i0; i < i1; ++i) {
              double a0 = diag[i] * p0[i], a1 = diag[i] * p1[i];
              for (int e = rp[i]; e < rp[i + 1]; ++e) {
                a0 += val[e] * p0[ci[e]];
                a1 += val[e] * p1[ci[e]];
              }
              Ap[0][i] = a0;
              Ap[1][i] = a1;
              acc0 += p0[i] * a0;
              acc1 += p1[i] * a1;
            }
            part[6 * (size_t)q] = acc0;
            part[6 * (size_t)q + 1] = acc1;
            continue;
          }
          for (int k = 0; k < 2; ++k) {
            if (!on[k]) continue;
            const double* pk = pp[k].data();
            double* ak = Ap[k].data();
            double acc = 0;
            for (int i = i0; i < i1; ++i) {
              double a = diag[i] * pk[i];
              for (int e = rp[i]; e < rp[i + 1]; ++e) a += val[e] * pk[ci[e]];
              ak[i] = a;
              acc += pk[i] * a;
            }
            part[6 * (size_t)q + k] = acc;
          }
        }
        bar.wait();
        double alpha[2] = {0, 0};
        for (int k = 0; k < 2; ++k)
          if (on[k]) alpha[k] = rzl[k] / total(k);
        // (2) x += alpha p, r -= alpha Ap, z = r / diag; partials of r.r and r.z
        for (int q = me; q < nb; q += nt) {
          const int i0 = q * kBlock, i1 = std::min(ni, i0 + kBlock);
          for (int k = 0; k < 2; ++k) {
            if (!on[k]) continue;
            double rn = 0, rzn = 0;
            for (int i = i0; i < i1; ++i) {
              sol[k][i] += alpha[k] * pp[k][i];
              res[k][i] -= alpha[k] * Ap[k][i];
              rn += res[k][i] * res[k][i];
            }
            for (int i = i0; i < i1; ++i) {
              zz[k][i] = res[k][i] / diag[i];
              rzn += res[k][i] * zz[k][i];
            }
            part[6 * (size_t)q + 2 + k] = rn;
            part[6 * (size_t)q + 4 + k] = rzn;
          }
        }
        bar.wait();
        double beta[2] = {0, 0};
        for (int k = 0; k < 2; ++k) {
          if (!on[k]) continue;
          if (std::sqrt(total(2 + k)) <= kUvTol * bn[k]) {
            on[k] = false;
            continue;
          }
          const double rz2 = total(4 + k);
          beta[k] = rz2 / rzl[k];
          rzl[k] = rz2;
        }
        if (!on[0] && !on[1]) {
          if (me == 0) converged = true;
          return;
        }
        // (3) p = z + beta p (the partial slots of phases 1 and 2 are disjoint, and a barrier
        // separates each phase's writes from the previous reads of its slots)
        for (int q = me; q < nb; q += nt) {
          const int i0 = q * kBlock, i1 = std::min(ni, i0 + kBlock);
          for (int k = 0; k < 2; ++k)
            if (on[k])
              for (int i = i0; i < i1; ++i) pp[k][i] = zz[k][i] + beta[k] * pp[k][i];
        }
        bar.wait();
      }
    });
  }
  if (!converged) throw std::runtime_error("cut: UV solve did not converge");
  const std::vector<double>& sx = sol[0];
  const std::vector<double>& sy = sol[1];
  for (int v = 0; v < nv; ++v)
    if (!fixed[v]) {
      uv[2 * v] = f32(std::clamp(sx[idx[v]], 0.0, 1.0));
      uv[2 * v + 1] = f32(std::clamp(sy[idx[v]], 0.0, 1.0));
    }

  // orientation: every UV triangle must have the same, nonzero orientation; swapping u and v
  // flips it and maps the gluing (t,0)~(0,t), (1,t)~(t,1) onto itself
  int npos = 0, nneg = 0;
  for (int f = 0; f < F; ++f) {
    const double* a = &uv[2 * faces[3 * f]];
    const double* b = &uv[2 * faces[3 * f + 1]];
    const double* c = &uv[2 * faces[3 * f + 2]];
    double ar = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
    if (ar > 0) ++npos;
    else if (ar < 0) ++nneg;
  }
  if (npos != F && nneg != F) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "cut: UV map folds (%d positive, %d negative, %d degenerate triangles)", npos,
                  nneg, F - npos - nneg);
    throw std::runtime_error(buf);
  }
  if (nneg == F)
    for (int v = 0; v < nv; ++v) std::swap(uv[2 * v], uv[2 * v + 1]);
  out.uv = uv;
  return out;
}

// ------------------------------------------------------------------ face buckets

void build_face_buckets(UVMesh& m, int grid) {
  // automatic: 1.5 sqrt(F) cells a side, about 3.6 candidate faces per cell (measured on MI355X at
  // 20,480 faces: 1.0 sqrt(F) cells, 5.2 faces, batched step 186.6 us; 1.5 sqrt(F), 182.0 us; 2.0,
  // 180.1 us with the single-system step 0.2 % slower; the 100,000-particle surface unchanged)
  if (grid <= 0) grid = std::max(1, (int)std::ceil(1.5 * std::sqrt(1.0 * m.nf)));
  m.grid = grid;
  const double eps = 1e-9;
  std::vector<std::vector<int32_t>> cells((size_t)grid * grid);
  // threads take bands of cell rows; each walks every face in order and fills only its band's
  // cells, so a cell lists its faces in ascending order whatever the number of threads
  const int threads = std::max(1, std::min(build_threads(), grid));
  run_team(threads, [&](int me, int nt, SpinBarrier&) {
    const int band0 = (int)((long)grid * me / nt), band1 = (int)((long)grid * (me + 1) / nt);
    for (int f = 0; f < m.nf; ++f) {
      double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
      for (int q = 0; q < 3; ++q)
        for (int c = 0; c < 2; ++c) {
          double x = m.uv[2 * m.faces[3 * f + q] + c];
          lo[c] = std::min(lo[c], x);
          hi[c] = std::max(hi[c], x);
        }
      int c0[2], c1[2];
      for (int c = 0; c < 2; ++c) {
        c0[c] = std::clamp((int)std::floor((lo[c] - eps) * grid), 0, grid - 1);
        c1[c] = std::clamp((int)std::floor((hi[c] + eps) * grid), 0, grid - 1);
      }
      c0[1] = std::max(c0[1], band0);
      c1[1] = std::min(c1[1], band1 - 1);
      if (c0[1] > c1[1]) continue;
      // of the cells under the face's bounding box, only those the triangle meets (each cell grown
      // by eps): the box's own axes are settled by the bounding box, so a cell is skipped when one of
      // the triangle's edge normals separates them (about a quarter of the bounding-box cells)
      const double* P[3] = {&m.uv[2 * m.faces[3 * f]], &m.uv[2 * m.faces[3 * f + 1]], &m.uv[2 * m.faces[3 * f + 2]]};
      double nx[3], ny[3], tlo[3], thi[3];
      for (int e = 0; e < 3; ++e) {
        const double* a = P[e];
        const double* b = P[(e + 1) % 3];
        nx[e] = -(b[1] - a[1]);
        ny[e] = b[0] - a[0];
        tlo[e] = 1e300;
        thi[e] = -1e300;
        for (int q = 0; q < 3; ++q) {
          const double t = nx[e] * P[q][0] + ny[e] * P[q][1];
          tlo[e] = std::min(tlo[e], t);
          thi[e] = std::max(thi[e], t);
        }
      }
      for (int cy = c0[1]; cy <= c1[1]; ++cy)
        for (int cx = c0[0]; cx <= c1[0]; ++cx) {
          const double bx[2] = {(double)cx / grid - eps, (double)(cx + 1) / grid + eps};
          const double by[2] = {(double)cy / grid - eps, (double)(cy + 1) / grid + eps};
          bool separated = false;
          for (int e = 0; e < 3 && !separated; ++e) {
            double blo = 1e300, bhi = -1e300;
            for (int i = 0; i < 2; ++i)
              for (int j = 0; j < 2; ++j) {
                const double t = nx[e] * bx[i] + ny[e] * by[j];
                blo = std::min(blo, t);
                bhi = std::max(bhi, t);
              }
            separated = bhi < tlo[e] || thi[e] < blo;
          }
          if (!separated) cells[(size_t)cy * grid + cx].push_back(f);
        }
    }
  });
  m.cell_start.assign((size_t)grid * grid + 1, 0);
  m.cell_faces.clear();
  for (size_t c = 0; c < cells.size(); ++c) {
    m.cell_faces.insert(m.cell_faces.end(), cells[c].begin(), cells[c].end());
    m.cell_start[c + 1] = (int32_t)m.cell_faces.size();
  }
}

// ------------------------------------------------------------------ geodesic table

namespace {

}  // namespace

// worker threads of the mesh build (UV solve, geodesic table): the host's cores, at most 16, or
// the environment variable TISSUE_HOST_THREADS when set to a positive count (at most 64); the
// results do not depend on it
int build_threads() {
  if (const char* e = std::getenv("TISSUE_HOST_THREADS")) {
    const int t = std::atoi(e);
    if (t > 0) return std::min(t, 64);
  }
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::clamp<unsigned>(h == 0 ? 1u : h, 1u, 16u);
}

void build_distance_table(UVMesh& m, double cutoff) {
  if (!(cutoff > 0)) throw std::runtime_error("distance table: cutoff must be positive");
  const int nv = m.nv;
  const double lim = cutoff * (1.0 + 1e-6);
  const geo::WindowMesh wm = geo::window_mesh(m.nv, m.nf, m.xyz.data(), m.faces.data());
  if (!wm.manifold) throw std::runtime_error("distance table: an edge has more than two faces");
  // one propagation per source vertex; sources are independent, so threads take them in turn,
  // each with its own scratch (the rows come out the same whatever the thread count)
  std::vector<std::vector<std::pair<int, double>>> rows(nv);
  std::atomic<int> next_source{0};
  auto propagate = [&]() {
    geo::WindowPropagation wp(wm, lim);
    for (int s = next_source++; s < nv; s = next_source++) wp.run(s, rows[s]);
  };
  {
    std::vector<std::thread> pool;
    const int nt = build_threads();
    for (int k = 1; k < nt; ++k) pool.emplace_back(propagate);
    propagate();
    for (auto& t : pool) t.join();
  }
  // symmetrize with the minimum of both directions (transform_into_symmetric_matrix,
  // src/simulation/Locomotion.cpp:127-138) and round to FP32: row v gathers its own entries and
  // the transposed entries of the rows that reach v, keeps the minimum per column, and sorts by
  // (FP32 distance, column)
  std::vector<int32_t> in_ptr(nv + 1, 0);
  for (int s = 0; s < nv; ++s)
    for (auto [v, t] : rows[s]) ++in_ptr[v + 1];
  for (int v = 0; v < nv; ++v) in_ptr[v + 1] += in_ptr[v];
  std::vector<std::pair<int, double>> in(in_ptr[nv]);
  {
    std::vector<int32_t> fill(in_ptr.begin(), in_ptr.end() - 1);
    for (int s = 0; s < nv; ++s)
      for (auto [v, t] : rows[s]) in[fill[v]++] = {s, t};
  }
  std::vector<std::vector<std::pair<float, int>>> out(nv);
  std::atomic<int> next_row{0};
  auto merge = [&]() {
    std::vector<std::pair<int, double>> r;
    for (int v = next_row++; v < nv; v = next_row++) {
      r.assign(rows[v].begin(), rows[v].end());
      r.insert(r.end(), in.begin() + in_ptr[v], in.begin() + in_ptr[v + 1]);
      std::sort(r.begin(), r.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      auto& o = out[v];
      o.clear();
      for (size_t k = 0; k < r.size();) {
        const int w = r[k].first;
        double t = r[k].second;
        for (++k; k < r.size() && r[k].first == w; ++k) t = std::min(t, r[k].second);
        o.push_back({(float)(w == v ? 0.0 : t), w});
      }
      std::sort(o.begin(), o.end());
    }
  };
  {
    std::vector<std::thread> pool;
    const int nt = build_threads();
    for (int k = 1; k < nt; ++k) pool.emplace_back(merge);
    merge();
    for (auto& t : pool) t.join();
  }
  m.cutoff = cutoff;
  m.row_ptr.assign(nv + 1, 0);
  m.col.clear();
  m.dist.clear();
  for (int v = 0; v < nv; ++v) {
    for (auto [d, w] : out[v]) { m.col.push_back(w); m.dist.push_back(d); }
    m.row_ptr[v + 1] = (int32_t)m.col.size();
  }
}

}  // namespace tissue
