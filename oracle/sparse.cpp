// FP64 step with a sparse neighbour search (test infrastructure only; see oracle.h).
//
// or_step (oracle.cpp) restates the reference literally: dense N x N distance, displacement and
// force matrices (Locomotion.cpp:57-89) and the restart tiling loop over all particles
// (EuclideanTiling.cpp:31-69). Both are O(N^2) and infeasible at 10^5 particles. This variant
// computes the same quantities with the same floating-point operations in the same order:
//   - the particles j with D[v_i, v_j] < 2 sigma are collected from the geodesic-table row of v_i
//     and a vertex -> particle index, then visited in ascending j, which is the order in which
//     the dense loops meet them (all other j are skipped there);
//   - the seam wrap runs particle by particle: a particle's wrap result depends only on its own
//     position, old position and heading, so the restart loop and the per-particle loop agree.
// So its results equal or_step's bit for bit (tests/test_oracle.py checks that).
#include <algorithm>
#include <cmath>
#include <vector>

#include "oracle.h"

namespace {

constexpr double kPi = M_PI;

inline bool in_unit(double x, double y) { return x >= 0.0 && x <= 1.0 && y >= 0.0 && y <= 1.0; }

// particles on each vertex, ascending (counting sort)
void vertex_index(const or_mesh* m, const int32_t* vid, int32_t n, std::vector<int32_t>& start,
                  std::vector<int32_t>& list) {
  start.assign((size_t)m->nv + 1, 0);
  for (int32_t i = 0; i < n; ++i) start[vid[i] + 1] += 1;
  for (int32_t v = 0; v < m->nv; ++v) start[v + 1] += start[v];
  list.resize(n);
  std::vector<int32_t> fill(start.begin(), start.end() - 1);
  for (int32_t i = 0; i < n; ++i) list[fill[vid[i]]++] = i;
}

// (j, D[v_i, v_j]) for D < limit, ascending j
void neighbours(const or_mesh* m, const std::vector<int32_t>& start, const std::vector<int32_t>& list, int32_t v,
                double limit, std::vector<std::pair<int32_t, double>>& out) {
  out.clear();
  for (int32_t e = m->row_ptr[v]; e < m->row_ptr[v + 1]; ++e) {
    const double d = (double)m->dist[e];
    if (!(d < limit)) continue;
    const int32_t w = m->col[e];
    for (int32_t q = start[w]; q < start[w + 1]; ++q) out.push_back({list[q], d});
  }
  std::sort(out.begin(), out.end());
}

}  // namespace

extern "C" {

int32_t or_neighbor_sets_sparse(const or_mesh* m, const int32_t* vid, int32_t n, double sigma, int32_t* row_ptr,
                                int32_t* cols, int32_t cap) {
  std::vector<int32_t> start, list;
  vertex_index(m, vid, n, start, list);
  std::vector<std::pair<int32_t, double>> nb;
  int32_t k = 0;
  row_ptr[0] = 0;
  for (int32_t i = 0; i < n; ++i) {
    neighbours(m, start, list, vid[i], 2 * sigma, nb);
    for (auto& e : nb) {
      if (k >= cap) return OR_E_INVALID;
      cols[k++] = e.first;
    }
    row_ptr[i + 1] = k;
  }
  return OR_OK;
}

int32_t or_step_sparse(const or_mesh* m, const or_params* p, int32_t n, double* x, double* y, double* theta,
                       int32_t* vid, const double* noise, double* vx, double* vy, double* r3d, int32_t* face,
                       int32_t* nbr, double* phi) {
  if (n <= 0) return OR_E_INVALID;
  std::vector<int32_t> start, list;
  vertex_index(m, vid, n, start, list);
  const double rate = p->dt / p->tau;
  std::vector<double> c(n), sn(n), r(2 * (size_t)n), r_old(2 * (size_t)n), th_new(n);
  for (int32_t i = 0; i < n; ++i) { c[i] = std::cos(theta[i]); sn[i] = std::sin(theta[i]); }
  const double count_limit = std::nextafter(2.4 * p->sigma, INFINITY);  // d <= 2.4 sigma
#pragma omp parallel num_threads(or_get_threads())
  {
    std::vector<std::pair<int32_t, double>> nb, cnt;
#pragma omp for schedule(dynamic, 64)
    for (int32_t i = 0; i < n; ++i) {
      neighbours(m, start, list, vid[i], 2 * p->sigma, nb);
      // ForceHelper::calculate_forces_between_particles (ForceHelper.cpp:34-73)
      double fx = 0.0, fy = 0.0;
      for (auto [j, dist] : nb) {
        if (j == i) continue;
        if (dist == 0) dist += 0.001;
        double o[2];
        or_repulsive_adhesion_motion(p->k, p->sigma, dist, 1.0, 0.75, x[i] - x[j], y[i] - y[j], o);
        fx += o[0];
        fy += o[1];
      }
      // Locomotion.cpp:71-84
      const double absF = std::sqrt(fx * fx + fy * fy);
      const double spd = p->v0 + p->mu * absF;
      vx[i] = c[i] * spd;
      vy[i] = sn[i] * spd;
      r[2 * i] = x[i] + vx[i] * p->dt;
      r[2 * i + 1] = y[i] + vy[i] * p->dt;
      r_old[2 * i] = x[i];
      r_old[2 * i + 1] = y[i];
      // continuous heading update over {j : d_ij < 2 sigma}, self included
      double mx = 0.0, my = 0.0;
      for (auto& e : nb) { mx += c[e.first]; my += sn[e.first]; }
      const double nn = std::sqrt(mx * mx + my * my);
      if (nn > 0) { mx /= nn; my /= nn; }
      double delta = std::atan2(my, mx) - theta[i];
      if (delta > kPi) delta -= 2 * kPi;
      else if (delta <= -kPi) delta += 2 * kPi;
      th_new[i] = theta[i] + rate * delta + (noise ? noise[i] : 0.0);
      // count_particle_neighbors (2DTissue.cpp:254-268): d != 0 && d <= 2.4 sigma
      neighbours(m, start, list, vid[i], count_limit, cnt);
      int32_t k = 0;
      for (auto& e : cnt) k += (e.second != 0) ? 1 : 0;
      nbr[i] = k;
    }
  }
  // the dense step's finish (oracle.cpp finish_step): a non-finite value stops before the wrap
  // and leaves x, y as they were; the headings are wrapped and phi_polar taken either way
  auto finish = [&]() -> int32_t {
    for (int32_t i = 0; i < n; ++i)
      if (!std::isfinite(r[2 * i]) || !std::isfinite(r[2 * i + 1]) || !std::isfinite(th_new[i])) return OR_E_NONFINITE;
    int32_t status = OR_OK;
    for (int32_t i = 0; i < n; ++i) {
      const int32_t st = or_seam_wrap(&r[2 * i], &r_old[2 * i], &th_new[i], 1, 0, p->max_crossings, nullptr);
      if (st != OR_OK && status == OR_OK) status = st;
    }
    or_remap(m, r.data(), n, 1, r3d, vid, face);
    for (int32_t i = 0; i < n; ++i) {
      if (!std::isfinite(r[2 * i]) || !std::isfinite(r[2 * i + 1])) return OR_E_NONFINITE;
      if (!in_unit(r[2 * i], r[2 * i + 1]) && status == OR_OK) status = OR_E_LOST_PARTICLE;
    }
    std::vector<double> rdot(2 * (size_t)n);
    for (int32_t i = 0; i < n; ++i) { rdot[2 * i] = vx[i]; rdot[2 * i + 1] = vy[i]; }
    phi[0] = or_order_parameter(r.data(), rdot.data(), n);
    for (int32_t i = 0; i < n; ++i) { x[i] = r[2 * i]; y[i] = r[2 * i + 1]; }
    return status;
  };
  const int32_t status = finish();
  double px = 0.0, py = 0.0;
  for (int32_t i = 0; i < n; ++i) {
    double t = th_new[i];
    if (std::isfinite(t)) t = t - 2 * kPi * std::floor((t + kPi) / (2 * kPi));
    theta[i] = t;
    px += std::cos(t);
    py += std::sin(t);
  }
  phi[1] = std::sqrt(px * px + py * py) / n;
  return status;
}

}  // extern "C"
