// FP64 reverse pass of the continuous step (test infrastructure only; see oracle.h).
//
// The reference has no derivative of its step (it is not differentiated anywhere); this is the
// reverse-mode derivative of the step or_step computes (Locomotion.cpp:57-89, ForceHelper.cpp:34-104,
// OrientationHelper.cpp:29-74 with real-valued headings, EuclideanTiling.cpp:31-208), taken with
// the step's discrete decisions held fixed: the neighbour sets {j : D[v_i, v_j] < 2 sigma} from the
// step's vertex ids, the d == 0 rule, the seam-wrap maps (a rotation by q quarter turns plus a
// translation) and the heading wrap. It works along a given trajectory (states 0..T), so it can
// be evaluated on the GPU's own states; everything inside is FP64. The derivation, per step and
// particle i with incoming adjoints (lx, ly, lt) of the new state:
//   F = sum_{j != i} c(d_ij) (r_i - r_j), c(d) = f(d~)/d~, f(d) = -k (2 sigma - d)/(2 sigma);
//   m = |F|, s = v0 + mu m, e = (cos, sin) theta, rdot = s e, p = r + dt rdot, r' = R_q p + b;
//   theta' = theta + rho wrap(atan2(S, C) - theta) + eta + const, (C, S) = sum_{j in A_i} (cos, sin);
//   loss += w_v |rdot - u(r)|^2.
#include <cmath>
#include <vector>

#include "oracle.h"

namespace {

constexpr double kPi = M_PI;

struct FieldVal {
  double ux = 0, uy = 0, jxx = 0, jxy = 0, jyx = 0, jyy = 0;
};

FieldVal field(const or_loss* L, double x, double y) {
  FieldVal f;
  for (int32_t m = 0; m < L->n_modes; ++m) {
    const double a = L->kx[m] * x + L->ky[m] * y + L->phase[m];
    const double v = L->amp[m] * std::sin(a), d = L->amp[m] * std::cos(a);
    if (L->component[m] == 0) { f.ux += v; f.jxx += d * L->kx[m]; f.jxy += d * L->ky[m]; }
    else { f.uy += v; f.jyx += d * L->kx[m]; f.jyy += d * L->ky[m]; }
  }
  return f;
}

// R_q^T v, R_q = rotation by q counter-clockwise quarter turns
void rot_t(int q, double vx, double vy, double* ox, double* oy) {
  switch (q & 3) {
    case 1: *ox = vy; *oy = -vx; break;
    case 2: *ox = -vx; *oy = -vy; break;
    case 3: *ox = -vy; *oy = vx; break;
    default: *ox = vx; *oy = vy;
  }
}

}  // namespace

extern "C" {

double or_field_eval(const or_loss* L, double x, double y, double* u /* 2 */, double* jac /* 4 */) {
  const FieldVal f = field(L, x, y);
  if (u) { u[0] = f.ux; u[1] = f.uy; }
  if (jac) { jac[0] = f.jxx; jac[1] = f.jxy; jac[2] = f.jyx; jac[3] = f.jyy; }
  return f.ux;
}

int32_t or_adjoint_reverse(const or_mesh* m, const or_params* p, int32_t n, int32_t T, const double* xs,
                           const double* ys, const double* ths, const int32_t* vids, const or_loss* L, double* out,
                           double* d_theta0, double* d_x0, double* d_y0) {
  if (n <= 0 || T <= 0) return OR_E_INVALID;
  const double two_sigma = 2 * p->sigma, rho = p->dt / p->tau;
  const double wv = L->w_velocity / ((double)T * n);
  // seeds: d/d theta_i of w_order (phi_T - target)^2, phi_T = |sum (cos, sin)| / n
  const double* thT = ths + (size_t)T * n;
  double Px = 0, Py = 0;
  for (int32_t i = 0; i < n; ++i) { Px += std::cos(thT[i]); Py += std::sin(thT[i]); }
  const double norm = std::sqrt(Px * Px + Py * Py), phi = norm / n;
  const double coef = norm > 0 ? L->w_order * 2 * (phi - L->target_order) / (n * norm) : 0.0;
  std::vector<double> lx(n, 0.0), ly(n, 0.0), lt(n);
  for (int32_t i = 0; i < n; ++i) lt[i] = coef * (-Px * std::sin(thT[i]) + Py * std::cos(thT[i]));
  double d_v0 = 0, d_mu = 0, d_k = 0, d_tau = 0, lv = 0;

  std::vector<std::vector<int32_t>> on_vertex(m->nv);
  std::vector<std::vector<std::pair<int32_t, double>>> nb(n);
  std::vector<double> Fx(n), Fy(n), lFx(n), lFy(n), lsin(n), lcos(n), plx(n), ply(n), plt(n);
  for (int32_t t = T - 1; t >= 0; --t) {
    const double* x = xs + (size_t)t * n;
    const double* y = ys + (size_t)t * n;
    const double* th = ths + (size_t)t * n;
    const int32_t* vid = vids + (size_t)t * n;
    for (auto& l : on_vertex) l.clear();
    for (int32_t i = 0; i < n; ++i) on_vertex[vid[i]].push_back(i);
    // frozen neighbour sets A_i (self included through the diagonal table entry)
    for (int32_t i = 0; i < n; ++i) {
      nb[i].clear();
      for (int32_t e = m->row_ptr[vid[i]]; e < m->row_ptr[vid[i] + 1]; ++e) {
        const double d = (double)m->dist[e];
        if (!(d < two_sigma)) continue;
        for (int32_t j : on_vertex[m->col[e]]) nb[i].push_back({j, d});
      }
    }
    for (int32_t i = 0; i < n; ++i) {
      double fx = 0, fy = 0, C = 0, S = 0;
      for (auto [j, d] : nb[i]) {
        C += std::cos(th[j]);
        S += std::sin(th[j]);
        if (j == i) continue;
        const double dd = d == 0.0 ? 0.001 : d;  // ForceHelper.cpp:60-63
        const double c = (-p->k * (two_sigma - dd)) / two_sigma / dd;
        fx += c * (x[i] - x[j]);
        fy += c * (y[i] - y[j]);
      }
      Fx[i] = fx;
      Fy[i] = fy;
      const double mg = std::sqrt(fx * fx + fy * fy);
      const double s = p->v0 + p->mu * mg;
      const double ex = std::cos(th[i]), ey = std::sin(th[i]);
      const double rdx = s * ex, rdy = s * ey;
      // the seam-wrap turns of this particle's step
      double r[2] = {x[i] + p->dt * rdx, y[i] + p->dt * rdy}, r_old[2] = {x[i], y[i]}, h = 0.0;
      or_seam_wrap(r, r_old, &h, 1, 0, p->max_crossings, nullptr);
      const int q = (int)(((std::llround(h / (0.5 * kPi)) % 4) + 4) % 4);
      double lpx, lpy;
      rot_t(q, lx[i], ly[i], &lpx, &lpy);
      const FieldVal u = field(L, x[i], y[i]);
      const double dvx = rdx - u.ux, dvy = rdy - u.uy;
      const double lrdx = p->dt * lpx + 2 * wv * dvx, lrdy = p->dt * lpy + 2 * wv * dvy;
      plx[i] = lpx - 2 * wv * (u.jxx * dvx + u.jyx * dvy);
      ply[i] = lpy - 2 * wv * (u.jxy * dvx + u.jyy * dvy);
      const double ls = lrdx * ex + lrdy * ey;
      plt[i] = (1 - rho) * lt[i] + s * (-lrdx * ey + lrdy * ex);
      lFx[i] = mg > 0 ? p->mu * ls * fx / mg : 0.0;
      lFy[i] = mg > 0 ? p->mu * ls * fy / mg : 0.0;
      const double Q = C * C + S * S;
      lsin[i] = Q > 0 ? rho * lt[i] * C / Q : 0.0;
      lcos[i] = Q > 0 ? -rho * lt[i] * S / Q : 0.0;
      double delta = std::atan2(S, C) - th[i];
      if (delta > kPi) delta -= 2 * kPi;
      else if (delta <= -kPi) delta += 2 * kPi;
      d_v0 += ls;
      d_mu += ls * mg;
      if (p->k != 0) d_k += (lFx[i] * fx + lFy[i] * fy) / p->k;
      d_tau += lt[i] * delta * (-p->dt / (p->tau * p->tau));
      lv += wv * (dvx * dvx + dvy * dvy);
    }
    // transposed neighbour sums (the relation is symmetric)
    for (int32_t i = 0; i < n; ++i) {
      double ax = 0, ay = 0, aS = 0, aC = 0;
      for (auto [j, d] : nb[i]) {
        aS += lsin[j];
        aC += lcos[j];
        if (j == i) continue;
        const double dd = d == 0.0 ? 0.001 : d;
        const double c = (-p->k * (two_sigma - dd)) / two_sigma / dd;
        ax += c * (lFx[i] - lFx[j]);
        ay += c * (lFy[i] - lFy[j]);
      }
      lx[i] = plx[i] + ax;
      ly[i] = ply[i] + ay;
      lt[i] = plt[i] + std::cos(th[i]) * aS - std::sin(th[i]) * aC;
    }
  }
  const double lo = L->w_order * (phi - L->target_order) * (phi - L->target_order);
  out[0] = lo + lv;
  out[1] = lo;
  out[2] = lv;
  out[3] = phi;
  out[4] = d_tau;
  out[5] = d_mu;
  out[6] = d_v0;
  out[7] = d_k;
  for (int32_t i = 0; i < n; ++i) {
    if (d_theta0) d_theta0[i] = lt[i];
    if (d_x0) d_x0[i] = lx[i];
    if (d_y0) d_y0[i] = ly[i];
  }
  return OR_OK;
}

}  // extern "C"
