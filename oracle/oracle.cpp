// FP64 CPU oracle (test infrastructure only; see oracle.h). Each function cites the reference
// function it restates (KAIZN-io/2DTissue, src/simulation/...). Built with -ffp-contract=off so
// every expression rounds as written.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = M_PI;
constexpr double DEG_TO_RAD = M_PI / 180.0;  // LinearAlgebra.h / OrientationHelper.h constants
constexpr double RAD_TO_DEG = 180.0 / M_PI;
constexpr double BORDER_THRESHOLD = 1e-3;    // EuclideanTiling.h
int g_threads = 1;

struct P2 { double x, y; };

// unit square polygon of the UV domain; boundary counts as inside
bool in_polygon(double x, double y) { return x >= 0.0 && x <= 1.0 && y >= 0.0 && y <= 1.0; }

// EuclideanTiling::is_point_on_segment, Locomotion/EuclideanTiling.cpp:156-168
bool is_point_on_segment(P2 P, P2 A, P2 B) {
  if (P.x < std::min(A.x, B.x) || P.x > std::max(A.x, B.x) || P.y < std::min(A.y, B.y) || P.y > std::max(A.y, B.y))
    return false;
  double crossProduct = (P.x - A.x) * (B.y - A.y) - (P.y - A.y) * (B.x - A.x);
  return std::fabs(crossProduct) < 1e-9;
}

// EuclideanTiling::intersection_point, Locomotion/EuclideanTiling.cpp:170-208; the border is a
// polyline, here the straight unit-square edge
bool intersection_point(P2 A, P2 B, const P2* border, int npts, P2* out) {
  for (int i = 0; i + 1 < npts; ++i) {
    P2 C = border[i], D = border[i + 1];
    if (is_point_on_segment(A, C, D)) continue;
    double det = (B.x - A.x) * (D.y - C.y) - (B.y - A.y) * (D.x - C.x);
    if (std::fabs(det) < 1e-9) continue;
    double t = ((C.x - A.x) * (D.y - C.y) - (C.y - A.y) * (D.x - C.x)) / det;
    double s = ((C.x - A.x) * (B.y - A.y) - (C.y - A.y) * (B.x - A.x)) / det;
    if (t >= 0 && t <= 1 && s >= 0 && s <= 1) {
      out->x = A.x + t * (B.x - A.x);
      out->y = A.y + t * (B.y - A.y);
      return true;
    }
  }
  return false;
}

// Tessellation::get_borders for the unit square: left, right, up, down (EuclideanTiling.cpp:34-38)
const P2 kLeft[2] = {{0, 0}, {0, 1}};
const P2 kRight[2] = {{1, 0}, {1, 1}};
const P2 kUp[2] = {{0, 1}, {1, 1}};
const P2 kDown[2] = {{0, 0}, {1, 0}};

enum Border { LEFT = 0, RIGHT = 1, UP = 2, DOWN = 3, NONE = 4 };

// EuclideanTiling::check_border_crossings, Locomotion/EuclideanTiling.cpp:71-101
Border check_border_crossings(P2 start, P2 end, P2* exit_point) {
  const P2* borders[4] = {kLeft, kRight, kUp, kDown};
  for (int b = 0; b < 4; ++b) {
    P2 e;
    if (intersection_point(start, end, borders[b], 2, &e)) {
      if (std::fabs(e.x) < BORDER_THRESHOLD) e.x = 0.0;
      if (std::fabs(e.y) < BORDER_THRESHOLD) e.y = 0.0;
      *exit_point = e;
      return (Border)b;
    }
  }
  *exit_point = start;
  return NONE;
}

// EuclideanTiling::processPoints, Locomotion/EuclideanTiling.cpp:112-154; rot90/rot270 are the
// heading decrements (90/270 degrees, or pi/2 and 3 pi/2 radians)
void process_points(P2 pointA, P2 point_outside, double n, double rot90, double rot270, P2* new_point, double* n_out,
                    P2* entry_point) {
  P2 np = {2, 1}, ep = {1, 1};
  if (!in_polygon(point_outside.x, point_outside.y)) {
    P2 exit_point;
    Border b = check_border_crossings(pointA, point_outside, &exit_point);
    ep = {exit_point.y, exit_point.x};
    if (b == LEFT) { np = {point_outside.y, -point_outside.x}; n -= rot90; }
    else if (b == RIGHT) { np = {point_outside.y, 2 - point_outside.x}; n -= rot90; }
    else if (b == UP) { np = {2 - point_outside.y, point_outside.x}; n -= rot270; }
    else { np = {-point_outside.y, point_outside.x}; n -= rot270; }
  } else {
    np = point_outside;
  }
  *new_point = np;
  *n_out = n;
  *entry_point = ep;
}

inline double sq(double a) { return a * a; }

// CellHelper::pointSegmentDistance, CellHelper.cpp:216-222 (2-D: the z components are zero)
double point_segment_distance(const double* p, const double* a, const double* b) {
  double abx = b[0] - a[0], aby = b[1] - a[1];
  double t = (abx * (p[0] - a[0]) + aby * (p[1] - a[1])) / (abx * abx + aby * aby);
  t = std::clamp(t, 0.0, 1.0);
  double qx = a[0] + abx * t - p[0], qy = a[1] + aby * t - p[1];
  return std::sqrt(qx * qx + qy * qy);
}

// D[v, w] from the truncated table; pairs beyond the cutoff are "far" (infinity)
double lookup(const or_mesh* m, int32_t v, int32_t w) {
  for (int32_t e = m->row_ptr[v]; e < m->row_ptr[v + 1]; ++e)
    if (m->col[e] == w) return (double)m->dist[e];
  return std::numeric_limits<double>::infinity();
}

// CellHelper::calculate_barycentric_3D_coord, CellHelper.cpp:104-160, for one point and a
// given candidate face list (all faces, or one bucket)
void remap_one(const or_mesh* m, double px, double py, const int32_t* cand, int32_t ncand, bool all_faces,
               double* r3d, int32_t* vid, int32_t* face) {
  const double p[2] = {px, py};
  double best_d = std::numeric_limits<double>::infinity();
  int32_t best_f = -1;
  for (int32_t q = 0; q < ncand; ++q) {
    const int32_t f = all_faces ? q : cand[q];
    const double* a = &m->uv[2 * m->faces[3 * f]];
    const double* b = &m->uv[2 * m->faces[3 * f + 1]];
    const double* c = &m->uv[2 * m->faces[3 * f + 2]];
    double d = or_point_triangle_distance(p, a, b, c);
    // std::min_element over (distance, face) pairs: lexicographic minimum
    if (best_f < 0 || d < best_d || (d == best_d && f < best_f)) { best_d = d; best_f = f; }
  }
  if (best_f < 0) best_f = 0;
  const int32_t ia = m->faces[3 * best_f], ib = m->faces[3 * best_f + 1], ic = m->faces[3 * best_f + 2];
  const double* ua = &m->uv[2 * ia];
  const double* ub = &m->uv[2 * ib];
  const double* uc = &m->uv[2 * ic];
  const double* A = &m->xyz[3 * ia];
  const double* B = &m->xyz[3 * ib];
  const double* C = &m->xyz[3 * ic];
  double w_a = std::sqrt(sq(px - ua[0]) + sq(py - ua[1]));
  double w_b = std::sqrt(sq(px - ub[0]) + sq(py - ub[1]));
  double w_c = std::sqrt(sq(px - uc[0]) + sq(py - uc[1]));
  // CellHelper::normalize_weights, CellHelper.cpp:224-230
  double sum_weights = w_a + w_b + w_c;
  w_a /= sum_weights;
  w_b /= sum_weights;
  w_c /= sum_weights;
  double np[3];
  for (int k = 0; k < 3; ++k) np[k] = w_a * A[k] + w_b * B[k] + w_c * C[k];
  double da = std::sqrt(sq(np[0] - A[0]) + sq(np[1] - A[1]) + sq(np[2] - A[2]));
  double db = std::sqrt(sq(np[0] - B[0]) + sq(np[1] - B[1]) + sq(np[2] - B[2]));
  double dc = std::sqrt(sq(np[0] - C[0]) + sq(np[1] - C[1]) + sq(np[2] - C[2]));
  double min_dist = std::min({da, db, dc});
  int32_t closest;
  if (min_dist == da) closest = ia;
  else if (min_dist == db) closest = ib;
  else closest = ic;
  for (int k = 0; k < 3; ++k) r3d[k] = np[k];
  *vid = closest;
  *face = best_f;
}

// shared body of both step variants: everything except the heading rule
struct StepScratch {
  std::vector<double> dx, dy, dist_length, F, r, r_old;
};

}  // namespace

extern "C" {

void or_set_threads(int32_t n) { g_threads = n < 1 ? 1 : n; }
int32_t or_get_threads(void) { return g_threads; }

// LinearAlgebra::angles_to_unit_vectors, LinearAlgebra.cpp:25-46
int32_t or_angles_to_unit_vectors(const int32_t* n_deg, int32_t n, double* out) {
  for (int32_t i = 0; i < n; ++i) {
    double angle_degrees = n_deg[i];
    double angle_radians = angle_degrees * DEG_TO_RAD;
    out[2 * i] = std::cos(angle_radians);
    out[2 * i + 1] = std::sin(angle_radians);
  }
  return OR_OK;
}

// OrientationHelper::mean_unit_circle_vector_angle_degrees, OrientationHelper.cpp:84-116;
// an empty input is the reference's std::invalid_argument (returned as OR_E_INVALID)
int32_t or_mean_unit_circle_vector_angle_degrees(const double* angles, int32_t n, double* out) {
  if (n <= 0) return OR_E_INVALID;
  double mx = 0.0, my = 0.0;
  for (int32_t i = 0; i < n; ++i) {
    double angle_radians = angles[i] * DEG_TO_RAD;
    mx += std::cos(angle_radians);
    my += std::sin(angle_radians);
  }
  double nn = std::sqrt(mx * mx + my * my);  // Eigen normalize(): unchanged when the norm is 0
  if (nn > 0) { mx /= nn; my /= nn; }
  double angle_degrees = std::atan2(my, mx) * RAD_TO_DEG;
  if (angle_degrees < 0) angle_degrees += 360.0;
  *out = angle_degrees;
  return OR_OK;
}

// ForceHelper::repulsive_adhesion_motion, ForceHelper.cpp:84-104
void or_repulsive_adhesion_motion(double k, double sigma, double dist, double r_adh, double k_adh, double dvx,
                                  double dvy, double* out) {
  double Fij_rep = 0, Fij_adh = 0;
  if (dist < 2 * sigma) Fij_rep = (-k * (2 * sigma - dist)) / (2 * sigma);
  if (dist >= 2 * sigma && dist <= r_adh) Fij_adh = (k_adh * (2 * sigma - dist)) / (2 * sigma - r_adh);
  double Fij = Fij_rep + Fij_adh;
  out[0] = Fij * (dvx / dist);
  out[1] = Fij * (dvy / dist);
}

// Locomotion::transform_into_symmetric_matrix, Locomotion.cpp:127-138
void or_transform_into_symmetric_matrix(double* a, int32_t n) {
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = i + 1; j < n; ++j) a[(size_t)i * n + j] = a[(size_t)j * n + i] = std::min(a[(size_t)i * n + j], a[(size_t)j * n + i]);
}

// Locomotion::get_dist_vect, Locomotion.cpp:143-162: diff(i, j) = r_i - r_j
void or_get_dist_vect(const double* r, int32_t n, double* dx, double* dy) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = 0; j < n; ++j) {
      dx[(size_t)i * n + j] = r[2 * i] - r[2 * j];
      dy[(size_t)i * n + j] = r[2 * i + 1] - r[2 * j + 1];
    }
}

// ForceHelper::calculate_forces_between_particles, ForceHelper.cpp:34-73
void or_calculate_forces_between_particles(const double* dist_length, const double* dx, const double* dy, int32_t n,
                                           double k, double sigma, double r_adh, double k_adh, double* f) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int32_t i = 0; i < n; ++i) {
    double fx = 0.0, fy = 0.0;
    for (int32_t j = 0; j < n; ++j) {
      if (i == j) continue;
      double dist = dist_length[(size_t)i * n + j];
      if (dist >= 2 * sigma) continue;
      if (dist == 0) dist += 0.001;
      double o[2];
      or_repulsive_adhesion_motion(k, sigma, dist, r_adh, k_adh, dx[(size_t)i * n + j], dy[(size_t)i * n + j], o);
      fx += o[0];
      fy += o[1];
    }
    f[2 * i] = fx;
    f[2 * i + 1] = fy;
  }
}

// OrientationHelper::calculate_average_n_within_distance, OrientationHelper.cpp:29-74
// (int headings; the reference's noise is uniform(0, 0), i.e. zero)
int32_t or_average_n_within_distance_deg(const double* dist_length, int32_t n, int32_t* n_deg, double sigma) {
  std::vector<int32_t> avg_n(n, 0);
  std::vector<double> angles;
  for (int32_t i = 0; i < n; ++i) {
    angles.clear();
    for (int32_t j = 0; j < n; ++j)
      if (dist_length[(size_t)i * n + j] < 2 * sigma) angles.push_back(n_deg[j]);
    double mean;
    if (or_mean_unit_circle_vector_angle_degrees(angles.data(), (int32_t)angles.size(), &mean) != OR_OK)
      return OR_E_INVALID;
    avg_n[i] = (int32_t)(mean + 0.0);  // double -> int conversion truncates toward zero
  }
  for (int32_t i = 0; i < n; ++i) n_deg[i] = avg_n[i];
  return OR_OK;
}

double or_dist_lookup(const or_mesh* m, int32_t v, int32_t w) { return lookup(m, v, w); }

// Locomotion::get_distances_between_particles, Locomotion.cpp:94-111
void or_distances_between_particles(const or_mesh* m, const int32_t* vid, int32_t n, double* dist_length) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = 0; j < n; ++j) dist_length[(size_t)i * n + j] = lookup(m, vid[i], vid[j]);
  for (int32_t i = 0; i < n; ++i) dist_length[(size_t)i * n + i] = 0.0;
  or_transform_into_symmetric_matrix(dist_length, n);
}

// _2DTissue::count_particle_neighbors, 2DTissue.cpp:254-268
void or_count_particle_neighbors(const double* dist_length, int32_t n, double sigma, int32_t* out) {
  for (int32_t i = 0; i < n; ++i) {
    int32_t c = 0;
    for (int32_t j = 0; j < n; ++j) {
      double d = dist_length[(size_t)i * n + j];
      if (d != 0 && d <= 2.4 * sigma) c += 1;
    }
    out[i] = c;
  }
}

// EuclideanTiling::diagonal_seam_edges_square_border, Locomotion/EuclideanTiling.cpp:31-69:
// restart from particle 0 whenever a mapped point is still outside. Every mapping of an outside
// point is one border crossing; max_crossings bounds them per particle (the reference loops
// without bound). A particle over the bound is left unmapped and the call reports
// OR_E_TILING_NO_CONVERGE.
int32_t or_seam_wrap(double* r, double* r_old, double* heading, int32_t n, int32_t unit_deg, int32_t max_crossings,
                     int32_t* crossings) {
  const double rot90 = unit_deg ? 90.0 : 0.5 * kPi;
  const double rot270 = unit_deg ? 270.0 : 1.5 * kPi;
  std::vector<int32_t> cnt(n, 0);
  int32_t status = OR_OK;
  bool valid;
  do {
    valid = true;
    for (int32_t i = 0; i < n; ++i) {
      P2 pointA = {r_old[2 * i], r_old[2 * i + 1]};
      P2 point_outside = {r[2 * i], r[2 * i + 1]};
      if (!in_polygon(point_outside.x, point_outside.y)) {
        if (cnt[i] >= max_crossings) { status = OR_E_TILING_NO_CONVERGE; continue; }
        cnt[i] += 1;
      }
      P2 new_point, entry_point;
      double n_new;
      process_points(pointA, point_outside, heading[i], rot90, rot270, &new_point, &n_new, &entry_point);
      heading[i] = n_new;
      if (in_polygon(new_point.x, new_point.y)) {
        r[2 * i] = new_point.x;
        r[2 * i + 1] = new_point.y;
      } else {
        r_old[2 * i] = entry_point.x;
        r_old[2 * i + 1] = entry_point.y;
        r[2 * i] = new_point.x;
        r[2 * i + 1] = new_point.y;
        valid = false;
        break;
      }
    }
  } while (!valid);
  if (crossings)
    for (int32_t i = 0; i < n; ++i) crossings[i] = cnt[i];
  return status;
}

// CellHelper::pointTriangleDistance, CellHelper.cpp:162-214 (Ericson's closest-point regions),
// with the particle's z taken as 0 (the reference reads it out of bounds)
double or_point_triangle_distance(const double* p, const double* a, const double* b, const double* c) {
  double abx = b[0] - a[0], aby = b[1] - a[1];
  double acx = c[0] - a[0], acy = c[1] - a[1];
  double apx = p[0] - a[0], apy = p[1] - a[1];
  double bpx = p[0] - b[0], bpy = p[1] - b[1];
  double cpx = p[0] - c[0], cpy = p[1] - c[1];
  double d_ab_ap = abx * apx + aby * apy;
  double d_ac_ap = acx * apx + acy * apy;
  double d_ab_bp = abx * bpx + aby * bpy;
  double d_ac_bp = acx * bpx + acy * bpy;
  double d_ab_cp = abx * cpx + aby * cpy;
  double d_ac_cp = acx * cpx + acy * cpy;
  if (d_ab_ap <= 0.0 && d_ac_ap <= 0.0) return std::sqrt(apx * apx + apy * apy);
  if (d_ab_bp >= 0.0 && d_ac_bp <= d_ab_bp) return std::sqrt(bpx * bpx + bpy * bpy);
  if (d_ac_cp >= 0.0 && d_ab_cp <= d_ac_cp) return std::sqrt(cpx * cpx + cpy * cpy);
  double vc = d_ab_ap * d_ac_bp - d_ab_bp * d_ac_ap;
  if (vc <= 0.0 && d_ab_ap >= 0.0 && d_ab_bp <= 0.0) return point_segment_distance(p, a, b);
  double vb = d_ab_cp * d_ac_ap - d_ab_ap * d_ac_cp;
  if (vb <= 0.0 && d_ac_ap >= 0.0 && d_ac_cp <= 0.0) return point_segment_distance(p, a, c);
  double va = d_ab_bp * d_ac_cp - d_ab_cp * d_ac_bp;
  if (va <= 0.0 && (d_ac_bp - d_ab_bp) >= 0.0 && (d_ab_cp - d_ac_cp) >= 0.0) return point_segment_distance(p, b, c);
  double denom = 1.0 / (va + vb + vc);
  double v = vb * denom;
  double w = vc * denom;
  double qx = a[0] + abx * v + acx * w - p[0];
  double qy = a[1] + aby * v + acy * w - p[1];
  return std::sqrt(qx * qx + qy * qy);
}

// CellHelper::get_r3d, CellHelper.cpp:71-85
void or_remap(const or_mesh* m, const double* r, int32_t n, int32_t mode, double* r3d, int32_t* vid, int32_t* face) {
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 16)
  for (int32_t i = 0; i < n; ++i) {
    const double px = r[2 * i], py = r[2 * i + 1];
    if (mode == 1 && m->grid > 0) {
      const int32_t G = m->grid;
      int32_t cx = (px >= 0.0) ? (int32_t)std::min<double>(std::floor(px * G), G - 1) : 0;
      int32_t cy = (py >= 0.0) ? (int32_t)std::min<double>(std::floor(py * G), G - 1) : 0;
      const int32_t c = cy * G + cx;
      remap_one(m, px, py, m->cell_faces + m->cell_start[c], m->cell_start[c + 1] - m->cell_start[c], false,
                &r3d[3 * i], &vid[i], &face[i]);
    } else {
      remap_one(m, px, py, nullptr, m->nf, true, &r3d[3 * i], &vid[i], &face[i]);
    }
  }
}

// LinearAlgebra::calculate_order_parameter with calculate_3D_cross_product, LinearAlgebra.cpp:67-105;
// the N x 2 rows are read as 3-vectors with z = 0
double or_order_parameter(const double* r, const double* rdot, int32_t n) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int32_t i = 0; i < n; ++i) {
    const double ax = r[2 * i], ay = r[2 * i + 1], az = 0.0;
    const double bx = rdot[2 * i], by = rdot[2 * i + 1], bz = 0.0;
    double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    double nn = std::sqrt(cx * cx + cy * cy + cz * cz);
    if (nn > 0) { cx /= nn; cy /= nn; cz /= nn; }  // rowwise().normalized(); zero rows stay zero
    sx += cx; sy += cy; sz += cz;
  }
  return (1.0 / n) * std::sqrt(sx * sx + sy * sy + sz * sz);
}

}  // extern "C"

namespace {

// U1 a-f (Locomotion::simulate_flight up to the position update) + distances; returns the
// dense distance matrix in s.dist_length and writes r_dot into vx, vy.
void flight_common(const or_mesh* m, const or_params* p, int32_t n, const double* x, const double* y,
                   const double* cos_t, const double* sin_t, const int32_t* vid, double mu, StepScratch& s, double* vx,
                   double* vy) {
  const size_t nn = (size_t)n * n;
  s.r.resize(2 * (size_t)n);
  for (int32_t i = 0; i < n; ++i) { s.r[2 * i] = x[i]; s.r[2 * i + 1] = y[i]; }
  s.r_old = s.r;
  s.dx.resize(nn);
  s.dy.resize(nn);
  s.dist_length.resize(nn);
  s.F.resize(2 * (size_t)n);
  or_get_dist_vect(s.r.data(), n, s.dx.data(), s.dy.data());                  // Locomotion.cpp:62
  or_distances_between_particles(m, vid, n, s.dist_length.data());            // Locomotion.cpp:64
  or_calculate_forces_between_particles(s.dist_length.data(), s.dx.data(), s.dy.data(), n, p->k, p->sigma, 1.0, 0.75,
                                        s.F.data());                          // Locomotion.cpp:67-69
  for (int32_t i = 0; i < n; ++i) {
    double absF = std::sqrt(s.F[2 * i] * s.F[2 * i] + s.F[2 * i + 1] * s.F[2 * i + 1]);  // :71
    double spd = p->v0 + mu * absF;                                                      // :77
    vx[i] = cos_t[i] * spd;                                                              // :81
    vy[i] = sin_t[i] * spd;
    s.r[2 * i] += vx[i] * p->dt;                                                         // :84
    s.r[2 * i + 1] += vy[i] * p->dt;
  }
}

int32_t finish_step(const or_mesh* m, const or_params* p, int32_t n, StepScratch& s, double* heading, int32_t unit_deg,
                    double* x, double* y, int32_t* vid, int32_t remap_mode, const double* vx, const double* vy,
                    double* r3d, int32_t* face, int32_t* nbr, double* phi) {
  int32_t status = OR_OK;
  for (int32_t i = 0; i < n; ++i)
    if (!std::isfinite(s.r[2 * i]) || !std::isfinite(s.r[2 * i + 1]) || !std::isfinite(heading[i]))
      return OR_E_NONFINITE;  // the reference's tiling loop would not end on NaN; reported instead
  // U2: seam wrap (2DTissue.cpp:228)
  int32_t st = or_seam_wrap(s.r.data(), s.r_old.data(), heading, n, unit_deg, p->max_crossings, nullptr);
  if (st != OR_OK) status = st;
  // U3: remap (2DTissue.cpp:230-232)
  or_remap(m, s.r.data(), n, remap_mode, r3d, vid, face);
  // U4: validation (2DTissue.cpp:235-236; Validation.cpp:40-72)
  for (int32_t i = 0; i < n; ++i) {
    if (!std::isfinite(s.r[2 * i]) || !std::isfinite(s.r[2 * i + 1])) return OR_E_NONFINITE;
    if (!in_polygon(s.r[2 * i], s.r[2 * i + 1]) && status == OR_OK) status = OR_E_LOST_PARTICLE;
  }
  // U5: neighbour count from the step's distance matrix (2DTissue.cpp:239)
  or_count_particle_neighbors(s.dist_length.data(), n, p->sigma, nbr);
  // U6: order parameter from post-wrap positions and the unrotated r_dot (2DTissue.cpp:242)
  std::vector<double> rdot(2 * (size_t)n);
  for (int32_t i = 0; i < n; ++i) { rdot[2 * i] = vx[i]; rdot[2 * i + 1] = vy[i]; }
  phi[0] = or_order_parameter(s.r.data(), rdot.data(), n);
  for (int32_t i = 0; i < n; ++i) { x[i] = s.r[2 * i]; y[i] = s.r[2 * i + 1]; }
  return status;
}

}  // namespace

extern "C" {

int32_t or_step(const or_mesh* m, const or_params* p, int32_t n, double* x, double* y, double* theta, int32_t* vid,
                const double* noise, int32_t remap_mode, double* vx, double* vy, double* r3d, int32_t* face,
                int32_t* nbr, double* phi) {
  if (n <= 0) return OR_E_INVALID;
  StepScratch s;
  std::vector<double> c(n), sn(n);
  for (int32_t i = 0; i < n; ++i) { c[i] = std::cos(theta[i]); sn[i] = std::sin(theta[i]); }
  flight_common(m, p, n, x, y, c.data(), sn.data(), vid, p->mu, s, vx, vy);
  // U1 g, continuous: theta_i <- theta_i + (dt/tau) wrap(mean_i - theta_i) + eta_i, mean over
  // {j : d_ij < 2 sigma} including i (OrientationHelper.cpp:29-74 with real-valued headings)
  const double rate = p->dt / p->tau;
  std::vector<double> th_new(n);
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int32_t i = 0; i < n; ++i) {
    double mx = 0.0, my = 0.0;
    for (int32_t j = 0; j < n; ++j)
      if (s.dist_length[(size_t)i * n + j] < 2 * p->sigma) { mx += c[j]; my += sn[j]; }
    double nn = std::sqrt(mx * mx + my * my);
    if (nn > 0) { mx /= nn; my /= nn; }
    double mean = std::atan2(my, mx);
    double delta = mean - theta[i];
    if (delta > kPi) delta -= 2 * kPi;
    else if (delta <= -kPi) delta += 2 * kPi;
    th_new[i] = theta[i] + rate * delta + (noise ? noise[i] : 0.0);
  }
  int32_t st = finish_step(m, p, n, s, th_new.data(), 0, x, y, vid, remap_mode, vx, vy, r3d, face, nbr, phi);
  // headings are kept in [-pi, pi) after the seam rotation
  double px = 0.0, py = 0.0;
  for (int32_t i = 0; i < n; ++i) {
    double t = th_new[i];
    if (std::isfinite(t)) t = t - 2 * kPi * std::floor((t + kPi) / (2 * kPi));
    theta[i] = t;
    px += std::cos(t);
    py += std::sin(t);
  }
  phi[1] = std::sqrt(px * px + py * py) / n;
  return st;
}

int32_t or_step_int_degrees(const or_mesh* m, const or_params* p, int32_t n, double* x, double* y, int32_t* n_deg,
                            int32_t* vid, int32_t remap_mode, double* vx, double* vy, double* r3d, int32_t* face,
                            int32_t* nbr, double* phi) {
  if (n <= 0) return OR_E_INVALID;
  StepScratch s;
  std::vector<double> nv(2 * (size_t)n), c(n), sn(n);
  or_angles_to_unit_vectors(n_deg, n, nv.data());                          // Locomotion.cpp:72
  for (int32_t i = 0; i < n; ++i) { c[i] = nv[2 * i]; sn[i] = nv[2 * i + 1]; }
  flight_common(m, p, n, x, y, c.data(), sn.data(), vid, 1.0, s, vx, vy);
  if (or_average_n_within_distance_deg(s.dist_length.data(), n, n_deg, p->sigma) != OR_OK) return OR_E_INVALID;
  std::vector<double> h(n);
  for (int32_t i = 0; i < n; ++i) h[i] = n_deg[i];
  int32_t st = finish_step(m, p, n, s, h.data(), 1, x, y, vid, remap_mode, vx, vy, r3d, face, nbr, phi);
  for (int32_t i = 0; i < n; ++i) n_deg[i] = (int32_t)h[i];  // EuclideanTiling.cpp:52
  double px = 0.0, py = 0.0;
  for (int32_t i = 0; i < n; ++i) { px += std::cos(n_deg[i] * DEG_TO_RAD); py += std::sin(n_deg[i] * DEG_TO_RAD); }
  phi[1] = std::sqrt(px * px + py * py) / n;
  return st;
}

int32_t or_neighbor_sets(const or_mesh* m, const int32_t* vid, int32_t n, double sigma, int32_t* row_ptr,
                         int32_t* cols, int32_t cap) {
  int32_t k = 0;
  row_ptr[0] = 0;
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t j = 0; j < n; ++j) {
      double d = (i == j) ? 0.0 : std::min(lookup(m, vid[i], vid[j]), lookup(m, vid[j], vid[i]));
      if (d < 2 * sigma) {
        if (k >= cap) return OR_E_INVALID;
        cols[k++] = j;
      }
    }
    row_ptr[i + 1] = k;
  }
  return OR_OK;
}

}  // extern "C"
