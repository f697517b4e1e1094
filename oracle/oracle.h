/* FP64 CPU oracle for the 2DTissue per-step particle update.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it. The product (libtissue) never links or calls it.
 *
 * It restates, function by function and in FP64, the reference's step (file:line citations
 * are into KAIZN-io/2DTissue, src/simulation/...):
 *   U1  Locomotion::simulate_flight            Locomotion.cpp:57-89 (dense N x N, as the reference)
 *   U2  EuclideanTiling seam wrap              Locomotion/EuclideanTiling.cpp:31-208 (restart loop)
 *   U3  CellHelper::get_r3d remap              CellHelper.cpp:71-85,104-230 (brute force over all faces)
 *   U4  Validation                             Validation.cpp:25-72
 *   U5  count_particle_neighbors               2DTissue.cpp:254-268
 *   U6  LinearAlgebra::calculate_order_parameter LinearAlgebra.cpp:67-105 (z := 0 reading)
 *
 * Two heading modes:
 *   - int degrees (reference-literal, Eigen::VectorXi n): or_step_int_degrees, used to pin the
 *     oracle against the reference's own test vectors;
 *   - continuous radians (the model the GPU implements): theta <- theta + (dt/tau) wrap(mean - theta)
 *     + eta, speed v0 + mu |F|, headings kept in [-pi, pi). With tau = dt, eta = 0, mu = 1 it is
 *     the reference's update with real-valued headings.
 *
 * Pinning (see tests/test_oracle.py): every primitive is checked against the reference's test
 * vectors (test_LinearAlgebra.cpp, test_EuclideanTiling.cpp, the vectors kept in the commented
 * test_Locomotion.cpp) or, where the reference holds none, against closed forms, brute force or
 * invariants. Parity unpinned: the table values produced by MeshCartographyLib (UV map, geodesic
 * distances), which the reference cannot produce here; the oracle takes those tables as inputs.
 */
#ifndef TISSUE_ORACLE_H
#define TISSUE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t nv, nf;
  const double* uv;        /* nv*2 */
  const double* xyz;       /* nv*3 */
  const int32_t* faces;    /* nf*3 */
  const int32_t* row_ptr;  /* nv+1: truncated distance table, CSR */
  const int32_t* col;
  const float* dist;
  int32_t grid;            /* optional face buckets (or_remap with mode 1) */
  const int32_t* cell_start;
  const int32_t* cell_faces;
} or_mesh;

typedef struct {
  double sigma, v0, k, mu, tau, dt;
  int32_t max_crossings;   /* seam crossings allowed per particle per step */
} or_params;

enum { OR_OK = 0, OR_E_LOST_PARTICLE = 2, OR_E_NONFINITE = 3, OR_E_TILING_NO_CONVERGE = 4, OR_E_INVALID = 1 };

void or_set_threads(int32_t n);

/* ---- primitives, reference-literal ---- */
int32_t or_angles_to_unit_vectors(const int32_t* n_deg, int32_t n, double* out /* n*2 */);
int32_t or_mean_unit_circle_vector_angle_degrees(const double* angles, int32_t n, double* out);
void or_repulsive_adhesion_motion(double k, double sigma, double dist, double r_adh, double k_adh, double dvx,
                                  double dvy, double* out /* 2 */);
void or_transform_into_symmetric_matrix(double* a /* n*n row-major */, int32_t n);
void or_get_dist_vect(const double* r /* n*2 */, int32_t n, double* dx, double* dy /* n*n */);
void or_calculate_forces_between_particles(const double* dist_length, const double* dx, const double* dy, int32_t n,
                                           double k, double sigma, double r_adh, double k_adh, double* f /* n*2 */);
int32_t or_average_n_within_distance_deg(const double* dist_length, int32_t n, int32_t* n_deg, double sigma);
double or_dist_lookup(const or_mesh* m, int32_t v, int32_t w);
void or_distances_between_particles(const or_mesh* m, const int32_t* vid, int32_t n, double* dist_length);
void or_count_particle_neighbors(const double* dist_length, int32_t n, double sigma, int32_t* out);
/* heading: per-particle value in degrees (unit_deg = 1) or radians (0); crossings: per particle (may be NULL) */
int32_t or_seam_wrap(double* r /* n*2 */, double* r_old /* n*2 */, double* heading, int32_t n, int32_t unit_deg,
                     int32_t max_crossings, int32_t* crossings);
double or_point_triangle_distance(const double* p, const double* a, const double* b, const double* c);
/* mode 0: brute force over all faces (CellHelper.cpp:106-117); 1: face-bucket candidates */
void or_remap(const or_mesh* m, const double* r /* n*2 */, int32_t n, int32_t mode, double* r3d /* n*3 */,
              int32_t* vid, int32_t* face);
double or_order_parameter(const double* r /* n*2 */, const double* rdot /* n*2 */, int32_t n);

/* ---- one full step, continuous radians ----
 * in/out: x, y, theta, vid (vertex ids from the previous remap); noise: eta per particle (may be NULL)
 * out: vx, vy (r_dot), r3d (n*3), face, nbr; phi[0] = reference order parameter, phi[1] = polar order */
int32_t or_step(const or_mesh* m, const or_params* p, int32_t n, double* x, double* y, double* theta, int32_t* vid,
                const double* noise, int32_t remap_mode, double* vx, double* vy, double* r3d, int32_t* face,
                int32_t* nbr, double* phi);

/* ---- one full step, int degrees, reference-literal (mu, tau, noise unused) ---- */
int32_t or_step_int_degrees(const or_mesh* m, const or_params* p, int32_t n, double* x, double* y, int32_t* n_deg,
                            int32_t* vid, int32_t remap_mode, double* vx, double* vy, double* r3d, int32_t* face,
                            int32_t* nbr, double* phi);

/* neighbour sets of one step (d_ij < 2 sigma, self included), CSR over particles, ascending j */
int32_t or_neighbor_sets(const or_mesh* m, const int32_t* vid, int32_t n, double sigma, int32_t* row_ptr /* n+1 */,
                         int32_t* cols, int32_t cap);

int32_t or_get_threads(void);

/* ---- the same step with a sparse neighbour search (oracle/sparse.cpp): identical operations in
 * identical order (neighbours in ascending index, per-particle seam wrap), so bitwise equal to
 * or_step with remap_mode 1, and O(N) per step instead of O(N^2) ---- */
int32_t or_step_sparse(const or_mesh* m, const or_params* p, int32_t n, double* x, double* y, double* theta,
                       int32_t* vid, const double* noise, double* vx, double* vy, double* r3d, int32_t* face,
                       int32_t* nbr, double* phi);
int32_t or_neighbor_sets_sparse(const or_mesh* m, const int32_t* vid, int32_t n, double sigma, int32_t* row_ptr,
                                int32_t* cols, int32_t cap);

/* ---- reverse pass (oracle/adjoint.cpp) ----
 * loss = w_order (phi_T - target_order)^2 + w_velocity mean_{t<T, i} |rdot_i^t - u(r_i^t)|^2,
 * u_c(x, y) = sum over modes with component c of amp sin(kx x + ky y + phase) */
typedef struct {
  double target_order, w_order, w_velocity;
  int32_t n_modes;
  const int32_t* component;
  const double *amp, *kx, *ky, *phase;
} or_loss;
/* u (2) and its Jacobian (du_x/dx, du_x/dy, du_y/dx, du_y/dy) at (x, y) */
double or_field_eval(const or_loss* L, double x, double y, double* u, double* jac);
/* gradients along a trajectory: xs, ys, ths, vids are states 0..T, each [T+1][n]; decisions of
 * step t come from state t. out[8] = loss, order loss, velocity loss, phi_T, d tau, d mu, d v0,
 * d k; d_theta0, d_x0, d_y0 [n] (may be NULL) */
int32_t or_adjoint_reverse(const or_mesh* m, const or_params* p, int32_t n, int32_t T, const double* xs,
                           const double* ys, const double* ths, const int32_t* vids, const or_loss* L, double* out,
                           double* d_theta0, double* d_x0, double* d_y0);

#ifdef __cplusplus
}
#endif
#endif
