/* tissue.h - C-ABI of the MI355X (gfx950) per-step particle update of 2DTissue.
 *
 * The reference has no FFI: its boundary is the C++ class _2DTissue (src/simulation/2DTissue.h:34-129)
 * and the two strategy interfaces LocomotionHelperInterface / OrientationHelperInterface
 * (src/simulation/LocomotionHelperInterface.h:3-13). This header is the plain-C layer under a
 * drop-in replacement of that class (include/2DTissue.h); every entry point below names the
 * reference code it replaces.
 *
 * Conventions
 *  - Every call returns an int32 status (TISSUE_OK = 0). No exceptions cross the boundary.
 *    tissue_error_string() returns the calling thread's last message.
 *  - Host pointers are borrowed for the duration of a call only. "_device" entry points take
 *    device pointers (memory on the simulation's GPU).
 *  - A tissue_sim / tissue_batch owns its device memory and one HIP stream (default priority).
 *    A handle is not thread-safe; distinct handles are independent. A mesh may be shared by any
 *    number of handles and must outlive them.
 *  - Arrays are structure-of-arrays, FP32 for state, int32 for ids; r3d is n*3 row-major.
 */
#ifndef TISSUE_H
#define TISSUE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define TISSUE_OK 0
#define TISSUE_E_INVALID_ARG 1
#define TISSUE_E_LOST_PARTICLE 2      /* Validation::error_lost_particles throws   (Validation.cpp:66-72)  */
#define TISSUE_E_NONFINITE 3          /* Validation::error_invalid_values exits     (Validation.cpp:40-46)  */
#define TISSUE_E_TILING_NO_CONVERGE 4 /* seam wrap over its crossing bound; the reference loops forever
                                         (Locomotion/EuclideanTiling.cpp:41-68)                          */
#define TISSUE_E_HIP 5                /* HIP runtime error (no device, launch failure, out of memory)      */
#define TISSUE_E_IO 6
#define TISSUE_E_MESH 7               /* mesh is not a closed genus-0 surface, or its UV map folds         */

/* flag bits reported by tissue_get_flags (a step may raise several) */
#define TISSUE_FLAG_LOST_PARTICLE 1u
#define TISSUE_FLAG_NONFINITE 2u
#define TISSUE_FLAG_TILING 4u
#define TISSUE_FLAG_BAD_INPUT 8u      /* tissue_step_noalloc: a vertex id outside the mesh was given      */

typedef struct tissue_mesh tissue_mesh;
typedef struct tissue_sim tissue_sim;

/* Physics of one simulation. Replaces the _2DTissue constructor arguments (2DTissue.h:37-54).
 * speed = v0 + mu |F|; heading += (dt/tau) wrap(mean - heading) + eta. With mu = 1 and tau = dt
 * this is the reference's update (which has neither parameter). */
typedef struct {
  double sigma;          /* particle radius (2DTissue.h:49)                                   */
  double v0;             /* self-propulsion speed (2DTissue.h:44)                             */
  double k;              /* repulsion stiffness (2DTissue.h:46)                               */
  double mu;             /* mobility (2DTissue.h:50)                                          */
  double tau;            /* alignment time                                                   */
  double dt;             /* step size (2DTissue.h:53, step_size)                               */
  int32_t max_crossings; /* seam crossings allowed per particle per step; <= 0 selects 16384   */
  int32_t reserved;
} tissue_params;

const char* tissue_error_string(void);

/* ---- mesh: replaces the construction-time MeshCartographyLib calls (2DTissue.cpp:85-107) ---- */
/* icosphere of frequency freq (V = 10 freq^2 + 2) scaled to semi-axes (a, b, c) */
int32_t tissue_mesh_ellipsoid(double a, double b, double c, int32_t freq, tissue_mesh** out);
/* unit icosphere with radius 1 + amp * Y_3^2(theta, phi) */
int32_t tissue_mesh_perturbed_sphere(double amp, int32_t freq, tissue_mesh** out);
/* closed genus-0 triangle mesh from an OFF file (loadMeshFaces + create_uv_surface, 2DTissue.cpp:85-91) */
int32_t tissue_mesh_load_off(const char* path, tissue_mesh** out);
/* face buckets (grid <= 0: automatic) and the geodesic table on the open mesh, truncated at
 * 2.4 * sigma_max (CachedGeodesicDistanceHelper(<stem>_open.off), 2DTissue.cpp:102-105): the exact
 * polyhedral geodesic distance between vertices (shortest path on the piecewise-flat surface, by
 * window propagation on the host's threads), rounded to FP32 */
int32_t tissue_mesh_build_tables(tissue_mesh* m, double sigma_max, int32_t grid);
/* where tissue_mesh_build_tables_on builds the table */
enum { TISSUE_BUILD_AUTO = -1, TISSUE_BUILD_HOST = 0, TISSUE_BUILD_GPU = 1 };
/* tissue_mesh_build_tables with the place chosen: TISSUE_BUILD_HOST and TISSUE_BUILD_AUTO build on
 * the host's threads; TISSUE_BUILD_GPU is rejected (TISSUE_E_INVALID_ARG): the table has no GPU
 * build. Replaces the same reference call as tissue_mesh_build_tables. */
int32_t tissue_mesh_build_tables_on(tissue_mesh* m, double sigma_max, int32_t grid, int32_t where);
/* Where the mesh constructors (tissue_mesh_ellipsoid, tissue_mesh_perturbed_sphere,
 * tissue_mesh_load_off) solve the UV map's harmonic system, process-wide: TISSUE_BUILD_AUTO (the
 * default: the current GPU when one is visible and the system has 100,000 free vertices or more,
 * else the host's threads), TISSUE_BUILD_HOST, or TISSUE_BUILD_GPU (a constructor then fails with
 * TISSUE_E_HIP when no GPU is visible). The GPU runs the host's
 * iteration operation for operation, so the UV map is the same bit for bit either way. Part of
 * the UV chart that replaces MeshCartographyLib's create_uv_surface (2DTissue.cpp:85-91), whose
 * source is absent. */
int32_t tissue_set_uv_solve(int32_t where);
int32_t tissue_get_uv_solve(int32_t* where);
void tissue_mesh_destroy(tissue_mesh* m);
int32_t tissue_mesh_info(const tissue_mesh* m, int32_t* nv, int32_t* nf, int32_t* nv3d, int64_t* nnz, int32_t* grid,
                         int64_t* n_cell_faces);
/* copies of the tables the step reads; any NULL pointer is skipped */
int32_t tissue_mesh_export(const tissue_mesh* m, double* uv /* nv*2 */, double* xyz /* nv*3 */, int32_t* faces,
                           int32_t* row_ptr, int32_t* col, float* dist, int32_t* cell_start, int32_t* cell_faces);
/* the 3-D mesh and its UV chart as OFF files (loadMeshFaces' format) */
int32_t tissue_mesh_save_off(const tissue_mesh* m, const char* path_3d, const char* path_uv);

/* ---- one simulation on the current GPU ---- */
/* replaces _2DTissue::_2DTissue (2DTissue.cpp:20-111) for the per-step state */
int32_t tissue_create(const tissue_mesh* m, const tissue_params* p, int32_t n, tissue_sim** out);
void tissue_destroy(tissue_sim* s);
/* replaces _2DTissue::start (2DTissue.cpp:117-134): positions uniform by area, headings uniform
 * on [0, 2 pi), then the UV->3D remap for the vertex ids */
int32_t tissue_init_random(tissue_sim* s, uint64_t seed);
/* state upload; vid == NULL computes vertex ids with the remap, as start() does */
int32_t tissue_set_state(tissue_sim* s, const float* x, const float* y, const float* theta, const int32_t* vid);
int32_t tissue_get_state(tissue_sim* s, float* x, float* y, float* theta, int32_t* vid);
/* per-particle outputs of the last step (System/Particle, Struct.h:9-29; 2DTissue.cpp:148-166) */
int32_t tissue_get_outputs(tissue_sim* s, float* vx, float* vy, float* r3d, int32_t* face, int32_t* nbr);
/* replaces n_steps calls of _2DTissue::update (2DTissue.cpp:136-206): noise is eta per step and
 * particle, host memory [n_steps][n] (NULL: zero, as the reference's uniform(0,0) noise) */
int32_t tissue_step(tissue_sim* s, int32_t n_steps, const float* noise);
int32_t tissue_step_device(tissue_sim* s, int32_t n_steps, const float* noise_device);
/* copy rows [n_rows][n] of eta into the handle's device buffer once, then step from row first_row */
int32_t tissue_stage_noise(tissue_sim* s, const float* noise, int32_t n_rows);
int32_t tissue_step_staged(tissue_sim* s, int32_t n_steps, int32_t first_row);
/* order parameter history (v_order, 2DTissue.cpp:242; get_order_parameter, :213-216):
 * phi_ref = reference reading, phi_polar = |sum (cos, sin)| / n; either pointer may be NULL */
int32_t tissue_get_order(tissue_sim* s, double* phi_ref, double* phi_polar, int32_t cap, int32_t* count);
int32_t tissue_reset_order(tissue_sim* s);
/* validation flags raised since the last call (clears them) and the first step that raised one */
int32_t tissue_get_flags(tissue_sim* s, uint32_t* flags, int32_t* first_step);
/* neighbour sets {j : d_ij < 2 sigma} (self included) of the current state, at most cap per
 * particle: counts[n], cols[n*cap], each list ascending */
int32_t tissue_neighbor_lists(tissue_sim* s, int32_t* counts, int32_t* cols, int32_t cap);
/* Visit check (no counterpart in the reference; a diagnostic for parity tests of the step's own
 * neighbour walk, which the reference performs as the dense loops of Locomotion.cpp:94-111 and
 * ForceHelper.cpp:34-73). While enabled, every step of the step kernel (TISSUE_MODE_STEP_KERNEL)
 * records per particle the number of particles its walk visited within 2 sigma (self included) and
 * the sum, modulo 2^32, of h(j) over their indices j (counted over all replicas' particles), with
 * h the MurmurHash3 32-bit finaliser; tissue_get_visits returns those of the last step of the step
 * kernel (-1 for a particle it has not stepped since the check was enabled, or whose replica was
 * stopped; steps of the persistent kernel, TISSUE_MODE_RESIDENT, record nothing). */
int32_t tissue_set_visit_check(tissue_sim* s, int32_t enable);
int32_t tissue_get_visits(tissue_sim* s, int32_t* na, uint32_t* hash);
/* launch shape: lanes per particle (1, 2, 4, 8, 16, 32, 64) and threads per block (64..512); 0 = auto */
int32_t tissue_set_launch(tissue_sim* s, int32_t lanes, int32_t block);
/* execution mode of tissue_step*: one kernel launch per step over the whole GPU, or the
 * persistent LDS-resident kernel (one workgroup, many steps per launch; needs
 * n <= tissue_resident_capacity). AUTO picks the resident kernel only for n <= 256 (one pass of
 * a workgroup), the step kernel otherwise, unless tissue_autotune chose for this size. */
#define TISSUE_MODE_AUTO 0
#define TISSUE_MODE_STEP_KERNEL 1
#define TISSUE_MODE_RESIDENT 2
int32_t tissue_set_mode(tissue_sim* s, int32_t mode);
int32_t tissue_resident_capacity(const tissue_mesh* m, int32_t* max_particles);
/* the launch shape the next step uses, shape[5] = {mode, lanes, block, steps per launch,
 * xcd order (1: each replica's tiles on one XCD)} */
int32_t tissue_get_launch(tissue_sim* s, int32_t* shape);
/* time n_steps of every mode and launch shape that fits (state restored after each trial), keep
 * the fastest for AUTO; shape[5] as tissue_get_launch (may be NULL) */
int32_t tissue_autotune(tissue_sim* s, int32_t n_steps, int32_t* shape);
int32_t tissue_synchronize(tissue_sim* s);
/* GPU time of the step kernels of the last tissue_step call (HIP events on the handle's stream) */
int32_t tissue_last_step_ms(tissue_sim* s, double* ms);

/* ---- many independent replicas, one hipGraph per run ----
 * Replica r owns particles [off_r, off_r + n[r]) of the concatenated arrays (off_r = sum of the
 * earlier n). Noise rows are [steps][sum n] with replica r's columns at off_r. Two modes: the
 * tiled step kernel (one launch per step covers every replica, workgroups are tiles of one
 * replica, optionally placed so that each replica stays on one XCD) or the persistent kernel (one
 * workgroup per replica, state resident in LDS, many steps per launch). The launches of a run are
 * captured in a hipGraph and replayed. A replica that raises a validation flag stops (its later
 * steps are no-ops, also in later runs, until its state is set again). */
typedef struct tissue_batch tissue_batch;
int32_t tissue_batch_create(const tissue_mesh* m, int32_t n_replicas, const int32_t* n, const tissue_params* params,
                            tissue_batch** out);
void tissue_batch_destroy(tissue_batch* b);
int32_t tissue_batch_init_random(tissue_batch* b, const uint64_t* seeds /* [n_replicas] */);
int32_t tissue_batch_set_state(tissue_batch* b, const float* x, const float* y, const float* theta, const int32_t* vid);
int32_t tissue_batch_get_state(tissue_batch* b, float* x, float* y, float* theta, int32_t* vid);
int32_t tissue_batch_get_outputs(tissue_batch* b, float* vx, float* vy, float* r3d, int32_t* face, int32_t* nbr);
/* eta for replica r, step t: stds[r] * N(0,1) of the stream seeds[r] (as tissue_gen_noise) */
int32_t tissue_batch_stage_noise(tissue_batch* b, const float* noise, int32_t n_rows);
int32_t tissue_batch_gen_noise(tissue_batch* b, const uint64_t* seeds, const double* stds, int32_t n_rows);
/* n_steps steps from staged noise row first_row (first_row < 0: no noise) */
int32_t tissue_batch_run(tissue_batch* b, int32_t n_steps, int32_t first_row);
int32_t tissue_batch_get_order(tissue_batch* b, int32_t replica, double* phi_ref, double* phi_polar, int32_t cap,
                               int32_t* count);
int32_t tissue_batch_get_flags(tissue_batch* b, uint32_t* flags /* [n_replicas] */, int32_t* first_step);
/* lanes per particle (1, 2, 4, 8, 16, 32, 64; the persistent kernel uses at most 8), threads per
 * workgroup (64..512), steps per launch; 0 = auto */
int32_t tissue_batch_set_launch(tissue_batch* b, int32_t lanes, int32_t block, int32_t steps_per_launch);
/* TISSUE_MODE_* and the tiled mode's workgroup order: xcd_order 1 keeps each replica on one XCD,
 * 0 is replica order, -1 automatic */
int32_t tissue_batch_set_mode(tissue_batch* b, int32_t mode, int32_t xcd_order);
/* shape[5] = {mode, lanes, block, steps per launch, xcd order} of the next run */
int32_t tissue_batch_get_launch(tissue_batch* b, int32_t* shape);
/* time n_steps of every mode and launch shape that fits (state restored after each trial) and
 * keep the fastest for AUTO; results are cached per (replicas, largest replica, particles, mesh)
 * for the process */
int32_t tissue_batch_autotune(tissue_batch* b, int32_t n_steps, int32_t* lanes, int32_t* block);
/* GPU time of the last tissue_batch_run (HIP events around the graph launch) */
int32_t tissue_batch_last_run_ms(tissue_batch* b, double* ms);
/* the visit check of tissue_set_visit_check over all replicas' particles ([total particles]; the
 * indices hashed are the global ones, off_r + i); no counterpart in the reference */
int32_t tissue_batch_set_visit_check(tissue_batch* b, int32_t enable);
int32_t tissue_batch_get_visits(tissue_batch* b, int32_t* na, uint32_t* hash);

/* ---- adjoint: gradients of a trajectory loss (no counterpart in the reference, which has no
 * differentiation; it differentiates n_steps of tissue_step) ----
 * loss = w_order (phi_T - target_order)^2 + w_velocity mean_{t<T, i} |rdot_i^t - u(r_i^t)|^2
 * phi_T = |sum_i (cos, sin)(theta_i^T)| / n (polar order of the final state); rdot_i^t and r_i^t
 * are step t's velocity and pre-step position; u is a smooth field on the UV square, a sum of
 * modes u_c(x, y) += amp sin(kx x + ky y + phase) with c = component (0: x, 1: y).
 * The reverse pass runs through the forward's frozen decisions (neighbour sets, seam-wrap maps,
 * vertex ids), recomputing each segment of ckpt_every steps from a checkpoint. */
#define TISSUE_MAX_MODES 16
typedef struct {
  double target_order;
  double w_order;
  double w_velocity;
  int32_t n_modes;
  int32_t reserved;
  int32_t component[TISSUE_MAX_MODES];
  double amp[TISSUE_MAX_MODES], kx[TISSUE_MAX_MODES], ky[TISSUE_MAX_MODES], phase[TISSUE_MAX_MODES];
} tissue_loss;
typedef struct {
  double loss, loss_order, loss_velocity, phi_final;
  double d_tau, d_mu, d_v0, d_k;  /* dloss / d parameter */
  double forward_ms, reverse_ms;  /* GPU time of the checkpointed forward and of the reverse pass */
} tissue_grads;
/* runs n_steps from the current state (noise from staged row first_row, < 0: none), leaves the
 * final state, and returns the gradients; d_theta0 / d_x0 / d_y0 [n] (may be NULL) are dloss with
 * respect to the initial headings and positions. The adjoint state is FP32, rescaled per replica
 * by powers of two after every step, so gradients far beyond the FP32 range (the d == 0 repulsion
 * rule makes them grow geometrically with the horizon) come back as FP64. */
int32_t tissue_adjoint(tissue_sim* s, int32_t n_steps, int32_t first_row, int32_t ckpt_every, const tissue_loss* loss,
                       tissue_grads* g, double* d_theta0, double* d_x0, double* d_y0);
/* every replica at once, each with its own loss and gradients: g[n_replicas], d_*[sum n] */
int32_t tissue_batch_adjoint(tissue_batch* b, int32_t n_steps, int32_t first_row, int32_t ckpt_every,
                             const tissue_loss* loss, tissue_grads* g, double* d_theta0, double* d_x0, double* d_y0);

/* ---- per-step output stream ---- */
/* From the next step on, every `every`-th step's UV positions, 3-D positions and neighbour counts
 * are written to the reference's per-step CSV files in dir (which must exist):
 * r_data_<step>.csv (x,y), r_data_3D_<step>.csv (x,y,z), particles_color_<step>.csv (count),
 * appended, 15 significant digits, <step> = steps done since this call (1 after the first); replaces the
 * per-step writes of _2DTissue::update (2DTissue.cpp:168-184,270-280, IO.h:40-69). The outputs are
 * copied on the GPU into a ring, to the host after each run, and written by a background thread
 * while the GPU runs on; runs use the per-step kernel while recording. */
int32_t tissue_record_start(tissue_sim* s, const char* dir, int32_t every);
/* finishes writing every recorded step and stops the stream (TISSUE_E_IO if a write failed);
 * destroying the simulation also finishes the files */
int32_t tissue_record_stop(tissue_sim* s);

/* ---- seam wrap on its own ---- */
/* Applies the seam wrap of the step to n moves on the GPU, as
 * EuclideanTiling::diagonal_seam_edges_square_border (EuclideanTiling.cpp:31-69) does to r_UV_old
 * -> r_UV: start_xy[2n] are the positions before the move (in the unit square), end_xy[2n] the
 * positions after it, overwritten with the wrapped positions; dtheta[n] receives the heading
 * rotation (radians), quarter[n] (may be NULL) the quarter turns of the composed map. lanes (1, 2,
 * 4, 8, 16, 32, 64) is the number of threads per move, as in the step kernel; max_crossings <= 0
 * selects 16384; wave = 1 runs the loop of crowded systems (a wave's last unfinished move is taken
 * over by all 64 lanes of the wave, three crossings per pass), 0 the plain loop: same results.
 * Returns TISSUE_E_TILING_NO_CONVERGE when a move needs more crossings (its outputs are those
 * reached at the bound). Host arrays; runs synchronously on the current device. */
int32_t tissue_seam_wrap(int32_t n, const double* start_xy, double* end_xy, double* dtheta, int32_t* quarter,
                         int32_t max_crossings, int32_t lanes, int32_t wave);

/* ---- one step without allocation ----
 * No counterpart in the reference. For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. One step of one system
 * of n particles, as one tissue_step of a simulation on the same mesh with the same parameters
 * and state (default launch shape of the per-step kernel), bit for bit. Every pointer except
 * mesh and p is device memory on the current GPU: the state (x_in, y_in, theta_in, vid_in [n]),
 * the noise eta [n] (NULL: zero), and the outputs (the new state x_out, y_out, theta_out,
 * vid_out [n]; r_dot vx, vy; the 3-D position r3x, r3y, r3z; face; neighbour count nbr [n];
 * phi[2] = (phi_ref, phi_polar) of the step; flags[1] = TISSUE_FLAG_* raised, which a call clears
 * first). Outputs must not alias inputs. The mesh's device tables must already be on this GPU
 * (any tissue_create on the mesh puts them there). The work is enqueued on `stream` (a
 * hipStream_t; NULL: the default stream) and the call returns without waiting for it.
 * Every buffer beyond the caller's lives in a static device array of
 * TISSUE_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the process ends:
 * calls on one GPU share it, so calls that overlap on one GPU race. Sizes beyond
 * TISSUE_NOALLOC_MAX_PARTICLES particles or TISSUE_NOALLOC_MAX_UV_VERTICES UV vertices are
 * rejected with TISSUE_E_INVALID_ARG. */
#define TISSUE_NOALLOC_MAX_PARTICLES 16384
#define TISSUE_NOALLOC_MAX_UV_VERTICES 32768
#define TISSUE_NOALLOC_SCRATCH_BYTES (6u << 20)
int32_t tissue_step_noalloc(const tissue_mesh* mesh, const tissue_params* p, int32_t n, const float* x_in,
                            const float* y_in, const float* theta_in, const int32_t* vid_in, const float* noise,
                            float* x_out, float* y_out, float* theta_out, int32_t* vid_out, float* vx, float* vy,
                            float* r3x, float* r3y, float* r3z, int32_t* face, int32_t* nbr, double* phi,
                            uint32_t* flags, void* stream);

/* ---- host helpers (no GPU) ---- */
/* The area-uniform initial state that tissue_init_random(seed) draws on the GPU, computed on the
 * host: face by 3-D area, uniform point of its UV triangle, heading uniform on [-pi, pi), from
 * Philox4x32-10 streams of the seed (replaces the reference's unseeded draw, CellHelper.cpp:43-67). */
int32_t tissue_random_init_host(const tissue_mesh* m, uint64_t seed, int32_t n, float* x, float* y, float* theta);
/* eta[t][i] = std * N(0,1) from Philox4x32-10 (Salmon et al., SC'11) + Box-Muller */
int32_t tissue_gen_noise(uint64_t seed, double std, int32_t n_steps, int32_t n, float* out);
/* Philox4x32-10 block for (counter, key); exposed for known-answer tests */
void tissue_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
