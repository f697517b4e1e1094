// 2DTissue.h - drop-in replacement of the reference's simulation class on top of the C-ABI
// (include/tissue.h, libtissue.so).
//
// Mirrors KAIZN-io/2DTissue src/simulation/2DTissue.h:34-59 (constructor arguments and defaults,
// start / update / is_finished / get_order_parameter) and src/simulation/Struct.h:9-29 (Particle,
// System). Differences a caller sees, each a decision recorded with the build:
//   - get_order_parameter returns std::vector<double> (the reference returns an Eigen::VectorXd;
//     this header does not depend on Eigen);
//   - headings are real-valued: Particle::alignment_UV is the heading in degrees in [0, 360),
//     not truncated to an integer;
//   - start() places particles uniformly by area from a seed (set_seed, default 0); the
//     reference's placement is not reproducible;
//   - the step's angular noise is zero, as in the reference (uniform(0, 0)); set_noise adds
//     N(0, std^2) per particle and step from a seed;
//   - the alignment time is step_size, which reproduces the reference's full re-alignment each
//     step; mu multiplies |F| as in the speed v0 + mu |F| (mu = 1 is the reference);
//   - validation failures throw std::runtime_error from update() (the reference throws for lost
//     particles and calls exit(1) for NaN);
//   - Kafka streaming is not available (use_kafka = true throws std::invalid_argument);
//     particle_innenleben, free_boundary, k_next, v0_next, r_adh, k_adh and map_cache_count are
//     accepted and unused, as they are in the reference's step;
//   - mesh_path is an OFF file of a closed genus-0 surface, or "ellipsoid:a,b,c,freq" /
//     "perturbed_sphere:amp,freq" for the built-in generators; the geodesic table is computed
//     from the mesh itself (the reference reads it from MeshCartographyLib's <stem>_open.off).
// With save_data the per-step CSV files of the reference (2DTissue.cpp:270-280, IO.h:40-69) are
// appended to data_dir (default "data"): r_data_<step>.csv, r_data_3D_<step>.csv,
// particles_color_<step>.csv, values written with 15 significant digits, by the library's output
// stream (tissue_record_start) in the background; they are complete once the last step has run.
#pragma once

#include <cmath>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tissue.h"

// Individual particle information (Struct.h:9-22)
struct Particle {
  double x_UV;
  double y_UV;
  double x_velocity_UV;
  double y_velocity_UV;
  double alignment_UV;
  double compass_north_pole_UV;
  double compass_south_pole_UV;
  double x_3D;
  double y_3D;
  double z_3D;
  int neighbor_count;
};

// System information (Struct.h:25-29)
struct System {
  double order_parameter;
  std::vector<Particle> particles;
};

class _2DTissue {
 public:
  _2DTissue(bool save_data, bool particle_innenleben, bool free_boundary, std::string mesh_path, int particle_count,
            int step_count = 1, double v0 = 0.1, bool use_kafka = false, double k = 1, double k_next = 10,
            double v0_next = 0.1, double sigma = 0.4166666666666667, double mu = 1, double r_adh = 1,
            double k_adh = 0.75, double step_size = 0.001, int map_cache_count = 30)
      : save_data_(save_data), particle_count_(particle_count), step_count_(step_count) {
    (void)particle_innenleben; (void)free_boundary; (void)k_next; (void)v0_next; (void)r_adh; (void)k_adh;
    (void)map_cache_count;
    if (use_kafka) throw std::invalid_argument("Kafka streaming is not supported by this library");
    if (particle_count < 1 || step_count < 0) throw std::invalid_argument("particle_count and step_count");
    try {
      check(open_mesh(mesh_path, &mesh_));
      check(tissue_mesh_build_tables(mesh_, sigma, 0));
      tissue_params p{sigma, v0, k, mu, /*tau=*/step_size, step_size, /*max_crossings=*/0, 0};
      check(tissue_create(mesh_, &p, particle_count, &sim_));
    } catch (...) {
      if (sim_) tissue_destroy(sim_);
      if (mesh_) tissue_mesh_destroy(mesh_);
      throw;
    }
    v_order_.assign((size_t)step_count, 0.0);
  }

  ~_2DTissue() {
    if (sim_) (void)tissue_record_stop(sim_);
    if (sim_) tissue_destroy(sim_);
    if (mesh_) tissue_mesh_destroy(mesh_);
  }
  _2DTissue(const _2DTissue&) = delete;
  _2DTissue& operator=(const _2DTissue&) = delete;

  // 2DTissue.cpp:117-134
  void start() {
    check(tissue_init_random(sim_, seed_));
    // the library's per-step output stream writes the reference's CSV files in the background
    if (save_data_) check(tissue_record_start(sim_, data_dir_.c_str(), 1));
    current_step_ = 0;
    finished_ = step_count_ == 0;
    noise_rows_.clear();
    if (noise_std_ > 0.0 && step_count_ > 0) {
      noise_rows_.resize((size_t)step_count_ * particle_count_);
      check(tissue_gen_noise(noise_seed_, noise_std_, step_count_, particle_count_, noise_rows_.data()));
    }
  }

  // 2DTissue.cpp:136-206: one step, then the System record of the new state
  System update() {
    const float* noise = nullptr;
    if (!noise_rows_.empty() && current_step_ < step_count_) noise = &noise_rows_[(size_t)current_step_ * particle_count_];
    const int32_t st = tissue_step(sim_, 1, noise);
    if (st == TISSUE_E_LOST_PARTICLE) throw std::runtime_error("Lost particles detected");
    if (st == TISSUE_E_NONFINITE) throw std::runtime_error("Invalid values detected (NaN or Inf)");
    check(st);
    const int n = particle_count_;
    std::vector<float> x(n), y(n), th(n), vx(n), vy(n), r3d(3 * (size_t)n);
    std::vector<int32_t> vid(n), face(n), nbr(n);
    check(tissue_get_state(sim_, x.data(), y.data(), th.data(), vid.data()));
    check(tissue_get_outputs(sim_, vx.data(), vy.data(), r3d.data(), face.data(), nbr.data()));
    // the library's history holds this step only (it is reset after every read)
    double phi_ref = 0.0;
    int32_t count = 0;
    check(tissue_get_order(sim_, &phi_ref, nullptr, 1, &count));
    check(tissue_reset_order(sim_));
    if (current_step_ < step_count_) v_order_[(size_t)current_step_] = phi_ref;

    System system;
    system.particles.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
      Particle& p = system.particles[(size_t)i];
      p.x_UV = x[i];
      p.y_UV = y[i];
      p.x_velocity_UV = vx[i];
      p.y_velocity_UV = vy[i];
      double deg = (double)th[i] * (180.0 / M_PI);
      if (deg < 0) deg += 360.0;
      p.alignment_UV = deg;
      p.compass_north_pole_UV = 0.0;  // never set by the reference either
      p.compass_south_pole_UV = 0.0;
      p.x_3D = r3d[3 * (size_t)i];
      p.y_3D = r3d[3 * (size_t)i + 1];
      p.z_3D = r3d[3 * (size_t)i + 2];
      p.neighbor_count = nbr[i];
    }
    // the reference reports the last slot of v_order, which fills on the final step (2DTissue.cpp:165)
    system.order_parameter = v_order_.empty() ? 0.0 : v_order_.back();
    ++current_step_;
    if (current_step_ >= step_count_) {
      finished_ = true;
      if (save_data_) check(tissue_record_stop(sim_));  // every CSV file complete
    }
    return system;
  }

  bool is_finished() const { return finished_; }
  std::vector<double> get_order_parameter() const { return v_order_; }

  // additions: reproducibility and access to the C-ABI handle
  void set_seed(uint64_t seed) { seed_ = seed; }
  // (call before start(); the rows are those of tissue_gen_noise(seed, std_dev, step_count, n))
  void set_noise(double std_dev, uint64_t seed) { noise_std_ = std_dev; noise_seed_ = seed; }
  void set_data_dir(const std::string& dir) { data_dir_ = dir; }
  tissue_sim* handle() { return sim_; }

 private:
  static void check(int32_t st) {
    if (st != TISSUE_OK) throw std::runtime_error(std::string("libtissue: ") + tissue_error_string());
  }

  static int32_t open_mesh(const std::string& path, tissue_mesh** out) {
    auto numbers = [](const std::string& s) {
      std::vector<double> v;
      std::stringstream ss(s);
      std::string tok;
      while (std::getline(ss, tok, ',')) v.push_back(std::stod(tok));
      return v;
    };
    if (path.rfind("ellipsoid:", 0) == 0) {
      const auto v = numbers(path.substr(10));
      if (v.size() != 4) throw std::invalid_argument("ellipsoid:a,b,c,freq");
      return tissue_mesh_ellipsoid(v[0], v[1], v[2], (int32_t)v[3], out);
    }
    if (path.rfind("perturbed_sphere:", 0) == 0) {
      const auto v = numbers(path.substr(17));
      if (v.size() != 2) throw std::invalid_argument("perturbed_sphere:amp,freq");
      return tissue_mesh_perturbed_sphere(v[0], (int32_t)v[1], out);
    }
    return tissue_mesh_load_off(path.c_str(), out);
  }

  bool save_data_;
  int particle_count_, step_count_;
  int current_step_ = 0;
  bool finished_ = false;
  uint64_t seed_ = 0;
  double noise_std_ = 0.0;
  uint64_t noise_seed_ = 0;
  std::string data_dir_ = "data";
  std::vector<double> v_order_;
  std::vector<float> noise_rows_;  // [step_count][particle_count] when set_noise was called
  tissue_mesh* mesh_ = nullptr;
  tissue_sim* sim_ = nullptr;
};
