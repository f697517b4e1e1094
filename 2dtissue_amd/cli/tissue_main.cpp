// Command-line driver with the reference's flags (src/main.cpp:20-101), on the drop-in class of
// include/2DTissue.h. Flags and their mapping are the reference's: --step-time is passed as v0
// (main.cpp:76-84). Additions: --sigma, --seed, --noise-std, --data-dir. Without --mesh-path the
// mesh is the built-in ellipsoid (2, 1, 1) at icosphere frequency 22 (the reference's default
// mesh lives in MeshCartographyLib, which is not part of this library).
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>

#include "2DTissue.h"

namespace {

void usage() {
  std::cerr << "usage: tissue_main [--step-count N] [--save-data] [--particle-innenleben]\n"
               "                   [--optimized-monotile-boundary] [--mesh-path PATH] [--particle-count N]\n"
               "                   [--step-time T] [--kafka] [--sigma S] [--seed N] [--noise-std S]\n"
               "                   [--data-dir DIR]\n";
}

}  // namespace

int main(int argc, char** argv) {
  int step_count = 20, particle_count = 1000;
  bool save_data = false, innenleben = false, monotile = false, kafka = false;
  std::string mesh_path = "ellipsoid:2,1,1,22", data_dir = "data";
  double step_time = 0.01, sigma = 0.4166666666666667, noise_std = 0.0;
  unsigned long long seed = 0;
  for (int a = 1; a < argc; ++a) {
    const std::string f = argv[a];
    auto value = [&]() -> std::string {
      if (a + 1 >= argc) { usage(); std::exit(EXIT_FAILURE); }
      return argv[++a];
    };
    try {
      if (f == "--step-count") step_count = std::stoi(value());
      else if (f == "--save-data") save_data = true;
      else if (f == "--particle-innenleben") innenleben = true;
      else if (f == "--optimized-monotile-boundary") monotile = true;
      else if (f == "--mesh-path") mesh_path = value();
      else if (f == "--particle-count") particle_count = std::stoi(value());
      else if (f == "--step-time") step_time = std::stod(value());
      else if (f == "--kafka") kafka = true;
      else if (f == "--sigma") sigma = std::stod(value());
      else if (f == "--seed") seed = std::stoull(value());
      else if (f == "--noise-std") noise_std = std::stod(value());
      else if (f == "--data-dir") data_dir = value();
      else { usage(); return EXIT_FAILURE; }
    } catch (const std::exception&) {
      usage();
      return EXIT_FAILURE;
    }
  }
  try {
    _2DTissue tissue(save_data, innenleben, monotile, mesh_path, particle_count, step_count, step_time, kafka, 1, 10,
                     0.1, sigma);
    tissue.set_seed(seed);
    tissue.set_data_dir(data_dir);
    if (noise_std > 0) tissue.set_noise(noise_std, seed);
    tissue.start();
    const auto t0 = std::chrono::steady_clock::now();
    while (!tissue.is_finished()) {
      System data = tissue.update();
      (void)data;
    }
    const auto t1 = std::chrono::steady_clock::now();
    std::cout << std::setprecision(15);
    for (double v : tissue.get_order_parameter()) std::cout << v << '\n';
    std::cout << "Time taken: " << std::chrono::duration<double>(t1 - t0).count() << " seconds" << '\n';
  } catch (const std::exception& e) {
    std::cerr << "tissue_main: " << e.what() << '\n';
    return EXIT_FAILURE;
  }
  return 0;
}
