// Per-step output stream of a simulation: the step kernels' outputs of every `every`-th step are
// copied on the GPU into a ring of `depth` slots; after each run (split so that no slot is
// overwritten before it is read) the filled slots are copied to the host and handed to a writer
// thread, which appends the reference's per-step CSV files while the GPU runs on:
// dir/r_data_<step>.csv (x,y), dir/r_data_3D_<step>.csv (x,y,z), dir/particles_color_<step>.csv
// (neighbour count), 15 significant digits (2DTissue.cpp:270-280, IO.h:40-69). <step> counts the
// steps done since the stream started, 1 after the first step.
#pragma once
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace tissue {

struct Recorder {
  std::string dir;
  int every = 1, depth = 1, ntot = 0;
  long origin = 0;  // steps done when the stream started (recorded steps count from there)
  float* d_ring = nullptr;    // [depth][5][ntot]: x, y, X, Y, Z
  int* d_ring_nbr = nullptr;  // [depth][ntot]

  struct Job {
    std::vector<long> steps;  // global step numbers, one slot of f / nb each
    std::vector<float> f;     // [steps][5][ntot]
    std::vector<int> nb;      // [steps][ntot]
  };
  std::thread writer;
  std::mutex mu;
  std::condition_variable cv;       // a job or the stop request arrived
  std::condition_variable cv_idle;  // the queue drained
  std::deque<Job> queue;
  bool busy = false, quit = false;
  std::string error;  // first write failure

  void start();
  void push(Job&& j);
  void drain();  // waits until every queued job is written
  void stop();   // drains, then joins the writer
  void write(const Job& j);
};

}  // namespace tissue
