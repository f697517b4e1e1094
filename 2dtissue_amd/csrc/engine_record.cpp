// Per-step output stream of the engine (record.h).
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>

#include "engine.h"
#include "record.h"

namespace tissue {

void Recorder::start() {
  writer = std::thread([this] {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv.wait(lk, [this] { return quit || !queue.empty(); });
      if (queue.empty()) break;  // quit with nothing left
      Job j = std::move(queue.front());
      queue.pop_front();
      busy = true;
      lk.unlock();
      write(j);
      lk.lock();
      busy = false;
      if (queue.empty()) cv_idle.notify_all();
    }
    cv_idle.notify_all();
  });
}

void Recorder::push(Job&& j) {
  {
    std::lock_guard<std::mutex> lk(mu);
    queue.push_back(std::move(j));
  }
  cv.notify_one();
}

void Recorder::drain() {
  std::unique_lock<std::mutex> lk(mu);
  cv_idle.wait(lk, [this] { return queue.empty() && !busy; });
}

void Recorder::stop() {
  {
    std::lock_guard<std::mutex> lk(mu);
    quit = true;
  }
  cv.notify_one();
  if (writer.joinable()) writer.join();
}

void Recorder::write(const Job& j) {
  char buf[96];
  for (size_t k = 0; k < j.steps.size(); ++k) {
    const std::string step = std::to_string(j.steps[k]);
    const float* f = j.f.data() + k * 5 * (size_t)ntot;
    const int* nb = j.nb.data() + k * (size_t)ntot;
    FILE* a = std::fopen((dir + "/r_data_" + step + ".csv").c_str(), "a");
    FILE* b = std::fopen((dir + "/r_data_3D_" + step + ".csv").c_str(), "a");
    FILE* c = std::fopen((dir + "/particles_color_" + step + ".csv").c_str(), "a");
    bool ok = a && b && c;
    for (int i = 0; ok && i < ntot; ++i) {
      int len = std::snprintf(buf, sizeof buf, "%.15g,%.15g\n", (double)f[i], (double)f[ntot + i]);
      ok = std::fwrite(buf, 1, (size_t)len, a) == (size_t)len;
      len = std::snprintf(buf, sizeof buf, "%.15g,%.15g,%.15g\n", (double)f[2 * (size_t)ntot + i],
                          (double)f[3 * (size_t)ntot + i], (double)f[4 * (size_t)ntot + i]);
      ok = ok && std::fwrite(buf, 1, (size_t)len, b) == (size_t)len;
      len = std::snprintf(buf, sizeof buf, "%d\n", nb[i]);
      ok = ok && std::fwrite(buf, 1, (size_t)len, c) == (size_t)len;
    }
    if (a && std::fclose(a) != 0) ok = false;
    if (b && std::fclose(b) != 0) ok = false;
    if (c && std::fclose(c) != 0) ok = false;
    if (!ok) {
      std::lock_guard<std::mutex> lk(mu);
      if (error.empty()) error = "cannot write the CSV files of step " + step + " in " + dir;
    }
  }
}

void Engine::record_start(const char* dir, int every) {
  if (!dir || !*dir) throw ApiError(TISSUE_E_INVALID_ARG, "null or empty output directory");
  if (every < 1) throw ApiError(TISSUE_E_INVALID_ARG, "every must be >= 1");
  struct stat st {};
  if (stat(dir, &st) != 0 || !S_ISDIR(st.st_mode)) throw ApiError(TISSUE_E_IO, std::string("no such directory: ") + dir);
  if (rec) (void)record_stop();
  HIP_OK(hipStreamSynchronize(stream));
  drop_graphs();  // graphs captured from now on carry the copy into the ring
  auto r = std::make_unique<Recorder>();
  r->dir = dir;
  r->every = every;
  r->ntot = ntot;
  r->origin = steps_done;
  // ring of up to 256 slots within about 64 MB
  const size_t slot_bytes = (size_t)ntot * (5 * sizeof(float) + sizeof(int));
  r->depth = (int)std::clamp<size_t>((64u << 20) / std::max<size_t>(slot_bytes, 1), 1, 256);
  r->d_ring = dalloc<float>((size_t)r->depth * 5 * ntot);
  r->d_ring_nbr = dalloc<int>((size_t)r->depth * ntot);
  r->start();
  rec = std::move(r);
}

int32_t Engine::record_stop() {
  if (!rec) return TISSUE_OK;
  rec->stop();
  const std::string err = rec->error;
  (void)hipStreamSynchronize(stream);
  drop_graphs();
  dfree(rec->d_ring);
  dfree(rec->d_ring_nbr);
  rec.reset();
  if (!err.empty()) throw ApiError(TISSUE_E_IO, err);
  return TISSUE_OK;
}

// after a run of n steps that started with `base` steps done: the recorded steps to the host
void Engine::record_collect(long base, int n_steps) {
  Recorder::Job j;
  const int every = rec->every;
  for (long g = base + 1; g <= base + n_steps; ++g)
    if (g % every == 0) j.steps.push_back(g);
  if (j.steps.empty()) return;
  const size_t nf = 5 * (size_t)ntot;
  j.f.resize(j.steps.size() * nf);
  j.nb.resize(j.steps.size() * (size_t)ntot);
  for (size_t k = 0; k < j.steps.size(); ++k) {
    const size_t slot = (size_t)((j.steps[k] / every) % rec->depth);
    HIP_OK(hipMemcpyAsync(j.f.data() + k * nf, rec->d_ring + slot * nf, nf * sizeof(float), hipMemcpyDeviceToHost,
                          stream));
    HIP_OK(hipMemcpyAsync(j.nb.data() + k * (size_t)ntot, rec->d_ring_nbr + slot * (size_t)ntot, ntot * sizeof(int),
                          hipMemcpyDeviceToHost, stream));
  }
  HIP_OK(hipStreamSynchronize(stream));
  rec->push(std::move(j));
}

}  // namespace tissue
