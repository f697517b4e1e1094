// Host engine: see engine.h.
#include "engine.h"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <mutex>
#include <numeric>
#include <thread>

#include "record.h"
#include "rng.h"

namespace tissue {

namespace {

// particles per run of the one-kernel-per-step strategy above which AUTO never considers the
// persistent kernel for a single system (each extra pass of its one workgroup pays the slowest
// lane's seam-wrap / point-location chain again)
constexpr int kAutoResidentMax = 256;
constexpr int kDefaultSpl = 100;          // steps per persistent launch / per phi reduction
constexpr size_t kMaxGraphs = 32;         // cached graph executables per engine
constexpr int kXcds = 8;                  // MI355X: 8 XCDs, workgroups dispatched round-robin

// autotuned shapes, keyed by (replicas, largest replica, particles, mesh vertices)
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, Shape> g_tuned;

unsigned host_threads() {
  unsigned t = std::thread::hardware_concurrency();
  return std::max(1u, std::min(t ? t : 1u, 16u));
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

int32_t status_from_flags(unsigned f) {
  if (f & TISSUE_FLAG_NONFINITE) return TISSUE_E_NONFINITE;
  if (f & TISSUE_FLAG_TILING) return TISSUE_E_TILING_NO_CONVERGE;
  if (f & TISSUE_FLAG_LOST_PARTICLE) return TISSUE_E_LOST_PARTICLE;
  return TISSUE_OK;
}

}  // namespace

Engine::Engine(tissue_mesh* m, int R_, const int* n_, const tissue_params* p_) {
  if (!m || !n_ || !p_ || R_ < 1) throw ApiError(TISSUE_E_INVALID_ARG, "bad engine arguments");
  try {
    // every argument is checked before the first device call
    R = R_;
    long tot = 0;
    for (int r = 0; r < R; ++r) {
      if (n_[r] < 1) throw ApiError(TISSUE_E_INVALID_ARG, "particle count must be positive");
      check_params(&p_[r], m->sigma_max);
      n.push_back(n_[r]);
      off.push_back((int)tot);
      p.push_back(p_[r]);
      reps.push_back(ReplicaDesc{(int)tot, n_[r], make_step_params(p_[r])});
      tot += n_[r];
      nmax = std::max(nmax, n_[r]);
    }
    if (tot > INT_MAX / 16) throw ApiError(TISSUE_E_INVALID_ARG, "too many particles");
    ntot = (int)tot;
    HIP_OK(hipGetDevice(&device));
    mesh = m;
    md = m->device_view();
    resident_cap = resident_max_particles(md.nv);
    // raw or sorted vertex buckets (raw_buckets_for); TISSUE_SORTED_BUCKETS = 1 / 0 forces the choice
    raw_buckets = raw_buckets_for(nmax, md.nv);
    if (const char* e = std::getenv("TISSUE_SORTED_BUCKETS")) {
      if (e[0] == '1') raw_buckets = false;
      if (e[0] == '0') raw_buckets = true;
    }
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&ev0));
    HIP_OK(hipEventCreate(&ev1));
    d_row0 = dalloc<int>(2);
    HIP_OK(hipHostMalloc((void**)&h_row0, 2 * sizeof(int), hipHostMallocDefault));
    h_row0[0] = h_row0[1] = 0;
    if (nmax <= resident_cap) HIP_OK(resident_prepare(md.nv, nmax));
    d_reps = dalloc<ReplicaDesc>(R);
    upload(d_reps, reps.data(), reps.size());
    std::vector<int> rid(ntot);
    for (int r = 0; r < R; ++r) std::fill(rid.begin() + off[r], rid.begin() + off[r] + n[r], r);
    d_rid = dalloc<int>(ntot);
    upload(d_rid, rid.data(), rid.size());
    for (int k = 0; k < 2; ++k) {
      xycs[k] = dalloc<float4>(ntot);
      th[k] = dalloc<float>(ntot);
      vid[k] = dalloc<int>(ntot);
    }
    vx = dalloc<float>(ntot); vy = dalloc<float>(ntot);
    r3x = dalloc<float>(ntot); r3y = dalloc<float>(ntot); r3z = dalloc<float>(ntot);
    face = dalloc<int>(ntot); nbr = dalloc<int>(ntot);
    d_halt = dalloc<unsigned>(R);
    d_rows = dalloc<unsigned>(3 * (size_t)R);
    d_first_bad = dalloc<int>(R);
    HIP_OK(hipMemsetAsync(d_halt, 0, R * sizeof(unsigned), stream));
    // the flag rows as a run expects them at its start (each run's end leaves them so again)
    HIP_OK(launch_flags_begin(R, d_halt, d_rows, d_first_bad, stream));
    HIP_OK(hipStreamSynchronize(stream));
    hist_ref.resize(R);
    hist_polar.resize(R);
    flags_acc.assign(R, 0u);
    first_bad_acc.assign(R, INT_MAX);
  } catch (...) {
    release();
    throw;
  }
}

Engine::~Engine() { release(); }

void Engine::drop_graphs() {
  for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
  graphs.clear();
}

void Engine::release() {
  (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  if (rec) {  // finish the files; an I/O error cannot be reported from here
    rec->stop();
    dfree(rec->d_ring);
    dfree(rec->d_ring_nbr);
    rec.reset();
  }
  drop_graphs();
  dfree(d_reps); dfree(d_rid); dfree(d_visit);
  for (int k = 0; k < 2; ++k) { dfree(xycs[k]); dfree(th[k]); dfree(vid[k]); }
  dfree(vx); dfree(vy); dfree(r3x); dfree(r3y); dfree(r3z); dfree(face); dfree(nbr);
  dfree(d_halt); dfree(d_rows); dfree(d_first_bad);
  for (int k = 0; k < 3; ++k) free_bucket_set(bs[k]);
  dfree(d_tiles); dfree(d_tile_begin); dfree(partial); dfree(noise); dfree(call_noise); dfree(d_row0);
  if (phi) (void)hipHostFree(phi);
  phi = phi_dev = nullptr;
  phi_cap = 0;
  if (h_row0) (void)hipHostFree(h_row0);
  h_row0 = nullptr;
  adjoint_release();
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
  stream = nullptr;
  ev0 = ev1 = nullptr;
}

void Engine::synchronize() { HIP_OK(hipStreamSynchronize(stream)); }

// ------------------------------------------------------------------ state

void Engine::set_state(const float* x, const float* y, const float* theta, const int32_t* vid_in) {
  if (!x || !y || !theta) throw ApiError(TISSUE_E_INVALID_ARG, "null state array");
  for (int i = 0; i < ntot; ++i) {
    if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(theta[i]))
      throw ApiError(TISSUE_E_INVALID_ARG, "non-finite state value");
    if (x[i] < 0.f || x[i] > 1.f || y[i] < 0.f || y[i] > 1.f)
      throw ApiError(TISSUE_E_INVALID_ARG, "position outside the UV square");
    if (vid_in && (vid_in[i] < 0 || vid_in[i] >= md.nv)) throw ApiError(TISSUE_E_INVALID_ARG, "vertex id out of range");
  }
  float *dx = dalloc<float>(ntot), *dy = dalloc<float>(ntot), *dt = dalloc<float>(ntot);
  try {
    HIP_OK(hipMemcpyAsync(dx, x, ntot * sizeof(float), hipMemcpyHostToDevice, stream));
    HIP_OK(hipMemcpyAsync(dy, y, ntot * sizeof(float), hipMemcpyHostToDevice, stream));
    HIP_OK(hipMemcpyAsync(dt, theta, ntot * sizeof(float), hipMemcpyHostToDevice, stream));
    adopt_state(dx, dy, dt, vid_in);
  } catch (...) {
    dfree(dx); dfree(dy); dfree(dt);
    throw;
  }
  dfree(dx); dfree(dy); dfree(dt);
}

// the state from device arrays x, y, theta (and host vertex ids, or located when null)
void Engine::adopt_state(const float* dx, const float* dy, const float* dt, const int32_t* vid_in) {
  const int cur = sphase;
  HIP_OK(launch_prepare(ntot, dx, dy, dt, xycs[cur], th[cur], stream));
  HIP_OK(launch_locate(md, ntot, xycs[cur], vid[cur], face, r3x, r3y, r3z, stream));
  if (vid_in) HIP_OK(hipMemcpyAsync(vid[cur], vid_in, ntot * sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_OK(hipMemsetAsync(vx, 0, ntot * sizeof(float), stream));
  HIP_OK(hipMemsetAsync(vy, 0, ntot * sizeof(float), stream));
  HIP_OK(hipMemsetAsync(nbr, 0, ntot * sizeof(int), stream));
  HIP_OK(hipMemsetAsync(d_halt, 0, R * sizeof(unsigned), stream));
  HIP_OK(hipStreamSynchronize(stream));
  buckets_stale = true;
  has_state = true;
}

// area-uniform initial state drawn on the GPU (the same draws as the host's random_init)
void Engine::init_random(const uint64_t* seeds) {
  if (!seeds) throw ApiError(TISSUE_E_INVALID_ARG, "null seeds");
  uint64_t* ds = dalloc<uint64_t>(R);
  float *dx = dalloc<float>(ntot), *dy = dalloc<float>(ntot), *dt = dalloc<float>(ntot);
  try {
    HIP_OK(hipMemcpyAsync(ds, seeds, R * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    HIP_OK(launch_init_random(ntot, d_rid, d_reps, ds, mesh->d_cdf, md, dx, dy, dt, stream));
    adopt_state(dx, dy, dt, nullptr);
  } catch (...) {
    dfree(ds); dfree(dx); dfree(dy); dfree(dt);
    throw;
  }
  dfree(ds); dfree(dx); dfree(dy); dfree(dt);
}

void Engine::get_state(float* x, float* y, float* theta, int32_t* vid_out) {
  const int cur = sphase;
  std::vector<float4> q(ntot);
  HIP_OK(hipMemcpyAsync(q.data(), xycs[cur], ntot * sizeof(float4), hipMemcpyDeviceToHost, stream));
  if (theta) HIP_OK(hipMemcpyAsync(theta, th[cur], ntot * sizeof(float), hipMemcpyDeviceToHost, stream));
  if (vid_out) HIP_OK(hipMemcpyAsync(vid_out, vid[cur], ntot * sizeof(int), hipMemcpyDeviceToHost, stream));
  HIP_OK(hipStreamSynchronize(stream));
  for (int i = 0; i < ntot; ++i) {
    if (x) x[i] = q[i].x;
    if (y) y[i] = q[i].y;
  }
}

void Engine::get_outputs(float* ovx, float* ovy, float* r3d, int32_t* oface, int32_t* onbr) {
  std::vector<float> a(ntot), b(ntot), c(ntot);
  if (ovx) HIP_OK(hipMemcpyAsync(ovx, vx, ntot * sizeof(float), hipMemcpyDeviceToHost, stream));
  if (ovy) HIP_OK(hipMemcpyAsync(ovy, vy, ntot * sizeof(float), hipMemcpyDeviceToHost, stream));
  if (oface) HIP_OK(hipMemcpyAsync(oface, face, ntot * sizeof(int), hipMemcpyDeviceToHost, stream));
  if (onbr) HIP_OK(hipMemcpyAsync(onbr, nbr, ntot * sizeof(int), hipMemcpyDeviceToHost, stream));
  if (r3d) {
    HIP_OK(hipMemcpyAsync(a.data(), r3x, ntot * sizeof(float), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(b.data(), r3y, ntot * sizeof(float), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(c.data(), r3z, ntot * sizeof(float), hipMemcpyDeviceToHost, stream));
  }
  HIP_OK(hipStreamSynchronize(stream));
  if (r3d)
    for (int i = 0; i < ntot; ++i) { r3d[3 * i] = a[i]; r3d[3 * i + 1] = b[i]; r3d[3 * i + 2] = c[i]; }
}

// ------------------------------------------------------------------ buckets and tiles

void Engine::alloc_bucket_set(BucketSet& b) {
  const size_t cells = (size_t)R * md.nv;
  b.sorted = raw_buckets ? 0 : 1;
  b.occ = dalloc<unsigned>((size_t)R * tdev::occ_words(md.nv));
  b.count = dalloc<int>(cells);
  b.slot = dalloc<int>(cells * tdev::kSlots);
  b.ovf_n = dalloc<int>(R);
  b.ovf = dalloc<int2>(ntot);
  b.multi_n = dalloc<int>(R);
  b.multi = dalloc<int>(ntot);
  b.big_n = dalloc<int>(R);
  b.big = dalloc<int>(ntot);
}

void Engine::free_bucket_set(BucketSet& b) {
  dfree(b.occ); dfree(b.count); dfree(b.slot); dfree(b.ovf_n); dfree(b.ovf);
  dfree(b.multi_n); dfree(b.multi); dfree(b.big_n); dfree(b.big);
}

size_t Engine::bucket_set_bytes() const {
  return (size_t)R * md.nv * (1 + tdev::kSlots) * sizeof(int) + (size_t)ntot * (sizeof(int2) + 2 * sizeof(int)) +
         3 * (size_t)R * sizeof(int) + (size_t)R * tdev::occ_words(md.nv) * sizeof(unsigned);
}

// outside any capture: clear set b (count and list lengths), on stream st (null: the engine's)
void Engine::clear_bucket_set(BucketSet& b, hipStream_t st) {
  if (!st) st = stream;
  HIP_OK(hipMemsetAsync(b.count, 0, (size_t)R * md.nv * sizeof(int), st));
  HIP_OK(hipMemsetAsync(b.occ, 0, (size_t)R * tdev::occ_words(md.nv) * sizeof(unsigned), st));
  HIP_OK(hipMemsetAsync(b.ovf_n, 0, R * sizeof(int), st));
  HIP_OK(hipMemsetAsync(b.multi_n, 0, R * sizeof(int), st));
  HIP_OK(hipMemsetAsync(b.big_n, 0, R * sizeof(int), st));
}

// outside any capture: clear set b, bucket the vertex ids v into it and sort its buckets, on
// stream st (null: the engine's)
void Engine::bucket_state_into(const int* v, BucketSet& b, hipStream_t st) {
  if (!st) st = stream;
  if (!d_tiles) ensure_tiles(resolve_tiled());
  clear_bucket_set(b, st);
  HIP_OK(launch_insert(ntot, v, d_rid, d_reps, md.nv, b, st));
  if (!raw_buckets)
    HIP_OK(launch_bucket_finalize(std::get<1>(tiles_key), d_reps, d_tiles, n_phys, md.nv, R == 1 ? ntot : 0, b, st));
}

void Engine::ensure_buckets() {
  if (buckets_ready) return;
  for (int k = 0; k < 3; ++k) alloc_bucket_set(bs[k]);
  buckets_ready = true;
  buckets_stale = true;
}

// outside any capture: clear set k, bucket the vertex ids v into it and sort its buckets
void Engine::bucket_state(const int* v, int k) { bucket_state_into(v, bs[k]); }

// outside any capture: clear every set, bucket the current state into set bphase
void Engine::rebuild_buckets() {
  ensure_buckets();
  const size_t cells = (size_t)R * md.nv;
  for (int k = 0; k < 3; ++k) {
    HIP_OK(hipMemsetAsync(bs[k].count, 0, cells * sizeof(int), stream));
    HIP_OK(hipMemsetAsync(bs[k].occ, 0, (size_t)R * tdev::occ_words(md.nv) * sizeof(unsigned), stream));
    HIP_OK(hipMemsetAsync(bs[k].ovf_n, 0, R * sizeof(int), stream));
    HIP_OK(hipMemsetAsync(bs[k].multi_n, 0, R * sizeof(int), stream));
    HIP_OK(hipMemsetAsync(bs[k].big_n, 0, R * sizeof(int), stream));
  }
  bucket_state(vid[sphase], bphase);
  buckets_stale = false;
}

// Tiles of block / lanes particles. Logical order: replica by replica (a replica's partials are
// consecutive). With the XCD order, replicas are spread over the 8 XCDs by descending size
// (longest-processing-time first) and workgroup b = x + 8 k runs the k-th tile of XCD x, so each
// replica's state and buckets stay in one XCD's L2 (workgroups are dispatched round-robin over
// the XCDs).
void Engine::ensure_tiles(const Shape& s) {
  const auto key = std::make_tuple(s.lanes, s.block, s.xcd);
  if (d_tiles && tiles_key == key) return;
  const int per = s.block / s.lanes;
  std::vector<TileDesc> logical;
  std::vector<int> begin(R + 1, 0);
  for (int r = 0; r < R; ++r) {
    begin[r] = (int)logical.size();
    const int nq = cdiv(n[r], per);
    for (int q = 0; q < nq; ++q) logical.push_back(TileDesc{r, q * per, q, nq, (int)logical.size(), 0});
  }
  begin[R] = (int)logical.size();
  std::vector<TileDesc> phys;
  if (s.xcd == 1 && R >= kXcds) {
    std::vector<int> order(R);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
    std::vector<long> load(kXcds, 0);
    std::vector<std::vector<int>> bins(kXcds);
    for (int r : order) {
      const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[x] += n[r];
      for (int L = begin[r]; L < begin[r + 1]; ++L) bins[x].push_back(L);
    }
    size_t kmax = 0;
    for (auto& b : bins) kmax = std::max(kmax, b.size());
    phys.assign(kmax * kXcds, TileDesc{-1, 0, 0, 0, 0, 0});
    for (int x = 0; x < kXcds; ++x)
      for (size_t k = 0; k < bins[x].size(); ++k) phys[x + kXcds * k] = logical[bins[x][k]];
  } else {
    phys = logical;
  }
  HIP_OK(hipStreamSynchronize(stream));
  drop_graphs();
  dfree(d_tiles);
  dfree(d_tile_begin);
  d_tiles = dalloc<TileDesc>(phys.size());
  upload(d_tiles, phys.data(), phys.size());
  d_tile_begin = dalloc<int>(begin.size());
  upload(d_tile_begin, begin.data(), begin.size());
  n_phys = (int)phys.size();
  n_logical = (int)logical.size();
  tiles_key = key;
}

// ------------------------------------------------------------------ noise

void Engine::stage_noise(const float* h, int rows) {
  if (!h || rows < 1) throw ApiError(TISSUE_E_INVALID_ARG, "bad noise rows");
  if (rows > noise_cap) {
    HIP_OK(hipStreamSynchronize(stream));
    drop_graphs();
    dfree(noise);
    noise_rows = 0;
    noise = dalloc<float>((size_t)rows * ntot);
    noise_cap = rows;
  }
  HIP_OK(hipMemcpyAsync(noise, h, (size_t)rows * ntot * sizeof(float), hipMemcpyHostToDevice, stream));
  HIP_OK(hipStreamSynchronize(stream));
  noise_rows = rows;  // rows past this count hold an earlier staging: out of range from now on
}

void Engine::gen_noise(const uint64_t* seeds, const double* stds, int rows) {
  if (!seeds || !stds || rows < 1) throw ApiError(TISSUE_E_INVALID_ARG, "bad noise arguments");
  std::vector<float> h((size_t)rows * ntot);
  const unsigned nt = host_threads();
  std::vector<std::thread> pool;
  for (unsigned w = 0; w < nt; ++w)
    pool.emplace_back([&, w] {
      for (int t = (int)w; t < rows; t += (int)nt)
        for (int r = 0; r < R; ++r)
          for (int i = 0; i < n[r]; ++i)
            h[(size_t)t * ntot + off[r] + i] = (float)(stds[r] * normal_draw(seeds[r], (uint32_t)i, (uint32_t)t));
    });
  for (auto& x : pool) x.join();
  stage_noise(h.data(), rows);
}

const float* Engine::staged_rows(int first_row, int n_steps) const {
  if (first_row < 0 || n_steps < 0 || (long)first_row + n_steps > noise_rows)
    throw ApiError(TISSUE_E_INVALID_ARG, "staged noise rows out of range");
  return noise + (size_t)first_row * ntot;
}

int Engine::checked_first_row(int first_row, int n_steps) const {
  (void)staged_rows(first_row, n_steps);
  return first_row;
}

const float* Engine::upload_call_noise(const float* h, int n_steps) {
  const size_t need = (size_t)std::max(n_steps, 0) * ntot;
  if (call_noise_cap < need) {
    HIP_OK(hipStreamSynchronize(stream));
    drop_graphs();
    dfree(call_noise);
    call_noise = dalloc<float>(need);
    call_noise_cap = need;
  }
  HIP_OK(hipMemcpyAsync(call_noise, h, need * sizeof(float), hipMemcpyHostToDevice, stream));
  return call_noise;
}

// ------------------------------------------------------------------ runs

void Engine::set_request(const Shape& s) { req = s; }

Shape Engine::resolve() const {
  Shape s = req;
  if (s.mode == TISSUE_MODE_AUTO) {
    if (s.lanes == 0 && s.block == 0) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tuned.find({R, nmax, ntot, md.nv});
      if (it != g_tuned.end()) return it->second;
    }
    s.mode = (R == 1 && ntot <= kAutoResidentMax && nmax <= resident_cap) ? TISSUE_MODE_RESIDENT
                                                                           : TISSUE_MODE_STEP_KERNEL;
  }
  return resolve_for(s);
}

Shape Engine::resolve_tiled() const {
  Shape s = req;
  if (s.mode == TISSUE_MODE_AUTO && s.lanes == 0 && s.block == 0) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find({R, nmax, ntot, md.nv});
    if (it != g_tuned.end() && it->second.mode == TISSUE_MODE_STEP_KERNEL) return it->second;
  }
  s.mode = TISSUE_MODE_STEP_KERNEL;
  return resolve_for(s);
}

Shape Engine::resolve_for(Shape s) const {
  if (s.mode == TISSUE_MODE_RESIDENT) {
    if (nmax > resident_cap) throw ApiError(TISSUE_E_INVALID_ARG, "replica too large for the resident kernel on this mesh");
    s.lanes = s.lanes > 0 ? std::min(s.lanes, 8) : resident_default_lanes(nmax);
    s.block = s.block > 0 ? std::min(s.block, kResidentMaxBlock) : kResidentMaxBlock;
    s.spl = s.spl > 0 ? s.spl : kDefaultSpl;
    s.xcd = 0;
  } else {
    int lanes = 0, block = 0;
    tiled_default_shape(R, ntot, &lanes, &block);
    s.lanes = s.lanes > 0 ? s.lanes : lanes;
    s.block = s.block > 0 ? s.block : block;
    s.spl = s.spl > 0 ? s.spl : kDefaultSpl;
    s.xcd = s.xcd >= 0 ? s.xcd : (R >= 2 * kXcds ? 1 : 0);
  }
  return s;
}

void Engine::enqueue(const Shape& s, int n_steps, const float* nz, size_t stride) {
  // flag rows rotate by step (t % 3); the first step reads the flags raised before the run from
  // d_halt itself, and finds its output row cleared and first_bad reset by the previous run's end
  // (or the construction)
  if (s.mode == TISSUE_MODE_RESIDENT) {
    ResidentIO io;
    const int cur = sphase;
    io.xycs = xycs[cur]; io.th = th[cur]; io.vid = vid[cur];
    io.stage_xycs = xycs[1 - cur]; io.stage_th = th[1 - cur]; io.stage_vid = vid[1 - cur];
    io.vx = vx; io.vy = vy; io.r3x = r3x; io.r3y = r3y; io.r3z = r3z;
    io.face = face; io.nbr = nbr;
    io.noise_stride = stride;
    io.phi = phi_dev;
    io.phi_stride = n_steps;
    io.first_bad = d_first_bad;
    for (int t0 = 0; t0 < n_steps; t0 += s.spl) {
      const int k = std::min(s.spl, n_steps - t0);
      io.noise = nz ? nz + (size_t)t0 * stride : nullptr;
      io.noise_row0 = d_row0;
      io.t_base = t0;
      io.flags_in = t0 == 0 ? d_halt : d_rows + (size_t)(t0 % 3) * R;
      io.flags_out = d_rows + (size_t)((t0 + k) % 3) * R;
      HIP_OK(launch_resident(s.lanes, s.block, md, d_reps, R, nmax, k, io, stream));
    }
  } else {
    int sp = sphase, bp = bphase;
    for (int t0 = 0; t0 < n_steps; t0 += s.spl) {
      const int k = std::min(s.spl, n_steps - t0);
      for (int q = 0; q < k; ++q) {
        const int t = t0 + q;
        TiledIO io{};
        io.xycs_in = xycs[sp]; io.th_in = th[sp]; io.vid_in = vid[sp];
        io.xycs_out = xycs[1 - sp]; io.th_out = th[1 - sp]; io.vid_out = vid[1 - sp];
        io.vx = vx; io.vy = vy; io.r3x = r3x; io.r3y = r3y; io.r3z = r3z;
        io.face = face; io.nbr = nbr;
        io.noise = nz ? nz + (size_t)t * stride : nullptr;
        io.noise_row0 = d_row0;
        io.noise_stride = stride;
        io.partial = partial + (size_t)q * n_logical * partials_per_tile(s.block);
        io.halt_in = t == 0 ? d_halt : d_rows + (size_t)(t % 3) * R;
        io.halt_out = d_rows + (size_t)((t + 1) % 3) * R;
        io.halt_clr = d_rows + (size_t)((t + 2) % 3) * R;
        io.first_bad = d_first_bad;
        io.step = t;
        io.visit = d_visit;
        BucketSet out = bs[(bp + 1) % 3];
        if (raw_buckets) out.multi = nullptr;  // no finalize pass reads the sort list
        HIP_OK(launch_step_tiled(s.lanes, s.block, false, !raw_buckets, md, d_reps, d_tiles, n_phys, bs[bp], out, bs[(bp + 2) % 3],
                                 io, stream));
        if (rec)
          HIP_OK(launch_record(ntot, d_row0, t, rec->every, rec->depth, io.xycs_out, r3x, r3y, r3z, nbr, rec->d_ring,
                               rec->d_ring_nbr, stream));
        if (!raw_buckets)
          HIP_OK(launch_bucket_finalize(s.block, d_reps, d_tiles, n_phys, md.nv, R == 1 ? ntot : 0, bs[(bp + 1) % 3],
                                        stream));
        sp = 1 - sp;
        bp = (bp + 1) % 3;
      }
      HIP_OK(launch_phi_tiles(R, k, d_tile_begin, n_logical, partials_per_tile(s.block), partial, d_reps, phi_dev,
                              n_steps, t0, stream));
    }
  }
  // the flags and first failing steps go right after the run's order history, in the pinned
  // buffer that run_once reads
  unsigned* tail = reinterpret_cast<unsigned*>(phi_dev + 2 * (size_t)R * n_steps);
  HIP_OK(launch_flags_end(R, d_rows + (size_t)(n_steps % 3) * R, d_halt, d_first_bad, tail, d_rows + R, stream));
}

int32_t Engine::run(int n_steps, const float* nz, size_t stride, int row0) {
  Shape s = resolve();
  if (rec && s.mode == TISSUE_MODE_RESIDENT) s = resolve_tiled();  // the stream records every step
  return run_shape(s, n_steps, nz, stride, row0);
}

int32_t Engine::run_shape(const Shape& s, int n_steps, const float* nz, size_t stride, int row0) {
  if (!rec) return run_once(s, n_steps, nz, stride, row0);
  if (s.mode != TISSUE_MODE_STEP_KERNEL)
    throw ApiError(TISSUE_E_INVALID_ARG, "the per-step output stream needs the per-step kernel");
  // runs short enough that no ring slot is overwritten before it is copied out
  const int chunk = rec->depth * rec->every;
  int32_t status = TISSUE_OK;
  for (int done = 0; done < n_steps; done += chunk) {
    const int c = std::min(chunk, n_steps - done);
    const long base = steps_done - rec->origin;
    const int32_t st = run_once(s, c, nz, stride, row0 + done);
    record_collect(base, c);
    if (status == TISSUE_OK) status = st;
  }
  return status;
}

int32_t Engine::run_once(const Shape& s, int n_steps, const float* nz, size_t stride, int row0) {
  if (!has_state) throw ApiError(TISSUE_E_INVALID_ARG, "no state: call init_random or set_state first");
  if (n_steps < 0) throw ApiError(TISSUE_E_INVALID_ARG, "negative step count");
  if (n_steps == 0) return TISSUE_OK;
  const size_t need_phi = 2 * (size_t)R * n_steps;
  const size_t need_back = need_phi + (size_t)R;  // + the flag words (2R of 4 bytes) written by flags_end
  if (phi_cap < need_back) {
    HIP_OK(hipStreamSynchronize(stream));
    drop_graphs();
    if (phi) (void)hipHostFree(phi);
    phi = phi_dev = nullptr;
    phi_cap = 0;
    HIP_OK(hipHostMalloc((void**)&phi, need_back * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    phi_cap = need_back;
    HIP_OK(hipHostGetDevicePointer((void**)&phi_dev, phi, 0));
  }
  if (s.mode == TISSUE_MODE_STEP_KERNEL) {
    ensure_buckets();
    ensure_tiles(s);
    const size_t need = (size_t)s.spl * n_logical * partials_per_tile(s.block);
    if (partial_cap < need) {
      HIP_OK(hipStreamSynchronize(stream));
      drop_graphs();
      dfree(partial);
      partial = dalloc<OrderPartial>(need);
      partial_cap = need;
    }
    if (buckets_stale) rebuild_buckets();
  }
  const auto key = std::make_tuple(n_steps, nz, stride, s.key(), sphase, bphase);
  auto it = graphs.find(key);
  if (it == graphs.end()) {
    if (graphs.size() >= kMaxGraphs) {
      HIP_OK(hipStreamSynchronize(stream));
      drop_graphs();
    }
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue(s, n_steps, nz, stride);
    } catch (...) {
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_OK(hipStreamEndCapture(stream, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_OK(e);
    it = graphs.emplace(key, ge).first;
  }
  // the previous run has completed (synchronised below), so the pinned words are free
  const int w0 = nz ? row0 : 0;
  const int w1 = rec ? (int)std::min<long>(steps_done - rec->origin, INT_MAX) : 0;
  if (!row0_known || h_row0[0] != w0 || h_row0[1] != w1) {
    h_row0[0] = w0;
    h_row0[1] = w1;
    HIP_OK(hipMemcpyAsync(d_row0, h_row0, 2 * sizeof(int), hipMemcpyHostToDevice, stream));
    row0_known = true;
  }
  HIP_OK(hipEventRecord(ev0, stream));
  HIP_OK(hipGraphLaunch(it->second, stream));
  HIP_OK(hipEventRecord(ev1, stream));
  HIP_OK(hipStreamSynchronize(stream));
  // the history, flags and first failing steps as the run's kernels wrote them
  const double* ph = phi;
  const unsigned* fl = reinterpret_cast<const unsigned*>(phi + need_phi);
  const int* fb = reinterpret_cast<const int*>(phi + need_phi) + R;
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
  last_ms = ms;
  if (s.mode == TISSUE_MODE_STEP_KERNEL) {
    sphase = (sphase + n_steps) & 1;
    bphase = (int)((bphase + (long)n_steps) % 3);
  } else {
    buckets_stale = true;  // the persistent kernel leaves the state in place and no global buckets
  }
  unsigned all = 0;
  for (int r = 0; r < R; ++r) {
    for (int t = 0; t < n_steps; ++t) {
      hist_ref[r].push_back(ph[2 * ((size_t)r * n_steps + t)]);
      hist_polar[r].push_back(ph[2 * ((size_t)r * n_steps + t) + 1]);
    }
    if (fb[r] != INT_MAX) {  // raised in this run (a replica stopped earlier raises nothing new)
      flags_acc[r] |= fl[r];
      const int abs_step = (int)std::min<long>(steps_done + fb[r], INT_MAX - 1);
      first_bad_acc[r] = std::min(first_bad_acc[r], abs_step);
      all |= fl[r];
    }
  }
  steps_done += n_steps;
  const int32_t st = status_from_flags(all);
  if (st != TISSUE_OK) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "validation flags 0x%x raised during the run", all);
    g_err = buf;
  }
  return st;
}

// ------------------------------------------------------------------ autotuning

Shape Engine::autotune(int n_steps) {
  if (!has_state) throw ApiError(TISSUE_E_INVALID_ARG, "no state to tune on");
  if (rec) throw ApiError(TISSUE_E_INVALID_ARG, "stop the output stream before autotuning");
  if (n_steps < 1) throw ApiError(TISSUE_E_INVALID_ARG, "autotune needs at least one step");
  std::vector<Shape> cands;
  for (int l : {1, 2, 4, 8, 16, 32, 64})
    for (int b : {64, 128, 256, 512}) {
      if (l >= 16 && ntot > 16384) continue;  // wide groups only pay for small systems
      if (b < l) continue;
      for (int x : {0, 1}) {
        if (x == 1 && R < kXcds) continue;
        Shape s;
        s.mode = TISSUE_MODE_STEP_KERNEL; s.lanes = l; s.block = b; s.spl = kDefaultSpl; s.xcd = x;
        cands.push_back(s);
      }
    }
  if (nmax <= resident_cap)
    for (int l : {1, 2, 4, 8})
      for (int b : {256, 512}) {
        Shape s;
        s.mode = TISSUE_MODE_RESIDENT; s.lanes = l; s.block = b; s.spl = kDefaultSpl; s.xcd = 0;
        cands.push_back(s);
      }
  // save what a trial changes
  const int cur = sphase;
  float4* sx = dalloc<float4>(ntot);
  float* st = dalloc<float>(ntot);
  int* sv = dalloc<int>(ntot);
  unsigned* sh = dalloc<unsigned>(R);
  const Shape keep_req = req;
  const long keep_steps = steps_done;
  std::vector<size_t> keep_hist(R);
  for (int r = 0; r < R; ++r) keep_hist[r] = hist_ref[r].size();
  const std::vector<unsigned> keep_fl = flags_acc;
  const std::vector<int> keep_fb = first_bad_acc;
  auto restore = [&] {
    HIP_OK(hipMemcpyAsync(xycs[sphase], sx, ntot * sizeof(float4), hipMemcpyDeviceToDevice, stream));
    HIP_OK(hipMemcpyAsync(th[sphase], st, ntot * sizeof(float), hipMemcpyDeviceToDevice, stream));
    HIP_OK(hipMemcpyAsync(vid[sphase], sv, ntot * sizeof(int), hipMemcpyDeviceToDevice, stream));
    HIP_OK(hipMemcpyAsync(d_halt, sh, R * sizeof(unsigned), hipMemcpyDeviceToDevice, stream));
    buckets_stale = true;
  };
  Shape best;
  double best_ms = 1e300;
  try {
    HIP_OK(hipMemcpyAsync(sx, xycs[cur], ntot * sizeof(float4), hipMemcpyDeviceToDevice, stream));
    HIP_OK(hipMemcpyAsync(st, th[cur], ntot * sizeof(float), hipMemcpyDeviceToDevice, stream));
    HIP_OK(hipMemcpyAsync(sv, vid[cur], ntot * sizeof(int), hipMemcpyDeviceToDevice, stream));
    HIP_OK(hipMemcpyAsync(sh, d_halt, R * sizeof(unsigned), hipMemcpyDeviceToDevice, stream));
    const float* nz = noise_rows >= n_steps ? noise : nullptr;
    for (const Shape& c : cands) {
      req = c;
      try {
        run(n_steps, nz, (size_t)ntot);
      } catch (const ApiError&) {
        restore();
        continue;  // this shape does not launch here (registers or LDS); skip it
      }
      if (last_ms < best_ms) { best_ms = last_ms; best = c; }
      restore();
    }
    HIP_OK(hipStreamSynchronize(stream));
  } catch (...) {
    req = keep_req;
    dfree(sx); dfree(st); dfree(sv); dfree(sh);
    throw;
  }
  dfree(sx); dfree(st); dfree(sv); dfree(sh);
  req = keep_req;
  steps_done = keep_steps;
  for (int r = 0; r < R; ++r) {
    hist_ref[r].resize(keep_hist[r]);
    hist_polar[r].resize(keep_hist[r]);
  }
  flags_acc = keep_fl;
  first_bad_acc = keep_fb;
  drop_graphs();
  if (best_ms >= 1e300) throw ApiError(TISSUE_E_HIP, "no launch shape ran");
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[{R, nmax, ntot, md.nv}] = best;
  }
  return best;
}

// ------------------------------------------------------------------ neighbour lists

void Engine::set_visit_check(bool on) {
  if (on == (d_visit != nullptr)) return;
  HIP_OK(hipStreamSynchronize(stream));
  drop_graphs();  // the captured steps launch the other kernel variant
  if (on) {
    d_visit = dalloc<int2>(ntot);
    HIP_OK(hipMemsetAsync(d_visit, 0xff, ntot * sizeof(int2), stream));  // (-1, -1): no step recorded yet
  } else {
    dfree(d_visit);
    d_visit = nullptr;
  }
}

void Engine::get_visits(int32_t* na, uint32_t* hash) {
  if (!na || !hash) throw ApiError(TISSUE_E_INVALID_ARG, "bad visit buffers");
  if (!d_visit) throw ApiError(TISSUE_E_INVALID_ARG, "the visit check is off (tissue_set_visit_check)");
  std::vector<int2> h(ntot);
  HIP_OK(hipMemcpyAsync(h.data(), d_visit, ntot * sizeof(int2), hipMemcpyDeviceToHost, stream));
  HIP_OK(hipStreamSynchronize(stream));
  for (int i = 0; i < ntot; ++i) {
    na[i] = h[i].x;
    hash[i] = (uint32_t)h[i].y;
  }
}

void Engine::neighbor_lists(int32_t* counts, int32_t* cols, int32_t cap) {
  if (!counts || !cols || cap < 1) throw ApiError(TISSUE_E_INVALID_ARG, "bad neighbour-list buffers");
  if (buckets_stale) rebuild_buckets();
  int* dc = dalloc<int>(ntot);
  int* dl = dalloc<int>((size_t)ntot * cap);
  try {
    HIP_OK(launch_neighbor_lists(md, d_reps, d_rid, ntot, bs[bphase], vid[sphase], cap, dc, dl, stream));
    HIP_OK(hipMemcpyAsync(counts, dc, ntot * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(cols, dl, (size_t)ntot * cap * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
  } catch (...) {
    dfree(dc); dfree(dl);
    throw;
  }
  dfree(dc); dfree(dl);
  for (int i = 0; i < ntot; ++i) std::sort(cols + (size_t)i * cap, cols + (size_t)i * cap + std::min(counts[i], cap));
}

}  // namespace tissue
