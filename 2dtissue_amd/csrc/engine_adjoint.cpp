// Adjoint of n_steps of the step (engine side) and its C-ABI entry points.
//
// Forward: the tiled step kernel, with the state copied to a checkpoint every ckpt_every steps.
// Reverse, segment by segment from the last: the segment is recomputed from its checkpoint by
// the recording variant of the same kernel (bitwise the forward's states), which stores every
// pre-step state, the force and alignment sums and the seam-wrap turns, and keeps the vertex
// buckets of every state of the segment (when they fit in about 16 GB; otherwise each reverse
// step rebuilds its state's buckets from the recorded vertex ids); then, step by step backwards,
// the two reverse kernels (per-particle factors, then the transposed neighbour gather over the
// step's buckets) map the adjoint of the new state to the adjoint of the old one. With several
// tapes the recomputes of the next segments run on their own streams during that reverse pass.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdlib>
#include <cmath>

#include "engine.h"

namespace tissue {

namespace {

LossDev make_loss(const tissue_loss& l, int n_steps) {
  if (l.n_modes < 0 || l.n_modes > kMaxModes) throw ApiError(TISSUE_E_INVALID_ARG, "n_modes out of range");
  const double sc[] = {l.target_order, l.w_order, l.w_velocity};
  for (double x : sc)
    if (!std::isfinite(x)) throw ApiError(TISSUE_E_INVALID_ARG, "non-finite loss weight");
  LossDev d{};
  d.n_modes = l.n_modes;
  d.n_steps = n_steps;
  d.target = (float)l.target_order;
  d.w_order = (float)l.w_order;
  d.w_velocity = (float)l.w_velocity;
  for (int m = 0; m < l.n_modes; ++m) {
    if (l.component[m] != 0 && l.component[m] != 1) throw ApiError(TISSUE_E_INVALID_ARG, "mode component must be 0 or 1");
    const double v[] = {l.amp[m], l.kx[m], l.ky[m], l.phase[m]};
    for (double x : v)
      if (!std::isfinite(x)) throw ApiError(TISSUE_E_INVALID_ARG, "non-finite field mode");
    d.comp[m] = l.component[m];
    d.amp[m] = (float)l.amp[m];
    d.kx[m] = (float)l.kx[m];
    d.ky[m] = (float)l.ky[m];
    d.phase[m] = (float)l.phase[m];
  }
  return d;
}

}  // namespace

void Engine::free_tape(Tape& t) {
  dfree(t.xycs); dfree(t.th); dfree(t.vid); dfree(t.fs); dfree(t.q);
  for (BucketSet& b : t.sbs) free_bucket_set(b);
  t.sbs.clear();
  dfree(t.vx); dfree(t.vy); dfree(t.r3x); dfree(t.r3y); dfree(t.r3z); dfree(t.face); dfree(t.nbr);
  dfree(t.rows); dfree(t.first_bad); dfree(t.partial);
  t.partial_n = 0;
}

void Engine::adjoint_release() {
  AdjointBuffers& a = adj;
  for (hipStream_t st : a.rec_stream)
    if (st) (void)hipStreamSynchronize(st);
  dfree(a.ck_xycs); dfree(a.ck_th); dfree(a.ck_vid);
  for (Tape& t : a.tp) free_tape(t);
  for (int k = 0; k < 2; ++k) { dfree(a.lr[k]); dfree(a.lt[k]); }
  dfree(a.gF); dfree(a.gSC); dfree(a.part); dfree(a.acc); dfree(a.out_seed); dfree(a.out_red);
  dfree(a.sexp); dfree(a.mx);
  if (a.ev_fwd) (void)hipEventDestroy(a.ev_fwd);
  for (int k = 0; k < kMaxTapes; ++k) {
    if (a.ev_ck[k]) (void)hipEventDestroy(a.ev_ck[k]);
    if (a.ev_rec[k]) (void)hipEventDestroy(a.ev_rec[k]);
    if (a.ev_rev[k]) (void)hipEventDestroy(a.ev_rev[k]);
  }
  for (hipStream_t st : a.rec_stream)
    if (st) (void)hipStreamDestroy(st);
  a = AdjointBuffers{};
}

void Engine::adjoint_buffers(int K, int C) {
  AdjointBuffers& a = adj;
  const size_t N = ntot;
  if (!a.lr[0]) {
    for (int k = 0; k < 2; ++k) { a.lr[k] = dalloc<float2>(N); a.lt[k] = dalloc<float>(N); }
    a.gF = dalloc<float2>(N); a.gSC = dalloc<float2>(N); a.part = dalloc<float4>(N);
    a.acc = dalloc<double>(kAdjAcc * N);
    a.out_seed = dalloc<double>(2 * (size_t)R);
    a.out_red = dalloc<double>(kAdjAcc * (size_t)R);
    a.sexp = dalloc<int>(2 * (size_t)R);
    a.mx = dalloc<unsigned>(2 * (size_t)R);
    for (hipStream_t& st : a.rec_stream) HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&a.ev_fwd, hipEventDisableTiming));
    for (int k = 0; k < kMaxTapes; ++k) {
      HIP_OK(hipEventCreateWithFlags(&a.ev_ck[k], hipEventDisableTiming));
      HIP_OK(hipEventCreateWithFlags(&a.ev_rec[k], hipEventDisableTiming));
      HIP_OK(hipEventCreateWithFlags(&a.ev_rev[k], hipEventDisableTiming));
    }
  }
  if (C > a.C) {
    dfree(a.ck_xycs); dfree(a.ck_th); dfree(a.ck_vid);
    a.ck_xycs = dalloc<float4>((size_t)C * N); a.ck_th = dalloc<float>((size_t)C * N); a.ck_vid = dalloc<int>((size_t)C * N);
    a.C = C;
  }
  // as many tapes (up to three) as have their bucket sets within about 16 GB together: the
  // recomputes of two segments then run side by side while the reverse pass of a third runs;
  // else one tape, with its states' bucket sets when they fit (otherwise each reverse step
  // re-buckets its state), and everything in order on the engine's stream
  // (the budget is TISSUE_ADJOINT_BUCKET_KB kilobytes when that is set: a smaller one selects the
  // one-tape and re-bucketing forms, which give the same gradients)
  const size_t sets = (size_t)K + 2;
  size_t budget = size_t(16) << 30;
  if (const char* e = std::getenv("TISSUE_ADJOINT_BUCKET_KB")) {
    const long long kb = std::atoll(e);
    if (kb >= 0) budget = (size_t)kb << 10;
  }
  int ntape = 1;
  while (ntape < kMaxTapes && (size_t)(ntape + 1) * sets * bucket_set_bytes() <= budget) ++ntape;
  const bool with_sets = sets * bucket_set_bytes() <= budget;
  if (K > a.K || ntape != a.ntape || with_sets != !a.tp[0].sbs.empty()) {
    for (Tape& t : a.tp) free_tape(t);
    for (int k = 0; k < ntape; ++k) {
      Tape& t = a.tp[k];
      t.xycs = dalloc<float4>((size_t)(K + 1) * N);
      t.th = dalloc<float>((size_t)(K + 1) * N);
      t.vid = dalloc<int>((size_t)(K + 1) * N);
      t.fs = dalloc<float4>((size_t)K * N);
      t.q = dalloc<int>((size_t)K * N);
      if (with_sets) {
        t.sbs.resize(sets);
        for (BucketSet& b : t.sbs) alloc_bucket_set(b);
      }
      t.vx = dalloc<float>(N); t.vy = dalloc<float>(N);
      t.r3x = dalloc<float>(N); t.r3y = dalloc<float>(N); t.r3z = dalloc<float>(N);
      t.face = dalloc<int>(N); t.nbr = dalloc<int>(N);
      t.rows = dalloc<unsigned>(3 * (size_t)R);
      t.first_bad = dalloc<int>(R);
    }
    a.K = K;
    a.ntape = ntape;
  }
  for (int k = 0; k < ntape; ++k) {
    Tape& t = a.tp[k];
    const int np = n_logical * partials_per_tile(kStepMaxBlock);  // written, not read: sized for any shape
    if (t.partial_n < np) {
      dfree(t.partial);
      t.partial = dalloc<OrderPartial>(np);
      t.partial_n = np;
    }
  }
}

int32_t Engine::adjoint(int T, const float* nz, size_t stride, int K, const tissue_loss& loss, tissue_grads* g,
                        double* d_theta0, double* d_x0, double* d_y0) {
  if (!has_state) throw ApiError(TISSUE_E_INVALID_ARG, "no state: call init_random or set_state first");
  if (T < 1 || K < 1) throw ApiError(TISSUE_E_INVALID_ARG, "n_steps and ckpt_every must be positive");
  if (!g) throw ApiError(TISSUE_E_INVALID_ARG, "null gradient output");
  const LossDev L = make_loss(loss, T);
  K = std::min(K, T);
  const int C = (T + K - 1) / K;
  const Shape s = resolve_tiled();
  // the reverse gather walks the same rows as the forward with at most 16 lanes per particle: wider
  // groups shorten the forward's critical path but not the gather's (1,000 particles: reverse
  // 16.0 us per step at 16 lanes, 18.2 us at 32)
  const int gather_lanes = std::min(s.lanes, 16);
  ensure_buckets();
  ensure_tiles(s);
  adjoint_buffers(K, C);
  AdjointBuffers& a = adj;
  const size_t N = ntot;

  // recomputes: with nt > 1 tapes, segment c goes to tape c % nt on recompute stream c % (nt - 1),
  // nt - 1 segments ahead of the reverse passes (tapes are handed over by events); with one tape
  // everything runs in order on the engine's stream
  // (the last nt - 1 segments are recomputed as soon as their checkpoints are written, while the
  // forward still runs)
  const int nt = a.ntape, ns = nt - 1;
  const bool overlap = nt > 1;
  auto rs_of = [&](int c) { return overlap ? a.rec_stream[c % ns] : stream; };
  auto segment_set = [&](Tape& tp, int kc, int q) -> BucketSet& {
    return tp.sbs.size() >= (size_t)kc + 2 ? tp.sbs[q] : bs[q % 3];
  };
  // recompute segment c from its checkpoint into tape c % nt, recording the tape; ready: an
  // event on the engine's stream after which checkpoint c is written
  auto recompute = [&](int c, hipEvent_t ready) {
    const int t0 = c * K, kc = std::min(K, T - t0);
    Tape& tp = a.tp[c % nt];
    hipStream_t rs = rs_of(c);
    if (overlap) {
      HIP_OK(hipStreamWaitEvent(rs, ready, 0));
      if (c + nt <= C - 1) HIP_OK(hipStreamWaitEvent(rs, a.ev_rev[c % nt], 0));  // segment c + nt is done with it
    }
    HIP_OK(hipMemcpyAsync(tp.xycs, a.ck_xycs + c * N, N * sizeof(float4), hipMemcpyDeviceToDevice, rs));
    HIP_OK(hipMemcpyAsync(tp.th, a.ck_th + c * N, N * sizeof(float), hipMemcpyDeviceToDevice, rs));
    HIP_OK(hipMemcpyAsync(tp.vid, a.ck_vid + c * N, N * sizeof(int), hipMemcpyDeviceToDevice, rs));
    clear_bucket_set(segment_set(tp, kc, 1), rs);
    clear_bucket_set(segment_set(tp, kc, 2), rs);
    bucket_state_into(tp.vid, segment_set(tp, kc, 0), rs);
    HIP_OK(launch_flags_begin(R, d_halt, tp.rows, tp.first_bad, rs));
    for (int q = 0; q < kc; ++q) {
      TiledIO io{};
      io.xycs_in = tp.xycs + q * N; io.th_in = tp.th + q * N; io.vid_in = tp.vid + q * N;
      io.xycs_out = tp.xycs + (q + 1) * N; io.th_out = tp.th + (q + 1) * N; io.vid_out = tp.vid + (q + 1) * N;
      io.vx = tp.vx; io.vy = tp.vy; io.r3x = tp.r3x; io.r3y = tp.r3y; io.r3z = tp.r3z; io.face = tp.face;
      io.nbr = tp.nbr;
      io.noise = nz ? nz + (size_t)(t0 + q) * stride : nullptr;
      io.partial = tp.partial;
      io.halt_in = tp.rows + (size_t)(q % 3) * R;
      io.halt_out = tp.rows + (size_t)((q + 1) % 3) * R;
      io.halt_clr = tp.rows + (size_t)((q + 2) % 3) * R;
      io.first_bad = tp.first_bad;
      io.step = q;
      io.tape_fs = tp.fs + q * N;
      io.tape_q = tp.q + q * N;
      HIP_OK(launch_step_tiled(s.lanes, s.block, true, !raw_buckets, md, d_reps, d_tiles, n_phys, segment_set(tp, kc, q),
                               segment_set(tp, kc, q + 1), segment_set(tp, kc, q + 2), io, rs));
      if (!raw_buckets)
        HIP_OK(launch_bucket_finalize(s.block, d_reps, d_tiles, n_phys, md.nv, R == 1 ? ntot : 0,
                                      segment_set(tp, kc, q + 1), rs));
    }
    if (overlap) HIP_OK(hipEventRecord(a.ev_rec[c % nt], rs));
  };
  // lr[cur], lt[cur]: the adjoint of the current state; slot cur of sexp and mx: its exponent and max
  int cur = 0;
  // backwards through segment c
  auto reverse = [&](int c) {
    const int t0 = c * K, kc = std::min(K, T - t0);
    Tape& tp = a.tp[c % nt];
    const bool keep = tp.sbs.size() >= (size_t)kc + 2;
    if (overlap) HIP_OK(hipStreamWaitEvent(stream, a.ev_rec[c % nt], 0));
    for (int q = kc - 1; q >= 0; --q) {
      if (!keep) bucket_state_into(tp.vid + q * N, bs[0], stream);
      HIP_OK(launch_adj_local(d_reps, d_rid, ntot, L, d_halt, a.sexp + cur * R, a.sexp + (1 - cur) * R, a.mx + cur * R,
                              a.mx + (1 - cur) * R, tp.xycs + q * N, tp.th + q * N, tp.fs + q * N, tp.q + q * N,
                              a.lr[cur], a.lt[cur], a.gF, a.gSC, a.part, a.acc, stream));
      HIP_OK(launch_adj_gather(gather_lanes, md, d_reps, d_rid, ntot, keep ? tp.sbs[q] : bs[0], d_halt, tp.xycs + q * N,
                               tp.vid + q * N, a.gF, a.gSC, a.part, a.lr[1 - cur], a.lt[1 - cur],
                               a.mx + (1 - cur) * R, stream));
      cur = 1 - cur;
    }
    if (overlap) HIP_OK(hipEventRecord(a.ev_rev[c % nt], stream));
  };
  // forward, with a checkpoint at the start of every segment
  double fwd_ms = 0.0;
  int32_t status = TISSUE_OK;
  for (int c = 0; c < C; ++c) {
    const int t0 = c * K, kc = std::min(K, T - t0);
    HIP_OK(hipMemcpyAsync(a.ck_xycs + c * N, xycs[sphase], N * sizeof(float4), hipMemcpyDeviceToDevice, stream));
    HIP_OK(hipMemcpyAsync(a.ck_th + c * N, th[sphase], N * sizeof(float), hipMemcpyDeviceToDevice, stream));
    HIP_OK(hipMemcpyAsync(a.ck_vid + c * N, vid[sphase], N * sizeof(int), hipMemcpyDeviceToDevice, stream));
    if (overlap && c >= C - ns) {
      HIP_OK(hipEventRecord(a.ev_ck[c % nt], stream));
      recompute(c, a.ev_ck[c % nt]);
    }
    const int32_t st = run_shape(s, kc, nz, stride, t0);
    fwd_ms += last_ms;
    if (st != TISSUE_OK && status == TISSUE_OK) status = st;
  }


  // reverse
  HIP_OK(hipEventRecord(ev0, stream));
  HIP_OK(launch_adj_seed(R, d_reps, xycs[sphase], L, d_halt, a.lr[0], a.lt[0], a.out_seed, stream));
  HIP_OK(hipMemsetAsync(a.acc, 0, kAdjAcc * N * sizeof(double), stream));
  HIP_OK(hipMemsetAsync(a.sexp, 0, 2 * R * sizeof(int), stream));
  HIP_OK(hipMemsetAsync(a.mx, 0, 2 * R * sizeof(unsigned), stream));
  if (overlap) {
    HIP_OK(hipEventRecord(a.ev_fwd, stream));  // every checkpoint is written
    for (int c = C - 1; c >= 0; --c) {
      reverse(c);
      if (c - ns >= 0) recompute(c - ns, a.ev_fwd);
    }
  } else {
    recompute(C - 1, nullptr);
    for (int c = C - 1; c >= 0; --c) {
      reverse(c);
      if (c >= 1) recompute(c - 1, nullptr);
    }
  }
  HIP_OK(launch_adj_reduce(R, d_reps, ntot, a.acc, a.out_red, stream));
  HIP_OK(hipEventRecord(ev1, stream));
  std::vector<double> seed(2 * (size_t)R), red(kAdjAcc * (size_t)R);
  std::vector<float2> lr(N);
  std::vector<float> lt(N);
  std::vector<int> sexp(R);
  HIP_OK(hipMemcpyAsync(seed.data(), a.out_seed, seed.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  HIP_OK(hipMemcpyAsync(red.data(), a.out_red, red.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  HIP_OK(hipMemcpyAsync(lr.data(), a.lr[cur], N * sizeof(float2), hipMemcpyDeviceToHost, stream));
  HIP_OK(hipMemcpyAsync(lt.data(), a.lt[cur], N * sizeof(float), hipMemcpyDeviceToHost, stream));
  HIP_OK(hipMemcpyAsync(sexp.data(), a.sexp + cur * R, R * sizeof(int), hipMemcpyDeviceToHost, stream));
  HIP_OK(hipStreamSynchronize(stream));
  float rev = 0.f;
  HIP_OK(hipEventElapsedTime(&rev, ev0, ev1));
  buckets_stale = true;  // the reverse pass used the bucket sets
  for (int r = 0; r < R; ++r) {
    tissue_grads& o = g[r];
    o.phi_final = seed[2 * r];
    o.loss_order = seed[2 * r + 1];
    o.loss_velocity = red[kAdjAcc * r + 4];
    o.loss = o.loss_order + o.loss_velocity;
    o.d_v0 = red[kAdjAcc * r + 0];
    o.d_mu = red[kAdjAcc * r + 1];
    o.d_k = red[kAdjAcc * r + 2];
    o.d_tau = red[kAdjAcc * r + 3];
    o.forward_ms = fwd_ms;
    o.reverse_ms = rev;
  }
  for (int r = 0; r < R; ++r) {
    const double sc = std::ldexp(1.0, sexp[r]);
    for (int i = off[r]; i < off[r] + n[r]; ++i) {
      if (d_theta0) d_theta0[i] = sc * (double)lt[i];
      if (d_x0) d_x0[i] = sc * (double)lr[i].x;
      if (d_y0) d_y0[i] = sc * (double)lr[i].y;
    }
  }
  last_ms = fwd_ms + rev;
  return status;
}

}  // namespace tissue

extern "C" {

int32_t tissue_adjoint(tissue_sim* s, int32_t n_steps, int32_t first_row, int32_t ckpt_every, const tissue_loss* loss,
                       tissue_grads* g, double* d_theta0, double* d_x0, double* d_y0) {
  return guarded([&] {
    if (!s || !s->e) throw ApiError(TISSUE_E_INVALID_ARG, "null simulation handle");
    if (!loss) throw ApiError(TISSUE_E_INVALID_ARG, "null loss");
    auto& e = *s->e;
    HIP_OK(hipSetDevice(e.device));
    const float* nz = first_row >= 0 ? e.staged_rows(first_row, n_steps) : nullptr;
    return e.adjoint(n_steps, nz, (size_t)e.ntot, ckpt_every, *loss, g, d_theta0, d_x0, d_y0);
  });
}

int32_t tissue_batch_adjoint(tissue_batch* b, int32_t n_steps, int32_t first_row, int32_t ckpt_every,
                             const tissue_loss* loss, tissue_grads* g, double* d_theta0, double* d_x0, double* d_y0) {
  return guarded([&] {
    if (!b || !b->e) throw ApiError(TISSUE_E_INVALID_ARG, "null batch handle");
    if (!loss) throw ApiError(TISSUE_E_INVALID_ARG, "null loss");
    auto& e = *b->e;
    HIP_OK(hipSetDevice(e.device));
    const float* nz = first_row >= 0 ? e.staged_rows(first_row, n_steps) : nullptr;
    return e.adjoint(n_steps, nz, (size_t)e.ntot, ckpt_every, *loss, g, d_theta0, d_x0, d_y0);
  });
}

}  // extern "C"
