// C-ABI of one simulation (include/tissue.h, tissue_create .. tissue_last_step_ms): a single
// replica of the shared engine.
#include <hip/hip_runtime.h>

#include <climits>
#include <memory>

#include "engine.h"

namespace {

tissue::Engine& eng(tissue_sim* s) {
  if (!s || !s->e) throw ApiError(TISSUE_E_INVALID_ARG, "null simulation handle");
  HIP_OK(hipSetDevice(s->e->device));
  return *s->e;
}

void shape_out(const tissue::Shape& s, int32_t* out) {
  if (!out) return;
  out[0] = s.mode; out[1] = s.lanes; out[2] = s.block; out[3] = s.spl; out[4] = s.xcd;
}

}  // namespace

extern "C" {

int32_t tissue_create(const tissue_mesh* mc, const tissue_params* p, int32_t n, tissue_sim** out) {
  return guarded([&] {
    if (!mc || !out) throw ApiError(TISSUE_E_INVALID_ARG, "null argument");
    if (n < 1) throw ApiError(TISSUE_E_INVALID_ARG, "particle count must be positive");
    auto s = std::make_unique<tissue_sim>();
    s->e = std::make_unique<tissue::Engine>(const_cast<tissue_mesh*>(mc), 1, &n, p);
    *out = s.release();
    return TISSUE_OK;
  });
}

void tissue_destroy(tissue_sim* s) { delete s; }

int32_t tissue_set_state(tissue_sim* s, const float* x, const float* y, const float* theta, const int32_t* vid) {
  return guarded([&] {
    eng(s).set_state(x, y, theta, vid);
    return TISSUE_OK;
  });
}

int32_t tissue_init_random(tissue_sim* s, uint64_t seed) {
  return guarded([&] {
    eng(s).init_random(&seed);
    return TISSUE_OK;
  });
}

int32_t tissue_get_state(tissue_sim* s, float* x, float* y, float* theta, int32_t* vid) {
  return guarded([&] {
    eng(s).get_state(x, y, theta, vid);
    return TISSUE_OK;
  });
}

int32_t tissue_get_outputs(tissue_sim* s, float* vx, float* vy, float* r3d, int32_t* face, int32_t* nbr) {
  return guarded([&] {
    eng(s).get_outputs(vx, vy, r3d, face, nbr);
    return TISSUE_OK;
  });
}

int32_t tissue_step(tissue_sim* s, int32_t n_steps, const float* noise) {
  return guarded([&] {
    auto& e = eng(s);
    if (n_steps < 0) throw ApiError(TISSUE_E_INVALID_ARG, "negative step count");
    if (!noise) return e.run(n_steps, nullptr, 0);
    return e.run(n_steps, e.upload_call_noise(noise, n_steps), (size_t)e.ntot);
  });
}

int32_t tissue_step_device(tissue_sim* s, int32_t n_steps, const float* noise_device) {
  return guarded([&] {
    auto& e = eng(s);
    return e.run(n_steps, noise_device, (size_t)e.ntot);
  });
}

int32_t tissue_stage_noise(tissue_sim* s, const float* noise, int32_t n_rows) {
  return guarded([&] {
    eng(s).stage_noise(noise, n_rows);
    return TISSUE_OK;
  });
}

int32_t tissue_record_start(tissue_sim* s, const char* dir, int32_t every) {
  return guarded([&] {
    eng(s).record_start(dir, every);
    return TISSUE_OK;
  });
}

int32_t tissue_record_stop(tissue_sim* s) {
  return guarded([&] { return eng(s).record_stop(); });
}

int32_t tissue_step_staged(tissue_sim* s, int32_t n_steps, int32_t first_row) {
  return guarded([&] {
    auto& e = eng(s);
    return e.run(n_steps, e.staged_rows(0, 0), (size_t)e.ntot, e.checked_first_row(first_row, n_steps));
  });
}

int32_t tissue_get_order(tissue_sim* s, double* phi_ref, double* phi_polar, int32_t cap, int32_t* count) {
  return guarded([&] {
    auto& e = eng(s);
    const auto& a = e.hist_ref[0];
    const auto& c = e.hist_polar[0];
    const int k = (int)std::min<size_t>(a.size(), (size_t)std::max(cap, 0));
    if (phi_ref) std::copy(a.begin(), a.begin() + k, phi_ref);
    if (phi_polar) std::copy(c.begin(), c.begin() + k, phi_polar);
    if (count) *count = (int32_t)a.size();
    return TISSUE_OK;
  });
}

int32_t tissue_reset_order(tissue_sim* s) {
  return guarded([&] {
    auto& e = eng(s);
    e.hist_ref[0].clear();
    e.hist_polar[0].clear();
    return TISSUE_OK;
  });
}

int32_t tissue_get_flags(tissue_sim* s, uint32_t* flags, int32_t* first_step) {
  return guarded([&] {
    auto& e = eng(s);
    if (flags) *flags = e.flags_acc[0];
    if (first_step) *first_step = e.flags_acc[0] ? e.first_bad_acc[0] : -1;
    e.flags_acc[0] = 0;
    e.first_bad_acc[0] = INT_MAX;
    return TISSUE_OK;
  });
}

int32_t tissue_neighbor_lists(tissue_sim* s, int32_t* counts, int32_t* cols, int32_t cap) {
  return guarded([&] {
    eng(s).neighbor_lists(counts, cols, cap);
    return TISSUE_OK;
  });
}

int32_t tissue_set_visit_check(tissue_sim* s, int32_t enable) {
  return guarded([&] {
    eng(s).set_visit_check(enable != 0);
    return TISSUE_OK;
  });
}

int32_t tissue_get_visits(tissue_sim* s, int32_t* na, uint32_t* hash) {
  return guarded([&] {
    eng(s).get_visits(na, hash);
    return TISSUE_OK;
  });
}

int32_t tissue_set_launch(tissue_sim* s, int32_t lanes, int32_t block) {
  return guarded([&] {
    auto& e = eng(s);
    if (lanes != 0 && !valid_lanes(lanes))
      throw ApiError(TISSUE_E_INVALID_ARG, "lanes must be 0, 1, 2, 4, 8, 16, 32 or 64");
    if (block != 0 && (block < 64 || block > tissue::kStepMaxBlock || block % 64))
      throw ApiError(TISSUE_E_INVALID_ARG, "block must be a multiple of 64 in [64, 512]");
    tissue::Shape r = e.req;
    r.lanes = lanes;
    r.block = block;
    e.set_request(r);
    return TISSUE_OK;
  });
}

int32_t tissue_set_mode(tissue_sim* s, int32_t mode) {
  return guarded([&] {
    auto& e = eng(s);
    if (mode != TISSUE_MODE_AUTO && mode != TISSUE_MODE_STEP_KERNEL && mode != TISSUE_MODE_RESIDENT)
      throw ApiError(TISSUE_E_INVALID_ARG, "unknown mode");
    if (mode == TISSUE_MODE_RESIDENT && e.nmax > e.resident_cap)
      throw ApiError(TISSUE_E_INVALID_ARG, "too many particles for the resident kernel on this mesh");
    tissue::Shape r = e.req;
    r.mode = mode;
    e.set_request(r);
    return TISSUE_OK;
  });
}

int32_t tissue_get_launch(tissue_sim* s, int32_t* shape) {
  return guarded([&] {
    if (!shape) throw ApiError(TISSUE_E_INVALID_ARG, "null argument");
    shape_out(eng(s).resolve(), shape);
    return TISSUE_OK;
  });
}

int32_t tissue_autotune(tissue_sim* s, int32_t n_steps, int32_t* shape) {
  return guarded([&] {
    shape_out(eng(s).autotune(n_steps), shape);
    return TISSUE_OK;
  });
}

int32_t tissue_synchronize(tissue_sim* s) {
  return guarded([&] {
    eng(s).synchronize();
    return TISSUE_OK;
  });
}

int32_t tissue_last_step_ms(tissue_sim* s, double* ms) {
  return guarded([&] {
    if (!ms) throw ApiError(TISSUE_E_INVALID_ARG, "null argument");
    *ms = eng(s).last_ms;
    return TISSUE_OK;
  });
}

}  // extern "C"
