// C-ABI of the batched replicas (include/tissue.h, tissue_batch_*): R independent replicas of
// the shared engine, run together in one hipGraph-captured launch sequence.
#include <hip/hip_runtime.h>

#include <climits>
#include <memory>

#include "engine.h"

namespace {

tissue::Engine& eng(tissue_batch* b) {
  if (!b || !b->e) throw ApiError(TISSUE_E_INVALID_ARG, "null batch handle");
  HIP_OK(hipSetDevice(b->e->device));
  return *b->e;
}

}  // namespace

extern "C" {

int32_t tissue_batch_create(const tissue_mesh* mc, int32_t n_replicas, const int32_t* n, const tissue_params* params,
                            tissue_batch** out) {
  return guarded([&] {
    if (!mc || !n || !params || !out || n_replicas < 1) throw ApiError(TISSUE_E_INVALID_ARG, "bad batch arguments");
    auto b = std::make_unique<tissue_batch>();
    b->e = std::make_unique<tissue::Engine>(const_cast<tissue_mesh*>(mc), n_replicas, n, params);
    *out = b.release();
    return TISSUE_OK;
  });
}

void tissue_batch_destroy(tissue_batch* b) { delete b; }

int32_t tissue_batch_init_random(tissue_batch* b, const uint64_t* seeds) {
  return guarded([&] {
    eng(b).init_random(seeds);
    return TISSUE_OK;
  });
}

int32_t tissue_batch_set_state(tissue_batch* b, const float* x, const float* y, const float* theta, const int32_t* vid) {
  return guarded([&] {
    eng(b).set_state(x, y, theta, vid);
    return TISSUE_OK;
  });
}

int32_t tissue_batch_get_state(tissue_batch* b, float* x, float* y, float* theta, int32_t* vid) {
  return guarded([&] {
    eng(b).get_state(x, y, theta, vid);
    return TISSUE_OK;
  });
}

int32_t tissue_batch_get_outputs(tissue_batch* b, float* vx, float* vy, float* r3d, int32_t* face, int32_t* nbr) {
  return guarded([&] {
    eng(b).get_outputs(vx, vy, r3d, face, nbr);
    return TISSUE_OK;
  });
}

int32_t tissue_batch_stage_noise(tissue_batch* b, const float* noise, int32_t n_rows) {
  return guarded([&] {
    eng(b).stage_noise(noise, n_rows);
    return TISSUE_OK;
  });
}

int32_t tissue_batch_gen_noise(tissue_batch* b, const uint64_t* seeds, const double* stds, int32_t n_rows) {
  return guarded([&] {
    eng(b).gen_noise(seeds, stds, n_rows);
    return TISSUE_OK;
  });
}

int32_t tissue_batch_run(tissue_batch* b, int32_t n_steps, int32_t first_row) {
  return guarded([&] {
    auto& e = eng(b);
    if (first_row < 0) return e.run(n_steps, nullptr, 0);
    return e.run(n_steps, e.staged_rows(0, 0), (size_t)e.ntot, e.checked_first_row(first_row, n_steps));
  });
}

int32_t tissue_batch_get_order(tissue_batch* b, int32_t replica, double* phi_ref, double* phi_polar, int32_t cap,
                               int32_t* count) {
  return guarded([&] {
    auto& e = eng(b);
    if (replica < 0 || replica >= e.R) throw ApiError(TISSUE_E_INVALID_ARG, "bad replica index");
    const auto& a = e.hist_ref[replica];
    const auto& c = e.hist_polar[replica];
    const int k = (int)std::min<size_t>(a.size(), (size_t)std::max(cap, 0));
    if (phi_ref) std::copy(a.begin(), a.begin() + k, phi_ref);
    if (phi_polar) std::copy(c.begin(), c.begin() + k, phi_polar);
    if (count) *count = (int32_t)a.size();
    return TISSUE_OK;
  });
}

int32_t tissue_batch_get_flags(tissue_batch* b, uint32_t* flags, int32_t* first_step) {
  return guarded([&] {
    auto& e = eng(b);
    for (int r = 0; r < e.R; ++r) {
      if (flags) flags[r] = e.flags_acc[r];
      if (first_step) first_step[r] = e.flags_acc[r] ? e.first_bad_acc[r] : -1;
      e.flags_acc[r] = 0;
      e.first_bad_acc[r] = INT_MAX;
    }
    return TISSUE_OK;
  });
}

int32_t tissue_batch_set_launch(tissue_batch* b, int32_t lanes, int32_t block, int32_t steps_per_launch) {
  return guarded([&] {
    auto& e = eng(b);
    if (lanes != 0 && !valid_lanes(lanes))
      throw ApiError(TISSUE_E_INVALID_ARG, "lanes must be 0, 1, 2, 4, 8, 16, 32 or 64");
    if (block != 0 && (block < 64 || block > tissue::kStepMaxBlock || block % 64))
      throw ApiError(TISSUE_E_INVALID_ARG, "block must be a multiple of 64 in [64, 512]");
    if (steps_per_launch < 0) throw ApiError(TISSUE_E_INVALID_ARG, "negative steps per launch");
    tissue::Shape r = e.req;
    r.lanes = lanes;
    r.block = block;
    r.spl = steps_per_launch;
    e.set_request(r);
    return TISSUE_OK;
  });
}

int32_t tissue_batch_set_mode(tissue_batch* b, int32_t mode, int32_t xcd_order) {
  return guarded([&] {
    auto& e = eng(b);
    if (mode != TISSUE_MODE_AUTO && mode != TISSUE_MODE_STEP_KERNEL && mode != TISSUE_MODE_RESIDENT)
      throw ApiError(TISSUE_E_INVALID_ARG, "unknown mode");
    if (mode == TISSUE_MODE_RESIDENT && e.nmax > e.resident_cap)
      throw ApiError(TISSUE_E_INVALID_ARG, "a replica is too large for the resident kernel on this mesh");
    if (xcd_order < -1 || xcd_order > 1) throw ApiError(TISSUE_E_INVALID_ARG, "xcd_order must be -1, 0 or 1");
    tissue::Shape r = e.req;
    r.mode = mode;
    r.xcd = xcd_order;
    e.set_request(r);
    return TISSUE_OK;
  });
}

int32_t tissue_batch_get_launch(tissue_batch* b, int32_t* shape) {
  return guarded([&] {
    if (!shape) throw ApiError(TISSUE_E_INVALID_ARG, "null argument");
    const tissue::Shape s = eng(b).resolve();
    shape[0] = s.mode; shape[1] = s.lanes; shape[2] = s.block; shape[3] = s.spl; shape[4] = s.xcd;
    return TISSUE_OK;
  });
}

int32_t tissue_batch_autotune(tissue_batch* b, int32_t n_steps, int32_t* lanes_out, int32_t* block_out) {
  return guarded([&] {
    const tissue::Shape s = eng(b).autotune(n_steps);
    if (lanes_out) *lanes_out = s.lanes;
    if (block_out) *block_out = s.block;
    return TISSUE_OK;
  });
}

int32_t tissue_batch_set_visit_check(tissue_batch* b, int32_t enable) {
  return guarded([&] {
    eng(b).set_visit_check(enable != 0);
    return TISSUE_OK;
  });
}

int32_t tissue_batch_get_visits(tissue_batch* b, int32_t* na, uint32_t* hash) {
  return guarded([&] {
    eng(b).get_visits(na, hash);
    return TISSUE_OK;
  });
}

int32_t tissue_batch_last_run_ms(tissue_batch* b, double* ms) {
  return guarded([&] {
    if (!ms) throw ApiError(TISSUE_E_INVALID_ARG, "null argument");
    *ms = eng(b).last_ms;
    return TISSUE_OK;
  });
}

}  // extern "C"
