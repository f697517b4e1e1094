// Counter-based host RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC'11), with Box-Muller for normal variates. Every draw is addressed by
// (particle, step, stream), so a stream is reproducible whatever the order of generation.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace tissue {

// (host and device: the particle initialisation runs on the GPU and is checked against the host)
__host__ __device__ inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;   // round multipliers
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;   // Weyl key increments
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 53-bit uniform in [0, 1) from two 32-bit words
__host__ __device__ inline double u53(uint32_t a, uint32_t b) { return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0; }

struct Draw4 { uint32_t w[4]; };

__host__ __device__ inline Draw4 philox_draw(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2) {
  const uint32_t ctr[4] = {c0, c1, c2, 0u};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  Draw4 d;
  philox4x32_10(ctr, key, d.w);
  return d;
}

// standard normal for (particle i, step t) of a noise stream
inline double normal_draw(uint64_t seed, uint32_t i, uint32_t t) {
  const Draw4 d = philox_draw(seed, i, t, 1u);
  const double ua = 1.0 - u53(d.w[0], d.w[1]);  // (0, 1]
  const double ub = u53(d.w[2], d.w[3]);
  return std::sqrt(-2.0 * std::log(ua)) * std::cos(2.0 * M_PI * ub);
}

// Particle i of an area-uniform initialisation (streams 0 and 2 of the seed): a face chosen with
// probability proportional to its 3-D area (the first index whose cumulative area exceeds
// u * total; cdf[nf] holds the cumulative FP64 face areas), folded barycentric weights (r1, r2)
// of a uniform point of the face, and a heading uniform on [-pi, pi). The same code runs on the
// host (the reference for the tests) and in the GPU initialisation kernel.
struct InitDraw {
  int face;
  double r1, r2;
  float th;
};

__host__ __device__ inline InitDraw area_uniform_draw(uint64_t seed, int i, const double* cdf, int nf) {
  const double kPiD = 3.141592653589793;
  const Draw4 d = philox_draw(seed, (uint32_t)i, 0u, 0u);
  const double target = u53(d.w[0], d.w[1]) * cdf[nf - 1];
  int lo = 0, hi = nf;  // first f with cdf[f] > target (upper bound)
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (target < cdf[mid]) hi = mid;
    else lo = mid + 1;
  }
  InitDraw r;
  r.face = lo < nf - 1 ? lo : nf - 1;
  const Draw4 e = philox_draw(seed, (uint32_t)i, 0u, 2u);
  r.r1 = u53(d.w[2], d.w[3]);
  r.r2 = u53(e.w[0], e.w[1]);
  if (r.r1 + r.r2 > 1.0) {
    r.r1 = 1.0 - r.r1;
    r.r2 = 1.0 - r.r2;
  }
  double t = 2.0 * kPiD * u53(e.w[2], e.w[3]);
  t = t - 2.0 * kPiD * floor((t + kPiD) / (2.0 * kPiD));
  r.th = (float)t;
  return r;
}

// the point a + r1 (b - a) + r2 (c - a) of a UV triangle, clamped to the unit square
__host__ __device__ inline void barycentric_uv(double ax, double ay, double bx, double by, double cx, double cy,
                                               double r1, double r2, float* x, float* y) {
  const double px = ax + r1 * (bx - ax) + r2 * (cx - ax);
  const double py = ay + r1 * (by - ay) + r2 * (cy - ay);
  *x = (float)(px < 0.0 ? 0.0 : (1.0 < px ? 1.0 : px));
  *y = (float)(py < 0.0 ? 0.0 : (1.0 < py ? 1.0 : py));
}

}  // namespace tissue
