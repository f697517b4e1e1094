// The local marches of the geodesic distance table on the GPU (gfx950): one wave per source
// vertex, the same acceptance order and the same FP64 front updates (geodesic.h) as the host
// march in mesh.cpp, so the rows come out identical.
//
// A march keeps the vertices it has touched in an LDS hash table (vertex id -> arrival time bits,
// accepted flag). Each round the wave picks the unaccepted vertex of least (time, id) by a
// butterfly reduction, accepts it, and relaxes its neighbours: every (face, target) pair and every
// unfolding across the face's far edge is one lane's item, and the candidate times meet in the
// table through a 64-bit atomic minimum (non-negative doubles order like their bit patterns).
// Within a round the relaxations are independent (they read only accepted vertices, whose times
// are final), so their order does not matter, as on the host where they run one after another.
// A march that outgrows the table, or accepts more than the row capacity, is left to the host.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "geodesic.h"
#include "mesh_gpu.h"

namespace tissue {

namespace {

constexpr int kMarchCap = 512;  // hash slots per march (vertices it may touch)
constexpr int kMarchLog = 9;
constexpr int kMarchWaves = 4;  // marches per workgroup
constexpr int kRowCap = 128;    // accepted vertices kept per source
constexpr unsigned long long kInfBits = 0x7ff0000000000000ull;

struct MarchArgs {
  const double* xyz;
  const int32_t* faces;
  const int32_t* vf_ptr;
  const int32_t* vf_idx;
  const int32_t* across;
  int nv;
  double lim;
  int* out_n;     // [nv] accepted count, -1: left to the host
  int* out_v;     // [kRowCap][nv]: entry k of every source together, so the entries in use are
  double* out_t;  // one leading range
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int home_slot(int v) { return (int)(((unsigned)v * 2654435761u) >> (32 - kMarchLog)); }

// slot of vertex v, or -1 when absent (vertices are never removed)
__device__ int find_slot(const int* keys, int v) {
  const int h = home_slot(v);
  for (int p = 0; p < kMarchCap; ++p) {
    const int slot = (h + p) & (kMarchCap - 1);
    const int k = __atomic_load_n(&keys[slot], __ATOMIC_RELAXED);
    if (k == v) return slot;
    if (k < 0) return -1;
  }
  return -1;
}

// slot of vertex v, inserted when absent; -1 when the table is full
__device__ int insert_slot(int* keys, int v) {
  const int h = home_slot(v);
  for (int p = 0; p < kMarchCap; ++p) {
    const int slot = (h + p) & (kMarchCap - 1);
    const int prev = atomicCAS(&keys[slot], -1, v);
    if (prev == -1 || prev == v) return slot;
  }
  return -1;
}

__global__ void __launch_bounds__(64 * kMarchWaves) k_geodesic_march(MarchArgs a) {
  __shared__ int keys[kMarchWaves][kMarchCap];
  __shared__ unsigned long long tbits[kMarchWaves][kMarchCap];
  __shared__ int accepted[kMarchWaves][kMarchCap];
  __shared__ int overflow[kMarchWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int s = blockIdx.x * kMarchWaves + w;
  if (s >= a.nv) return;  // whole wave
  int* K = keys[w];
  unsigned long long* TB = tbits[w];
  int* AC = accepted[w];
  for (int q = lane; q < kMarchCap; q += 64) {
    K[q] = -1;
    TB[q] = kInfBits;
    AC[q] = 0;
  }
  if (lane == 0) overflow[w] = 0;
  wave_sync();
  if (lane == 0) TB[insert_slot(K, s)] = 0ull;  // the source, at time 0
  wave_sync();
  int count = 0;
  for (int round = 0; round < kMarchCap; ++round) {
    // unaccepted vertex of least (time, id)
    unsigned long long bt = kInfBits;
    int bv = 0x7fffffff, bslot = -1;
    for (int q = lane; q < kMarchCap; q += 64) {
      const int k = K[q];
      const unsigned long long t = TB[q];
      if (k >= 0 && !AC[q] && t != kInfBits && (t < bt || (t == bt && k < bv))) {
        bt = t;
        bv = k;
        bslot = q;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long ot = __shfl_xor(bt, off, 64);
      const int ov = __shfl_xor(bv, off, 64), os = __shfl_xor(bslot, off, 64);
      if (ot < bt || (ot == bt && ov < bv)) {
        bt = ot;
        bv = ov;
        bslot = os;
      }
    }
    if (bt == kInfBits) break;
    const double tv = __longlong_as_double((long long)bt);
    if (tv > a.lim) break;
    const int v = bv;
    if (lane == 0) {
      AC[bslot] = 1;
      if (count < kRowCap) {
        a.out_v[(size_t)count * a.nv + s] = v;
        a.out_t[(size_t)count * a.nv + s] = tv;
      } else {
        overflow[w] = 1;
      }
    }
    ++count;
    wave_sync();
    // relax: item 3 i + c of face i around v: c = 0, 1 the two other corners, c = 2 the unfolding
    // across the edge opposite v
    const double X[3] = {a.xyz[3 * v], a.xyz[3 * v + 1], a.xyz[3 * v + 2]};
    const int fp = a.vf_ptr[v], nfv = a.vf_ptr[v + 1] - fp;
    for (int it = lane; it < 3 * nfv; it += 64) {
      const int f = a.vf_idx[fp + it / 3], c = it % 3;
      const int f0 = a.faces[3 * f], f1 = a.faces[3 * f + 1], f2 = a.faces[3 * f + 2];
      const int qv = f0 == v ? 0 : (f1 == v ? 1 : 2);
      int tgt;
      double cand;
      if (c < 2) {
        const int fc[3] = {f0, f1, f2};
        tgt = fc[(qv + 1 + c) % 3];
        const int oth = f0 + f1 + f2 - v - tgt;
        const int ts = find_slot(K, tgt);
        if (ts >= 0 && AC[ts]) continue;
        double d[3];
        geo::sub3(&a.xyz[3 * tgt], X, d);
        cand = tv + geo::norm3(d);
        const int os = find_slot(K, oth);
        if (os >= 0 && AC[os]) {
          const double to = __longlong_as_double((long long)TB[os]);
          const double tu = geo::tri_update(X, tv, &a.xyz[3 * oth], to, &a.xyz[3 * tgt]);
          cand = tu < cand ? tu : cand;
        }
      } else {
        const int g2 = a.across[3 * f + qv];
        if (g2 < 0) continue;
        const int ea = f0 == v ? f1 : f0;
        const int eb = f0 + f1 + f2 - v - ea;
        tgt = a.faces[3 * g2] + a.faces[3 * g2 + 1] + a.faces[3 * g2 + 2] - ea - eb;
        const int ts = find_slot(K, tgt);
        if (ts >= 0 && AC[ts]) continue;
        cand = tv + geo::unfold_across(X, &a.xyz[3 * ea], &a.xyz[3 * eb], &a.xyz[3 * tgt]);
      }
      if (!(cand < geo::inf())) continue;  // an infinite candidate touches nothing (as on the host)
      const int slot = insert_slot(K, tgt);
      if (slot < 0) {
        overflow[w] = 1;
        continue;
      }
      atomicMin(&TB[slot], (unsigned long long)__double_as_longlong(cand));
    }
    wave_sync();
    if (overflow[w]) break;
  }
  if (lane == 0) a.out_n[s] = overflow[w] ? -1 : count;
}

#define MG_OK(expr)                                                                              \
  do {                                                                                           \
    const hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
T* to_device(const T* h, size_t n, hipStream_t st) {
  void* p = nullptr;
  MG_OK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
  if (n) MG_OK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
  return (T*)p;
}

}  // namespace

void march_on_gpu(const UVMesh& m, const MarchMesh& mm, double lim, MarchRows& rows, std::vector<char>& done) {
  const int nv = m.nv;
  if (nv == 0) return;
  hipStream_t st = nullptr;
  MG_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  void* bufs[8] = {};
  try {
    MarchArgs a{};
    a.xyz = to_device(m.xyz.data(), m.xyz.size(), st);
    bufs[0] = (void*)a.xyz;
    a.faces = to_device(m.faces.data(), m.faces.size(), st);
    bufs[1] = (void*)a.faces;
    a.vf_ptr = to_device(mm.vf_ptr.data(), mm.vf_ptr.size(), st);
    bufs[2] = (void*)a.vf_ptr;
    a.vf_idx = to_device(mm.vf_idx.data(), mm.vf_idx.size(), st);
    bufs[3] = (void*)a.vf_idx;
    a.across = to_device(mm.across.data(), mm.across.size(), st);
    bufs[4] = (void*)a.across;
    MG_OK(hipMalloc((void**)&a.out_n, (size_t)nv * sizeof(int)));
    bufs[5] = a.out_n;
    MG_OK(hipMalloc((void**)&a.out_v, (size_t)nv * kRowCap * sizeof(int)));
    bufs[6] = a.out_v;
    MG_OK(hipMalloc((void**)&a.out_t, (size_t)nv * kRowCap * sizeof(double)));
    bufs[7] = a.out_t;
    a.nv = nv;
    a.lim = lim;
    hipLaunchKernelGGL(k_geodesic_march, dim3((nv + kMarchWaves - 1) / kMarchWaves), dim3(64 * kMarchWaves), 0, st, a);
    MG_OK(hipGetLastError());
    std::vector<int> n(nv);
    MG_OK(hipMemcpyAsync(n.data(), a.out_n, n.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    MG_OK(hipStreamSynchronize(st));
    int used = 0;
    for (int s = 0; s < nv; ++s) used = n[s] > used ? n[s] : used;
    std::vector<int> vv((size_t)used * nv);
    std::vector<double> tt((size_t)used * nv);
    MG_OK(hipMemcpyAsync(vv.data(), a.out_v, vv.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    MG_OK(hipMemcpyAsync(tt.data(), a.out_t, tt.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    MG_OK(hipStreamSynchronize(st));
    for (int s = 0; s < nv; ++s) {
      if (n[s] < 0) continue;
      auto& r = rows[s];
      r.resize(n[s]);
      for (int k = 0; k < n[s]; ++k) r[k] = {vv[(size_t)k * nv + s], tt[(size_t)k * nv + s]};
      done[s] = 1;
    }
  } catch (...) {
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    (void)hipStreamDestroy(st);
    throw;
  }
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  MG_OK(hipStreamDestroy(st));
}

}  // namespace tissue
