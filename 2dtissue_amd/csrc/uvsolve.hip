// The UV map's interior solve (cut_and_parameterize, mesh.cpp) on the GPU: the same
// Jacobi-preconditioned CG iteration, operation for operation, so the solution equals the host's
// bit for bit (FP64, no contraction: -ffp-contract=off). Row-wise work (the matrix product, the
// updates) runs one thread per row; each dot product is formed as on the host: per block of
// kUvBlock consecutive rows, the products (formed side by side in LDS) added in row order by one
// thread, then the block sums added in block order by one thread, which also takes the step
// lengths and the convergence decisions.
// Iterations run as a captured graph of a batch of them, replayed until the done flag is set
// (iterations after convergence change nothing: every kernel skips a stopped coordinate).
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "mesh.h"
#include "uvsolve.h"

namespace tissue {

namespace {

struct CgState {
  double rz[2], bn[2], alpha[2], beta[2];
  int on[2];
  int done;
  int pad;
};

constexpr int kRowBlock = 256;

// Ap = A p for the live coordinates (A = diag - off-diagonal weights, rows in CSR order)
__global__ void k_cg_apply(int ni, const int* rp, const int* ci, const double* val, const double* diag,
                           const double* p0, const double* p1, double* ap0, double* ap1, const CgState* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  const double* p[2] = {p0, p1};
  double* ap[2] = {ap0, ap1};
  for (int k = 0; k < 2; ++k) {
    if (!st->on[k]) continue;
    double a = diag[i] * p[k][i];
    for (int e = rp[i]; e < rp[i + 1]; ++e) a += val[e] * p[k][ci[e]];
    ap[k][i] = a;
  }
}

// Per block q and live coordinate k (one workgroup each): the sums over the block's rows, in row
// order, of x[i] * y[i] into part[6 q + slot + k] and, when x2 is given, of x2[i] * y2[i] into
// part[6 q + slot2 + k]. The workgroup forms the products side by side in LDS, then one thread per
// sum adds them in order (the host's sequence of roundings).
constexpr int kDotThreads = 256;
__global__ void __launch_bounds__(kDotThreads)
    k_cg_block_dot(int ni, const double* x0, const double* x1, const double* y0, const double* y1, int slot,
                   const double* x20, const double* x21, const double* y20, const double* y21, int slot2, double* part,
                   const CgState* st) {
  __shared__ double prod[2][kUvBlock];
  const int q = blockIdx.x, k = blockIdx.y;
  if (!st->on[k]) return;
  const int i0 = q * kUvBlock, n = min(ni, i0 + kUvBlock) - i0;
  const double* x = k ? x1 : x0;
  const double* y = k ? y1 : y0;
  const double* x2 = k ? x21 : x20;
  const double* y2 = k ? y21 : y20;
  for (int j = threadIdx.x; j < n; j += kDotThreads) {
    prod[0][j] = x[i0 + j] * y[i0 + j];
    if (x2) prod[1][j] = x2[i0 + j] * y2[i0 + j];
  }
  __syncthreads();
  const int which = threadIdx.x == 0 ? 0 : (threadIdx.x == 64 && x2 ? 1 : -1);  // two waves, two sums
  if (which < 0) return;
  const double* pr = prod[which];
  double acc = 0.0;
  int j = 0;
  for (; j + 8 <= n; j += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = pr[j + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; j < n; ++j) acc += pr[j];
  part[6 * (size_t)q + (which ? slot2 : slot) + k] = acc;
}

__device__ double block_total(const double* part, int nb, int slot) {
  double t = 0.0;
  for (int q = 0; q < nb; ++q) t += part[6 * (size_t)q + slot];
  return t;
}

// alpha = r.z / p.Ap
__global__ void k_cg_alpha(int nb, const double* part, CgState* st) {
  for (int k = 0; k < 2; ++k)
    if (st->on[k]) st->alpha[k] = st->rz[k] / block_total(part, nb, k);
}

// x += alpha p, r -= alpha Ap, z = r / diag
__global__ void k_cg_update(int ni, const double* diag, double* x0, double* x1, double* r0, double* r1, double* z0,
                            double* z1, const double* p0, const double* p1, const double* ap0, const double* ap1,
                            const CgState* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  double* x[2] = {x0, x1};
  double* r[2] = {r0, r1};
  double* z[2] = {z0, z1};
  const double* p[2] = {p0, p1};
  const double* ap[2] = {ap0, ap1};
  for (int k = 0; k < 2; ++k) {
    if (!st->on[k]) continue;
    const double a = st->alpha[k];
    x[k][i] += a * p[k][i];
    r[k][i] -= a * ap[k][i];
    z[k][i] = r[k][i] / diag[i];
  }
}

// convergence (|r| <= tol |b|) and beta = r.z (new) / r.z (old)
__global__ void k_cg_beta(int nb, const double* part, CgState* st) {
  for (int k = 0; k < 2; ++k) {
    if (!st->on[k]) continue;
    if (sqrt(block_total(part, nb, 2 + k)) <= kUvTol * st->bn[k]) {
      st->on[k] = 0;
      continue;
    }
    const double rz2 = block_total(part, nb, 4 + k);
    st->beta[k] = rz2 / st->rz[k];
    st->rz[k] = rz2;
  }
  if (!st->on[0] && !st->on[1]) st->done = 1;
}

// p = z + beta p
__global__ void k_cg_direction(int ni, double* p0, double* p1, const double* z0, const double* z1, const CgState* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  double* p[2] = {p0, p1};
  const double* z[2] = {z0, z1};
  for (int k = 0; k < 2; ++k)
    if (st->on[k]) p[k][i] = z[k][i] + st->beta[k] * p[k][i];
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("UV solve on the GPU: ") + what + ": " + hipGetErrorString(e));
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  explicit DevBuf(size_t n) { check(hipMalloc((void**)&p, n * sizeof(T)), "hipMalloc"); }
  ~DevBuf() { if (p) (void)hipFree(p); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace

bool uv_solve_gpu_available() {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

bool uv_solve_gpu(const UvSystem& s, std::vector<double>* sol) {
  const int ni = s.ni, nb = s.nb;
  if (ni < 1 || nb != (ni + kUvBlock - 1) / kUvBlock) throw std::runtime_error("UV solve on the GPU: bad system");
  const size_t nnz = (size_t)s.rp[ni];
  hipStream_t st = nullptr;
  check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  DevBuf<int> rp(ni + 1), ci(nnz > 0 ? nnz : 1);
  DevBuf<double> val(nnz > 0 ? nnz : 1), diag(ni);
  DevBuf<double> x(2 * (size_t)ni), r(2 * (size_t)ni), z(2 * (size_t)ni), p(2 * (size_t)ni), ap(2 * (size_t)ni);
  DevBuf<double> part(6 * (size_t)nb);
  DevBuf<CgState> state(1);
  auto up = [&](void* d, const void* h, size_t bytes) { check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st), "upload"); };
  up(rp.p, s.rp, (ni + 1) * sizeof(int));
  if (nnz) {
    up(ci.p, s.ci, nnz * sizeof(int));
    up(val.p, s.val, nnz * sizeof(double));
  }
  up(diag.p, s.diag, ni * sizeof(double));
  check(hipMemsetAsync(x.p, 0, 2 * (size_t)ni * sizeof(double), st), "clear");
  for (int k = 0; k < 2; ++k) {
    up(r.p + (size_t)k * ni, s.b[k], ni * sizeof(double));
    up(z.p + (size_t)k * ni, s.z0[k], ni * sizeof(double));
    up(p.p + (size_t)k * ni, s.z0[k], ni * sizeof(double));
  }
  CgState h{};
  for (int k = 0; k < 2; ++k) {
    h.rz[k] = s.rz[k];
    h.bn[k] = s.bn[k];
    h.on[k] = s.live[k] ? 1 : 0;
  }
  h.done = (!h.on[0] && !h.on[1]) ? 1 : 0;
  up(state.p, &h, sizeof h);
  double *x0 = x.p, *x1 = x.p + ni, *r0 = r.p, *r1 = r.p + ni, *z0 = z.p, *z1 = z.p + ni;
  double *p0 = p.p, *p1 = p.p + ni, *a0 = ap.p, *a1 = ap.p + ni;
  const dim3 rows((ni + kRowBlock - 1) / kRowBlock), blk(kRowBlock);
  const dim3 dots(nb, 2), dblk(kDotThreads);
  // a batch of iterations captured once as a graph and replayed until the done flag is set
  // (the graph holds kernels only)
  constexpr int kBatch = 32;
  auto enqueue = [&](int n) {
    for (int q = 0; q < n; ++q) {
      hipLaunchKernelGGL(k_cg_apply, rows, blk, 0, st, ni, rp.p, ci.p, val.p, diag.p, p0, p1, a0, a1, state.p);
      hipLaunchKernelGGL(k_cg_block_dot, dots, dblk, 0, st, ni, p0, p1, a0, a1, 0, nullptr, nullptr, nullptr,
                         nullptr, 0, part.p, state.p);
      hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(1), 0, st, nb, part.p, state.p);
      hipLaunchKernelGGL(k_cg_update, rows, blk, 0, st, ni, diag.p, x0, x1, r0, r1, z0, z1, p0, p1, a0, a1, state.p);
      hipLaunchKernelGGL(k_cg_block_dot, dots, dblk, 0, st, ni, r0, r1, r0, r1, 2, r0, r1, z0, z1, 4, part.p,
                         state.p);
      hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(1), 0, st, nb, part.p, state.p);
      hipLaunchKernelGGL(k_cg_direction, rows, blk, 0, st, ni, p0, p1, z0, z1, state.p);
    }
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  check(hipStreamSynchronize(st), "sync");
  check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture");
  enqueue(kBatch);
  check(hipStreamEndCapture(st, &graph), "capture");
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  check(ie, "graph");
  struct ExecGuard {
    hipGraphExec_t e;
    ~ExecGuard() { (void)hipGraphExecDestroy(e); }
  } exec_guard{exec};
  int done = h.done, it = 0;
  while (!done && it < s.max_it) {
    if (s.max_it - it >= kBatch) {
      check(hipGraphLaunch(exec, st), "graph launch");
      it += kBatch;
    } else {
      enqueue(s.max_it - it);
      check(hipGetLastError(), "launch");
      it = s.max_it;
    }
    check(hipMemcpyAsync(&done, &state.p->done, sizeof(int), hipMemcpyDeviceToHost, st), "flag");
    check(hipStreamSynchronize(st), "sync");
  }
  if (!done) throw std::runtime_error("cut: UV solve did not converge");
  for (int k = 0; k < 2; ++k) {
    sol[k].resize(ni);
    check(hipMemcpyAsync(sol[k].data(), x.p + (size_t)k * ni, ni * sizeof(double), hipMemcpyDeviceToHost, st),
          "download");
  }
  check(hipStreamSynchronize(st), "sync");
  return true;
}

}  // namespace tissue
