// tissue_step_noalloc: one step of one system for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The caller passes the state and the outputs as device pointers. Everything else the step
// kernel needs (the packed state, three vertex-bucket sets, the tile and replica descriptors, the
// order-parameter partials and the flag rows) lives in `scratch`, a file-scope device array of
// kScratchBytes that each GPU the library runs on holds until the process ends. The entry point
// calls no hipMalloc, hipFree or allocating library routine, and no hipMemsetAsync: buffers that
// must start at zero are cleared by the set-up kernel. It launches the same step kernel as
// tissue_step with the same default launch shape, so its results equal one tissue_step bit for
// bit. Calls that overlap on one GPU share its scratch and race; the array keeps the last call's
// contents.
#include <hip/hip_runtime.h>

#include <array>
#include <climits>
#include <mutex>

#include "capi_common.h"

namespace {

using tissue::BucketSet;
using tissue::OrderPartial;
using tissue::ReplicaDesc;
using tissue::TileDesc;
using tissue::TiledIO;

constexpr int kMaxParticles = TISSUE_NOALLOC_MAX_PARTICLES;
constexpr int kMaxVertices = TISSUE_NOALLOC_MAX_UV_VERTICES;
// tiles of the default launch shape: 2 particles per tile (32 lanes in 64 threads) up to 4,096
// particles, 32 up to 8,192 and 64 beyond, so at most 2,048 tiles
constexpr int kMaxTiles = 2048;

constexpr size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// byte offsets of the scratch regions
struct Layout {
  static constexpr size_t xycs_in = 0;
  static constexpr size_t xycs_out = xycs_in + al(sizeof(float4) * kMaxParticles);
  static constexpr size_t vid = xycs_out + al(sizeof(float4) * kMaxParticles);
  static constexpr size_t rid = vid + al(sizeof(int) * kMaxParticles);
  // one bucket set: occupancy mask, count[V], slot[V * kSlots], ovf[N] (int2), multi[N], big[N],
  // counters[3]
  static constexpr size_t set_occ = 0;
  static constexpr size_t set_count = set_occ + al(sizeof(unsigned) * tdev::occ_words(kMaxVertices));
  static constexpr size_t set_slot = set_count + al(sizeof(int) * kMaxVertices);
  static constexpr size_t set_ovf = set_slot + al(sizeof(int) * kMaxVertices * tdev::kSlots);
  static constexpr size_t set_multi = set_ovf + al(sizeof(int2) * kMaxParticles);
  static constexpr size_t set_big = set_multi + al(sizeof(int) * kMaxParticles);
  static constexpr size_t set_ctr = set_big + al(sizeof(int) * kMaxParticles);
  static constexpr size_t set_bytes = set_ctr + al(sizeof(int) * 3);
  static constexpr size_t sets = rid + al(sizeof(int) * kMaxParticles);
  static constexpr size_t partial = sets + 3 * set_bytes;
  static constexpr size_t tiles = partial + al(sizeof(OrderPartial) * kMaxTiles * (tissue::kStepMaxBlock / 64));
  static constexpr size_t rep = tiles + al(sizeof(TileDesc) * kMaxTiles);
  static constexpr size_t tile_begin = rep + al(sizeof(ReplicaDesc));
  static constexpr size_t rows = tile_begin + al(sizeof(int) * 2);
  static constexpr size_t first_bad = rows + al(sizeof(unsigned) * 3);
  static constexpr size_t total = first_bad + al(sizeof(int));
};
constexpr size_t kScratchBytes = TISSUE_NOALLOC_SCRATCH_BYTES;
static_assert(Layout::total <= kScratchBytes, "scratch regions exceed the declared scratch size");

// The scratch array. In host code its name is a host-side stand-in, not a device address: host
// code uses only the per-device address from hipGetSymbolAddress (scratch_base).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
std::array<unsigned char*, 64> g_base{};  // per device, from hipGetSymbolAddress

unsigned char* scratch_base(int device) {
  if (device < 0 || device >= (int)g_base.size()) throw ApiError(TISSUE_E_INVALID_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_base[device]) {
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw ApiError(TISSUE_E_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_base[device] = (unsigned char*)p;
  }
  return g_base[device];
}

BucketSet set_at(unsigned char* base, int k, int sorted) {
  unsigned char* s = base + Layout::sets + (size_t)k * Layout::set_bytes;
  int* ctr = (int*)(s + Layout::set_ctr);
  BucketSet b;
  b.occ = (unsigned*)(s + Layout::set_occ);
  b.count = (int*)(s + Layout::set_count);
  b.slot = (int*)(s + Layout::set_slot);
  b.ovf = (int2*)(s + Layout::set_ovf);
  b.multi = (int*)(s + Layout::set_multi);
  b.big = (int*)(s + Layout::set_big);
  b.ovf_n = ctr;
  b.multi_n = ctr + 1;
  b.big_n = ctr + 2;
  b.sorted = sorted;
  return b;
}

struct SetupArgs {
  int n, nv, n_tiles, per_tile;
  const float* x;
  const float* y;
  const float* th;
  const int* vid_in;
  float4* xycs;
  int* vid;
  int* rid;
  int* counts[3];
  unsigned* occ[3];
  int* ctrs[3];
  TileDesc* tiles;
  ReplicaDesc* rep;
  int* tile_begin;
  unsigned* rows;
  int* first_bad;
  unsigned* flags_out;
  ReplicaDesc rep_value;
};

// packs the state (x, y, cos theta, sin theta), copies the vertex ids (an id outside the mesh is
// replaced by 0 and reported with TISSUE_FLAG_BAD_INPUT, so no later access leaves its table),
// clears the bucket sets (counts, occupancy masks, list counters) and the flag rows, and writes
// the tile and replica descriptors
__global__ void k_noalloc_setup(SetupArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < a.n) {
    float s, c;
    sincosf(a.th[t], &s, &c);
    a.xycs[t] = make_float4(a.x[t], a.y[t], c, s);
    const int v = a.vid_in[t];
    const bool ok = v >= 0 && v < a.nv;
    a.vid[t] = ok ? v : 0;
    a.rid[t] = 0;
    if (!ok) atomicOr(a.flags_out, TISSUE_FLAG_BAD_INPUT);
  }
  if (t < a.nv) {
    a.counts[0][t] = 0;
    a.counts[1][t] = 0;
    a.counts[2][t] = 0;
  }
  if (t < tdev::occ_words(a.nv)) {
    a.occ[0][t] = 0u;
    a.occ[1][t] = 0u;
    a.occ[2][t] = 0u;
  }
  if (t < a.n_tiles) a.tiles[t] = TileDesc{0, t * a.per_tile, t, a.n_tiles, t, 0};
  if (t < 3) {
    a.ctrs[t][0] = 0;
    a.ctrs[t][1] = 0;
    a.ctrs[t][2] = 0;
    a.rows[t] = 0u;
  }
  if (t == 0) {
    *a.rep = a.rep_value;
    a.tile_begin[0] = 0;
    a.tile_begin[1] = a.n_tiles;
    *a.first_bad = INT_MAX;
  }
}

// flags_out must start at zero before the set-up kernel can report bad vertex ids into it
__global__ void k_noalloc_clear_flags(unsigned* flags_out) { *flags_out = 0u; }

// splits the new packed positions into the caller's x and y and reports the step's flags
__global__ void k_noalloc_finish(int n, const float4* xycs, float* x, float* y, const unsigned* halt_out,
                                 unsigned* flags_out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const float4 q = xycs[t];
    x[t] = q.x;
    y[t] = q.y;
  }
  if (t == 0) atomicOr(flags_out, *halt_out);
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace

extern "C" int32_t tissue_step_noalloc(const tissue_mesh* mc, const tissue_params* p, int32_t n, const float* x_in,
                                       const float* y_in, const float* theta_in, const int32_t* vid_in,
                                       const float* noise, float* x_out, float* y_out, float* theta_out,
                                       int32_t* vid_out, float* vx, float* vy, float* r3x, float* r3y, float* r3z,
                                       int32_t* face, int32_t* nbr, double* phi, uint32_t* flags, void* stream) {
  return guarded([&] {
    if (!mc || !p) throw ApiError(TISSUE_E_INVALID_ARG, "null mesh or params");
    if (n < 1 || n > kMaxParticles)
      throw ApiError(TISSUE_E_INVALID_ARG, "tissue_step_noalloc: particle count must be in 1.." +
                                               std::to_string(kMaxParticles));
    const float* ins[] = {x_in, y_in, theta_in};
    float* outs[] = {x_out, y_out, theta_out, vx, vy, r3x, r3y, r3z};
    for (const float* q : ins)
      if (!q) throw ApiError(TISSUE_E_INVALID_ARG, "null state pointer");
    for (float* q : outs)
      if (!q) throw ApiError(TISSUE_E_INVALID_ARG, "null output pointer");
    if (!vid_in || !vid_out || !face || !nbr || !phi || !flags)
      throw ApiError(TISSUE_E_INVALID_ARG, "null output pointer");
    auto* m = const_cast<tissue_mesh*>(mc);
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    const tdev::MeshDev md = m->resident_view(dev);
    if (md.nv > kMaxVertices)
      throw ApiError(TISSUE_E_INVALID_ARG, "tissue_step_noalloc: the mesh has more than " +
                                               std::to_string(kMaxVertices) + " UV vertices");
    check_params(p, m->sigma_max);
    unsigned char* base = scratch_base(dev);
    auto st = (hipStream_t)stream;

    int lanes = 0, block = 0;
    tissue::tiled_default_shape(1, n, &lanes, &block);
    const int per = block / lanes;
    const int n_tiles = cdiv(n, per);
    if (n_tiles > kMaxTiles) throw ApiError(TISSUE_E_INVALID_ARG, "tissue_step_noalloc: too many tiles");
    const int sorted = tissue::raw_buckets_for(n, md.nv) ? 0 : 1;
    BucketSet b[3] = {set_at(base, 0, sorted), set_at(base, 1, sorted), set_at(base, 2, sorted)};

    SetupArgs a{};
    a.n = n;
    a.nv = md.nv;
    a.n_tiles = n_tiles;
    a.per_tile = per;
    a.x = x_in;
    a.y = y_in;
    a.th = theta_in;
    a.vid_in = vid_in;
    a.xycs = (float4*)(base + Layout::xycs_in);
    a.vid = (int*)(base + Layout::vid);
    a.rid = (int*)(base + Layout::rid);
    for (int k = 0; k < 3; ++k) {
      a.counts[k] = b[k].count;
      a.occ[k] = b[k].occ;
      a.ctrs[k] = b[k].ovf_n;
    }
    a.tiles = (TileDesc*)(base + Layout::tiles);
    a.rep = (ReplicaDesc*)(base + Layout::rep);
    a.tile_begin = (int*)(base + Layout::tile_begin);
    a.rows = (unsigned*)(base + Layout::rows);
    a.first_bad = (int*)(base + Layout::first_bad);
    a.flags_out = flags;
    a.rep_value = ReplicaDesc{0, n, make_step_params(*p)};
    hipLaunchKernelGGL(k_noalloc_clear_flags, dim3(1), dim3(1), 0, st, flags);
    HIP_OK(hipGetLastError());
    const int span = std::max(std::max(n, md.nv), n_tiles);
    hipLaunchKernelGGL(k_noalloc_setup, dim3(cdiv(span, 256)), dim3(256), 0, st, a);
    HIP_OK(hipGetLastError());

    // buckets of the input state (set 0), sorted when the engine's rule sorts them
    HIP_OK(tissue::launch_insert(n, a.vid, a.rid, a.rep, md.nv, b[0], st));
    if (sorted) HIP_OK(tissue::launch_bucket_finalize(block, a.rep, a.tiles, n_tiles, md.nv, n, b[0], st));

    TiledIO io{};
    io.xycs_in = a.xycs;
    io.th_in = theta_in;
    io.vid_in = a.vid;
    io.xycs_out = (float4*)(base + Layout::xycs_out);
    io.th_out = theta_out;
    io.vid_out = vid_out;
    io.vx = vx;
    io.vy = vy;
    io.r3x = r3x;
    io.r3y = r3y;
    io.r3z = r3z;
    io.face = face;
    io.nbr = nbr;
    io.noise = noise;
    io.noise_row0 = nullptr;
    io.noise_stride = 0;
    io.partial = (OrderPartial*)(base + Layout::partial);
    io.halt_in = a.rows;
    io.halt_out = a.rows + 1;
    io.halt_clr = a.rows + 2;
    io.first_bad = a.first_bad;
    io.step = 0;
    HIP_OK(tissue::launch_step_tiled(lanes, block, false, sorted != 0, md, a.rep, a.tiles, n_tiles, b[0], b[1], b[2], io, st));
    HIP_OK(tissue::launch_phi_tiles(1, 1, a.tile_begin, n_tiles, tissue::partials_per_tile(block), io.partial, a.rep,
                                    phi, 1, 0, st));
    hipLaunchKernelGGL(k_noalloc_finish, dim3(cdiv(n, 256)), dim3(256), 0, st, n, io.xycs_out, x_out, y_out,
                       io.halt_out, flags);
    HIP_OK(hipGetLastError());
    return TISSUE_OK;
  });
}
