// gfx950 kernels of the tiled per-step update (one launch per step, all replicas at once).
//
// k_step_tiled<LANES>: each workgroup is a tile of one replica (TileDesc); LANES lanes of a wave64
// cooperate on one particle. They stride over the particle's row of the truncated geodesic table
// (vertices within 2.4 sigma of its vertex, ascending in distance), visit the particles bucketed on
// those vertices, and reduce force, alignment sum and neighbour count with butterfly shuffles.
// Every lane then runs the per-particle epilogue (speed, Euler, heading, seam wrap) and the lanes
// split the point-location candidates of the face bucket. Lane 0 writes the SoA outputs, inserts
// the particle into next step's vertex buckets and contributes to the tile's order-parameter
// partials. A single system is the case of one replica.
#include <hip/hip_runtime.h>

#include <climits>

#include "launch.h"
#include "rng.h"

using namespace tdev;

namespace tissue {

namespace {

__device__ __forceinline__ void bucket_insert(const BucketsOut& b, int v, int i) {
  const int old = atomicAdd(&b.count[v], 1);
  if (old == 0) atomicOr(&b.occ[v >> 5], 1u << (v & 31));
  if (old < kSlots) {
    b.slot[(size_t)v * kSlots + old] = i;
  } else {
    const int o = atomicAdd(b.ovf_n, 1);
    b.ovf[o] = make_int2(v, i);
  }
  if (old == 1 && b.multi) b.multi[atomicAdd(b.multi_n, 1)] = v;  // the bucket needs sorting
}

__device__ __forceinline__ BucketsOut replica_out(const BucketSet& s, int r, int nv, int off) {
  return {s.count + (size_t)r * nv, s.slot + (size_t)r * nv * kSlots, s.ovf_n + r, s.ovf + off, s.multi_n + r,
          s.multi ? s.multi + off : nullptr, s.occ + (size_t)r * occ_words(nv)};
}

// orders the LDS / global accesses of one wave (its lanes exchange data through memory)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// the wave's sums of (sgn, cs, sn) in fixed point, written by its lane 0 to out: each wave keeps
// its own partial (no workgroup barrier, so a finished wave leaves at once); integer sums make the
// result independent of their order
__device__ __forceinline__ void wave_order_partial(int sgn, long long cs, long long sn, OrderPartial* out) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sgn += __shfl_xor(sgn, off, 64);
    cs += __shfl_xor(cs, off, 64);
    sn += __shfl_xor(sn, off, 64);
  }
  if ((threadIdx.x & 63) == 0) *out = OrderPartial{cs, sn, sgn, 0};
}

template <int LANES, bool TAPE, bool OCC, bool VISITS, bool WAVE>
__global__ void __launch_bounds__(kStepMaxBlock) k_step_tiled(MeshDev mesh, const ReplicaDesc* reps,
                                                              const TileDesc* tiles, BucketSet bin, BucketSet bout,
                                                              BucketSet bclr, TiledIO io) {
  extern __shared__ unsigned s_occ[];  // OCC: occ_words(nv), the input buckets' occupancy mask
  const TileDesc tile = tiles[blockIdx.x];
  if (tile.r < 0) return;  // idle padding workgroup (whole block)
  const int r = tile.r;
  const ReplicaDesc& rep = reps[r];  // uniform: read through the scalar cache where used
  const int nv = mesh.nv, tid = threadIdx.x, nw = occ_words(nv);
  const int lane = tid & (LANES - 1);
  const int i = tile.first + tid / LANES;
  const bool active = i < rep.n;
  const int g = rep.off + i;
  // the particle's inputs first: their loads are in flight during the set-up below
  float4 me = make_float4(0.f, 0.f, 0.f, 0.f);
  float th = 0.f, eta = 0.f;
  int v = 0;
  if (active) {
    me = io.xycs_in[g];
    th = io.th_in[g];
    v = io.vid_in[g];
    const size_t shift = io.noise_row0 ? (size_t)io.noise_row0[0] * io.noise_stride : 0;
    eta = io.noise ? io.noise[shift + g] : 0.0f;
  }
  if constexpr (OCC) {  // stage this replica's occupancy mask of the input buckets in LDS
    const unsigned* occ_in = bin.occ + (size_t)r * nw;
    for (int k = tid; k < nw; k += (int)blockDim.x) s_occ[k] = occ_in[k];
  }
  // clear this replica's bucket set consumed two steps ago (it is next step's output)
  int* clr = bclr.count + (size_t)r * nv;
  for (int w = tile.q * (int)blockDim.x + tid; w < nv; w += tile.nq * (int)blockDim.x) clr[w] = 0;
  unsigned* clr_occ = bclr.occ + (size_t)r * nw;
  for (int k = tile.q * (int)blockDim.x + tid; k < nw; k += tile.nq * (int)blockDim.x) clr_occ[k] = 0u;
  __syncthreads();
  // a replica flagged before this step passes its state through unchanged (the reference stops
  // at the first failure); the flag rows rotate like the bucket sets
  const unsigned halt = io.halt_in[r];
  if (tile.q == 0 && tid == 0) {
    bclr.ovf_n[r] = 0;
    bclr.multi_n[r] = 0;
    bclr.big_n[r] = 0;
    io.halt_clr[r] = 0u;
    if (halt) atomicOr(&io.halt_out[r], halt);
  }
  const BucketsOut bo = replica_out(bout, r, nv, rep.off);

  int sgn = 0;
  long long pcs = 0, psn = 0;
  if (active) {
    if (halt) {
      if (lane == 0) {
        io.xycs_out[g] = me;
        io.th_out[g] = th;
        io.vid_out[g] = v;
        bucket_insert(bo, v, g);
        if constexpr (VISITS) io.visit[g] = make_int2(-1, 0);
      }
    } else {
      const GlobalSource<OCC> src{replica_in(bin, r, nv, rep.off), io.xycs_in, io.th_in, s_occ};
      const StepOut o = particle_step<LANES, WAVE>(mesh, rep.prm, src, g, me, th, v, eta, lane);
      if (lane == 0) {
        io.xycs_out[g] = make_float4(o.xo, o.yo, o.co, o.so);
        io.th_out[g] = o.to;
        io.vid_out[g] = o.L.vid;
        io.vx[g] = o.vx;
        io.vy[g] = o.vy;
        io.r3x[g] = (float)o.L.X;
        io.r3y[g] = (float)o.L.Y;
        io.r3z[g] = (float)o.L.Z;
        io.face[g] = o.L.face;
        io.nbr[g] = o.nb;  // count_particle_neighbors (2DTissue.cpp:254-268)
        bucket_insert(bo, o.L.vid, g);
        if (o.fl) {
          atomicOr(&io.halt_out[r], o.fl);
          atomicMin(&io.first_bad[r], io.step);
        }
        if constexpr (TAPE) {
          io.tape_fs[g] = make_float4(o.Fx, o.Fy, o.Sx, o.Sy);
          io.tape_q[g] = o.q;
        }
        if constexpr (VISITS) io.visit[g] = make_int2(o.na, (int)o.vh);
        sgn = o.sgn;
        pcs = __double2ll_rn((double)o.co * kPhiScale);
        psn = __double2ll_rn((double)o.so * kPhiScale);
      }
    }
  }
  const int wpb = (int)blockDim.x >> 6;
  wave_order_partial(sgn, pcs, psn, io.partial + (size_t)tile.logical * wpb + (tid >> 6));
}

// crowded buckets up to this many particles are sorted in LDS (larger ones in global memory)
constexpr int kSortCap = 512;

// Sort every bucket of set bs that holds two or more particles (listed in multi by the insertion).
// One thread per listed bucket: up to kSlots ids are sorted in place with the register network.
// A crowded bucket (more than kSlots) is then taken by its whole wave: its slots and its entries
// of the overflow list are gathered (ballot compaction), ranked, and written in ascending order to
// a fresh range of the replica's big area, whose offset replaces the bucket's first slot.
// The kernel is a chain of dependent loads, so independent ones are issued together: the list
// length, the overflow count and the thread's list entry (read speculatively within the
// replica's range), then a bucket's count and all of its slots, and for a crowded bucket the
// big-area reservation with the overflow scan. single_n > 0: one replica of single_n particles,
// whose tiles need not be read (block b covers list entries b * blockDim.x ...).
__global__ void __launch_bounds__(kStepMaxBlock) k_bucket_finalize(const ReplicaDesc* reps, const TileDesc* tiles,
                                                                   int nv, int single_n, BucketSet bs) {
  __shared__ __align__(16) int lds[kStepMaxBlock / 64][kSortCap];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int r = 0, off = 0, k = blockIdx.x * (int)blockDim.x + tid, cap = single_n;
  if (single_n <= 0) {
    const TileDesc tile = tiles[blockIdx.x];
    if (tile.r < 0) return;
    r = tile.r;
    off = reps[r].off;
    cap = reps[r].n;
    k = tile.q * (int)blockDim.x + tid;
  }
  const int n_multi = bs.multi_n[r];
  const int novf = bs.ovf_n[r];
  const int v_listed = k < cap ? bs.multi[off + k] : -1;
  int* slot = bs.slot + (size_t)r * nv * kSlots;
  int v = -1, c = 0;
  int4 s0 = make_int4(0, 0, 0, 0), s1 = s0;
  if (k < n_multi) {
    v = v_listed;
    int* sl = slot + (size_t)v * kSlots;
    s0 = reinterpret_cast<const int4*>(sl)[0];
    s1 = reinterpret_cast<const int4*>(sl)[1];
    c = bs.count[(size_t)r * nv + v];
    if (c <= kSlots) {
      int a[kSlots] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
      for (int q = 0; q < kSlots; ++q) a[q] = q < c ? a[q] : 0x7fffffff;
      sort8(a);
#pragma unroll
      for (int q = 0; q < kSlots; ++q)
        if (q < c) sl[q] = a[q];
    }
  }
  unsigned long long crowded = __ballot(c > kSlots);
  if (crowded == 0ull) return;  // uniform over the wave
  const int2* ovf = bs.ovf + off;
  int* big = bs.big + off;
  // the ranges of all the wave's crowded buckets reserved with one atomic, in lane order (a
  // returning device-scope atomic is a memory-side round trip; one per bucket serialised them)
  int pre = c > kSlots ? c : 0;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(pre, d, 64);
    if (lane >= d) pre += o;
  }
  const int total = __shfl(pre, 63, 64);
  int base0 = 0;
  if (lane == 0) base0 = atomicAdd(&bs.big_n[r], total);
  base0 = __shfl(base0, 0, 64);
  const int my_start = base0 + pre - (c > kSlots ? c : 0);
  // the overflow list's first 256 entries, loaded once for all the wave's crowded buckets
  int2 e0[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int q = 64 * u + lane;
    e0[u] = q < novf ? ovf[q] : make_int2(-1, 0);
  }
  while (crowded) {
    const int src = __ffsll((long long)crowded) - 1;
    crowded &= crowded - 1ull;
    const int w = __shfl(v, src, 64);
    const int cw = __shfl(c, src, 64);
    const int reserved = __shfl(my_start, src, 64);
    int* sl = slot + (size_t)w * kSlots;
    const bool in_lds = cw <= kSortCap - 4;  // room for the padding below
    if (in_lds) {
      int* buf = lds[wave];
      // the bucket's slots as its lane loaded them above
      const int f0 = __builtin_amdgcn_readlane(s0.x, src), f1 = __builtin_amdgcn_readlane(s0.y, src);
      const int f2 = __builtin_amdgcn_readlane(s0.z, src), f3 = __builtin_amdgcn_readlane(s0.w, src);
      const int f4 = __builtin_amdgcn_readlane(s1.x, src), f5 = __builtin_amdgcn_readlane(s1.y, src);
      const int f6 = __builtin_amdgcn_readlane(s1.z, src), f7 = __builtin_amdgcn_readlane(s1.w, src);
      if (lane < kSlots)
        buf[lane] = lane < 4 ? (lane < 2 ? (lane == 0 ? f0 : f1) : (lane == 2 ? f2 : f3))
                             : (lane < 6 ? (lane == 4 ? f4 : f5) : (lane == 6 ? f6 : f7));
      int fill = kSlots;
      // the overflow entries in chunks of 256 (four loads per lane in flight together; the first
      // chunk loaded before the loop), compacted in order
      for (int q0 = 0; q0 < novf; q0 += 256) {
        int2 e[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int q = q0 + 64 * u + lane;
          e[u] = q0 == 0 ? e0[u] : (q < novf ? ovf[q] : make_int2(-1, 0));
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool m = e[u].x == w;
          const unsigned long long bm = __ballot(m);
          if (m) buf[fill + __popcll(bm & ((1ull << lane) - 1ull))] = e[u].y;
          fill += __popcll(bm);
        }
      }
      if (lane < 4) buf[cw + lane] = 0x7fffffff;  // pad to a multiple of four (never below an id)
      wave_sync();
      int* dst = big + reserved;
      // rank of each id among the bucket's (distinct) ids = its position in ascending order; the
      // ids are read four at a time, several reads in flight
      const int cw4 = (cw + 3) & ~3;
      for (int q = lane; q < cw; q += 64) {
        const int val = buf[q];
        int rank = 0;
#pragma unroll 4
        for (int j = 0; j < cw4; j += 4) {
          const int4 b4 = *reinterpret_cast<const int4*>(buf + j);
          rank += (int)(b4.x < val) + (int)(b4.y < val) + (int)(b4.z < val) + (int)(b4.w < val);
        }
        dst[rank] = val;
      }
    } else {
      // rare: gathered and sorted in place in global memory
      int* dst = big + reserved;
      if (lane < kSlots) dst[lane] = sl[lane];
      int fill = kSlots;
      for (int q0 = 0; q0 < novf; q0 += 64) {
        const int q = q0 + lane;
        const int2 e = q < novf ? ovf[q] : make_int2(-1, 0);
        const bool m = e.x == w;
        const unsigned long long bm = __ballot(m);
        if (m) dst[fill + __popcll(bm & ((1ull << lane) - 1ull))] = e.y;
        fill += __popcll(bm);
      }
      wave_sync();
      if (lane == 0) {
        for (int a = 1; a < cw; ++a) {
          const int val = dst[a];
          int b = a - 1;
          while (b >= 0 && dst[b] > val) { dst[b + 1] = dst[b]; --b; }
          dst[b + 1] = val;
        }
      }
    }
    wave_sync();
    if (lane == 0) sl[0] = reserved;
  }
}

// order parameters of replica r at step s, one workgroup per (r, s): the threads add the
// replica's wave partials of the step in a strided split, then a wave and a workgroup sum (integer,
// so the split does not matter)
constexpr int kPhiBlock = 256;
__global__ void __launch_bounds__(kPhiBlock) k_phi_tiles(int steps, const int* tile_begin, int n_logical, int per_tile,
                                                         const OrderPartial* partial, const ReplicaDesc* reps,
                                                         double* phi, int phi_stride, int t_base) {
  __shared__ long long s_cs[kPhiBlock / 64], s_sn[kPhiBlock / 64];
  __shared__ long s_sg[kPhiBlock / 64];
  const int r = blockIdx.x / steps, s = blockIdx.x % steps;
  long long cs = 0, sn = 0;
  long sg = 0;
  const OrderPartial* row = partial + (size_t)s * n_logical * per_tile;
  for (int b = tile_begin[r] * per_tile + (int)threadIdx.x; b < tile_begin[r + 1] * per_tile; b += kPhiBlock) {
    const OrderPartial p = row[b];
    cs += p.cs;
    sn += p.sn;
    sg += p.sgn;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    cs += __shfl_xor(cs, off, 64);
    sn += __shfl_xor(sn, off, 64);
    sg += __shfl_xor(sg, off, 64);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { s_cs[wave] = cs; s_sn[wave] = sn; s_sg[wave] = sg; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kPhiBlock / 64; ++w) { cs += s_cs[w]; sn += s_sn[w]; sg += s_sg[w]; }
    const int n = reps[r].n;
    const double c = (double)cs / kPhiScale, q = (double)sn / kPhiScale;
    double* out = phi + 2 * ((size_t)r * phi_stride + t_base + s);
    out[0] = (1.0 / n) * fabs((double)sg);
    out[1] = sqrt(c * c + q * q) / n;
  }
}

__global__ void k_flags_begin(int R, const unsigned* halt, unsigned* rows, int* first_bad) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  rows[r] = halt[r];
  rows[R + r] = 0u;
  first_bad[r] = INT_MAX;
}

__global__ void k_flags_end(int R, const unsigned* last_row, unsigned* halt, int* first_bad, unsigned* tail,
                            unsigned* next_out_row) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const unsigned f = last_row[r];
  const int fb = first_bad[r];
  halt[r] = f;
  if (tail) {  // the run's flags and first failing steps, beside its order history, for one read-back
    tail[r] = f;
    tail[R + r] = (unsigned)fb;
  }
  // what k_flags_begin would set for the next run (which reads halt itself at its first step):
  // the first step's output row cleared, no failure yet
  next_out_row[r] = 0u;
  first_bad[r] = INT_MAX;
}

__global__ void k_prepare(int n, const float* x, const float* y, const float* th, float4* xycs, float* th_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s, c;
  sincosf(th[i], &s, &c);
  xycs[i] = make_float4(x[i], y[i], c, s);
  th_out[i] = th[i];
}

// seam wrap of n moves, LANES lanes per move (the step kernel's lane groups); whole groups only
template <int LANES, bool WAVE>
__global__ void __launch_bounds__(256) k_seam_wrap(int n, const double2* start, double2* end, double* dtheta,
                                                   int* quarter, int* ok, int max_crossings) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / LANES, lane = t & (LANES - 1);
  if (i >= n) return;  // uniform over the group
  const double2 a = start[i];
  double px = end[i].x, py = end[i].y, dth = 0.0;
  int q = 0;
  const bool good = seam_wrap_group<LANES, WAVE>(a.x, a.y, px, py, dth, max_crossings, lane, &q);
  if (lane == 0) {
    end[i] = make_double2(px, py);
    dtheta[i] = dth;
    quarter[i] = q;
    ok[i] = good ? 1 : 0;
  }
}

// area-uniform initial state of every replica (replica r from seeds[r]); the draws are those of
// the host initialisation (rng.h), so the state equals it bit for bit
__global__ void k_init_random(int ntot, const int* rid, const ReplicaDesc* reps, const uint64_t* seeds,
                              const double* cdf, MeshDev mesh, float* x, float* y, float* th) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  const int r = rid[g];
  const tissue::InitDraw d = tissue::area_uniform_draw(seeds[r], g - reps[r].off, cdf, mesh.nf);
  const float2 a = mesh.uv[mesh.faces[3 * d.face]], b = mesh.uv[mesh.faces[3 * d.face + 1]],
               c = mesh.uv[mesh.faces[3 * d.face + 2]];
  tissue::barycentric_uv(a.x, a.y, b.x, b.y, c.x, c.y, d.r1, d.r2, &x[g], &y[g]);
  th[g] = d.th;
}

// per-step output stream: copies step t's UV position, 3-D position and neighbour count into the
// ring slot of its global step number g = ctl[1] + t + 1 when g is a multiple of `every`
__global__ void k_record(int ntot, const int* ctl, int t, int every, int depth, const float4* xycs, const float* r3x,
                         const float* r3y, const float* r3z, const int* nbr, float* ring, int* ring_nbr) {
  const long g = (long)ctl[1] + t + 1;
  if (g % every != 0) return;  // uniform over the grid
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntot) return;
  const size_t slot = (size_t)((g / every) % depth);
  float* f = ring + slot * 5 * (size_t)ntot;
  const float4 q = xycs[i];
  f[i] = q.x;
  f[(size_t)ntot + i] = q.y;
  f[2 * (size_t)ntot + i] = r3x[i];
  f[3 * (size_t)ntot + i] = r3y[i];
  f[4 * (size_t)ntot + i] = r3z[i];
  ring_nbr[slot * (size_t)ntot + i] = nbr[i];
}

__global__ void k_locate(MeshDev mesh, int n, const float4* xycs, int* vid, int* face, float* r3x, float* r3y,
                         float* r3z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = xycs[i];
  const Located L = locate<1>(mesh, (double)p.x, (double)p.y, 0);
  vid[i] = L.vid;
  if (face) face[i] = L.face;
  if (r3x) { r3x[i] = (float)L.X; r3y[i] = (float)L.Y; r3z[i] = (float)L.Z; }
}

__global__ void k_insert(int ntot, const int* vid, const int* rid, const ReplicaDesc* reps, int nv, BucketSet bs) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  const int r = rid[g];
  bucket_insert(replica_out(bs, r, nv, reps[r].off), vid[g], g);
}

__global__ void k_neighbor_lists(MeshDev mesh, const ReplicaDesc* reps, const int* rid, int ntot, BucketSet bs,
                                 const int* vid, int cap, int* counts, int* cols) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  const int r = rid[g];
  const int off = reps[r].off;
  const float t_near = reps[r].prm.t_near;
  const BucketsIn b = replica_in(bs, r, mesh.nv, off);
  const int v = vid[g];
  int k = 0;
  for (int e = mesh.row_ptr[v]; e < mesh.row_ptr[v + 1]; ++e) {
    const int2 ent = mesh.nbr[e];
    if (__int_as_float(ent.y) > t_near) break;
    for_bucket(b, ent.x, [&](int j) {
      if (k < cap) cols[(size_t)g * cap + k] = j - off;
      ++k;
    });
  }
  counts[g] = k;
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

template <int LANES, bool OCC, bool WAVE>
void launch_tiled_mode(int block, bool tape, const MeshDev& mesh, const ReplicaDesc* reps, const TileDesc* tiles,
                       int n_tiles, const BucketSet& bin, const BucketSet& bout, const BucketSet& bclr,
                       const TiledIO& io, hipStream_t st) {
  const size_t lds = OCC ? sizeof(unsigned) * (size_t)occ_words(mesh.nv) : 0;
  if (tape)
    hipLaunchKernelGGL((k_step_tiled<LANES, true, OCC, false, WAVE>), dim3(n_tiles), dim3(block), lds, st, mesh, reps,
                       tiles, bin, bout, bclr, io);
  else if (io.visit)  // the visit check: the same step, recording each particle's visited set
    hipLaunchKernelGGL((k_step_tiled<LANES, false, OCC, true, WAVE>), dim3(n_tiles), dim3(block), lds, st, mesh, reps,
                       tiles, bin, bout, bclr, io);
  else
    hipLaunchKernelGGL((k_step_tiled<LANES, false, OCC, false, WAVE>), dim3(n_tiles), dim3(block), lds, st, mesh, reps,
                       tiles, bin, bout, bclr, io);
}

// wave: the seam-wrap loop of crowded systems (seam_wrap_group), compiled for up to 8 lanes per
// particle (the shapes such systems run at); wider shapes take the plain loop (same results)
template <int LANES, bool OCC>
void launch_tiled_variant(int block, bool tape, bool wave, const MeshDev& mesh, const ReplicaDesc* reps,
                          const TileDesc* tiles, int n_tiles, const BucketSet& bin, const BucketSet& bout,
                          const BucketSet& bclr, const TiledIO& io, hipStream_t st) {
  if constexpr (LANES <= 8) {
    if (wave) {
      launch_tiled_mode<LANES, OCC, true>(block, tape, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st);
      return;
    }
  }
  (void)wave;
  launch_tiled_mode<LANES, OCC, false>(block, tape, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st);
}

template <int LANES>
void launch_tiled_lanes(int block, bool tape, bool wave, const MeshDev& mesh, const ReplicaDesc* reps,
                        const TileDesc* tiles, int n_tiles, const BucketSet& bin, const BucketSet& bout,
                        const BucketSet& bclr, const TiledIO& io, hipStream_t st) {
  if (stage_occupancy(mesh.nv, block, LANES))
    launch_tiled_variant<LANES, true>(block, tape, wave, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st);
  else
    launch_tiled_variant<LANES, false>(block, tape, wave, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st);
}

}  // namespace

hipError_t launch_step_tiled(int lanes, int block, bool tape, bool wave, const MeshDev& mesh, const ReplicaDesc* reps,
                             const TileDesc* tiles, int n_tiles, const BucketSet& bin, const BucketSet& bout,
                             const BucketSet& bclr, const TiledIO& io, hipStream_t st) {
  if (block < 64 || block > kStepMaxBlock || block % 64 || n_tiles < 1) return hipErrorInvalidValue;
  switch (lanes) {
    case 1: launch_tiled_lanes<1>(block, tape, wave, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st); break;
    case 2: launch_tiled_lanes<2>(block, tape, wave, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st); break;
    case 4: launch_tiled_lanes<4>(block, tape, wave, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st); break;
    case 8: launch_tiled_lanes<8>(block, tape, wave, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st); break;
    case 16: launch_tiled_lanes<16>(block, tape, wave, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st); break;
    case 32: launch_tiled_lanes<32>(block, tape, wave, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st); break;
    case 64: launch_tiled_lanes<64>(block, tape, wave, mesh, reps, tiles, n_tiles, bin, bout, bclr, io, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_bucket_finalize(int block, const ReplicaDesc* reps, const TileDesc* tiles, int n_tiles, int nv,
                                  int single_n, const BucketSet& bs, hipStream_t st) {
  hipLaunchKernelGGL(k_bucket_finalize, dim3(n_tiles), dim3(block), 0, st, reps, tiles, nv, single_n, bs);
  return hipGetLastError();
}

hipError_t launch_phi_tiles(int R, int steps, const int* tile_begin, int n_logical, int per_tile,
                            const OrderPartial* partial, const ReplicaDesc* reps, double* phi, int phi_stride,
                            int t_base, hipStream_t st) {
  hipLaunchKernelGGL(k_phi_tiles, dim3(R * steps), dim3(kPhiBlock), 0, st, steps, tile_begin, n_logical, per_tile,
                     partial, reps, phi, phi_stride, t_base);
  return hipGetLastError();
}

hipError_t launch_flags_begin(int R, const unsigned* halt, unsigned* rows, int* first_bad, hipStream_t st) {
  hipLaunchKernelGGL(k_flags_begin, dim3(cdiv(R, 64)), dim3(64), 0, st, R, halt, rows, first_bad);
  return hipGetLastError();
}

hipError_t launch_flags_end(int R, const unsigned* last_row, unsigned* halt, int* first_bad, unsigned* tail,
                            unsigned* next_out_row, hipStream_t st) {
  hipLaunchKernelGGL(k_flags_end, dim3(cdiv(R, 64)), dim3(64), 0, st, R, last_row, halt, first_bad, tail,
                     next_out_row);
  return hipGetLastError();
}

hipError_t launch_prepare(int n, const float* x, const float* y, const float* th, float4* xycs, float* th_out,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_prepare, dim3(cdiv(n, 256)), dim3(256), 0, st, n, x, y, th, xycs, th_out);
  return hipGetLastError();
}

hipError_t launch_locate(const MeshDev& mesh, int n, const float4* xycs, int* vid, int* face, float* r3x, float* r3y,
                         float* r3z, hipStream_t st) {
  hipLaunchKernelGGL(k_locate, dim3(cdiv(n, 256)), dim3(256), 0, st, mesh, n, xycs, vid, face, r3x, r3y, r3z);
  return hipGetLastError();
}

template <bool WAVE>
void launch_seam_wrap_mode(int lanes, dim3 grid, dim3 block, int n, const double2* start, double2* end, double* dtheta,
                           int* quarter, int* ok, int max_crossings, hipStream_t st) {
  switch (lanes) {
    case 1: hipLaunchKernelGGL((k_seam_wrap<1, WAVE>), grid, block, 0, st, n, start, end, dtheta, quarter, ok, max_crossings); break;
    case 2: hipLaunchKernelGGL((k_seam_wrap<2, WAVE>), grid, block, 0, st, n, start, end, dtheta, quarter, ok, max_crossings); break;
    case 4: hipLaunchKernelGGL((k_seam_wrap<4, WAVE>), grid, block, 0, st, n, start, end, dtheta, quarter, ok, max_crossings); break;
    case 8: hipLaunchKernelGGL((k_seam_wrap<8, WAVE>), grid, block, 0, st, n, start, end, dtheta, quarter, ok, max_crossings); break;
    case 16: hipLaunchKernelGGL((k_seam_wrap<16, WAVE>), grid, block, 0, st, n, start, end, dtheta, quarter, ok, max_crossings); break;
    case 32: hipLaunchKernelGGL((k_seam_wrap<32, WAVE>), grid, block, 0, st, n, start, end, dtheta, quarter, ok, max_crossings); break;
    case 64: hipLaunchKernelGGL((k_seam_wrap<64, WAVE>), grid, block, 0, st, n, start, end, dtheta, quarter, ok, max_crossings); break;
    default: break;
  }
}

hipError_t launch_seam_wrap(int lanes, int n, const double2* start, double2* end, double* dtheta, int* quarter, int* ok,
                            int max_crossings, int wave, hipStream_t st) {
  if (lanes < 1 || lanes > 64 || (lanes & (lanes - 1))) return hipErrorInvalidValue;
  const dim3 grid(cdiv(n * lanes, 256)), block(256);
  if (wave)
    launch_seam_wrap_mode<true>(lanes, grid, block, n, start, end, dtheta, quarter, ok, max_crossings, st);
  else
    launch_seam_wrap_mode<false>(lanes, grid, block, n, start, end, dtheta, quarter, ok, max_crossings, st);
  return hipGetLastError();
}

hipError_t launch_init_random(int ntot, const int* rid, const ReplicaDesc* reps, const uint64_t* seeds,
                              const double* cdf, const MeshDev& mesh, float* x, float* y, float* th, hipStream_t st) {
  hipLaunchKernelGGL(k_init_random, dim3(cdiv(ntot, 256)), dim3(256), 0, st, ntot, rid, reps, seeds, cdf, mesh, x, y, th);
  return hipGetLastError();
}

hipError_t launch_record(int ntot, const int* ctl, int t, int every, int depth, const float4* xycs, const float* r3x,
                         const float* r3y, const float* r3z, const int* nbr, float* ring, int* ring_nbr, hipStream_t st) {
  hipLaunchKernelGGL(k_record, dim3(cdiv(ntot, 256)), dim3(256), 0, st, ntot, ctl, t, every, depth, xycs, r3x, r3y,
                     r3z, nbr, ring, ring_nbr);
  return hipGetLastError();
}

hipError_t launch_insert(int ntot, const int* vid, const int* rid, const ReplicaDesc* reps, int nv,
                         const BucketSet& bs, hipStream_t st) {
  hipLaunchKernelGGL(k_insert, dim3(cdiv(ntot, 256)), dim3(256), 0, st, ntot, vid, rid, reps, nv, bs);
  return hipGetLastError();
}

hipError_t launch_neighbor_lists(const MeshDev& mesh, const ReplicaDesc* reps, const int* rid, int ntot,
                                 const BucketSet& bs, const int* vid, int cap, int* counts, int* cols,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_neighbor_lists, dim3(cdiv(ntot, 256)), dim3(256), 0, st, mesh, reps, rid, ntot, bs, vid, cap,
                     counts, cols);
  return hipGetLastError();
}

}  // namespace tissue
