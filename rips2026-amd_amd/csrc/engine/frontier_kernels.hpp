// Frontier kernels shared by the Bellman-Ford and delta-stepping drivers.
#pragma once

#include "graph.hpp"
#include "relax.hpp"

namespace mi355 {

struct FrontierIn {
  const int* q;          // frontier vertex ids
  const uint32_t* keys;  // Jacobi snapshot keys (nullptr: read the live label)
  const uint32_t* cnt;   // frontier size (device)
};

struct ChunkQueue {
  uint4* q;       // {vertex, key, first edge, end edge}
  uint32_t* cnt;  // appended chunk count (device)
  uint32_t cap;
};

// Book-keeping done by thread 0 of block 0 of a round's first kernel.
struct RoundHousekeeping {
  uint32_t* zero0;  // counters that the next rounds will append to
  uint32_t* zero1;
  uint32_t* log;    // log[0] = frontier size of this round (nullptr: none)
};

__device__ inline uint32_t atomic_load_key(const unsigned long long* p) {
  return (uint32_t)(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32);
}

// Bellman-Ford sink: improved vertices go to the next frontier, once per round
// (round stamp), staged per wave in LDS.
struct BfSink {
  int* qout;
  uint32_t* cnt;
  uint32_t cap;
  uint32_t* stamp;
  uint32_t stamp_val;
  uint32_t* err;
  int* stage = nullptr;  // LDS, set by the kernel per wave
  uint32_t staged = 0;
  __device__ inline void vertex(bool, int, uint32_t) {}
  __device__ inline void improved(bool imp, uint32_t v, uint32_t) {
    bool push = false;
    if (imp && stamp[v] != stamp_val) push = atomicExch(stamp + v, stamp_val) != stamp_val;
    stage_push(push, (int)v, stage, staged, qout, cnt, cap, err);
  }
  __device__ inline void flush() { stage_flush(stage, staged, qout, cnt, cap, err); }
};

// Small/medium-degree frontier vertices: 64 per wave, edges packed across
// lanes.  Frontier entries [first, n_in) are processed; cls selects all,
// light or heavy edges.  With `split` (rows sorted by weight, split[u] = light
// edges of u) a class is a prefix or suffix of the row and only that range is
// read.  ids[u] is the input id of internal vertex u, the parent written into
// the labels (nullptr: internal ids are input ids).  kWaveLocal: the calling
// wave alone relaxes entries [first, n_in) (a list of its own); otherwise the
// entries are dealt over every wave of the grid.
template <class W, bool JACOBI, class Sink, bool kWaveLocal = false>
__device__ inline void relax_frontier_body(const int* __restrict__ rowptr,
                                           const uint2* __restrict__ edges,
                                           unsigned long long* __restrict__ labels,
                                           const int* __restrict__ qin,
                                           const uint32_t* __restrict__ qkeys, uint32_t first,
                                           uint32_t n_in, ChunkQueue cq, int cls,
                                           uint32_t delta_bits, Sink& sink,
                                           unsigned long long* __restrict__ work,
                                           const int* __restrict__ ids,
                                           const int* __restrict__ split = nullptr) {
  __shared__ int s_incl[kWavesPerBlock][kWave];
  __shared__ int s_base[kWavesPerBlock][kWave];
  __shared__ uint32_t s_key[kWavesPerBlock][kWave];
  __shared__ int s_u[kWavesPerBlock][kWave];

  const unsigned lane = threadIdx.x & (kWave - 1);
  const unsigned wv = threadIdx.x / kWave;
  const uint32_t nwaves = kWaveLocal ? 1u : gridDim.x * kWavesPerBlock;
  unsigned long long edge_visits = 0;

  for (uint32_t base = first + (kWaveLocal ? 0u : (blockIdx.x * kWavesPerBlock + wv) * kWave); base < n_in;
       base += nwaves * kWave) {
    const uint32_t i = base + lane;
    int u = 0, beg = 0, deg = 0, pu = 0;
    uint32_t key = W::kInfKey;
    if (i < n_in) {
      u = qin[i];
      pu = ids ? ids[u] : u;
      key = JACOBI ? qkeys[i] : atomic_load_key(labels + u);
      beg = rowptr[u];
      int end = rowptr[u + 1];
      if (split && cls != kAllEdges) {
        const int l = split[u];
        if (cls == kLightEdges) end = beg + l;
        else beg += l;
      }
      if (key != W::kInfKey) {
        deg = end - beg;
        if (deg > kLargeDegree) {
          const uint32_t nch = (uint32_t)((deg + kChunk - 1) / kChunk);
          const uint32_t pos = atomicAdd(cq.cnt, nch);
          for (uint32_t c = 0; c < nch; ++c) {
            const int a = beg + (int)c * kChunk;
            const int b = min(end, a + kChunk);
            if (pos + c < cq.cap) cq.q[pos + c] = make_uint4((uint32_t)pu, key, (uint32_t)a, (uint32_t)b);
          }
          deg = 0;
        }
      }
    }
    sink.vertex(i < n_in && key != W::kInfKey, u, key);
    int incl = deg;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int t = __shfl_up(incl, off, kWave);
      if ((int)lane >= off) incl += t;
    }
    const int total = __shfl(incl, kWave - 1, kWave);
    s_incl[wv][lane] = incl;
    s_base[wv][lane] = beg - (incl - deg);
    s_key[wv][lane] = key;
    s_u[wv][lane] = pu;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    edge_visits += (unsigned long long)total;

    // two edges per lane per step (j and j + 64), relaxed as a pair
    for (int t = 0; t < total; t += 2 * kWave) {
      const int j0 = t + (int)lane, j1 = j0 + kWave;
      const bool ok0 = j0 < total, ok1 = j1 < total;
      int p0 = 0, p1 = 0;
#pragma unroll
      for (int s = 32; s > 0; s >>= 1) {
        if (s_incl[wv][p0 + s - 1] <= j0) p0 += s;
        if (s_incl[wv][p1 + s - 1] <= j1) p1 += s;
      }
      const uint2 e0 = edges[ok0 ? s_base[wv][p0] + j0 : 0];
      const uint2 e1 = edges[ok1 ? s_base[wv][p1] + j1 : 0];
      const bool sel0 = ok0 && edge_selected<W>(cls, e0.y, delta_bits);
      const bool sel1 = ok1 && edge_selected<W>(cls, e1.y, delta_bits);
      const uint32_t ck0 = sel0 ? W::relax(s_key[wv][p0], e0.y) : W::kInfKey;
      const uint32_t ck1 = sel1 ? W::relax(s_key[wv][p1], e1.y) : W::kInfKey;
      bool imp0, imp1;
      relax_edge_pair<W>(labels, e0.x, ck0, (uint32_t)s_u[wv][p0], e1.x, ck1, (uint32_t)s_u[wv][p1],
                         imp0, imp1);
      sink.improved(imp0, e0.x, ck0);
      sink.improved(imp1, e1.x, ck1);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (work && lane == 0 && edge_visits) atomicAdd(work, edge_visits);
}

// Chunks of high-degree vertices: one chunk per wave, lanes stream its edges
// ({parent id, key, first edge, end edge}).
template <class W, class Sink>
__device__ inline void relax_chunks_body(const uint2* __restrict__ edges,
                                         unsigned long long* __restrict__ labels, ChunkQueue cq,
                                         int cls, uint32_t delta_bits, Sink& sink,
                                         unsigned long long* __restrict__ work) {
  const uint32_t n_ch = min(*cq.cnt, cq.cap);
  const unsigned lane = threadIdx.x & (kWave - 1);
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  unsigned long long edge_visits = 0;
  for (uint32_t c = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; c < n_ch; c += nwaves) {
    const uint4 ch = cq.q[c];
    edge_visits += ch.w - ch.z;
    for (uint32_t t = ch.z; t < ch.w; t += 2 * kWave) {
      const uint32_t i0 = t + lane, i1 = i0 + kWave;
      const bool ok0 = i0 < ch.w, ok1 = i1 < ch.w;
      const uint2 e0 = edges[ok0 ? i0 : ch.z];
      const uint2 e1 = edges[ok1 ? i1 : ch.z];
      const bool sel0 = ok0 && edge_selected<W>(cls, e0.y, delta_bits);
      const bool sel1 = ok1 && edge_selected<W>(cls, e1.y, delta_bits);
      const uint32_t ck0 = sel0 ? W::relax(ch.y, e0.y) : W::kInfKey;
      const uint32_t ck1 = sel1 ? W::relax(ch.y, e1.y) : W::kInfKey;
      bool imp0, imp1;
      relax_edge_pair<W>(labels, e0.x, ck0, ch.x, e1.x, ck1, ch.x, imp0, imp1);
      sink.improved(imp0, e0.x, ck0);
      sink.improved(imp1, e1.x, ck1);
    }
  }
  if (work && lane == 0 && edge_visits) atomicAdd(work, edge_visits);
}

template <class W, bool JACOBI, class Sink>
__global__ __launch_bounds__(kBlock) void k_relax_frontier(
    const int* __restrict__ rowptr, const uint2* __restrict__ edges,
    unsigned long long* __restrict__ labels, FrontierIn in, ChunkQueue cq, Sink sink,
    RoundHousekeeping hk, unsigned long long* __restrict__ work, const int* __restrict__ ids) {
  __shared__ int s_stage[kWavesPerBlock][kStageCap];
  const uint32_t n_in = *in.cnt;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (hk.zero0) *hk.zero0 = 0;
    if (hk.zero1) *hk.zero1 = 0;
    if (hk.log) *hk.log = n_in;
  }
  sink.stage = s_stage[threadIdx.x / kWave];
  relax_frontier_body<W, JACOBI>(rowptr, edges, labels, in.q, in.keys, 0u, n_in, cq, kAllEdges,
                                 0u, sink, work, ids);
  sink.flush();
}

template <class W, class Sink>
__global__ __launch_bounds__(kBlock) void k_relax_chunks(const uint2* __restrict__ edges,
                                                         unsigned long long* __restrict__ labels,
                                                         ChunkQueue cq, Sink sink,
                                                         unsigned long long* __restrict__ work) {
  __shared__ int s_stage[kWavesPerBlock][kStageCap];
  sink.stage = s_stage[threadIdx.x / kWave];
  relax_chunks_body<W>(edges, labels, cq, kAllEdges, 0u, sink, work);
  sink.flush();
}

// In-place Bellman-Ford with chained relaxations.  A wave keeps the vertices it
// improves in an LDS list (up to kChainCap) and relaxes them itself, up to
// kChainPasses passes, before the round ends; what is left (a full list, or
// the last pass's improvements) goes to the next round's frontier as usual.
// A chain of improvements thus advances several hops per round instead of
// one.  In-place rounds promise only the fixpoint (every label drop is
// followed by a relaxation of the vertex with its new label, here or in the
// next round), not a round count, so the Jacobi mode keeps plain rounds.
// For graphs without chunked rows (every row at most kLargeDegree edges): the
// chunk queue holds a high-degree row once per round.
// 256 x 256 grid: 575 -> 63 host rounds, 6.1 -> 4.0 ms; 1024 x 1024: 43.7 -> 17.3 ms
// (lists of 128-512 entries and 8-32 passes measured; this pair was fastest)
constexpr uint32_t kChainCap = 128;
constexpr int kChainPasses = 16;

struct ChainSink {
  BfSink g;           // the next round's frontier
  int* list;          // LDS, kChainCap entries (this wave's)
  uint32_t n = 0;     // wave-uniform
  __device__ inline void vertex(bool, int, uint32_t) {}
  __device__ inline void improved(bool imp, uint32_t v, uint32_t key) {
    const unsigned long long mask = ballot64(imp);
    if (mask == 0) return;
    const uint32_t k = (uint32_t)__popcll(mask);
    if (n + k <= kChainCap) {
      if (imp) list[n + prefix_count(mask)] = (int)v;
      n += k;
    } else {
      g.improved(imp, v, key);
    }
  }
  __device__ inline void flush() { g.flush(); }
};

template <class W>
__global__ __launch_bounds__(kBlock) void k_relax_frontier_chain(
    const int* __restrict__ rowptr, const uint2* __restrict__ edges,
    unsigned long long* __restrict__ labels, FrontierIn in, ChunkQueue cq, BfSink sink,
    RoundHousekeeping hk, unsigned long long* __restrict__ work, const int* __restrict__ ids) {
  __shared__ int s_stage[kWavesPerBlock][kStageCap];
  __shared__ int s_list[2][kWavesPerBlock][kChainCap];
  const uint32_t n_in = *in.cnt;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (hk.zero0) *hk.zero0 = 0;
    if (hk.zero1) *hk.zero1 = 0;
    if (hk.log) *hk.log = n_in;
  }
  const unsigned wv = threadIdx.x / kWave;
  sink.stage = s_stage[wv];
  ChainSink cs{sink, s_list[0][wv], 0u};
  relax_frontier_body<W, false>(rowptr, edges, labels, in.q, nullptr, 0u, n_in, cq, kAllEdges, 0u, cs, work, ids);
  for (int pass = 0; pass < kChainPasses && cs.n > 0; ++pass) {
    // relax this pass's list; its improvements fill the other buffer
    const int* cur = cs.list;
    const uint32_t n_cur = cs.n;
    cs.list = s_list[(pass + 1) & 1][wv];
    cs.n = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    relax_frontier_body<W, false, ChainSink, true>(rowptr, edges, labels, cur, nullptr, 0u, n_cur, cq, kAllEdges,
                                                   0u, cs, work, ids);
  }
  // the last list goes to the next round
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const unsigned lane = threadIdx.x & (kWave - 1);
  for (uint32_t b = 0; b < cs.n; b += kWave) {
    const bool has = b + lane < cs.n;
    cs.g.improved(has, has ? (uint32_t)cs.list[b + lane] : 0u, 0u);
  }
  cs.g.flush();
}

// Light edges of every vertex: rows are sorted by weight, so they are the
// prefix with w <= delta (the test of edge_selected); binary search per row.
template <class W>
__global__ void k_light_count(const int* __restrict__ rowptr, const uint2* __restrict__ edges,
                              int64_t n, uint32_t delta_bits, int* __restrict__ split) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    const int beg = rowptr[v];
    int lo = beg, hi = rowptr[v + 1];  // first heavy edge in [lo, hi]
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (edge_selected<W>(kLightEdges, edges[mid].y, delta_bits)) lo = mid + 1;
      else hi = mid;
    }
    split[v] = lo - beg;
  }
}

// The same split at a width chosen on the device: sel[0] = its float bits,
// sel[1] = nonzero when `split` does not hold them yet (else nothing to do).
template <class W>
__global__ void k_light_count_sel(const int* __restrict__ rowptr, const uint2* __restrict__ edges,
                                  int64_t n, const uint32_t* __restrict__ sel, int* __restrict__ split) {
  if (sel[1] == 0u) return;
  const uint32_t delta_bits = sel[0];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    const int beg = rowptr[v];
    int lo = beg, hi = rowptr[v + 1];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (edge_selected<W>(kLightEdges, edges[mid].y, delta_bits)) lo = mid + 1;
      else hi = mid;
    }
    split[v] = lo - beg;
  }
}

// Jacobi snapshot: keys of the next frontier as they stand after the round.
__global__ void k_snapshot_keys(const int* __restrict__ q, const uint32_t* __restrict__ cnt,
                                const unsigned long long* __restrict__ labels,
                                uint32_t* __restrict__ keys);

}  // namespace mi355
