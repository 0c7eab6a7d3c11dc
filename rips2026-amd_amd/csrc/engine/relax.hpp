// Wave64 frontier relaxation: the hot loop shared by Bellman-Ford and
// delta-stepping.
//
// One wave takes 64 frontier vertices (one per lane), reads their offsets and
// distance keys, and packs their edge lists across the 64 lanes with a wave
// prefix sum, so a lane relaxes edge j of the packed range and consecutive
// lanes stream consecutive 8-byte edge records of the same vertex (coalesced).
// A lane finds the vertex owning packed edge j by a 6-step binary search over
// the wave's inclusive degree scan held in LDS.  Vertices with more than
// kLargeDegree edges are instead cut into kChunk-edge pieces and appended to a
// chunk queue that whole waves drain (relax_chunks), so R-MAT hubs spread over
// the grid instead of stalling one wave.
//
// Relaxation is a packed 64-bit atomicMin of (dist key << 32 | parent) after
// a plain load that filters candidates which cannot win.  A lane whose
// candidate strictly lowers the distance part pushes the target to the output
// queue once per round (round stamp), and the pushes of a wave are compacted
// with a ballot + mbcnt into one atomicAdd on the queue counter.
#pragma once

#include "common.hpp"

namespace mi355 {

__device__ inline unsigned lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ inline unsigned long long ballot64(bool p) { return __ballot(p); }

// Count of set bits of mask below this lane.
__device__ inline unsigned prefix_count(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave-aggregated append of `v` for the lanes with push == true.
__device__ inline void wave_append(bool push, int v, int* __restrict__ q, uint32_t* __restrict__ cnt,
                                   uint32_t cap, uint32_t* __restrict__ err) {
  const unsigned long long mask = ballot64(push);
  if (mask == 0) return;
  const unsigned lane = lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if ((int)lane == leader) base = atomicAdd(cnt, (uint32_t)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (push) {
    const uint32_t pos = base + prefix_count(mask);
    if (pos < cap) q[pos] = v;
    else atomicOr(err, 1u);
  }
}

// Per-wave LDS staging of queue appends.  A global atomicAdd on a queue
// counter is paid once per kStageCap staged entries (or at the end of the
// wave's work) instead of once per wave-iteration: queue counters are single
// addresses that every wave of the grid hits, and same-address atomics
// serialise.  `staged` is wave-uniform and lives in a register.
constexpr uint32_t kStageCap = 512;

__device__ inline void stage_flush(int* __restrict__ stage, uint32_t& staged, int* __restrict__ q,
                                   uint32_t* __restrict__ cnt, uint32_t cap,
                                   uint32_t* __restrict__ err) {
  if (staged == 0) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const unsigned lane = lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(cnt, staged);
  base = __shfl(base, 0, 64);
  for (uint32_t i = lane; i < staged; i += kWave) {
    if (base + i < cap) q[base + i] = stage[i];
    else atomicOr(err, 1u);
  }
  __builtin_amdgcn_wave_barrier();
  staged = 0;
}

__device__ inline void stage_push(bool push, int v, int* __restrict__ stage, uint32_t& staged,
                                  int* __restrict__ q, uint32_t* __restrict__ cnt, uint32_t cap,
                                  uint32_t* __restrict__ err) {
  const unsigned long long mask = ballot64(push);
  if (mask == 0) return;
  const uint32_t k = (uint32_t)__popcll(mask);
  if (staged + k > kStageCap) stage_flush(stage, staged, q, cnt, cap, err);
  if (push) stage[staged + prefix_count(mask)] = v;
  staged += k;
}

// Per-wave LDS staging of appends into kBins queues (bucket windows).  Each
// staged entry keeps its vertex and bin; a flush counts the bins in LDS,
// reserves every non-empty bin's range with one global atomicAdd on that
// bin's counter (counters are cnt_stride words apart, so they sit on
// separate cache lines), then writes the entries.
constexpr int kBins = 32;

struct BinStage {
  int* v;          // kStageCap vertices
  uint16_t* bin;   // kStageCap bins
  uint16_t* off;   // kStageCap offsets within the bin (flush only)
  uint32_t* hist;  // kBins counts, then bases
  uint32_t n;      // wave-uniform
};

__device__ inline void bin_flush(BinStage& st, uint32_t* __restrict__ cnt, int cnt_stride,
                                 int* __restrict__ q, int64_t q_stride, int64_t cap,
                                 uint32_t* __restrict__ err) {
  if (st.n == 0) return;
  const unsigned lane = lane_id();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < (unsigned)kBins) st.hist[lane] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (uint32_t i = lane; i < st.n; i += kWave) st.off[i] = (uint16_t)atomicAdd(st.hist + st.bin[i], 1u);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < (unsigned)kBins) {
    const uint32_t k = st.hist[lane];
    if (k) st.hist[lane] = atomicAdd(cnt + (int64_t)lane * cnt_stride, k);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (uint32_t i = lane; i < st.n; i += kWave) {
    const uint32_t b = st.bin[i];
    const int64_t pos = (int64_t)st.hist[b] + st.off[i];
    if (pos < cap) q[(int64_t)b * q_stride + pos] = st.v[i];
    else atomicOr(err, 2u);
  }
  __builtin_amdgcn_wave_barrier();
  st.n = 0;
}

__device__ inline void bin_push(bool push, int v, uint32_t bin, BinStage& st,
                                uint32_t* __restrict__ cnt, int cnt_stride, int* __restrict__ q,
                                int64_t q_stride, int64_t cap, uint32_t* __restrict__ err) {
  const unsigned long long mask = ballot64(push);
  if (mask == 0) return;
  const uint32_t k = (uint32_t)__popcll(mask);
  if (st.n + k > kStageCap) bin_flush(st, cnt, cnt_stride, q, q_stride, cap, err);
  if (push) {
    const uint32_t i = st.n + prefix_count(mask);
    st.v[i] = v;
    st.bin[i] = (uint16_t)bin;
  }
  st.n += k;
}

// Edge filter for delta-stepping's light/heavy split (w <= delta is light,
// DS/src/delta_stepping_hip_CSR.cpp:402, :440).
enum EdgeClass { kAllEdges = 0, kLightEdges = 1, kHeavyEdges = 2 };

template <class W>
__device__ inline bool edge_selected(int cls, uint32_t wbits, uint32_t delta_bits) {
  if (cls == kAllEdges) return true;
  bool light;
  if (W::kType == MI355_W_F32) light = __uint_as_float(wbits) <= __uint_as_float(delta_bits);
  else light = (float)wbits <= __uint_as_float(delta_bits);
  return cls == kLightEdges ? light : !light;
}

// The packed-label relaxation step for two edges at once.  A label is
// (distance key << 32 | parent): a candidate (ck << 32 | parent) replaces it
// when smaller, by one 64-bit atomicMin, so the smaller distance wins and
// among equal distances the smaller parent (the canonical predecessor); a
// plain load first filters candidates that cannot win.  imp = the distance
// part strictly decreased.  Candidate keys ck0 / ck1 (kInfKey = no edge):
// both label probes, then both atomicMins, are issued before either result
// is used, so the two dependent chains overlap.  A lane without an
// edge probes label 0 and discards it (a guarded load would wait inside its
// branch).
template <class W>
__device__ inline void relax_edge_pair(unsigned long long* __restrict__ labels, uint32_t v0,
                                       uint32_t ck0, uint32_t parent0, uint32_t v1, uint32_t ck1,
                                       uint32_t parent1, bool& imp0, bool& imp1) {
  constexpr unsigned long long kNone = ~0ull;
  const bool has0 = ck0 != W::kInfKey, has1 = ck1 != W::kInfKey;
  const unsigned long long cand0 = has0 ? ((unsigned long long)ck0 << 32) | parent0 : kNone;
  const unsigned long long cand1 = has1 ? ((unsigned long long)ck1 << 32) | parent1 : kNone;
  unsigned long long cur0 = labels[has0 ? v0 : 0u];
  unsigned long long cur1 = labels[has1 ? v1 : 0u];
  // both probes complete here, so no wait separates the two atomics below
  asm volatile("" : "+v"(cur0), "+v"(cur1));
  const bool try0 = cand0 < cur0, try1 = cand1 < cur1;
  unsigned long long old0 = kNone, old1 = kNone;
  if (try0) old0 = atomicMin(labels + v0, cand0);
  if (try1) old1 = atomicMin(labels + v1, cand1);
  imp0 = try0 && ck0 < (uint32_t)(old0 >> 32);
  imp1 = try1 && ck1 < (uint32_t)(old1 >> 32);
}

}  // namespace mi355
