// Persistent ADDS worklist (adds.hip) and its phase policy, including the
// learned delta/priority MLP evaluated on the device (mlp.hip).
#pragma once

#include "common.hpp"

namespace mi355 {

constexpr int kAddsSlots = 32;     // circular bucket window
constexpr int kMlpIn = 32;         // phase features
constexpr int kMlpHidden = 64;
// log2 delta scale, active buckets, log2 worker share, log2 light width (first phase)
constexpr int kMlpOut = 4;
constexpr int kTraceStride = 40;   // floats per traced phase

enum PolicyMode { kPolicyFixed = 0, kPolicyHeuristic = 1, kPolicyMlp = 2, kPolicySchedule = 3 };
constexpr int kSchedMax = 64;      // phases of an explicit schedule (the last entry repeats)

// Device control block of one ADDS run.  Every counter that many waves hit
// with atomics sits on its own 256-byte line (kPad words apart) so unrelated
// counters do not serialise on one line.
constexpr int kPad = 64;
struct AddsCtrl {
  // read-mostly parameters
  float delta;                   // current bucket width
  float base_delta;              // heuristic delta the policy scales
  float min_delta, max_delta;
  float util_low, util_high, growth, shrink;
  uint32_t active;               // buckets served concurrently (priority window)
  uint32_t workers;
  uint32_t phases;
  uint32_t err;
  uint32_t active_workers;       // workgroups that claim work (the rest only idle)
  float light_scale;             // network's light width / first bucket width (first phase)
  uint32_t pad0[kPad - 14];
  unsigned long long base_word;  // (T_base float bits << 32) | base bucket
  uint32_t pad1[kPad - 2];
  unsigned long long pending;    // items queued or being processed (two's-complement adds)
  uint32_t pad2[kPad - 2];
  uint32_t done;
  uint32_t pad3[kPad - 1];
  uint32_t hub_head;
  uint32_t pad4[kPad - 1];
  uint32_t hub_reserve;
  uint32_t pad5[kPad - 1];
  int hub_avail;                 // published, unclaimed hub chunks (may dip below 0)
  uint32_t pad5b[kPad - 1];
  unsigned long long edges, pushes, items;
  uint32_t pad6[kPad - 6];
  // diagnostics: claims, failed claims, idle polls, ring-entry waits, hub
  // chunks, stale entries dropped, re-queued entries, heavy-list items; then
  // workgroup-0-wave time (wall clock ticks) in successful claims, idle
  // polling, item claims, hub chunks, publishing, relaxing
  unsigned long long diag[16];
  uint32_t pad7[kPad - 32];
  // pull diagnostics (debug runs): workgroup-0-wave ticks in pull chunks,
  // in-edges read, vertices whose in-edge range was searched
  unsigned long long pdiag[4];
  uint32_t pad7b[kPad - 8];
  // Pulls.  At a phase boundary the heavy edges of the bucket just finished
  // are either pushed from its vertices (its heavy list) or pulled by every
  // vertex that an edge from it could still improve; the pull runs as
  // kPullChunk-vertex chunks while base == pull_base.  `decided` is
  // (b << 1) | pulled for the last base b whose boundary decision is
  // published: the base does not move past b before, nor is the heavy list of
  // bucket b - 1 claimed (pushed, or dropped when pulled).
  uint32_t decided;
  uint32_t pull_mode;            // 0: always push; 1: pull when estimated cheaper
  float pull_margin;             // log2(push / pull work estimate) above which to pull
  uint32_t pull_base;            // base during which the registered pull runs
  uint32_t pull_chunks;          // chunks of the registered pull
  uint32_t pull_count;           // heavy-list entries of the pulled bucket
  uint32_t pulls;                // buckets pulled
  uint32_t pad8[kPad - 7];
  int pull_avail;                // unclaimed chunks of the registered pull (may dip below 0)
  uint32_t pad9[kPad - 1];
  uint32_t pull_head;            // next chunk index
  uint32_t pad10[kPad - 1];
  uint32_t pull_done;            // finished chunks
  uint32_t pad11[kPad - 1];
  // the pulled bucket's bound: every heavy edge its list would have pushed
  // starts at this key or higher (the bucket's hv_min at the boundary); the
  // chunks start once pull_keyed reaches pull_count (set together with it)
  uint32_t pull_key;
  uint32_t pull_keyed;
  uint32_t pad12[kPad - 2];
  uint32_t head[kAddsSlots * kPad];  // pop counter of slot s at [s * kPad] (monotone)
  uint32_t tail[kAddsSlots * kPad];  // push counter of slot s
  int inflight[kAddsSlots * kPad];   // claims of slot s still being processed
  int avail[kAddsSlots * kPad];      // written, unclaimed entries of slot s (may dip below 0)
  // heavy lists: vertices processed in bucket b whose heavy edges are still
  // to be relaxed, in list slot b % kAddsSlots, claimable while base == b + 1
  uint32_t hv_head[kAddsSlots * kPad];
  uint32_t hv_tail[kAddsSlots * kPad];
  int hv_avail[kAddsSlots * kPad];
  // smallest key with which a vertex that has heavy edges was processed in
  // the light pass of bucket b (slot b % kAddsSlots; 0xFFFFFFFF: none), kept
  // by the pull kernel and reset when b is finished
  uint32_t hv_min[kAddsSlots * kPad];
};

// One 2048-edge piece of a high-degree vertex.
struct HubChunk {
  uint32_t u, key, first, cnt;
  uint32_t slot, bucket, seq, pad;
};

// Policy parameters in device memory (read-only during a run except trace).
struct AddsPolicy {
  int mode;
  uint32_t trace_cap;            // phases the trace can hold
  float* trace;                  // trace_cap x kTraceStride floats (may be null)
  float graph_feat[14];          // degree + weight quantile features
  float feat_mean[kMlpIn], feat_inv_std[kMlpIn];
  // bf16 weights (stored as raw 16-bit words), fp32 biases
  uint16_t w1[kMlpHidden * kMlpIn];
  float b1[kMlpHidden];
  uint16_t w2[kMlpOut * kMlpHidden];
  float b2[kMlpOut];
  // schedule mode: phase i runs with entry min(i, sched_len - 1): width as a
  // multiple of delta0, buckets served at once, share of claiming workgroups
  uint32_t sched_len;
  float sched[kSchedMax * 3];
};

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Feature i of the phase, held by lane i (< 32): *raw as computed, *feat
// normalised for the network.  occ = items queued in slot (base + lane) for
// lanes < kAddsSlots.
__device__ inline void mlp_features(const AddsPolicy* p, const AddsCtrl* c, uint32_t occ,
                                    uint32_t window_items, float delta, unsigned lane, float* feat,
                                    float* raw) {
  float x = 0.0f;
  if (lane < 14) x = p->graph_feat[lane];
  else if (lane == 14) x = log2f(1.0f + (float)window_items);
  else if (lane == 15) x = log2f(delta / c->base_delta);
  const float occ_log = log2f(1.0f + (float)occ);
  // lanes 16..31 take the occupancy of window offsets 0..15
  const float shifted = __shfl(occ_log, (int)(lane - 16) & 63, 64);
  if (lane >= 16 && lane < 32) x = shifted;
  *raw = lane < kMlpIn ? x : 0.0f;
  if (lane < kMlpIn) x = (x - p->feat_mean[lane]) * p->feat_inv_std[lane];
  *feat = lane < kMlpIn ? x : 0.0f;
}

// 32 -> 64 (ReLU) -> 4 on one wave: lane j owns hidden unit j.
__device__ inline void mlp_forward_wave(const AddsPolicy* p, float feat, unsigned lane, float* out) {
  float h = p->b1[lane];
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) h += bf16_to_f32(p->w1[lane * kMlpIn + i]) * __shfl(feat, i, 64);
  h = fmaxf(h, 0.0f);
#pragma unroll
  for (int k = 0; k < kMlpOut; ++k) {
    float o = bf16_to_f32(p->w2[k * kMlpHidden + lane]) * h;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) o += __shfl_xor(o, off, 64);
    out[k] = o + p->b2[k];
  }
}

}  // namespace mi355

struct mi355_graph;
namespace mi355 {
struct RunRequest;
struct RunResult;
void run_adds(mi355_graph& g, uint32_t source, const RunRequest& req, RunResult& res,
              hipStream_t stream, const AddsPolicy* d_policy);
}  // namespace mi355
