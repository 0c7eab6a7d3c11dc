// Device-resident graph handle and the per-graph SSSP workspace.
#pragma once

#include <vector>

#include "adds.hpp"
#include "common.hpp"

namespace mi355 {

// Control block shared by the pipelined round kernels.  All fields are
// written and read on the device; the host copies it back at sync points.
struct RoundCtrl {
  uint32_t cnt[3];    // frontier sizes, rotating (round r reads cnt[r%3])
  uint32_t lcnt[2];   // large-vertex chunk counts (round r uses lcnt[r&1])
  uint32_t err;       // capacity overflow flag
  uint32_t pad;
  unsigned long long edges;     // edge visits
  unsigned long long vertices;  // frontier vertices processed
};

// Delta-stepping / ADDS device state (see delta.hip).
struct BucketCtrl;

struct Workspace {
  DevBuf<unsigned long long> labels;  // n packed labels
  DevBuf<uint32_t> stamp;             // n round stamps (frontier dedup)
  DevBuf<int> q[2];                   // frontier queues (n each)
  DevBuf<uint32_t> qkey[2];           // Jacobi snapshot keys (n each)
  DevBuf<uint4> lq;                   // large-vertex chunk queue
  DevBuf<RoundCtrl> ctrl;
  DevBuf<uint32_t> log;               // per-round frontier sizes
  uint32_t stamp_epoch = 0;
  int64_t bf_rounds_done = 0;         // rounds of the last Bellman-Ford run
  // delta-stepping state (allocated on first use)
  DevBuf<int> slots;                  // bucket window queues
  DevBuf<uint32_t> bstamp;            // smallest bucket a vertex was queued in
  DevBuf<int> far;                    // (vertex, bucket) pairs beyond the window
  DevBuf<int> settled;                // vertices processed in the current bucket
  DevBuf<uint32_t> sstamp;            // settled-list dedup
  DevBuf<unsigned char> bctrl_mem;    // BucketCtrl
  DevBuf<int> split;                  // light edges per vertex for split_delta_bits
  uint32_t split_delta_bits = 0;
  bool split_valid = false;
  // ADDS state (allocated on first use)
  DevBuf<int> adds_ring;              // kAddsSlots bucket rings, then kAddsSlots heavy-list rings
  bool adds_ring_clean = false;       // rings hold only empty sentinels
  DevBuf<uint32_t> adds_qstamp;       // n queue stamps, then n heavy-list stamps
  DevBuf<HubChunk> adds_hub;
  DevBuf<unsigned char> adds_ctrl_mem;
  DevBuf<AddsPolicy> adds_policy;
  // batched multi-source queries (batch.hip)
  DevBuf<unsigned long long> batch_labels;
  DevBuf<uint32_t> batch_stamp;
  DevBuf<uint32_t> batch_fstamp;  // near-far: far stamp per label slot
  DevBuf<uint2> batch_q[2];
  DevBuf<uint2> batch_far[2];     // near-far: the two far piles
  DevBuf<unsigned char> batch_qinfo;
  DevBuf<uint2> batch_src;
  DevBuf<unsigned char> batch_ctrl;
  int64_t slot_cap = 0;
  int nslots = 0;
};

}  // namespace mi355

struct mi355_graph {
  int device = 0;
  int64_t n = 0;
  int64_t m = 0;
  int32_t weight_type = MI355_W_F32;
  mi355::DevBuf<int> rowptr;   // n+1 int32 offsets
  mi355::DevBuf<uint2> edges;  // m records {dst, weight bits}
  mi355_graph_info info{};
  std::vector<int32_t> out_degree_host;  // kept only when built from host data
  mi355::Workspace ws;
  // optional per-round log request
  int32_t* log_dst = nullptr;
  int32_t log_cap = 0;
  // graph features of the learned delta policy (computed on first use)
  bool graph_feat_ready = false;
  float graph_feat[14] = {};
};

namespace mi355 {
// Build a graph handle (graph_build.hip).
// allow_negative: bit 0 admits negative weights, bit 1 admits +inf weights.
mi355_graph* build_graph(const mi355_csr_desc& d, int allow_negative, hipStream_t stream);
}  // namespace mi355
