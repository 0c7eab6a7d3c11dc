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
  Workspace() = default;
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  ~Workspace() {
    for (hipEvent_t& e : adds_ev)
      if (e) (void)hipEventDestroy(e);
  }
  DevBuf<unsigned long long> labels;  // n packed labels
  DevBuf<unsigned long long> reach;   // reached vertices / edges of the last output conversion
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
  // light splits at the delta network's first width, chosen on the device
  // (adds.hip): out-rows, transposed in-rows, and their selection words
  DevBuf<int> mlp_split, mlp_tsplit;
  DevBuf<uint32_t> mlp_sel;
  bool mlp_sel_ready = false;
  DevBuf<AddsPolicy> adds_policy;
  std::vector<unsigned char> adds_policy_host;  // what adds_policy holds (skip identical uploads)
  std::vector<unsigned char> adds_ctrl_host;    // host image of the control block
  hipEvent_t adds_ev[3] = {nullptr, nullptr, nullptr};  // run start, kernel start, kernel end
  // batched multi-source queries (batch.hip)
  DevBuf<unsigned long long> batch_labels;
  DevBuf<uint32_t> batch_stamp;   // per slot: round stamp, far stamp (interleaved)
  DevBuf<uint2> batch_q[2];
  DevBuf<uint2> batch_far[2];     // near-far: the two far piles
  DevBuf<unsigned char> batch_qinfo;
  DevBuf<uint2> batch_src;
  DevBuf<unsigned char> batch_ctrl;
  DevBuf<uint2> batch_tgt;        // (query, local slot) of every target
  DevBuf<uint32_t> batch_tstate;  // target settled flags, targets left per query, stop thresholds
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
  // Vertex order.  A graph with a skewed degree distribution is renumbered by
  // descending degree at build time (stable by input id), so the labels,
  // stamps and row offsets of the vertices that most edges point to share a
  // few cache lines; internal vertex v is input vertex perm[v], input vertex
  // x is internal inv[x].  Labels store input ids as parents (so the
  // smallest-parent tie rule refers to input ids) and outputs are written in
  // input order.  Empty perm: internal ids are the input ids.
  bool relabeled = false;
  mi355::DevBuf<int> perm, inv;
  std::vector<int32_t> inv_host;
  // input-order copy of the arrays (built on first use by the paths that
  // read the input numbering's geometry: batched grid queries, download)
  mi355::DevBuf<int> in_rowptr;
  mi355::DevBuf<uint2> in_edges;
  mi355_graph_info info{};
  std::vector<int32_t> out_degree_host;  // kept only when built from host data
  mi355::Workspace ws;
  // optional per-round log request
  int32_t* log_dst = nullptr;
  int32_t log_cap = 0;
  // graph features of the learned delta policy (computed on first use)
  bool graph_feat_ready = false;
  float graph_feat[14] = {};
  // In-edge rows for ADDS pulls, built on first use: row v lists {u, w} for
  // every edge u -> v, weight-sorted.  pull_state 1: the graph is symmetric
  // (every row holds the same (neighbour, weight) multiset in both
  // orientations) and its own rows serve; 2: t_rowptr / t_edges hold the
  // transpose.
  int pull_state = 0;
  // 1 + the last vertex with an in-edge: pulls skip the vertices past it (in
  // degree order, the isolated vertices at the end)
  int64_t pull_n = 0;
  mi355::DevBuf<int> t_rowptr;
  mi355::DevBuf<uint2> t_edges;
  mi355::DevBuf<int> t_split;  // light in-edges per row of the transpose (for t_split_bits)
  uint32_t t_split_bits = 0;
  bool t_split_valid = false;
};

namespace mi355 {
// Vertex order of a graph handle (mi355_graph_create_ex).
enum VertexOrder : int32_t { kOrderAuto = 0, kOrderInput = 1, kOrderDegree = 2 };

// Build a graph handle (graph_build.hip).
// allow_negative: bit 0 admits negative weights, bit 1 admits +inf weights.
mi355_graph* build_graph(const mi355_csr_desc& d, int allow_negative, hipStream_t stream,
                         int32_t order = kOrderAuto);

// The 14 graph features of the learned delta policy (policy.hip), sampled in
// the input numbering (computed once per graph, before any renumbering).
void graph_features(mi355_graph& g, hipStream_t s, float* out14);

// Internal id of input vertex v.
inline uint32_t internal_vertex(const mi355_graph& g, int64_t v) {
  return g.relabeled ? (uint32_t)g.inv_host[(size_t)v] : (uint32_t)v;
}
// Input id of every internal vertex (nullptr when the input order is kept).
inline const int* input_ids(const mi355_graph& g) { return g.relabeled ? g.perm.p : nullptr; }
// Row offsets and edge records in the input numbering (the graph's own
// arrays, or a copy built on first use when the graph is renumbered).
void input_order_arrays(mi355_graph& g, const int** rowptr, const uint2** edges, hipStream_t stream);
// In-edge rows of the internal numbering, weight-sorted (the graph's own rows
// when it is symmetric; a transpose built on first use otherwise).
void pull_rows(mi355_graph& g, const int** rowptr, const uint2** edges, hipStream_t stream);
}  // namespace mi355
