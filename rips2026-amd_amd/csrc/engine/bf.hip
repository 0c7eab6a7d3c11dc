// Frontier Bellman-Ford on a graph handle, and the label <-> output kernels.
//
// Rounds are pipelined: the host enqueues a batch of rounds without waiting,
// every kernel reads its frontier size from device memory (an empty frontier
// makes a round a no-op), and queue counters rotate over three slots so a
// round can clear the counter two rounds ahead.  The host synchronises once
// per batch; the batch doubles while the frontier stays non-empty.
//
// Jacobi mode reproduces the reference's iteration semantics exactly
// (BF/src/bf_hip_CSR.cpp:60-95: d_k = min(d_{k-1}, A (x) d_{k-1})): the
// frontier of round k carries the keys its vertices had at the end of round
// k-1 (snapshot kernel), so every candidate of round k is computed from d_{k-1}
// while the atomicMin target accumulates d_k.  Vertices absent from the
// frontier did not change in round k-1, so their candidates are already
// folded in; the frontier after round k is exactly {v : d_k[v] < d_{k-1}[v]}.
#include <algorithm>

#include "frontier_kernels.hpp"
#include "sssp.hpp"

namespace mi355 {

__global__ void k_snapshot_keys(const int* __restrict__ q, const uint32_t* __restrict__ cnt,
                                const unsigned long long* __restrict__ labels,
                                uint32_t* __restrict__ keys) {
  const uint32_t n = *cnt;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    keys[i] = atomic_load_key(labels + q[i]);
}

namespace {

__global__ void k_init_bf(unsigned long long* __restrict__ labels, int64_t n, uint32_t src,
                          uint32_t inf_key, uint32_t zero_key, int* __restrict__ q0,
                          uint32_t* __restrict__ k0, RoundCtrl* __restrict__ ctrl) {
  const unsigned long long inf_label = ((unsigned long long)inf_key << 32) | kNoPred;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
    labels[v] = (v == (int64_t)src) ? (((unsigned long long)zero_key << 32) | kNoPred) : inf_label;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    q0[0] = (int)src;
    k0[0] = zero_key;
    ctrl->cnt[0] = 1;
    ctrl->cnt[1] = 0;
    ctrl->cnt[2] = 0;
    ctrl->lcnt[0] = 0;
    ctrl->lcnt[1] = 0;
    ctrl->err = 0;
    ctrl->edges = 0;
    ctrl->vertices = 0;
  }
}

// Output element x (input numbering) comes from internal vertex inv[x]
// (inv == nullptr: the same id); parents are stored as input ids.  src is the
// internal id of the source.  The reached counts walk the internal ids in
// order (label and row offsets of internal vertex x: coalesced reads; every
// vertex is counted once either way), not inv[x] (two more random lines per
// vertex).
template <class W>
__global__ void k_labels_to_output(const unsigned long long* __restrict__ labels, int64_t n,
                                   uint32_t src, int inf_mode, uint32_t* __restrict__ dist,
                                   int32_t* __restrict__ pred, const int* __restrict__ rowptr,
                                   const int* __restrict__ inv, unsigned long long* __restrict__ reach) {
  __shared__ unsigned long long s_r[kBlock], s_e[kBlock];
  unsigned long long r = 0, e = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
    const int64_t v = inv ? (int64_t)inv[x] : x;
    const unsigned long long l = labels[v];
    const uint32_t key = (uint32_t)(l >> 32);
    const bool reached = key != W::kInfKey;
    uint32_t out;
    if (W::kType == MI355_W_F32) {
      if (!reached) out = inf_mode == MI355_INF_FLT_MAX ? 0x7F7FFFFFu : 0x7F800000u;
      else out = __float_as_uint(key_to_f32(key));
    } else {
      out = key;
    }
    if (dist) dist[x] = out;
    if (pred) {
      const uint32_t p = (uint32_t)l;
      pred[x] = (v == (int64_t)src || !reached || p == kNoPred) ? -1 : (int32_t)p;
    }
    if (reach) {
      const bool rx = inv ? (uint32_t)(labels[x] >> 32) != W::kInfKey : reached;
      if (rx) {
        ++r;
        e += (unsigned long long)(rowptr[x + 1] - rowptr[x]);
      }
    }
  }
  if (!reach) return;
  s_r[threadIdx.x] = r;
  s_e[threadIdx.x] = e;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s_r[threadIdx.x] += s_r[threadIdx.x + s];
      s_e[threadIdx.x] += s_e[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(reach, s_r[0]);
    atomicAdd(reach + 1, s_e[0]);
  }
}

unsigned relax_grid() {
  static const unsigned g = [] {
    const DeviceProps p = device_props();
    return (unsigned)std::max(64, p.cus * 8);
  }();
  return g;
}

template <class W>
void run_bf_typed(mi355_graph& g, uint32_t src, const RunRequest& req, RunResult& res,
                  hipStream_t s, bool resume) {
  Workspace& ws = g.ws;
  const int64_t n = g.n;
  ws.labels.ensure((size_t)n);
  ws.stamp.ensure((size_t)n);
  if (ws.stamp_epoch == 0) MI_HIP(hipMemsetAsync(ws.stamp.p, 0, (size_t)n * 4, s));
  ws.q[0].ensure((size_t)n);
  ws.q[1].ensure((size_t)n);
  const bool jacobi = req.algorithm == MI355_ALG_BF_JACOBI;
  if (jacobi) {
    ws.qkey[0].ensure((size_t)n);
    ws.qkey[1].ensure((size_t)n);
  } else {
    ws.qkey[0].ensure(1);
  }
  const size_t lcap = (size_t)std::max<int64_t>(g.info.large_chunks, 1);
  ws.lq.ensure(lcap);
  ws.ctrl.ensure(1);
  int64_t max_rounds = req.max_iters < 0 ? std::max<int64_t>(n - 1, 0) : req.max_iters;
  // The log keeps |F_r| for every round so Jacobi iteration counts are exact.
  ws.log.ensure((size_t)std::min<int64_t>(max_rounds, n + 1) + 2);

  hipEvent_t ev0, ev1;
  MI_HIP(hipEventCreate(&ev0));
  MI_HIP(hipEventCreate(&ev1));
  MI_HIP(hipEventRecord(ev0, s));

  // resume: continue from the frontier the previous run on this handle left
  // (used for the negative-cycle probe round of bellman_ford_sssp_hip).
  int64_t r = resume ? ws.bf_rounds_done : 0;
  if (resume) max_rounds += r;
  if (!resume) {
    k_init_bf<<<std::min<unsigned>(div_up((uint64_t)n, kBlock), 4096u), kBlock, 0, s>>>(
        ws.labels.p, n, src, W::kInfKey, W::kZeroKey, ws.q[0].p, ws.qkey[0].p, ws.ctrl.p);
    MI_LAUNCH_CHECK();
  }

  RoundCtrl* c = ws.ctrl.p;
  const unsigned grid = relax_grid();
  unsigned long long* work = req.want_work ? &c->edges : nullptr;
  int batch = req.rounds_per_sync > 0 ? req.rounds_per_sync : 1;
  uint32_t frontier = 1;
  res.syncs = 0;
  // in-place rounds relax chains of improvements inside the round
  // (k_relax_frontier_chain); Jacobi rounds keep their iteration semantics.
  // Only graphs without chunked rows: the chunk queue holds each high-degree
  // row once per round, and a chain could chunk a row again.
  const bool chain = !jacobi && g.info.large_vertices == 0;
  while (r < max_rounds && frontier > 0) {
    const int64_t k = std::min<int64_t>(batch, max_rounds - r);
    for (int64_t i = 0; i < k; ++i, ++r) {
      // Round r: F_r = q[r&1] (size cnt[r%3]) -> F_{r+1} = q[(r+1)&1].
      if (ws.stamp_epoch == 0xFFFFFFFFu) {
        MI_HIP(hipMemsetAsync(ws.stamp.p, 0, (size_t)n * 4, s));
        ws.stamp_epoch = 0;
      }
      const uint32_t stamp_val = ++ws.stamp_epoch;
      const int qi = (int)(r & 1), qo = (int)((r + 1) & 1);
      FrontierIn in{ws.q[qi].p, jacobi ? ws.qkey[qi].p : nullptr, &c->cnt[r % 3]};
      ChunkQueue cq{ws.lq.p, &c->lcnt[r & 1], (uint32_t)lcap};
      BfSink sink{ws.q[qo].p, &c->cnt[(r + 1) % 3], (uint32_t)n, ws.stamp.p, stamp_val, &c->err};
      RoundHousekeeping hk{&c->cnt[(r + 2) % 3], &c->lcnt[(r + 1) & 1],
                           r < (int64_t)ws.log.n ? ws.log.p + r : nullptr};
      if (jacobi)
        k_relax_frontier<W, true, BfSink>
            <<<grid, kBlock, 0, s>>>(g.rowptr.p, g.edges.p, ws.labels.p, in, cq, sink, hk, work,
                                     input_ids(g));
      else if (chain)
        k_relax_frontier_chain<W>
            <<<grid, kBlock, 0, s>>>(g.rowptr.p, g.edges.p, ws.labels.p, in, cq, sink, hk, work,
                                     input_ids(g));
      else
        k_relax_frontier<W, false, BfSink>
            <<<grid, kBlock, 0, s>>>(g.rowptr.p, g.edges.p, ws.labels.p, in, cq, sink, hk, work,
                                     input_ids(g));
      MI_LAUNCH_CHECK();
      if (g.info.large_vertices > 0) {
        k_relax_chunks<W, BfSink>
            <<<grid, kBlock, 0, s>>>(g.edges.p, ws.labels.p, cq, sink, work);
        MI_LAUNCH_CHECK();
      }
      if (jacobi) {
        k_snapshot_keys<<<std::min<unsigned>(div_up((uint64_t)n, kBlock), 2048u), kBlock, 0, s>>>(
            ws.q[qo].p, &c->cnt[(r + 1) % 3], ws.labels.p, ws.qkey[qo].p);
        MI_LAUNCH_CHECK();
      }
    }
    RoundCtrl h;
    MI_HIP(hipMemcpyAsync(&h, c, sizeof(RoundCtrl), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    ++res.syncs;
    if (h.err) fail(MI355_ECAPACITY, "frontier queue overflow");
    frontier = h.cnt[r % 3];
    if (req.rounds_per_sync <= 0) batch = std::min(batch * 2, 64);
  }
  MI_HIP(hipEventRecord(ev1, s));
  MI_HIP(hipEventSynchronize(ev1));
  MI_HIP(hipEventElapsedTime(&res.device_ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);

  ws.bf_rounds_done = r;
  // Frontier log: |F_0..F_{r-1}| from the kernels, |F_r| from the counter.
  const int64_t logged = std::min<int64_t>(r, (int64_t)ws.log.n);
  res.frontier_log.assign((size_t)logged, 0);
  if (logged > 0)
    MI_HIP(hipMemcpyAsync(res.frontier_log.data(), ws.log.p, (size_t)logged * 4,
                          hipMemcpyDeviceToHost, s));
  RoundCtrl h;
  MI_HIP(hipMemcpyAsync(&h, c, sizeof(RoundCtrl), hipMemcpyDeviceToHost, s));
  MI_HIP(hipStreamSynchronize(s));
  res.frontier_log.push_back(h.cnt[r % 3]);
  res.rounds = r;
  res.edges = (int64_t)h.edges;
  // iterations_used / converged under "check every iteration": the first
  // round k (1-based) whose output frontier F_k is empty.
  res.converged = false;
  res.iterations_used = r;
  for (int64_t k = 1; k <= r; ++k) {
    if (res.frontier_log[(size_t)k] == 0) {
      res.iterations_used = k;
      res.converged = true;
      break;
    }
  }
  res.negative_cycle = false;
}

}  // namespace

float heuristic_delta(const mi355_graph& g) {
  const double d = 16.0 * g.info.avg_weight / std::max(1.0, g.info.avg_degree);
  return (float)std::max(1.0e-6, d);
}

void run_sssp(mi355_graph& g, uint32_t input_source, const RunRequest& req, RunResult& res,
              hipStream_t stream) {
  if ((int64_t)input_source >= g.n) fail(MI355_ERANGE, "source vertex is outside the graph");
  const uint32_t source = internal_vertex(g, input_source);
  switch (req.algorithm) {
    case MI355_ALG_BF:
    case MI355_ALG_BF_JACOBI:
      if (g.weight_type == MI355_W_F32) run_bf_typed<WF32>(g, source, req, res, stream, false);
      else run_bf_typed<WU32>(g, source, req, res, stream, false);
      if (g.log_dst) {
        // the caller's round log (mi355_sssp_round_log): |F_0| .. |F_rounds|,
        // for this run only
        const size_t k = std::min(res.frontier_log.size(), (size_t)std::max<int32_t>(g.log_cap, 0));
        for (size_t i = 0; i < k; ++i) g.log_dst[i] = (int32_t)res.frontier_log[i];
        g.log_dst = nullptr;
        g.log_cap = 0;
      }
      break;
    case MI355_ALG_DELTA:
      run_delta_stepping(g, source, req, res, stream);
      break;
    case MI355_ALG_ADDS: {
      int mode = kPolicyFixed;
      const AddsPolicy* pol = prepare_adds_policy(g, req, stream, &mode);
      RunRequest r = req;
      r.use_mlp = mode == kPolicyMlp || mode == kPolicySchedule;  // the policy sets the first phase
      run_adds(g, source, r, res, stream, pol);
      res.policy = mode;
      break;
    }
    default:
      fail(MI355_EINVAL, "unknown SSSP algorithm");
  }
}

void run_bf_continue(mi355_graph& g, uint32_t input_source, const RunRequest& req, RunResult& res,
                     hipStream_t stream) {
  const uint32_t source = internal_vertex(g, input_source);
  if (g.weight_type == MI355_W_F32) run_bf_typed<WF32>(g, source, req, res, stream, true);
  else run_bf_typed<WU32>(g, source, req, res, stream, true);
}

void labels_to_output(mi355_graph& g, uint32_t source, int inf_mode, void* d_dist, int32_t* d_pred,
                      unsigned long long* reach, hipStream_t s) {
  labels_to_output_internal(g, internal_vertex(g, source), inf_mode, d_dist, d_pred, reach, s);
}

void labels_to_output_internal(mi355_graph& g, uint32_t src, int inf_mode, void* d_dist, int32_t* d_pred,
                               unsigned long long* reach, hipStream_t s) {
  const unsigned grid = std::min<unsigned>(div_up((uint64_t)g.n, kBlock), 4096u);
  const int* inv = g.relabeled ? g.inv.p : nullptr;
  if (g.weight_type == MI355_W_F32)
    k_labels_to_output<WF32><<<grid, kBlock, 0, s>>>(g.ws.labels.p, g.n, src, inf_mode,
                                                     static_cast<uint32_t*>(d_dist), d_pred,
                                                     g.rowptr.p, inv, reach);
  else
    k_labels_to_output<WU32><<<grid, kBlock, 0, s>>>(g.ws.labels.p, g.n, src, inf_mode,
                                                     static_cast<uint32_t*>(d_dist), d_pred,
                                                     g.rowptr.p, inv, reach);
  MI_LAUNCH_CHECK();
}

}  // namespace mi355
