// Bucket-synchronous delta-stepping driven from the device.
//
// Semantics follow DS/src/delta_stepping_hip_CSR.cpp:648-815: buckets of
// width delta indexed by floorf(d * (1/delta)) (:233-239); within the current
// bucket, light edges (w <= delta) are relaxed until no vertex re-enters the
// bucket (:704-726), then heavy edges of every vertex processed in the bucket
// are relaxed once (:729-741); the next bucket is the smallest one holding a
// tentative distance (:743-760).  iterations_used counts processed buckets.
//
// What differs is where the decisions are made.  The reference crosses to the
// host six or more times per bucket; here a one-thread controller kernel
// chooses each step on the device and the host only enqueues a fixed
// [controller, work, chunks] triple per step in batches, syncing once per
// batch.  Future buckets are kept in a window of kSlots per-bucket queues
// (each vertex enters a bucket queue at most once, guarded by an atomicMin of
// the smallest bucket it was queued for) plus a far pile for buckets beyond
// the window, so finding the next bucket never rescans all pending vertices.
#include <algorithm>
#include <cmath>

#include "frontier_kernels.hpp"
#include "sssp.hpp"

namespace mi355 {

constexpr int kSlots = kBins;
constexpr int kSlotStride = 32;  // words between slot counters (one 128-byte line)
constexpr int kNoBucket = 0x3fffffff;

enum DsStep : uint32_t {
  kStepInit = 0,
  kStepLight = 1,
  kStepHeavy = 2,
  kStepFilter = 3,
  kStepRefillMin = 4,
  kStepRefillDist = 5,
  kStepDone = 6
};

struct BucketCtrl {
  // decided by the controller for the next work kernel
  uint32_t kind;
  int32_t bucket;       // current bucket B
  int32_t target;       // bucket being filtered
  uint32_t in_sel;      // light queue read this step
  uint32_t first;       // first frontier entry (heavy steps)
  uint32_t count;       // frontier entries to process
  uint32_t lstamp;      // light-queue dedup value of this step
  uint32_t far_sel;     // far pile read / written
  // results of work kernels; counters that the whole grid appends to sit on
  // cache lines of their own
  alignas(128) uint32_t light_cnt[2];
  alignas(128) uint32_t settled_cnt;
  uint32_t heavy_done;
  alignas(128) uint32_t chunk_cnt;
  alignas(128) uint32_t far_cnt[2];
  int32_t far_min;
  uint32_t err;
  alignas(128) uint32_t slot_cnt[kSlots * kSlotStride];  // slot s at [s * kSlotStride]
  // bookkeeping
  alignas(128) uint32_t buckets_done;
  uint32_t steps;
  uint32_t done;
  uint32_t converged;
  uint32_t prev_kind;
  unsigned long long edges;
};

#define SLOT_CNT(c, s) ((c)->slot_cnt[(s) * kSlotStride])

struct DsArrays {
  const int* rowptr;
  const uint2* edges;
  unsigned long long* labels;
  int* lq[2];
  int* settled;
  int* slots;
  int* far[2];
  uint32_t* bstamp;
  uint32_t* sstamp;
  uint32_t* lstamp;
  uint32_t* farflag;
  uint4* chunks;
  uint32_t chunk_cap;
  uint32_t cap;        // n
  int64_t slot_cap;    // entries per slot
  float inv_delta;
  uint32_t delta_bits;
  int want_work;
  const int* split;    // light edges per vertex (rows are sorted by weight)
  const int* ids;      // input id of every internal vertex (nullptr: the same)
};

__device__ inline int bucket_of(float d, float inv_delta) {
  // DS/src/delta_stepping_hip_CSR.cpp:233-239
  if (!isfinite(d)) return kNoBucket;
  const float b = floorf(d * inv_delta);
  if (b < 0.0f) return 0;
  if (b >= (float)kNoBucket) return kNoBucket - 1;
  return (int)b;
}

template <class W>
__device__ inline int bucket_of_key(uint32_t key, float inv_delta) {
  if (key == W::kInfKey) return kNoBucket;
  return bucket_of(W::bucket_value(key), inv_delta);
}

// LDS backing of one wave's BinStage.
struct DsSlotLds {
  int v[kStageCap];
  uint16_t bin[kStageCap];
  uint16_t off[kStageCap];
  uint32_t hist[kBins];
  __device__ BinStage stage() { return BinStage{v, bin, off, hist, 0u}; }
};

template <class W>
struct DsSink {
  const DsArrays* a;
  BucketCtrl* c;
  int B;
  uint32_t lstamp_val;
  int* lq_out;
  uint32_t* lq_cnt;
  bool record_settled;  // light steps: processed vertices join the settled list
  bool heavy;           // heavy steps: a same-bucket straggler must rejoin it
  uint32_t far_out;

  int* stage_light = nullptr;    // LDS per wave
  int* stage_settled = nullptr;  // LDS per wave
  uint32_t n_light = 0, n_settled = 0;
  BinStage slot_stage{};         // LDS per wave: future-bucket appends

  __device__ inline void vertex(bool valid, int u, uint32_t) {
    bool push = false;
    if (record_settled && valid) {
      const uint32_t tag = (uint32_t)B + 1u;
      if (a->sstamp[u] != tag) push = atomicExch(a->sstamp + u, tag) != tag;
    }
    stage_push(push, u, stage_settled, n_settled, a->settled, &c->settled_cnt, a->cap, &c->err);
  }

  __device__ inline void improved(bool imp, uint32_t v, uint32_t key) {
    bool to_light = false, to_slot = false;
    int b = 0;
    if (imp) {
      b = bucket_of_key<W>(key, a->inv_delta);
      if (b <= B) {
        // floorf(fl(d + w) / delta) can round back into bucket B for a heavy
        // edge; such a vertex is relaxed again and its heavy edges redone.
        if (heavy) a->sstamp[v] = 0u;
        if (a->lstamp[v] != lstamp_val) to_light = atomicExch(a->lstamp + v, lstamp_val) != lstamp_val;
      } else if (b < B + kSlots) {
        // bstamp starts at 0xFFFFFFFF: compare unsigned
        to_slot = atomicMin(a->bstamp + v, (uint32_t)b) > (uint32_t)b;
      } else if (b < kNoBucket) {
        if (atomicExch(a->farflag + v, 1u) == 0u) {
          const uint32_t pos = atomicAdd(&c->far_cnt[far_out], 1u);
          if (pos < a->cap) a->far[far_out][pos] = (int)v;
          else atomicOr(&c->err, 4u);
        }
      }
    }
    bin_push(to_slot, (int)v, (uint32_t)(b % kSlots), slot_stage, c->slot_cnt, kSlotStride, a->slots,
             a->slot_cap, a->slot_cap, &c->err);
    stage_push(to_light, (int)v, stage_light, n_light, lq_out, lq_cnt, a->cap, &c->err);
  }

  __device__ inline void flush() {
    stage_flush(stage_light, n_light, lq_out, lq_cnt, a->cap, &c->err);
    stage_flush(stage_settled, n_settled, a->settled, &c->settled_cnt, a->cap, &c->err);
    bin_flush(slot_stage, c->slot_cnt, kSlotStride, a->slots, a->slot_cap, a->slot_cap, &c->err);
  }
};

// ---- controller (one thread) ------------------------------------------------
__device__ inline void ds_advance(BucketCtrl* c) {
  // next bucket with a non-empty window slot after B
  for (int b = c->bucket + 1; b < c->bucket + kSlots; ++b) {
    if (SLOT_CNT(c, b % kSlots) > 0) {
      c->kind = kStepFilter;
      c->target = b;
      c->light_cnt[0] = 0;
      c->in_sel = 1;  // filter writes light queue 0
      return;
    }
  }
  if (c->far_cnt[c->far_sel] > 0) {
    c->kind = kStepRefillMin;
    c->far_min = kNoBucket;
    return;
  }
  c->kind = kStepDone;
  c->done = 1;
  c->converged = 1;
}

__device__ inline void ds_start_light(BucketCtrl* c, uint32_t in_sel) {
  c->kind = kStepLight;
  c->in_sel = in_sel;
  c->first = 0;
  c->count = c->light_cnt[in_sel];
  c->light_cnt[in_sel ^ 1u] = 0;
  c->lstamp += 1;
}

__device__ inline void ds_bucket_complete(BucketCtrl* c, int64_t max_buckets) {
  c->buckets_done += 1;
  (void)max_buckets;
  ds_advance(c);
}

__global__ void k_ds_control(BucketCtrl* __restrict__ c, int64_t max_buckets) {
  if (threadIdx.x != 0 || c->done) return;
  c->chunk_cnt = 0;
  const uint32_t prev = c->kind;
  c->steps += 1;
  switch (prev) {
    case kStepInit:
      c->bucket = 0;
      c->light_cnt[0] = 1;  // source placed by the init kernel
      c->settled_cnt = 0;
      c->heavy_done = 0;
      if (max_buckets == 0) {
        c->kind = kStepDone;
        c->done = 1;
        c->converged = 0;
        return;
      }
      ds_start_light(c, 0);
      return;
    case kStepLight: {
      const uint32_t out = c->in_sel ^ 1u;
      if (c->light_cnt[out] > 0) {
        ds_start_light(c, out);
      } else if (c->settled_cnt > c->heavy_done) {
        c->kind = kStepHeavy;
        c->first = c->heavy_done;
        c->count = c->settled_cnt;
        c->heavy_done = c->settled_cnt;
        c->in_sel = out;  // heavy writes same-bucket stragglers to light queue `out ^ 1`
        c->light_cnt[out ^ 1u] = 0;
        c->lstamp += 1;
      } else {
        ds_bucket_complete(c, max_buckets);
      }
      return;
    }
    case kStepHeavy: {
      const uint32_t out = c->in_sel ^ 1u;
      if (c->light_cnt[out] > 0) ds_start_light(c, out);
      else ds_bucket_complete(c, max_buckets);
      return;
    }
    case kStepFilter: {
      SLOT_CNT(c, c->target % kSlots) = 0;
      c->bucket = c->target;
      if (c->light_cnt[0] > 0) {
        if ((int64_t)c->buckets_done >= max_buckets && max_buckets >= 0) {
          c->kind = kStepDone;
          c->done = 1;
          c->converged = 0;
          return;
        }
        c->settled_cnt = 0;
        c->heavy_done = 0;
        ds_start_light(c, 0);
      } else {
        ds_advance(c);
      }
      return;
    }
    case kStepRefillMin:
      if (c->far_min >= kNoBucket) {
        c->far_cnt[c->far_sel] = 0;
        c->kind = kStepDone;
        c->done = 1;
        c->converged = 1;
      } else {
        c->bucket = c->far_min - 1;
        c->kind = kStepRefillDist;
        c->far_cnt[c->far_sel ^ 1u] = 0;
      }
      return;
    case kStepRefillDist:
      c->far_cnt[c->far_sel] = 0;
      c->far_sel ^= 1u;
      ds_advance(c);
      return;
    default:
      return;
  }
}

template <class W>
__global__ void k_ds_init(DsArrays a, BucketCtrl* __restrict__ c, int64_t n, uint32_t src,
                          const uint32_t* __restrict__ init) {
  const unsigned long long inf_label = ((unsigned long long)W::kInfKey << 32) | kNoPred;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    unsigned long long l = inf_label;
    if (init) l = ((unsigned long long)W::output_to_key(init[a.ids ? a.ids[v] : v]) << 32) | kNoPred;
    a.labels[v] = (v == (int64_t)src) ? (((unsigned long long)W::kZeroKey << 32) | kNoPred) : l;
    a.bstamp[v] = 0xFFFFFFFFu;
    a.sstamp[v] = 0u;
    a.lstamp[v] = 0u;
    a.farflag[v] = 0u;
  }
  if (blockIdx.x == 0) {
    uint32_t* words = reinterpret_cast<uint32_t*>(c);
    for (uint32_t i = threadIdx.x; i < sizeof(BucketCtrl) / 4; i += blockDim.x) words[i] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) {
      a.lq[0][0] = (int)src;
      c->kind = kStepInit;
    }
  }
}

template <class W>
__global__ __launch_bounds__(kBlock) void k_ds_work(DsArrays a, BucketCtrl* __restrict__ c) {
  __shared__ int s_stage_l[kWavesPerBlock][kStageCap];
  __shared__ int s_stage_s[kWavesPerBlock][kStageCap];
  __shared__ DsSlotLds s_slot[kWavesPerBlock];
  const uint32_t kind = c->kind;
  if (kind == kStepDone || kind == kStepInit) return;
  const int B = c->bucket;
  const unsigned wv = threadIdx.x / kWave;
  unsigned long long* work = a.want_work ? &c->edges : nullptr;
  if (kind == kStepLight || kind == kStepHeavy) {
    const uint32_t in_sel = c->in_sel;
    const uint32_t out_sel = in_sel ^ 1u;
    DsSink<W> sink{&a, c, B, c->lstamp, a.lq[out_sel], &c->light_cnt[out_sel], kind == kStepLight,
                   kind == kStepHeavy, c->far_sel};
    sink.stage_light = s_stage_l[wv];
    sink.stage_settled = s_stage_s[wv];
    sink.slot_stage = s_slot[wv].stage();
    ChunkQueue cq{a.chunks, &c->chunk_cnt, a.chunk_cap};
    const int* qin = kind == kStepLight ? a.lq[in_sel] : a.settled;
    relax_frontier_body<W, false>(a.rowptr, a.edges, a.labels, qin, nullptr, c->first, c->count, cq,
                                  kind == kStepLight ? kLightEdges : kHeavyEdges, a.delta_bits, sink,
                                  work, a.ids, a.split);
    sink.flush();
    return;
  }
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  if (kind == kStepFilter) {
    const int b = c->target;
    const int s = b % kSlots;
    const uint32_t cnt = SLOT_CNT(c, s);
    const int* slot = a.slots + (int64_t)s * a.slot_cap;
    const uint32_t rounds = div_up(cnt, stride) * stride;
    uint32_t staged = 0;
    for (uint32_t i = tid; i < rounds; i += stride) {
      bool push = false;
      int v = 0;
      if (i < cnt) {
        v = slot[i];
        push = bucket_of_key<W>(atomic_load_key(a.labels + v), a.inv_delta) == b;
      }
      stage_push(push, v, s_stage_l[wv], staged, a.lq[0], &c->light_cnt[0], a.cap, &c->err);
    }
    stage_flush(s_stage_l[wv], staged, a.lq[0], &c->light_cnt[0], a.cap, &c->err);
    return;
  }
  if (kind == kStepRefillMin) {
    const uint32_t sel = c->far_sel;
    const uint32_t cnt = c->far_cnt[sel];
    int local = kNoBucket;
    for (uint32_t i = tid; i < cnt; i += stride) {
      const int v = a.far[sel][i];
      const int b = bucket_of_key<W>(atomic_load_key(a.labels + v), a.inv_delta);
      if (b > B && b < local) local = b;
    }
    if (local < kNoBucket) atomicMin(&c->far_min, local);
    return;
  }
  if (kind == kStepRefillDist) {
    const uint32_t sel = c->far_sel;
    const uint32_t cnt = c->far_cnt[sel];
    for (uint32_t i = tid; i < cnt; i += stride) {
      const int v = a.far[sel][i];
      const int b = bucket_of_key<W>(atomic_load_key(a.labels + v), a.inv_delta);
      if (b <= B) {
        a.farflag[v] = 0u;
      } else if (b < B + kSlots) {
        a.farflag[v] = 0u;
        if (atomicMin(a.bstamp + v, (uint32_t)b) > (uint32_t)b) {
          const int s2 = b % kSlots;
          const uint32_t pos = atomicAdd(&SLOT_CNT(c, s2), 1u);
          if ((int64_t)pos < a.slot_cap) a.slots[(int64_t)s2 * a.slot_cap + pos] = v;
          else atomicOr(&c->err, 2u);
        }
      } else {
        const uint32_t pos = atomicAdd(&c->far_cnt[sel ^ 1u], 1u);
        if (pos < a.cap) a.far[sel ^ 1u][pos] = v;
        else atomicOr(&c->err, 4u);
      }
    }
  }
}

template <class W>
__global__ __launch_bounds__(kBlock) void k_ds_chunks(DsArrays a, BucketCtrl* __restrict__ c) {
  const uint32_t kind = c->kind;
  if (kind != kStepLight && kind != kStepHeavy) return;
  const uint32_t in_sel = c->in_sel;
  const uint32_t out_sel = in_sel ^ 1u;
  __shared__ int s_stage_l[kWavesPerBlock][kStageCap];
  __shared__ int s_stage_s[kWavesPerBlock][kStageCap];
  __shared__ DsSlotLds s_slot[kWavesPerBlock];
  DsSink<W> sink{&a, c, c->bucket, c->lstamp, a.lq[out_sel], &c->light_cnt[out_sel], false,
                 kind == kStepHeavy, c->far_sel};
  sink.stage_light = s_stage_l[threadIdx.x / kWave];
  sink.stage_settled = s_stage_s[threadIdx.x / kWave];
  sink.slot_stage = s_slot[threadIdx.x / kWave].stage();
  ChunkQueue cq{a.chunks, &c->chunk_cnt, a.chunk_cap};
  relax_chunks_body<W>(a.edges, a.labels, cq, kind == kStepLight ? kLightEdges : kHeavyEdges,
                       a.delta_bits, sink, a.want_work ? &c->edges : nullptr);
  sink.flush();
}

template <class W>
void run_ds_typed(mi355_graph& g, uint32_t src, const RunRequest& req, RunResult& res,
                  hipStream_t s) {
  Workspace& ws = g.ws;
  const int64_t n = g.n;
  float delta = req.delta > 0.0f ? req.delta : heuristic_delta(g);
  if (!(delta > 0.0f) || !std::isfinite(delta)) fail(MI355_EINVAL, "delta must be a finite positive float");
  ws.labels.ensure((size_t)n);
  ws.q[0].ensure((size_t)n);
  ws.q[1].ensure((size_t)n);
  ws.settled.ensure((size_t)n);
  ws.slot_cap = n;
  ws.slots.ensure((size_t)kSlots * (size_t)n);
  ws.far.ensure((size_t)n * 2);
  ws.bstamp.ensure((size_t)n);
  ws.sstamp.ensure((size_t)n * 3);  // sstamp | lstamp | farflag
  const size_t lcap = (size_t)std::max<int64_t>(g.info.large_chunks, 1);
  ws.lq.ensure(lcap);
  ws.bctrl_mem.ensure(sizeof(BucketCtrl));
  BucketCtrl* c = reinterpret_cast<BucketCtrl*>(ws.bctrl_mem.p);

  DsArrays a{};
  a.rowptr = g.rowptr.p;
  a.edges = g.edges.p;
  a.labels = ws.labels.p;
  a.lq[0] = ws.q[0].p;
  a.lq[1] = ws.q[1].p;
  a.settled = ws.settled.p;
  a.slots = ws.slots.p;
  a.far[0] = ws.far.p;
  a.far[1] = ws.far.p + n;
  a.bstamp = ws.bstamp.p;
  a.sstamp = ws.sstamp.p;
  a.lstamp = ws.sstamp.p + n;
  a.farflag = ws.sstamp.p + 2 * n;
  a.chunks = ws.lq.p;
  a.chunk_cap = (uint32_t)lcap;
  a.cap = (uint32_t)n;
  a.slot_cap = n;
  a.inv_delta = 1.0f / delta;
  a.delta_bits = __builtin_bit_cast(uint32_t, delta);
  a.want_work = req.want_work ? 1 : 0;
  a.split = light_split(g, delta, s);
  a.ids = input_ids(g);

  const int64_t max_buckets = req.max_iters < 0 ? (int64_t)INT64_MAX : req.max_iters;

  hipEvent_t ev0, ev1;
  MI_HIP(hipEventCreate(&ev0));
  MI_HIP(hipEventCreate(&ev1));
  MI_HIP(hipEventRecord(ev0, s));
  k_ds_init<W><<<std::min<unsigned>(div_up((uint64_t)n, kBlock), 4096u), kBlock, 0, s>>>(
      a, c, n, src, req.init_dist);
  MI_LAUNCH_CHECK();
  const unsigned grid = (unsigned)std::max(64, device_props().cus * 8);
  int batch = req.rounds_per_sync > 0 ? req.rounds_per_sync : 4;
  BucketCtrl h{};
  res.syncs = 0;
  int64_t steps = 0;
  for (;;) {
    for (int i = 0; i < batch; ++i, ++steps) {
      k_ds_control<<<1, kWave, 0, s>>>(c, max_buckets);
      k_ds_work<W><<<grid, kBlock, 0, s>>>(a, c);
      if (g.info.large_vertices > 0) k_ds_chunks<W><<<grid, kBlock, 0, s>>>(a, c);
    }
    MI_LAUNCH_CHECK();
    MI_HIP(hipMemcpyAsync(&h, c, sizeof(BucketCtrl), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    ++res.syncs;
    if (h.err) fail(MI355_ECAPACITY, "delta-stepping queue overflow");
    if (h.done) break;
    // every step settles or moves work; this bound is never reached by a run
    // that makes progress
    if (steps > 64 * (n + 1024)) fail(MI355_EDEVICE, "delta-stepping made no progress");
    if (req.rounds_per_sync <= 0) batch = std::min(batch * 2, 256);
  }
  MI_HIP(hipEventRecord(ev1, s));
  MI_HIP(hipEventSynchronize(ev1));
  MI_HIP(hipEventElapsedTime(&res.device_ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  res.iterations_used = h.buckets_done;
  res.converged = h.converged != 0;
  res.rounds = h.steps;
  res.edges = (int64_t)h.edges;
  res.initial_delta = res.final_delta = delta;
}

namespace {
void light_count(const int* rowptr, const uint2* edges, int64_t n, int32_t wtype, uint32_t bits, int* out,
                 hipStream_t s) {
  const unsigned grid = std::min<unsigned>(div_up((uint64_t)n, kBlock), 8192u);
  if (wtype == MI355_W_F32) k_light_count<WF32><<<grid, kBlock, 0, s>>>(rowptr, edges, n, bits, out);
  else k_light_count<WU32><<<grid, kBlock, 0, s>>>(rowptr, edges, n, bits, out);
  MI_LAUNCH_CHECK();
}
}  // namespace

const int* light_split(mi355_graph& g, float delta, hipStream_t s) {
  Workspace& ws = g.ws;
  const uint32_t bits = __builtin_bit_cast(uint32_t, delta);
  if (!ws.split_valid || ws.split_delta_bits != bits) {
    ws.split.ensure((size_t)g.n);
    light_count(g.rowptr.p, g.edges.p, g.n, g.weight_type, bits, ws.split.p, s);
    ws.split_valid = true;
    ws.split_delta_bits = bits;
  }
  return ws.split.p;
}

const int* pull_light_split(mi355_graph& g, float delta, hipStream_t s) {
  if (g.pull_state != 2) return light_split(g, delta, s);
  const uint32_t bits = __builtin_bit_cast(uint32_t, delta);
  if (!g.t_split_valid || g.t_split_bits != bits) {
    g.t_split.ensure((size_t)g.n);
    light_count(g.t_rowptr.p, g.t_edges.p, g.n, g.weight_type, bits, g.t_split.p, s);
    g.t_split_valid = true;
    g.t_split_bits = bits;
  }
  return g.t_split.p;
}

void run_delta_stepping(mi355_graph& g, uint32_t source, const RunRequest& req, RunResult& res,
                        hipStream_t stream) {
  if (g.info.has_negative) fail(MI355_EINVAL, "delta-stepping requires nonnegative edge weights");
  if (g.weight_type == MI355_W_F32) run_ds_typed<WF32>(g, source, req, res, stream);
  else run_ds_typed<WU32>(g, source, req, res, stream);
}

}  // namespace mi355
