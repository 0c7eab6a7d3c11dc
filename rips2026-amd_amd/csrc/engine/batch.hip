// Batched multi-source shortest paths: many routing queries (each a set of
// sources and a grid rectangle that limits the expansion, the "net bounding
// box") solved together as one min-plus SpMM frontier.
//
// Every query q owns a dense block of packed (dist key, parent) labels, one
// per vertex of its rectangle, at labels[base_q ..), laid out in 4 x 4 tiles
// (tile-major, row-major inside a tile; the block is padded to multiples of
// 4): one 128-byte line of labels is a 4 x 4 patch, so a frontier that
// sweeps across the rectangle touches about a quarter of the lines that a
// row-major block's vertical fronts would, and reuses each line for about
// four rounds.  The
// frontier is a list of (query, local index) pairs from all queries at once,
// so one round relaxes every query's frontier in the same launch:
// D_q <- min(D_q, A^T (x) D_q) restricted to the rectangle, for all q.  A
// round packs the edges of 64 entries per wave across lanes (four edges per
// lane in flight), skips out-of-rectangle targets and stages appends in LDS.
//
// Near-far frontier (bucketed label correcting): an improved slot whose new
// key is below the threshold T joins the next round's frontier ("near", at
// most once per round: per-slot round stamp); otherwise it joins the far pile
// (at most once per far epoch: per-slot far stamp).  When a round's input is
// empty the round becomes a far pass instead: T advances by delta (or past
// the smallest far key, skipping empty buckets), far entries below the new T
// become the next frontier, entries above it move to the other far pile, and
// entries below the old T are stale (their slot improved below T since and
// was relaxed as near).  Bellman-Ford rounds are the case T = infinity; on
// weighted routing grids they relax each slot 40-90 times, near-far a few.
//
// Every round is two kernel nodes: a one-thread decision kernel that reads
// the counters the previous round produced and writes this round's
// descriptor (relax / far pass / idle, threshold, pile sizes) and clears this
// round's output counters, then the work kernel, which reads only the
// descriptor and appends.  So no kernel reads a word that a workgroup of the
// same kernel may be changing.  K rounds (K even: the queue parity returns)
// plus a one-thread epoch node form a captured HIP graph (no memset nodes);
// the host replays it and reads the number of entries left once per replay.
// Distances are the exact fixpoint of the fp32 / uint32 left folds, identical
// to a multi-source Dijkstra on each query's induced subgraph.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "fastdiv.hpp"
#include "frontier_kernels.hpp"
#include "sssp.hpp"

namespace mi355 {

namespace {

struct QueryInfo {
  int64_t base;      // first label slot of the query (schedule order, 4 x 4 tiles)
  int64_t out_base;  // first output element of the query (input order)
  int32_t r0, c0;    // rectangle origin (grid row, column)
  int32_t bh, bw;    // rectangle height and width
};

constexpr int kCntStride = 32;  // counters on separate 128-byte lines
// Queries are split into kParts parts of about equal area; part p's frontier
// queues and far piles are segments of the shared arrays and only workgroups
// with blockIdx.x % kParts == p process them.  Workgroups are dealt to the 8
// XCDs round-robin by workgroup id, so a query's frontier stays on one XCD
// and its label lines in that XCD's L2 from round to round.
constexpr int kParts = 8;
__host__ __device__ constexpr int cidx(uint32_t parity, uint32_t part) {
  return (int)((parity * kParts + part) * kCntStride);
}

// Local slot of rectangle cell (dr, dc) in a bh x bw block: 4 x 4 tiles,
// tile-major, row-major inside the tile; a block less than 4 rows high (a
// 1 x n "rectangle" is how a general graph's query spans all its vertices)
// is plain row-major, so it carries no tile padding.
__host__ __device__ inline uint32_t tiled_slot(uint32_t dr, uint32_t dc, uint32_t bh, uint32_t bw) {
  if (bh < 4u) return dr * bw + dc;
  const uint32_t tw = (bw + 3u) >> 2;
  return ((((dr >> 2) * tw) + (dc >> 2)) << 4) | ((dr & 3u) << 2) | (dc & 3u);
}
__host__ __device__ inline void tiled_cell(uint32_t l, uint32_t bh, uint32_t bw, uint32_t& dr, uint32_t& dc) {
  if (bh < 4u) {
    dr = l / bw;
    dc = l - dr * bw;
    return;
  }
  const uint32_t tw = (bw + 3u) >> 2;
  const uint32_t t = l >> 4, tr = t / tw;
  dr = (tr << 2) | ((l >> 2) & 3u);
  dc = ((t - tr * tw) << 2) | (l & 3u);
}
// Slots of a bh x bw block (padded to whole tiles when tiled).
__host__ __device__ inline int64_t tiled_area(int64_t bh, int64_t bw) {
  return bh < 4 ? bh * bw : ((bh + 3) & ~3LL) * ((bw + 3) & ~3LL);
}
constexpr uint32_t kRoundIdle = 0, kRoundRelax = 1, kRoundFar = 2;
constexpr uint32_t kNearStage = 256;  // staged near appends per wave
constexpr uint32_t kFarStage = 128;   // staged far appends per wave
// Wave-local queue of a relax round: a slot that a wave improves below the
// threshold is relaxed again by the same wave later in the round, instead of
// waiting for the next round (its label line is still in this XCD's L2).  Such
// entries take no round stamp, so an improvement that arrives after the wave
// has relaxed the slot is pushed by its own producer as usual.  At most
// kLocalBudget local entries per wave and round; beyond that, and when the
// ring is full, improved slots go to the next round.
constexpr uint32_t kLocalQ = 512;
constexpr uint32_t kLocalBudget = 4096;
static_assert((kLocalQ & (kLocalQ - 1)) == 0, "local queue is a power-of-two ring");

struct BatchCtrl {
  alignas(128) uint32_t cnt[2 * kParts * kCntStride];   // near queue sizes (round parity, part)
  alignas(128) uint32_t fcnt[2 * kParts * kCntStride];  // far pile sizes (far epoch parity, part)
  alignas(128) uint32_t fmin[2 * kParts * kCntStride];  // smallest key in each far pile
  alignas(128) uint32_t epoch;                 // graph replays (round stamps)
  alignas(128) uint32_t err;
  alignas(128) unsigned long long edges;
  // written only by the decision kernel (and initialisation)
  alignas(128) uint32_t fepoch;  // far epoch: appends go to pile fepoch & 1, stamp fepoch + 1
  uint32_t tkey;                 // near threshold (key domain)
  uint32_t busy_rounds, far_passes;
  // this round's descriptor
  uint32_t mode, t_old, t_new;
  uint32_t n_in[kParts], far_n[kParts];
  // entries left after the last round of a replay (read by the host)
  uint32_t left;
};

struct BatchArgs {
  const int* rowptr;
  const uint2* edges;
  unsigned long long* labels;
  // per slot: round stamp of its last near append (.x) and far stamp of its
  // last far append (.y), side by side so that one 8-byte probe reads both
  uint2* stamps;
  const QueryInfo* qinfo;
  uint2* q[2];        // near queues of (query, local index)
  uint2* far[2];      // far piles
  uint32_t qoff[kParts + 1];  // part p's segment of every queue / pile: [qoff[p], qoff[p + 1])
  BatchCtrl* ctrl;
  int32_t cols;       // grid width: vertex v = row * cols + col
  FastDiv cdiv;       // v / cols without a divide instruction
  int32_t rounds_per_graph;
  int want_edges;
  float delta;        // near-far bucket width (value domain); <= 0: Bellman-Ford rounds
  // per query: kNoStop while it runs, else the threshold below which its labels
  // are final when it stopped (all its targets settled); nullptr: no targets
  const uint32_t* qstop;
};

constexpr uint32_t kNoStop = 0xFFFFFFFFu;

// Targets of a far pass: at a far pass the near frontier is empty, so every
// label below the closing threshold t_old is final.  A target found below it
// is settled (once: its flag); the query whose last target settles stops with
// threshold t_old: its entries are dropped from then on, and its outputs at or
// above t_old read as unreached.  Runs between the decision and the work
// kernel of every round, and does nothing in relax rounds.
__global__ void k_batch_targets(const unsigned long long* __restrict__ labels, const QueryInfo* __restrict__ qinfo,
                                const BatchCtrl* __restrict__ c, const uint2* __restrict__ tgt,
                                uint32_t* __restrict__ tflag, uint32_t ntgt, uint32_t* __restrict__ qleft,
                                uint32_t* __restrict__ qstop) {
  if (c->mode != kRoundFar) return;
  const uint32_t t_old = c->t_old;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ntgt; i += gridDim.x * blockDim.x) {
    if (tflag[i]) continue;
    const uint2 t = tgt[i];
    const uint32_t key = (uint32_t)(labels[qinfo[t.x].base + t.y] >> 32);
    if (key < t_old) {
      tflag[i] = 1u;
      if (atomicSub(qleft + t.x, 1u) == 1u) qstop[t.x] = t_old;
    }
  }
}

__device__ inline void batch_stage_flush(uint2* stage, uint32_t& staged, uint2* q, uint32_t* cnt,
                                         uint32_t cap, uint32_t* err) {
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

// Next near threshold (key domain) after the far pile whose smallest key is
// kmin: one bucket further, or the end of kmin's bucket when the buckets in
// between are empty, and always past kmin so that a far pass moves at least
// one entry.  Correctness does not depend on T; only the work per slot does.
template <class W>
__device__ inline uint32_t next_threshold(uint32_t t_old, uint32_t kmin, float delta) {
  uint32_t t;
  if (W::kType == MI355_W_F32) {
    const float told = key_to_f32(t_old), km = key_to_f32(kmin);
    t = f32_to_key(fmaxf(told + delta, (floorf(km / delta) + 1.0f) * delta));
  } else {
    const double d = fmax(1.0, (double)delta);
    const double tn = fmax((double)t_old + d, (floor((double)kmin / d) + 1.0) * d);
    t = tn >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)tn;
  }
  return max(t, kmin + 1u);  // far keys are finite: kmin + 1 does not wrap
}

// The round's decision (one thread): relax the near frontier, or (frontier
// empty) run a far pass with the next threshold, or idle (nothing left).
// Clears the counters this round appends to.
template <class W>
__global__ void k_batch_decide(BatchArgs a, int k) {
  BatchCtrl* c = a.ctrl;
  uint32_t n_in = 0;
  for (uint32_t p = 0; p < kParts; ++p) {
    const uint32_t np = min(c->cnt[cidx(k & 1, p)], a.qoff[p + 1] - a.qoff[p]);
    c->n_in[p] = np;
    n_in += np;
    c->cnt[cidx((k + 1) & 1, p)] = 0;  // this round's near output
  }
  uint32_t mode = kRoundIdle;
  if (n_in > 0) {
    mode = kRoundRelax;
  } else {
    const uint32_t e = c->fepoch;
    uint32_t nf = 0, kmin = W::kInfKey;
    for (uint32_t p = 0; p < kParts; ++p) {
      nf += min(c->fcnt[cidx(e & 1, p)], a.qoff[p + 1] - a.qoff[p]);
      kmin = min(kmin, c->fmin[cidx(e & 1, p)]);
    }
    if (nf > 0) {
      // the pass reads pile e & 1; entries above the new threshold go to the
      // other pile, which the previous pass emptied
      mode = kRoundFar;
      const uint32_t t_old = c->tkey;
      const uint32_t t_new = next_threshold<W>(t_old, kmin, a.delta);
      for (uint32_t p = 0; p < kParts; ++p) {
        c->far_n[p] = min(c->fcnt[cidx(e & 1, p)], a.qoff[p + 1] - a.qoff[p]);
        c->fcnt[cidx((e + 1) & 1, p)] = 0;
        c->fmin[cidx((e + 1) & 1, p)] = W::kInfKey;
      }
      c->t_old = t_old;
      c->t_new = t_new;
      c->tkey = t_new;
      c->fepoch = e + 1;
      c->far_passes += 1;
    }
  }
  if (mode != kRoundIdle) c->busy_rounds += 1;
  c->mode = mode;
}

// Stage one lane's entry (wave-uniform call) into a per-wave LDS stage.
__device__ inline void stage_entry(bool push, uint2 item, uint2* stage, uint32_t& staged,
                                   uint32_t stage_cap, uint2* q, uint32_t* cnt, uint32_t cap,
                                   uint32_t* err) {
  const unsigned long long m = __ballot(push);
  if (m == 0) return;
  const uint32_t kk = (uint32_t)__popcll(m);
  if (staged + kk > stage_cap) batch_stage_flush(stage, staged, q, cnt, cap, err);
  if (push) stage[staged + prefix_count(m)] = item;
  staged += kk;
}

template <class W>
__global__ __launch_bounds__(kBlock) void k_batch_work(BatchArgs a, int k) {
  __shared__ int s_incl[kWavesPerBlock][kWave];
  __shared__ int s_beg[kWavesPerBlock][kWave];
  __shared__ uint32_t s_key[kWavesPerBlock][kWave];
  __shared__ int s_u[kWavesPerBlock][kWave];
  __shared__ int s_q[kWavesPerBlock][kWave];
  // each entry's query rectangle (origin, size) and label base, read once
  // per entry instead of once per edge
  __shared__ int4 s_rect[kWavesPerBlock][kWave];
  __shared__ int64_t s_qbase[kWavesPerBlock][kWave];
  __shared__ uint2 s_near[kWavesPerBlock][kNearStage];
  __shared__ uint2 s_far[kWavesPerBlock][kFarStage];
  __shared__ uint2 s_loc[kWavesPerBlock][kLocalQ];
  BatchCtrl* c = a.ctrl;
  const uint32_t mode = c->mode;
  if (mode == kRoundIdle) return;
  const uint32_t stamp = c->epoch * (uint32_t)a.rounds_per_graph + (uint32_t)k + 1u;
  const uint32_t fe = c->fepoch;  // this epoch's pile: fe & 1, far stamp fe + 1
  const uint32_t fst = fe + 1u;
  const uint32_t part = blockIdx.x % kParts, pwg = blockIdx.x / kParts;
  const uint32_t seg = a.qoff[part], cap = a.qoff[part + 1] - seg;
  uint2* fout = a.far[fe & 1] + seg;
  uint32_t* fcnt_out = &c->fcnt[cidx(fe & 1, part)];
  uint2* qout = a.q[(k + 1) & 1] + seg;
  uint32_t* cnt_out = &c->cnt[cidx((k + 1) & 1, part)];
  const unsigned lane = threadIdx.x & (kWave - 1);
  const unsigned wv = threadIdx.x / kWave;
  uint2* nst = s_near[wv];
  uint2* fst_stage = s_far[wv];
  uint32_t n_staged = 0, f_staged = 0;
  uint32_t fmin_lane = W::kInfKey;
  unsigned long long edge_visits = 0;
  const uint32_t nwaves = gridDim.x / kParts * kWavesPerBlock;  // of this part

  if (mode == kRoundFar) {
    // far pass over the pile of the epoch just closed
    const uint2* fin = a.far[(fe - 1u) & 1] + seg;
    const uint32_t nf = c->far_n[part], t_old = c->t_old, t_new = c->t_new;
    for (uint32_t first = (pwg * kWavesPerBlock + wv) * kWave; first < nf; first += nwaves * kWave) {
      const uint32_t i = first + lane;
      const bool ok = i < nf;
      const uint2 it = fin[ok ? i : 0u];
      const int64_t slot = a.qinfo[it.x].base + it.y;
      const uint32_t key = (uint32_t)(a.labels[slot] >> 32);
      // below t_old: stale (the slot improved below the threshold since and
      // was relaxed as near); [t_old, t_new): next frontier; above: stays far.
      // Entries of a stopped query are dropped.
      const bool live = ok && (!a.qstop || a.qstop[it.x] == kNoStop);
      const bool to_near = live && key >= t_old && key < t_new;
      const bool keep = live && key >= t_new;
      uint32_t x = stamp, y = fst;
      const uint2 sf = (to_near || keep) ? a.stamps[slot] : make_uint2(stamp, fst);
      if (to_near && sf.x != stamp) x = atomicExch(&a.stamps[slot].x, stamp);
      if (keep && sf.y != fst) y = atomicExch(&a.stamps[slot].y, fst);
      if (keep) fmin_lane = min(fmin_lane, key);
      stage_entry(x != stamp, it, nst, n_staged, kNearStage, qout, cnt_out, cap, &c->err);
      stage_entry(y != fst, it, fst_stage, f_staged, kFarStage, fout, fcnt_out, cap, &c->err);
    }
  } else {
    const uint32_t n_in = c->n_in[part], tkey = c->tkey;
    const uint2* qin = a.q[k & 1] + seg;
    uint2* loc = s_loc[wv];
    uint32_t lhead = 0, ltail = 0, lbudget = kLocalBudget;  // wave-uniform
    uint32_t first = (pwg * kWavesPerBlock + wv) * kWave;
    for (;;) {
      // the wave's local queue first, then its next chunk of the round's input
      const bool from_local = lhead != ltail;
      if (!from_local && first >= n_in) break;
      uint32_t take = 0;
      if (from_local) {
        take = min(ltail - lhead, (uint32_t)kWave);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      const uint32_t i = from_local ? lane : first + lane;
      const bool has = from_local ? lane < take : i < n_in;
      int u = 0, beg = 0, deg = 0, qid = 0;
      uint32_t key = W::kInfKey;
      QueryInfo qi{};
      if (has) {
        const uint2 it = from_local ? loc[(lhead + lane) & (kLocalQ - 1u)] : qin[i];
        qid = (int)it.x;
        qi = a.qinfo[qid];
        uint32_t r, col;
        tiled_cell(it.y, (uint32_t)qi.bh, (uint32_t)qi.bw, r, col);
        u = (qi.r0 + (int)r) * a.cols + qi.c0 + (int)col;
        key = atomic_load_key(a.labels + qi.base + it.y);
        beg = a.rowptr[u];
        if (a.qstop && a.qstop[qid] != kNoStop) key = W::kInfKey;  // stopped query
        if (key != W::kInfKey) deg = a.rowptr[u + 1] - beg;
      }
      int incl = deg;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const int t = __shfl_up(incl, off, kWave);
        if ((int)lane >= off) incl += t;
      }
      const int total = __shfl(incl, kWave - 1, kWave);
      if (from_local) lhead += take;  // entries read; their ring cells may be reused
      else first += nwaves * kWave;
      s_incl[wv][lane] = incl;
      s_beg[wv][lane] = beg - (incl - deg);
      s_key[wv][lane] = key;
      s_u[wv][lane] = u;
      s_q[wv][lane] = qid;
      s_rect[wv][lane] = make_int4(qi.r0, qi.c0, qi.bh, qi.bw);
      s_qbase[wv][lane] = qi.base;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      edge_visits += (unsigned long long)total;
      // Four edges per lane per step (j, j + 64, j + 128, j + 192): four
      // rounds of independent accesses (edges; labels; atomicMins; stamp
      // exchanges of the improving edges), each issued for all four edges
      // before any result is used.  Lanes past the end read edge 0 / slot 0
      // and discard them (guarded loads would each wait inside their
      // branch).
      for (int t = 0; t < total; t += 4 * kWave) {
        int pos[4], qq[4];
        bool ok[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          pos[p] = 0;
          ok[p] = t + p * kWave + (int)lane < total;
        }
#pragma unroll
        for (int s2 = 32; s2 > 0; s2 >>= 1) {
#pragma unroll
          for (int p = 0; p < 4; ++p)
            if (s_incl[wv][pos[p] + s2 - 1] <= t + p * kWave + (int)lane) pos[p] += s2;
        }
        uint2 e[4];
        int4 rect[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          qq[p] = ok[p] ? s_q[wv][pos[p]] : 0;
          e[p] = a.edges[ok[p] ? s_beg[wv][pos[p]] + t + p * kWave + (int)lane : 0];
          rect[p] = s_rect[wv][pos[p]];
        }
        constexpr unsigned long long kNone = ~0ull;
        int64_t slot[4];
        uint32_t lv[4], ck[4];
        unsigned long long cand[4], cur[4], old[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          // target inside the query's rectangle -> its local slot
          const int rv = (int)fastdiv(e[p].x, a.cdiv), cv = (int)e[p].x - rv * a.cols;
          const uint32_t dr = (uint32_t)(rv - rect[p].x), dc = (uint32_t)(cv - rect[p].y);
          const bool in = ok[p] && dr < (uint32_t)rect[p].z && dc < (uint32_t)rect[p].w;
          lv[p] = in ? tiled_slot(dr, dc, (uint32_t)rect[p].z, (uint32_t)rect[p].w) : 0u;
          slot[p] = in ? s_qbase[wv][pos[p]] + lv[p] : 0;
          ck[p] = in ? W::relax(s_key[wv][pos[p]], e[p].y) : W::kInfKey;
          cand[p] = ck[p] != W::kInfKey ? ((unsigned long long)ck[p] << 32) | (uint32_t)s_u[wv][pos[p]] : kNone;
          // label probe only: the stamps (another line of HBM) are touched by
          // the exchanges of the improving edges alone, a minority
          cur[p] = a.labels[slot[p]];
        }
        asm volatile("" : "+v"(cur[0]), "+v"(cur[1]), "+v"(cur[2]), "+v"(cur[3]));
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          old[p] = kNone;
          if (cand[p] < cur[p]) old[p] = atomicMin(a.labels + slot[p], cand[p]);
        }
        uint32_t x[4], y[4];
        bool nr[4], fr[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const bool imp = cand[p] < cur[p] && ck[p] < (uint32_t)(old[p] >> 32);
          // below the threshold: next round (once per round); above: far
          // pile (once per far epoch)
          nr[p] = imp && ck[p] < tkey;
          fr[p] = imp && !nr[p];
          if (fr[p]) fmin_lane = min(fmin_lane, ck[p]);
          x[p] = stamp;  // already stamped: no push
          y[p] = fst;
        }
        // near improvements this wave relaxes again later in the round, while
        // its ring has room and its budget lasts
        bool lc[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const unsigned long long m = __ballot(nr[p]);
          const uint32_t kk = (uint32_t)__popcll(m);
          lc[p] = false;
          if (kk && ltail - lhead + kk <= kLocalQ && kk <= lbudget) {
            if (nr[p]) loc[(ltail + prefix_count(m)) & (kLocalQ - 1u)] = make_uint2((uint32_t)qq[p], lv[p]);
            lc[p] = nr[p];
            ltail += kk;
            lbudget -= kk;
          }
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          if (nr[p] && !lc[p]) x[p] = atomicExch(&a.stamps[slot[p]].x, stamp);
          if (fr[p]) y[p] = atomicExch(&a.stamps[slot[p]].y, fst);
        }
        asm volatile("" : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]), "+v"(y[0]), "+v"(y[1]),
                     "+v"(y[2]), "+v"(y[3]));
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const uint2 it = make_uint2((uint32_t)qq[p], lv[p]);
          stage_entry(x[p] != stamp, it, nst, n_staged, kNearStage, qout, cnt_out, cap, &c->err);
          stage_entry(y[p] != fst, it, fst_stage, f_staged, kFarStage, fout, fcnt_out, cap, &c->err);
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  batch_stage_flush(nst, n_staged, qout, cnt_out, cap, &c->err);
  batch_stage_flush(fst_stage, f_staged, fout, fcnt_out, cap, &c->err);
  // smallest far key of the pile (one atomic per wave)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) fmin_lane = min(fmin_lane, (uint32_t)__shfl_xor((int)fmin_lane, off, 64));
  if (lane == 0 && fmin_lane != W::kInfKey) atomicMin(&c->fmin[cidx(fe & 1, part)], fmin_lane);
  if (a.want_edges && lane == 0 && edge_visits) atomicAdd(&c->edges, edge_visits);
}

// End of a replay: the graph epoch (round stamps) and the entries left.
__global__ void k_batch_epoch(BatchCtrl* c, int K) {
  c->epoch += 1u;
  uint32_t left = 0;
  for (uint32_t p = 0; p < kParts; ++p) left += c->cnt[cidx(K & 1, p)] + c->fcnt[cidx(c->fepoch & 1, p)];
  c->left = left;
}

template <class W>
__global__ void k_batch_init(unsigned long long* __restrict__ labels, uint2* __restrict__ stamps,
                             int64_t total) {
  const unsigned long long inf_label = ((unsigned long long)W::kInfKey << 32) | kNoPred;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    labels[i] = inf_label;
    stamps[i] = make_uint2(0u, 0u);
  }
}

// Sources of part p are src[soff[p] .. soff[p + 1]); they go to the start of
// part p's segment of queue 0.
struct PartOffsets {
  uint32_t soff[kParts + 1];
  uint32_t qoff[kParts + 1];
};

template <class W>
__global__ void k_batch_sources(unsigned long long* __restrict__ labels, const QueryInfo* __restrict__ qinfo,
                                const uint2* __restrict__ src, uint32_t nsrc, uint2* __restrict__ q0,
                                BatchCtrl* __restrict__ c, uint32_t tkey0, PartOffsets po) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsrc) {
    const uint2 s = src[i];
    labels[qinfo[s.x].base + s.y] = ((unsigned long long)W::kZeroKey << 32) | kNoPred;
    uint32_t p = 0;
    while (p + 1 < kParts && i >= po.soff[p + 1]) ++p;
    q0[po.qoff[p] + (i - po.soff[p])] = s;
  }
  if (i == 0) {
    for (uint32_t k = 0; k < 2; ++k)
      for (uint32_t p = 0; p < kParts; ++p) {
        c->cnt[cidx(k, p)] = k == 0 ? po.soff[p + 1] - po.soff[p] : 0u;
        c->fcnt[cidx(k, p)] = 0;
        c->fmin[cidx(k, p)] = W::kInfKey;
      }
    c->epoch = 0;
    c->err = 0;
    c->edges = 0;
    c->fepoch = 0;
    c->tkey = tkey0;
    c->busy_rounds = 0;
    c->far_passes = 0;
    c->mode = kRoundIdle;
    c->left = nsrc;
  }
}

// Labels -> outputs, one query per blockIdx.y (its rectangle in chunks).
template <class W>
__global__ void k_batch_output(const unsigned long long* __restrict__ labels, const QueryInfo* __restrict__ qinfo,
                               int inf_mode, uint32_t* __restrict__ dist, int32_t* __restrict__ pred,
                               const uint32_t* __restrict__ qstop) {
  const QueryInfo qi = qinfo[blockIdx.y];
  const uint32_t stop = qstop ? qstop[blockIdx.y] : kNoStop;
  const int64_t area = (int64_t)qi.bh * qi.bw;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < area; i += stride) {
    const uint32_t dr = (uint32_t)(i / qi.bw), dc = (uint32_t)(i - (int64_t)dr * qi.bw);
    const unsigned long long l = labels[qi.base + tiled_slot(dr, dc, (uint32_t)qi.bh, (uint32_t)qi.bw)];
    // a stopped query's labels at or above its threshold may not be final
    const uint32_t key = (uint32_t)(l >> 32) >= stop ? W::kInfKey : (uint32_t)(l >> 32);
    const bool reached = key != W::kInfKey;
    if (dist) {
      uint32_t out = key;
      if (W::kType == MI355_W_F32)
        out = reached ? __float_as_uint(key_to_f32(key)) : (inf_mode == MI355_INF_FLT_MAX ? 0x7F7FFFFFu : 0x7F800000u);
      dist[qi.out_base + i] = out;
    }
    if (pred) {
      const uint32_t p = (uint32_t)l;
      pred[qi.out_base + i] = (!reached || p == kNoPred) ? -1 : (int32_t)p;
    }
  }
}

__global__ void k_batch_source_preds(const QueryInfo* __restrict__ qinfo, const uint2* __restrict__ src,
                                     uint32_t nsrc, int32_t* __restrict__ pred) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsrc && pred) {
    const QueryInfo qi = qinfo[src[i].x];
    uint32_t dr, dc;
    tiled_cell(src[i].y, (uint32_t)qi.bh, (uint32_t)qi.bw, dr, dc);
    pred[qi.out_base + (int64_t)dr * qi.bw + dc] = -1;
  }
}

}  // namespace

// One target of a path read: its query's output block and rectangle.
struct PathTarget {
  int64_t out_base;
  int32_t r0, c0, bh, bw;
  int32_t target;
  int32_t pad;
};

// Output index of global vertex v in target t's block; -1 outside its rectangle.
__device__ inline int64_t path_index(const PathTarget& t, int64_t v, int32_t cols) {
  const int64_t r = v / cols - t.r0, c = v % cols - t.c0;
  if (v < 0 || r < 0 || r >= t.bh || c < 0 || c >= t.bw) return -1;
  return t.out_base + r * t.bw + c;
}

// Pass 1 (write == false): each target's path length (vertices from its
// source to it; 0 when unreached: parent -1 and a nonzero distance) and its
// distance bits.  Pass 2: the vertices, source first, from off[i].  A walk
// longer than its rectangle's area, or one that leaves the rectangle, sets err.
__global__ void k_batch_paths(const PathTarget* __restrict__ tg, int64_t count, int32_t cols,
                              const int32_t* __restrict__ pred, const uint32_t* __restrict__ dist,
                              int64_t* __restrict__ len, uint32_t* __restrict__ tdist,
                              const int64_t* __restrict__ off, int32_t* __restrict__ path,
                              uint32_t* __restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const PathTarget t = tg[i];
  const int64_t i0 = path_index(t, t.target, cols);
  if (i0 < 0) {
    atomicOr(err, 1u);
    return;
  }
  const uint32_t d0 = dist[i0];
  int64_t n = 0;
  if (!(pred[i0] < 0 && d0 != 0u)) {
    const int64_t limit = (int64_t)t.bh * t.bw;
    int64_t cur = t.target, at = i0;
    const int64_t L = path ? len[i] : 0;  // pass 2: the length pass 1 found
    for (;;) {
      if (path && n < L) path[off[i] + L - 1 - n] = (int32_t)cur;
      ++n;
      const int32_t p = pred[at];
      if (p < 0) break;
      cur = p;
      at = path_index(t, cur, cols);
      if (at < 0 || n >= limit) {
        atomicOr(err, 1u);
        break;
      }
    }
  }
  if (!path) {
    len[i] = n;
    tdist[i] = d0;
  }
}

void batch_paths(const mi355_query_batch& qb, const int64_t* out_offsets, const int32_t* d_pred,
                 const uint32_t* d_dist, const int32_t* query, const int32_t* target, int64_t count,
                 int64_t* path_offsets, uint32_t* target_dist, int32_t* path, int64_t path_capacity,
                 hipStream_t s) {
  if (count < 0) fail(MI355_EINVAL, "negative target count");
  path_offsets[0] = 0;
  if (count == 0) return;
  if (qb.grid_cols <= 0 || !qb.boxes) fail(MI355_EINVAL, "query batch without rectangles");
  std::vector<PathTarget> h((size_t)count);
  for (int64_t i = 0; i < count; ++i) {
    const int32_t q = query[i];
    if (q < 0 || q >= qb.count) fail(MI355_ERANGE, "target's query index outside the batch");
    const int32_t* b = qb.boxes + 4 * (size_t)q;
    PathTarget& t = h[(size_t)i];
    t.out_base = out_offsets[q];
    t.r0 = b[0];
    t.c0 = b[1];
    t.bh = b[2] - b[0];
    t.bw = b[3] - b[1];
    t.target = target[i];
    t.pad = 0;
    if (out_offsets[q + 1] - out_offsets[q] != (int64_t)t.bh * t.bw)
      fail(MI355_EINVAL, "output offsets do not match the query rectangles");
    const int64_t r = target[i] / qb.grid_cols - b[0], c = target[i] % qb.grid_cols - b[1];
    if (target[i] < 0 || r < 0 || r >= t.bh || c < 0 || c >= t.bw)
      fail(MI355_EINVAL, "target outside its query rectangle");
  }
  DevBuf<PathTarget> d_t;
  DevBuf<int64_t> d_len, d_off;
  DevBuf<uint32_t> d_td, d_err;
  d_t.ensure((size_t)count);
  d_len.ensure((size_t)count);
  d_td.ensure((size_t)count);
  d_err.ensure(1);
  MI_HIP(hipMemcpyAsync(d_t.p, h.data(), h.size() * sizeof(PathTarget), hipMemcpyHostToDevice, s));
  MI_HIP(hipMemsetAsync(d_err.p, 0, 4, s));
  const unsigned grid = (unsigned)div_up((uint64_t)count, kBlock);
  k_batch_paths<<<grid, kBlock, 0, s>>>(d_t.p, count, qb.grid_cols, d_pred, d_dist, d_len.p, d_td.p, nullptr,
                                        nullptr, d_err.p);
  MI_LAUNCH_CHECK();
  std::vector<int64_t> len((size_t)count);
  uint32_t err = 0;
  MI_HIP(hipMemcpyAsync(len.data(), d_len.p, len.size() * 8, hipMemcpyDeviceToHost, s));
  MI_HIP(hipMemcpyAsync(&err, d_err.p, 4, hipMemcpyDeviceToHost, s));
  if (target_dist) MI_HIP(hipMemcpyAsync(target_dist, d_td.p, (size_t)count * 4, hipMemcpyDeviceToHost, s));
  MI_HIP(hipStreamSynchronize(s));
  if (err) fail(MI355_EINVAL, "parent chain leaves its rectangle or does not end (not a batched run's output?)");
  for (int64_t i = 0; i < count; ++i) path_offsets[i + 1] = path_offsets[i] + len[(size_t)i];
  const int64_t total = path_offsets[count];
  if (!path) return;
  if (path_capacity < total) fail(MI355_ECAPACITY, "path buffer smaller than path_offsets[count]");
  if (total == 0) return;
  DevBuf<int32_t> d_path;
  d_off.ensure((size_t)count);
  d_path.ensure((size_t)total);
  MI_HIP(hipMemcpyAsync(d_off.p, path_offsets, (size_t)count * 8, hipMemcpyHostToDevice, s));
  k_batch_paths<<<grid, kBlock, 0, s>>>(d_t.p, count, qb.grid_cols, d_pred, d_dist, d_len.p, d_td.p, d_off.p,
                                        d_path.p, d_err.p);
  MI_LAUNCH_CHECK();
  MI_HIP(hipMemcpyAsync(path, d_path.p, (size_t)total * 4, hipMemcpyDeviceToHost, s));
  MI_HIP(hipMemcpyAsync(&err, d_err.p, 4, hipMemcpyDeviceToHost, s));
  MI_HIP(hipStreamSynchronize(s));
  if (err) fail(MI355_EINVAL, "parent chain leaves its rectangle or does not end");
}

void batch_layout(const mi355_query_batch& qb, int64_t n, std::vector<int64_t>& offsets) {
  if (qb.count <= 0) fail(MI355_EINVAL, "batch needs at least one query");
  if (qb.count > 65535) fail(MI355_EINVAL, "at most 65535 queries per batch");
  if (qb.grid_cols <= 0 || n % qb.grid_cols != 0)
    fail(MI355_EINVAL, "grid_cols must divide the vertex count (vertex = row * grid_cols + col)");
  if (!qb.src_offsets || !qb.sources || !qb.boxes) fail(MI355_EINVAL, "null query array");
  const int64_t rows = n / qb.grid_cols;
  offsets.assign((size_t)qb.count + 1, 0);
  for (int32_t q = 0; q < qb.count; ++q) {
    const int32_t* b = qb.boxes + 4 * (size_t)q;
    if (b[0] < 0 || b[1] < 0 || b[2] <= b[0] || b[3] <= b[1] || b[2] > rows || b[3] > qb.grid_cols)
      fail(MI355_EINVAL, "query rectangle outside the grid or empty");
    const int32_t s0 = qb.src_offsets[q], s1 = qb.src_offsets[q + 1];
    if (s0 < 0 || s1 <= s0) fail(MI355_EINVAL, "every query needs at least one source");
    for (int32_t i = s0; i < s1; ++i) {
      const int64_t v = qb.sources[i];
      if (v < 0 || v >= n) fail(MI355_ERANGE, "source vertex is outside the graph");
      const int64_t r = v / qb.grid_cols, c = v % qb.grid_cols;
      if (r < b[0] || r >= b[2] || c < b[1] || c >= b[3]) fail(MI355_EINVAL, "source outside its query rectangle");
    }
    const int64_t area = (int64_t)(b[2] - b[0]) * (int64_t)(b[3] - b[1]);
    if (tiled_area(b[2] - b[0], b[3] - b[1]) > (int64_t)UINT32_MAX) fail(MI355_EOVERFLOW, "query rectangle too large");
    offsets[(size_t)q + 1] = offsets[(size_t)q] + area;
  }
}

template <class W>
static void run_batch_typed(mi355_graph& g, const mi355_query_batch& qb, const mi355_batch_options& o,
                            const int32_t* t_off, const int32_t* targets, uint32_t* d_dist, int32_t* d_pred,
                            mi355_batch_stats& st, hipStream_t s) {
  std::vector<int64_t> off;
  batch_layout(qb, g.n, off);
  const int64_t total = off.back();
  const int32_t nq = qb.count;
  // Label blocks are laid out in schedule order: input order, or by
  // rectangle position (row, then column) so that queries whose rectangles
  // overlap sit next to each other.
  std::vector<int32_t> order((size_t)nq);
  std::iota(order.begin(), order.end(), 0);
  if (o.order == 1)
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
      const int32_t* bx = qb.boxes + 4 * (size_t)x;
      const int32_t* by = qb.boxes + 4 * (size_t)y;
      return bx[0] != by[0] ? bx[0] < by[0] : bx[1] < by[1];
    });
  std::vector<int64_t> lbase((size_t)nq);
  int64_t acc = 0;
  for (int32_t q : order) {
    const int32_t* b = qb.boxes + 4 * (size_t)q;
    lbase[(size_t)q] = acc;
    acc += tiled_area(b[2] - b[0], b[3] - b[1]);
  }
  const int64_t slots = acc;  // label slots incl. tile padding
  // parts of about equal area: largest query first to the lightest part
  std::vector<int32_t> part((size_t)nq), by_area((size_t)nq);
  std::iota(by_area.begin(), by_area.end(), 0);
  std::stable_sort(by_area.begin(), by_area.end(), [&](int32_t x, int32_t y) {
    return off[(size_t)x + 1] - off[(size_t)x] > off[(size_t)y + 1] - off[(size_t)y];
  });
  int64_t part_area[kParts] = {};
  for (int32_t q : by_area) {
    int b = 0;
    for (int x = 1; x < kParts; ++x)
      if (part_area[x] < part_area[b]) b = x;
    part[(size_t)q] = b;
    part_area[b] += off[(size_t)q + 1] - off[(size_t)q];
  }
  std::vector<QueryInfo> qi((size_t)nq);
  std::vector<std::vector<uint2>> psrc(kParts);
  for (int32_t q = 0; q < nq; ++q) {
    const int32_t* b = qb.boxes + 4 * (size_t)q;
    qi[(size_t)q] = QueryInfo{lbase[(size_t)q], off[(size_t)q], b[0], b[1], b[2] - b[0], b[3] - b[1]};
    for (int32_t i = qb.src_offsets[q]; i < qb.src_offsets[q + 1]; ++i) {
      const int64_t v = qb.sources[i];
      const int64_t r = v / qb.grid_cols - b[0], c = v % qb.grid_cols - b[1];
      psrc[(size_t)part[(size_t)q]].push_back(
          make_uint2((uint32_t)q, tiled_slot((uint32_t)r, (uint32_t)c, (uint32_t)(b[2] - b[0]), (uint32_t)(b[3] - b[1]))));
    }
  }
  // targets of each query (local slots); a query without any runs to the end
  std::vector<uint2> tgt;
  std::vector<uint32_t> qleft;
  if (t_off) {
    if (t_off[0] != 0) fail(MI355_EINVAL, "target offsets must start at 0");
    qleft.assign((size_t)nq, 0xFFFFFFFFu);
    for (int32_t q = 0; q < nq; ++q) {
      const int32_t* b = qb.boxes + 4 * (size_t)q;
      const int32_t t0 = t_off[q], t1 = t_off[q + 1];
      if (t1 < t0) fail(MI355_EINVAL, "target offsets must not decrease");
      if (t1 > t0) qleft[(size_t)q] = (uint32_t)(t1 - t0);
      for (int32_t i = t0; i < t1; ++i) {
        const int64_t v = targets[i];
        if (v < 0 || v >= g.n) fail(MI355_ERANGE, "target vertex is outside the graph");
        const int64_t r = v / qb.grid_cols - b[0], c = v % qb.grid_cols - b[1];
        if (r < 0 || r >= b[2] - b[0] || c < 0 || c >= b[3] - b[1]) fail(MI355_EINVAL, "target outside its query rectangle");
        tgt.push_back(make_uint2((uint32_t)q, tiled_slot((uint32_t)r, (uint32_t)c, (uint32_t)(b[2] - b[0]),
                                                        (uint32_t)(b[3] - b[1]))));
      }
    }
    if (tgt.size() > (size_t)UINT32_MAX) fail(MI355_EINVAL, "too many targets");
  }
  const uint32_t ntgt = (uint32_t)tgt.size();
  const bool stopping = t_off && ntgt > 0;
  std::vector<uint2> src;
  PartOffsets po{};
  for (int x = 0; x < kParts; ++x) {
    po.soff[x] = (uint32_t)src.size();
    src.insert(src.end(), psrc[(size_t)x].begin(), psrc[(size_t)x].end());
  }
  po.soff[kParts] = (uint32_t)src.size();
  const int64_t nsrc = (int64_t)src.size();
  // Queue capacity (entries of all parts), split in proportion to area; a
  // part never gets less than its sources.
  int64_t cap = o.queue_capacity > 0 ? o.queue_capacity : std::max<int64_t>((total + nsrc) / 4, int64_t(1) << 24);
  cap = std::max<int64_t>(std::min<int64_t>(cap, total + nsrc), nsrc);
  int64_t qacc = 0;
  for (int x = 0; x < kParts; ++x) {
    const int64_t ns = (int64_t)psrc[(size_t)x].size();
    int64_t cx = (int64_t)((double)cap * (double)part_area[x] / (double)std::max<int64_t>(total, 1));
    cx = std::max<int64_t>(std::min<int64_t>(cx, part_area[x] + ns), ns);
    po.qoff[x] = (uint32_t)qacc;
    qacc += cx;
    if (qacc > (int64_t)UINT32_MAX) fail(MI355_ECAPACITY, "batched frontier queues exceed 2^32 entries");
  }
  po.qoff[kParts] = (uint32_t)qacc;
  const int64_t qtotal = qacc;
  // even: the queue and counter parity return to round 0's after K rounds
  const int K = std::max(2, (o.rounds_per_graph > 0 ? o.rounds_per_graph : 24) / 2 * 2);
  // near-far bucket width; Bellman-Ford rounds (threshold infinity) when
  // asked for or when the graph gives no positive width
  if (!std::isfinite(o.delta_scale)) fail(MI355_EINVAL, "delta_scale must be finite");
  const float delta = o.delta_scale < 0.0f ? 0.0f : heuristic_delta(g) * (o.delta_scale > 0.0f ? o.delta_scale : 1.0f);
  const bool bucketed = delta > 0.0f && std::isfinite(delta);
  uint32_t tkey0 = W::kInfKey;
  if (bucketed)
    tkey0 = W::kType == MI355_W_F32 ? f32_to_key(delta)
                                    : (uint32_t)std::min(4294967294.0, std::max(1.0, std::ceil((double)delta)));
  const DeviceProps p = device_props();
  // a whole number of workgroups per part
  const unsigned grid = (unsigned)(div_up((uint64_t)std::max(1, (o.blocks_per_cu > 0 ? o.blocks_per_cu : 4) * p.cus),
                                          kParts) * kParts);

  // state cached on the graph handle and grown as needed
  Workspace& ws = g.ws;
  ws.batch_labels.ensure((size_t)slots);
  ws.batch_stamp.ensure((size_t)slots * 2);  // uint2 per slot
  ws.batch_q[0].ensure((size_t)std::max<int64_t>(qtotal, 1));
  ws.batch_q[1].ensure((size_t)std::max<int64_t>(qtotal, 1));
  if (bucketed) {
    ws.batch_far[0].ensure((size_t)std::max<int64_t>(qtotal, 1));
    ws.batch_far[1].ensure((size_t)std::max<int64_t>(qtotal, 1));
  }
  ws.batch_qinfo.ensure((size_t)nq * sizeof(QueryInfo));
  if (stopping) {
    ws.batch_tgt.ensure(ntgt);
    ws.batch_tstate.ensure((size_t)ntgt + 2 * (size_t)nq);  // target flags | queries left | stop
  }
  ws.batch_src.ensure((size_t)std::max<int64_t>(nsrc, 1));
  ws.batch_ctrl.ensure(sizeof(BatchCtrl));
  struct {
    unsigned long long* labels;
    uint32_t* stamp;
    uint2* q[2];
    QueryInfo* qinfo;
    uint2* src;
  } bs{ws.batch_labels.p, ws.batch_stamp.p, {ws.batch_q[0].p, ws.batch_q[1].p},
       reinterpret_cast<QueryInfo*>(ws.batch_qinfo.p), ws.batch_src.p};
  BatchCtrl* ctrl = reinterpret_cast<BatchCtrl*>(ws.batch_ctrl.p);
  MI_HIP(hipMemcpyAsync(bs.qinfo, qi.data(), qi.size() * sizeof(QueryInfo), hipMemcpyHostToDevice, s));
  MI_HIP(hipMemcpyAsync(bs.src, src.data(), src.size() * sizeof(uint2), hipMemcpyHostToDevice, s));

  BatchArgs a{};
  // the rectangles are in the input numbering's grid geometry
  input_order_arrays(g, &a.rowptr, &a.edges, s);
  a.labels = bs.labels;
  a.stamps = reinterpret_cast<uint2*>(bs.stamp);
  // Bellman-Ford rounds never touch the far piles: they alias the queues
  // (never read there)
  a.qinfo = bs.qinfo;
  a.q[0] = bs.q[0];
  a.q[1] = bs.q[1];
  a.far[0] = bucketed ? ws.batch_far[0].p : bs.q[0];
  a.far[1] = bucketed ? ws.batch_far[1].p : bs.q[1];
  std::copy(po.qoff, po.qoff + kParts + 1, a.qoff);
  a.delta = bucketed ? delta : 0.0f;
  uint32_t* tflag = nullptr;
  uint32_t* d_qleft = nullptr;
  uint32_t* d_qstop = nullptr;
  if (stopping) {
    tflag = ws.batch_tstate.p;
    d_qleft = tflag + ntgt;
    d_qstop = d_qleft + nq;
    std::vector<uint32_t> init((size_t)ntgt + 2 * (size_t)nq, 0u);
    std::copy(qleft.begin(), qleft.end(), init.begin() + ntgt);
    std::fill(init.begin() + ntgt + nq, init.end(), kNoStop);
    MI_HIP(hipMemcpyAsync(ws.batch_tgt.p, tgt.data(), tgt.size() * sizeof(uint2), hipMemcpyHostToDevice, s));
    MI_HIP(hipMemcpyAsync(tflag, init.data(), init.size() * 4, hipMemcpyHostToDevice, s));
    MI_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
    a.qstop = d_qstop;
  }
  const unsigned tgrid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(div_up((uint64_t)std::max(ntgt, 1u), kBlock), 1024));
  auto round_nodes = [&](int k) {
    k_batch_decide<W><<<1, 1, 0, s>>>(a, k);
    if (stopping)
      k_batch_targets<<<tgrid, kBlock, 0, s>>>(bs.labels, bs.qinfo, ctrl, ws.batch_tgt.p, tflag, ntgt, d_qleft,
                                                d_qstop);
    k_batch_work<W><<<grid, kBlock, 0, s>>>(a, k);
  };
  a.ctrl = ctrl;
  a.cols = qb.grid_cols;
  a.cdiv = make_fastdiv((uint32_t)qb.grid_cols);
  a.rounds_per_graph = K;
  a.want_edges = o.count_edges ? 1 : 0;

  hipEvent_t ev0, ev1;
  MI_HIP(hipEventCreate(&ev0));
  MI_HIP(hipEventCreate(&ev1));
  MI_HIP(hipEventRecord(ev0, s));
  k_batch_init<W><<<(unsigned)std::min<int64_t>(div_up((uint64_t)slots, kBlock), 8192), kBlock, 0, s>>>(
      bs.labels, a.stamps, slots);
  k_batch_sources<W><<<(unsigned)div_up((uint64_t)std::max<int64_t>(nsrc, 1), kBlock), kBlock, 0, s>>>(
      bs.labels, bs.qinfo, bs.src, (uint32_t)nsrc, bs.q[0], ctrl, tkey0, po);
  MI_LAUNCH_CHECK();

  // K rounds (+ the epoch node) as one graph, or as plain launches
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  if (o.use_graph) {
    MI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < K; ++k) round_nodes(k);
    k_batch_epoch<<<1, 1, 0, s>>>(ctrl, K);
    MI_HIP(hipStreamEndCapture(s, &graph));
    MI_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  uint32_t* h_cnt = nullptr;
  MI_HIP(hipHostMalloc((void**)&h_cnt, 4 * sizeof(uint32_t), hipHostMallocDefault));
  int64_t replays = 0;
  const int64_t replay_limit = (total + nsrc) * 8 / K + 4096;  // far above any real run
  int batch = 1;
  bool done = false;
  try {
    while (!done) {
      for (int r = 0; r < batch; ++r, ++replays) {
        if (exec) {
          MI_HIP(hipGraphLaunch(exec, s));
        } else {
          for (int k = 0; k < K; ++k) round_nodes(k);
          k_batch_epoch<<<1, 1, 0, s>>>(ctrl, K);
          MI_LAUNCH_CHECK();
        }
      }
      // entries left (next frontier + far pile) after the replay's last round
      MI_HIP(hipMemcpyAsync(h_cnt, &ctrl->left, 4, hipMemcpyDeviceToHost, s));
      MI_HIP(hipMemcpyAsync(h_cnt + 1, &ctrl->err, 4, hipMemcpyDeviceToHost, s));
      MI_HIP(hipStreamSynchronize(s));
      if (h_cnt[1]) fail(MI355_ECAPACITY, "batched frontier queue or far pile overflow (raise queue_capacity)");
      done = h_cnt[0] == 0;
      if (!done && replays > replay_limit) fail(MI355_EDEVICE, "batched frontier made no progress");
      batch = std::min(batch * 2, 8);
    }
  } catch (...) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipHostFree(h_cnt);
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    throw;
  }
  MI_HIP(hipEventRecord(ev1, s));
  {
    int64_t max_area = 0;
    for (const QueryInfo& x : qi) max_area = std::max<int64_t>(max_area, (int64_t)x.bh * x.bw);
    const dim3 og((unsigned)std::min<int64_t>(div_up((uint64_t)max_area, kBlock), 64), (unsigned)nq);
    k_batch_output<W><<<og, kBlock, 0, s>>>(bs.labels, bs.qinfo, o.inf_mode, d_dist, d_pred, d_qstop);
  }
  if (d_pred && nsrc)
    k_batch_source_preds<<<(unsigned)div_up((uint64_t)nsrc, kBlock), kBlock, 0, s>>>(bs.qinfo, bs.src,
                                                                                    (uint32_t)nsrc, d_pred);
  MI_LAUNCH_CHECK();
  BatchCtrl hc;
  MI_HIP(hipMemcpyAsync(&hc, ctrl, sizeof(BatchCtrl), hipMemcpyDeviceToHost, s));
  MI_HIP(hipStreamSynchronize(s));
  float ms = 0.0f;
  MI_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipHostFree(h_cnt);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  st = mi355_batch_stats{};
  st.label_slots = total;
  st.sources = nsrc;
  st.rounds = (int32_t)hc.busy_rounds;
  st.far_passes = (int32_t)hc.far_passes;
  st.graph_replays = (int32_t)replays;
  st.rounds_per_graph = K;
  st.edges_relaxed = (int64_t)hc.edges;
  st.device_ms = ms;
  st.queue_capacity = cap;
  if (stopping) {
    std::vector<uint32_t> hs((size_t)nq);
    MI_HIP(hipMemcpy(hs.data(), d_qstop, (size_t)nq * 4, hipMemcpyDeviceToHost));
    st.stopped_queries = (int32_t)std::count_if(hs.begin(), hs.end(), [](uint32_t x) { return x != kNoStop; });
  }
}

void run_batch(mi355_graph& g, const mi355_query_batch& qb, const mi355_batch_options& o, const int32_t* t_off,
               const int32_t* targets, void* d_dist, int32_t* d_pred, mi355_batch_stats& st, hipStream_t caller) {
  if (g.info.has_negative) fail(MI355_EINVAL, "batched queries require nonnegative edge weights");
  // Graph capture needs a stream of its own (not the legacy default stream);
  // it is ordered after the caller's work and the call is blocking.
  hipStream_t s;
  MI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t after;
  MI_HIP(hipEventCreateWithFlags(&after, hipEventDisableTiming));
  try {
    MI_HIP(hipEventRecord(after, caller));
    MI_HIP(hipStreamWaitEvent(s, after, 0));
    if (g.weight_type == MI355_W_F32)
      run_batch_typed<WF32>(g, qb, o, t_off, targets, static_cast<uint32_t*>(d_dist), d_pred, st, s);
    else
      run_batch_typed<WU32>(g, qb, o, t_off, targets, static_cast<uint32_t*>(d_dist), d_pred, st, s);
    MI_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    (void)hipEventDestroy(after);
    throw;
  }
  (void)hipStreamDestroy(s);
  (void)hipEventDestroy(after);
}

void batch_autotune(mi355_graph& g, const mi355_query_batch& qb, int32_t stride, mi355_batch_options& best,
                    double* timings, hipStream_t caller) {
  std::vector<int64_t> off;
  batch_layout(qb, g.n, off);
  stride = std::max(1, stride);
  // the sampled sub-batch
  std::vector<int32_t> so{0}, srcs, boxes;
  for (int32_t q = 0; q < qb.count; q += stride) {
    for (int32_t i = qb.src_offsets[q]; i < qb.src_offsets[q + 1]; ++i) srcs.push_back(qb.sources[i]);
    so.push_back((int32_t)srcs.size());
    boxes.insert(boxes.end(), qb.boxes + 4 * (size_t)q, qb.boxes + 4 * (size_t)q + 4);
  }
  mi355_query_batch sub{qb.grid_cols, (int32_t)(so.size() - 1), so.data(), srcs.data(), boxes.data()};
  std::vector<int64_t> soff;
  batch_layout(sub, g.n, soff);
  DevBuf<uint32_t> dist((size_t)soff.back());
  const int bpc[3] = {2, 4, 8};
  const float scales[3] = {0.5f, 1.0f, 2.0f};
  double best_ms = 1e300;
  int idx = 0;
  mi355_batch_options chosen = best;
  for (int order = 0; order < 2; ++order)
    for (float sc : scales)
      for (int b : bpc) {
        mi355_batch_options o = best;
        o.blocks_per_cu = b;
        o.delta_scale = sc;
        o.order = order;
        o.use_graph = 1;
        o.count_edges = 0;
        mi355_batch_stats st{};
        run_batch(g, sub, o, nullptr, nullptr, dist.p, nullptr, st, caller);
        if (timings) timings[idx] = st.device_ms;
        if (st.device_ms < best_ms) {
          best_ms = st.device_ms;
          chosen = o;
        }
        ++idx;
      }
  best = chosen;
}

}  // namespace mi355
