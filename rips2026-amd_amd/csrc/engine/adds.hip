// Asynchronous dynamic delta-stepping (ADDS) as one persistent kernel.
//
// Design inputs: the paper-faithful ADDS of ADDSF/kernel.hip:66-221 +
// ADDSF/wl.h:145-1915 (a persistent manager kernel beside persistent worker
// kernels, 32 bags mapped onto 32 lanes, wave32 only) and the single-kernel
// variant of JAS/hip_adds_sssp.cpp:396-866 (manager block + worker blocks,
// paged pool).  Neither runs on gfx950 as written, so the worklist is
// re-designed for wave64 and has no manager at all:
//
//  * Buckets.  Work items (vertex ids) live in kAddsSlots circular bucket
//    queues (ring buffers with monotone head/tail counters); absolute bucket
//    b uses slot b % kAddsSlots.  The window [base, base + kAddsSlots) moves
//    forward one bucket per phase.  A candidate distance d maps to
//    base + floor((d - T_base) / delta), clamped into the window (far items
//    are clipped into the last slot and re-queued when they come up) and
//    never below the bucket of the vertex that produced it.
//  * Workers.  Every workgroup (8 waves) is a worker: wave 0 claims up to
//    512 items (1024 from a bucket holding 64K or more: fewer claims when
//    work is plentiful, more claimers when it is scarce) and, 512 at a time,
//    the eight waves relax them, their edges packed across the workgroup's
//    lanes (four edges per lane in flight); vertices with more than
//    kHubDegree edges in the range being relaxed are cut into 8192-edge hub
//    chunks any workgroup can claim (R-MAT scale 22: 2048-edge chunks cost
//    12 % more time in claims and publishes).  Claims go through counting semaphores
//    (`avail`: one fetch-sub, the surplus given back, then a fetch-add on the
//    queue head for the positions), so hundreds of workgroups never retry a
//    CAS on one queue head.  Order of preference: hub chunks, the heavy list
//    of the bucket just finished, the active buckets [base, base + A).
//  * Light and heavy edges.  Rows are sorted by weight, so the edges with
//    w <= delta0 (the starting bucket width) are a row prefix.  Processing a
//    vertex of bucket b relaxes that prefix and lists the vertex (once per
//    bucket, stamp b + 1) for its heavy suffix, which is relaxed after bucket
//    b is finished, while base == b + 1 (those claims hold slot b + 1 open).
//    The suffix is relaxed from the label current at that time; a vertex
//    improved later is processed and listed again, so every final label is
//    relaxed over all its edges (the stamp is cleared before the label is
//    read).  This is delta-stepping's light/heavy split made asynchronous:
//    heavy edges are not re-relaxed every time a vertex is re-processed.
//  * Pushes are staged per wave in LDS and published in batches: one atomic
//    on `pending`, then one reservation per distinct target slot (an LDS
//    histogram gives each entry its offset), then the entries, then `avail`.
//    A reserved entry is the sentinel -1 until written, so a claimer that
//    gets there first waits for the write.
//  * Dedup.  qstamp[v] = smallest bucket v is queued in (atomicMin on push).
//    A pop exchanges it with "none"; the returned stamp is the bucket the
//    entry is processed in (if it falls in this slot), so every live entry is
//    processed once and superseded entries are dropped.  The label is read
//    after the exchange has returned, and every improvement completes its
//    label atomicMin before it touches qstamp, so each final distance is
//    relaxed from.
//  * Phases.  The base advances (one 64-bit CAS of (base, T_base)) only when
//    slot base and the heavy list of bucket base - 1 are empty and no claim
//    holding slot base is in flight, so no push can land behind the window.
//    The advancing wave evaluates the delta policy: fixed, the feedback rule
//    of plain ADDS (grow delta when the window holds too little work for the
//    workers, shrink it when it holds too much; after
//    JAS/hip_adds_sssp.cpp:683-727), or the learned MLP, which reads 32
//    features of the phase and returns log2 of the delta scale and the active
//    bucket count.  Re-mapping changes priorities only; distances remain the
//    exact fixpoint of the fp32 / uint32 left folds.
//  * Termination.  `pending` counts items queued or being processed; a batch
//    adds its count before its entries exist and before its producer
//    finishes, a finished (or dropped) claim subtracts afterwards, so
//    pending == 0 means there is no work anywhere.  Only running workgroups
//    ever hold items, so idle workgroups (which back off exponentially) need
//    not be co-resident for the run to make progress.
//  * Memory ordering.  All shared words are device-scope atomics, and ring
//    entries and hub records are written and read with device-coherent
//    (relaxed atomic) stores and loads; where an order matters the wave waits
//    for its outstanding memory operations (vm_wait) instead of using
//    agent-scope release/acquire, which on this device writes back or
//    invalidates the whole L2 and made the worklist 10x slower.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "adds.hpp"
#include "frontier_kernels.hpp"
#include "sssp.hpp"

namespace mi355 {

namespace {

constexpr uint32_t kNoStamp = 0xFFFFFFFFu;
// the control block's header (everything before the slot arrays): what the
// host uploads before a run and reads back after it
constexpr size_t kCtrlHeader = offsetof(AddsCtrl, head);
// An ADDS workgroup is 8 waves: its claims are shared by more waves (more
// memory accesses in flight per claimer) while the number of claimers, which
// contend on the worklist's control words, stays that of 4-wave workgroups.
constexpr int kAB = 512;            // threads per ADDS workgroup
constexpr int kAW = kAB / kWave;    // waves per ADDS workgroup
constexpr uint32_t kClaim = 2 * kAB;  // most items per workgroup claim (two passes of kAB)
constexpr uint32_t kBigClaimAvail = 65536;  // slot size above which claims take kClaim
constexpr int kHubChunk = 8192;     // edges per hub chunk
constexpr int kHubDegree = 4096;  // rows with more edges become hub chunks
static_assert(9 + kAddsSlots <= 64,
              "claim probe: lanes 0-3, the window's semaphores, the worker share, the pull words");
// vertices per pull chunk (four passes of kAB)
constexpr uint32_t kPullChunk = 4 * 512;
// edge reads a pull charges an unreached vertex: four probed in-edges, their
// labels, and the prefix their best candidate bounds
constexpr float kPullUnreachedReads = 10.0f;
constexpr int kEmptyEntry = -1;     // ring sentinel: reserved, not yet written
constexpr uint32_t kPushStage = 512;  // staged pushes per wave
constexpr uint32_t kPublishAt = 64;   // staged entries of other buckets offered between passes
constexpr uint32_t kHeavyStage = 256; // staged heavy-list entries per wave
// Bound of every spin (ring entry, hub chunk, idle worker): far above any
// legitimate wait; reaching it aborts the run with an error instead of
// hanging the device.
constexpr uint32_t kSpinLimit = 1u << 22;
// Idle polls (each backs off up to ~16K cycles) before a run counts as stalled.
constexpr uint32_t kIdleLimit = 1u << 20;
constexpr uint32_t kErrRing = 1u, kErrHub = 2u, kErrStall = 4u;

#define HEAD(c, s) ((c)->head[(s) * kPad])
#define TAIL(c, s) ((c)->tail[(s) * kPad])
#define INFL(c, s) ((c)->inflight[(s) * kPad])
#define AVAIL(c, s) ((c)->avail[(s) * kPad])
#define HV_HEAD(c, s) ((c)->hv_head[(s) * kPad])
#define HV_TAIL(c, s) ((c)->hv_tail[(s) * kPad])
#define HV_AVAIL(c, s) ((c)->hv_avail[(s) * kPad])
#define HV_MIN(c, s) ((c)->hv_min[(s) * kPad])

// Wait until this wave's outstanding vector memory operations have completed
// (s_waitcnt vmcnt(0)).  Device-scope atomics, and loads and stores with the
// device coherence bit, are performed at the device's coherence point, so
// completion in issue order is all the ordering the worklist needs.  Release
// and acquire at agent scope would write back or invalidate the whole L2
// instead.
__device__ inline void vm_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Relaxed device-scope add that has completed when it returns.
__device__ inline int atomic_add_done(int* p, int v) {
  const int old = __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  vm_wait();
  return old;
}

__device__ inline unsigned long long pack_base(uint32_t base, float t) {
  return ((unsigned long long)__float_as_uint(t) << 32) | base;
}
__device__ inline uint32_t base_of(unsigned long long w) { return (uint32_t)w; }
__device__ inline float tbase_of(unsigned long long w) { return __uint_as_float((uint32_t)(w >> 32)); }

template <class T>
__device__ inline T ld_rlx(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline void st_rlx(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bucket of distance d under the mapping (base, T_base, 1/delta), clamped to
// [lo, base + kAddsSlots - 1] (lo is the producer's bucket).
__device__ inline uint32_t map_bucket(float d, uint32_t base, float tbase, float inv_delta,
                                      uint32_t lo) {
  const float rel = (d - tbase) * inv_delta;
  uint32_t j = 0;
  if (rel > 0.0f) j = rel >= (float)(kAddsSlots - 1) ? (uint32_t)(kAddsSlots - 1) : (uint32_t)rel;
  uint32_t b = base + j;
  if (b < lo) b = lo;
  const uint32_t hi = base + kAddsSlots - 1;
  return b > hi ? (lo > hi ? lo : hi) : b;
}

struct AddsArgs {
  const int* rowptr;
  const uint2* edges;
  unsigned long long* labels;
  uint32_t* qstamp;
  int* ring;          // kAddsSlots * (cap_mask + 1)
  int* hring;         // heavy lists, same shape
  uint32_t cap_mask;
  // most live entries one bucket ring may hold before the run aborts with a
  // capacity error: the ring size, or a caller's smaller per-bucket capacity
  uint32_t ring_limit;
  uint32_t local_keep;  // most vertices a workgroup keeps in its local queue (0: none)
  uint32_t* sstamp;   // heavy-list stamp: bucket + 1 of the list holding v, 0 none
  const int* split;   // light edges per vertex (weight-sorted rows)
  HubChunk* hub;      // hub_mask + 1 records
  uint32_t hub_mask;
  AddsCtrl* c;
  const AddsPolicy* policy;
  const int* ids;     // input id of every internal vertex, the parent in labels (nullptr: the same)
  // in-edge rows for pulls (weight-sorted; nullptr: pulls off) and their
  // light prefixes at the split width (a pull reads only heavy in-edges: the
  // light edges of a bucket's vertices were relaxed when they were processed)
  const int* trow;
  const uint2* tedges;
  const int* tsplit;
  uint32_t max_degree;  // largest out-degree (bounds the push work of a heavy list)
  uint32_t n;
  uint32_t pull_n;      // 1 + the last vertex with an in-edge: pulls cover [0, pull_n)
  int live_edges;     // keep the live edge counter (phase traces read it)
  // Upper bound on claims by one workgroup: far above any real run; reaching
  // it means a livelock and aborts the run with an error.
  unsigned long long claim_limit;
};

// Write a ring entry.  The cell must hold the empty sentinel: its previous
// occupant (one lap earlier) has been read and cleared by its consumer.  A
// consumer that reads a cell before its own producer wrote it may take the
// previous lap's value, and that lap's consumer then takes this one: each
// value is still consumed once.  Overwriting an unread value would lose it,
// so that case aborts the run (ring too small) instead.
__device__ inline void ring_put(int* cell, int v, uint32_t* err, uint32_t* done) {
  if (__hip_atomic_exchange(cell, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kEmptyEntry) {
    atomicOr(err, kErrRing);
    atomicExch(done, 1u);
  }
}
__device__ inline void ring_store(const AddsArgs& a, uint32_t slot, uint32_t pos, int v) {
  ring_put(a.ring + (size_t)slot * (a.cap_mask + 1u) + (pos & a.cap_mask), v, &a.c->err, &a.c->done);
}

// Workgroup-local queue (an LDS ring) of the bucket a claim holds.  Improved
// vertices that fall into that bucket are processed by the same workgroup in
// further passes of its claim instead of going through the global ring
// (publish, then someone's claim): a chain of light edges then advances by
// one pass per hop, not one publish-claim round trip per hop, which is what
// bounds a high-diameter graph.  The entries stay inside the claim (pending
// and the slot's inflight count already cover them); the stamp protocol is
// the ring's.  Entries [free, head) are being read by the current pass,
// [head, tail) are queued; appends need tail - free < kLocal.
constexpr uint32_t kLocal = 4096;
// Most vertices a workgroup keeps locally at a time: one pass.  Measured on
// the 4900 x 4900 routing grid / R-MAT scale 22 (plain ADDS, kernel ms):
// none 226 / 7.05, 32: 213 / 6.95, 128: 191 / 6.92, 512: 159 / 6.33, 4096:
// 167 / 10.8 (with 4096 a few workgroups hoard R-MAT's ramp-up while the
// others idle).
constexpr uint32_t kLocalKeep = 512;
constexpr uint32_t kNoLocal = 0xFFFFFFFFu;
constexpr uint32_t kTakenLocal = 0xFFFFu;  // stage entry moved to the local queue
static_assert((kLocal & (kLocal - 1)) == 0, "local queue is a power-of-two ring");
struct LocalQueue {
  int* q;          // kLocal entries
  uint32_t* free;  // LDS counters (monotone)
  uint32_t* head;
  uint32_t* tail;
  uint32_t slot;   // ring slot of the held bucket; kNoLocal: keep nothing local
  uint32_t keep;   // most entries queued locally at a time (<= kLocal)
};

// Per-wave push stage in LDS: vertex and target slot of each queued entry.
struct PushStage {
  int* v;          // kPushStage
  uint32_t* slot;  // kPushStage; reused for the entry offset during a flush
  uint32_t* hist;  // kAddsSlots counters
  uint32_t n;      // wave-uniform
  LocalQueue lq;   // where entries of the held bucket go first
};

// Move the staged entries of the held bucket to the local queue (as many as
// it has room for); returns how many moved (wave-uniform).
__device__ inline uint32_t stage_to_local(PushStage& st) {
  const LocalQueue& lq = st.lq;
  const unsigned lane = lane_id();
  uint32_t want = 0;
  for (uint32_t i0 = 0; i0 < st.n; i0 += kWave) {
    const uint32_t i = i0 + lane;
    want += (uint32_t)__popcll(__ballot(i < st.n && st.slot[i] == lq.slot));
  }
  if (want == 0) return 0;
  uint32_t pos0 = 0, take = 0;
  if (lane == 0) {
    uint32_t t = __hip_atomic_load(lq.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
      const uint32_t used = t - __hip_atomic_load(lq.free, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      take = used < lq.keep ? min(want, lq.keep - used) : 0u;
      if (take == 0) break;
      const uint32_t seen = atomicCAS(lq.tail, t, t + take);
      if (seen == t) break;
      t = seen;
    }
    pos0 = t;
  }
  take = __shfl(take, 0, kWave);
  pos0 = __shfl(pos0, 0, kWave);
  if (take == 0) return 0;
  uint32_t rank = 0;
  for (uint32_t i0 = 0; i0 < st.n; i0 += kWave) {
    const uint32_t i = i0 + lane;
    const bool mine = i < st.n && st.slot[i] == lq.slot;
    const unsigned long long m = __ballot(mine);
    const uint32_t r = rank + prefix_count(m);
    if (mine && r < take) {
      lq.q[(pos0 + r) & (kLocal - 1u)] = st.v[i];
      st.slot[i] = kTakenLocal;
    }
    rank += (uint32_t)__popcll(m);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  return take;
}

// Move the staged entries of the held bucket to the local queue and drop
// them from the stage (the rest keep their order).
__device__ inline void stage_keep_local(PushStage& st) {
  if (st.n == 0 || st.lq.slot == kNoLocal) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (stage_to_local(st) == 0) return;
  const unsigned lane = lane_id();
  uint32_t out = 0;
  for (uint32_t i0 = 0; i0 < st.n; i0 += kWave) {
    // an entry moves only left, below every entry this chunk reads
    const uint32_t i = i0 + lane;
    const bool keep = i < st.n && st.slot[i] != kTakenLocal;
    const int v = keep ? st.v[i] : 0;
    const uint32_t sl = keep ? st.slot[i] : 0u;
    const unsigned long long m = __ballot(keep);
    __builtin_amdgcn_wave_barrier();
    if (keep) {
      st.v[out + prefix_count(m)] = v;
      st.slot[out + prefix_count(m)] = sl;
    }
    out += (uint32_t)__popcll(m);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  st.n = out;
}

template <bool P>
__device__ inline void stage_publish(const AddsArgs& a, PushStage& st) {
  if (st.n == 0) return;
  const unsigned lane = lane_id();
  AddsCtrl* c = a.c;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint32_t n_local = st.lq.slot != kNoLocal ? stage_to_local(st) : 0u;
  if (n_local == st.n) {
    st.n = 0;
    return;
  }
  __builtin_amdgcn_wave_barrier();
  // pending covers the entries before they are offered (the wait before the
  // avail update below); the producer's own claim keeps pending above zero
  // meanwhile
  if (lane == 0) atomicAdd(&c->pending, (unsigned long long)(st.n - n_local));
  if (lane < (unsigned)kAddsSlots) st.hist[lane] = 0;
  __builtin_amdgcn_wave_barrier();
  for (uint32_t i = lane; i < st.n; i += kWave) {
    const uint32_t s = st.slot[i];
    if (s != kTakenLocal) st.slot[i] = (s << 16) | atomicAdd(st.hist + s, 1u);  // LDS atomic: offset in slot group
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // one reservation per non-empty slot; hist[s] becomes the slot's base
  uint32_t k = 0;
  if (lane < (unsigned)kAddsSlots) {
    k = st.hist[lane];
    if (k) {
      const uint32_t pos0 = atomicAdd(&TAIL(c, lane), k);
      const uint32_t h = ld_rlx(&HEAD(c, lane));
      // live entries after the push; head may already be past reserved but
      // unwritten positions, so the difference is signed
      if ((int)(pos0 + k - h) > (int)a.ring_limit) {
        atomicOr(&c->err, kErrRing);
        atomicExch(&c->done, 1u);
      }
      st.hist[lane] = pos0;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // four entries per lane in flight: the exchanges' old values (the overwrite
  // check of ring_put) are examined after all four are issued
  for (uint32_t i0 = 0; i0 < st.n; i0 += 4 * kWave) {
    int old[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t i = i0 + (uint32_t)j * kWave + lane;
      old[j] = kEmptyEntry;
      const uint32_t so = i < st.n ? st.slot[i] : kTakenLocal;
      if (so != kTakenLocal) {
        const uint32_t s = so >> 16;
        int* cell = a.ring + (size_t)s * (a.cap_mask + 1u) + ((st.hist[s] + (so & 0xFFFFu)) & a.cap_mask);
        old[j] = __hip_atomic_exchange(cell, st.v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (old[0] != kEmptyEntry || old[1] != kEmptyEntry || old[2] != kEmptyEntry || old[3] != kEmptyEntry) {
      atomicOr(&c->err, kErrRing);
      atomicExch(&c->done, 1u);
    }
  }
  // pending (and the entries: a claimer that is early waits on the sentinel
  // anyway) before the offer: a claimer could otherwise finish the entries
  // and take their count off pending before it was added, and pending could
  // read zero while this claim still works.  The offer itself is not waited
  // for here: every wave waits for its outstanding operations once, before
  // its claim is released.
  vm_wait();
  __builtin_amdgcn_wave_barrier();
  if (k) atomicAdd(&AVAIL(c, lane), (int)k);
  st.n = 0;
}

// Claim the queue stamp of an improved vertex: true when v is not yet queued
// in bucket bkt or below.  The label update it follows has completed (its
// returned value decided `imp`), so a popper that takes the stamp afterwards
// reads the new label.
__device__ inline bool claim_stamp(const AddsArgs& a, bool imp, uint32_t v, uint32_t bkt) {
  if (!imp) return false;
  return __hip_atomic_fetch_min(a.qstamp + v, bkt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > bkt;
}

// Stage the lanes that claimed a stamp.
template <bool P>
__device__ inline void stage_wanted(const AddsArgs& a, PushStage& st, bool want, uint32_t v,
                                    uint32_t bkt, unsigned long long& pushed) {
  const unsigned long long mask = __ballot(want);
  if (mask == 0) return;
  const uint32_t k = (uint32_t)__popcll(mask);
  if (st.n + k > kPushStage) stage_publish<P>(a, st);
  if (want) {
    const uint32_t i = st.n + prefix_count(mask);
    st.v[i] = (int)v;
    st.slot[i] = bkt % kAddsSlots;
  }
  st.n += k;
  pushed += k;
}

// Stage the lanes' improved vertices (dedup by qstamp).
template <bool P>
__device__ inline void push_improved(const AddsArgs& a, PushStage& st, bool imp, uint32_t v,
                                     uint32_t bkt, unsigned long long& pushed) {
  stage_wanted<P>(a, st, claim_stamp(a, imp, v, bkt), v, bkt, pushed);
}

// Relax four edges per lane (edge p: u[p] -> v[p], candidate key ck[p],
// kInfKey = none, lo[p] = producer bucket) and stage the improved targets.
// The four label probes, then the four atomicMins, then the four stamp
// claims are issued before any of their results is used: the loop is bound by
// the latency of these dependent accesses, so four independent chains per
// lane keep four times the accesses in flight.  A lane without an edge probes
// label 0 and discards it (a guarded load would wait inside its branch).  Per
// edge this is the step of relax_edge_pair followed by push_improved.
template <class W, bool P>
__device__ inline void relax_quad(const AddsArgs& a, PushStage& st, const uint32_t (&v)[4],
                                  const uint32_t (&ck)[4], const uint32_t (&u)[4],
                                  const uint32_t (&lo)[4], uint32_t base, float tbase, float inv,
                                  unsigned long long& pushed) {
  constexpr unsigned long long kNone = ~0ull;
  unsigned long long cand[4], cur[4], old[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const bool has = ck[p] != W::kInfKey;
    cand[p] = has ? ((unsigned long long)ck[p] << 32) | u[p] : kNone;
    cur[p] = a.labels[has ? v[p] : 0u];
  }
  // all four probes have arrived before the first atomic is issued
  asm volatile("" : "+v"(cur[0]), "+v"(cur[1]), "+v"(cur[2]), "+v"(cur[3]));
  bool imp[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    old[p] = kNone;
    if (cand[p] < cur[p]) old[p] = atomicMin(a.labels + v[p], cand[p]);
  }
  uint32_t bk[4], q[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    imp[p] = cand[p] < cur[p] && ck[p] < (uint32_t)(old[p] >> 32);
    bk[p] = imp[p] ? map_bucket(W::bucket_value(ck[p]), base, tbase, inv, lo[p]) : 0u;
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    q[p] = 0u;  // 0 never exceeds a bucket: no claim
    if (imp[p]) q[p] = __hip_atomic_fetch_min(a.qstamp + v[p], bk[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // keep the four claims in flight: the comparisons stay after the last issue
  asm volatile("" : "+v"(q[0]), "+v"(q[1]), "+v"(q[2]), "+v"(q[3]));
#pragma unroll
  for (int p = 0; p < 4; ++p) stage_wanted<P>(a, st, q[p] > bk[p], v[p], bk[p], pushed);
}

// Publish a wave's staged heavy-list entries (all of one bucket, list slot
// hs): pending first, then the reservation, the entries, and the count that
// offers them to claimers.
__device__ inline void heavy_publish(const AddsArgs& a, int* stage, uint32_t& n, uint32_t hs) {
  if (n == 0) return;
  const unsigned lane = lane_id();
  AddsCtrl* c = a.c;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint32_t pos0 = 0;
  if (lane == 0) {
    atomicAdd(&c->pending, (unsigned long long)n);  // completed before the avail update below
    pos0 = atomicAdd(&HV_TAIL(c, hs), n);
    if ((int)(pos0 + n - ld_rlx(&HV_HEAD(c, hs))) > (int)(a.cap_mask + 1u)) {
      atomicOr(&c->err, kErrRing);
      atomicExch(&c->done, 1u);
    }
  }
  pos0 = __shfl(pos0, 0, 64);
  int* ring = a.hring + (size_t)hs * (a.cap_mask + 1u);
  for (uint32_t i = lane; i < n; i += kWave) ring_put(ring + ((pos0 + i) & a.cap_mask), stage[i], &c->err, &c->done);
  // pending and the entries before the offer, as in stage_publish; the
  // offer is waited for before the claim's release
  vm_wait();
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) atomicAdd(&HV_AVAIL(c, hs), (int)n);
  n = 0;
}

// The network's decision for the next phase: delta = delta0 * 2^clamp(o0, -4, 4)
// (within the run's delta bounds), round(o1) active buckets in [1, 8], and
// the share 2^clamp(o2, -5, 0) of the workgroups that claim (at least 16);
// before the first claim also the light width, 2^clamp(o3, -5, 0) times the
// first bucket width (light_scale).
__device__ inline void mlp_decide(const AddsPolicy* p, const AddsCtrl* c, float feat, unsigned lane,
                                  float& delta, uint32_t& active, uint32_t& workers_now,
                                  float* light_scale = nullptr) {
  float o[kMlpOut];
  mlp_forward_wave(p, feat, lane, o);
  if (light_scale) *light_scale = exp2f(fminf(fmaxf(o[3], -5.0f), 0.0f));
  const float share = exp2f(fminf(fmaxf(o[2], -5.0f), 0.0f));
  workers_now = (uint32_t)fmaxf((float)c->workers * share, 16.0f);
  workers_now = min(workers_now, c->workers);
  const float scale = exp2f(fminf(fmaxf(o[0], -4.0f), 4.0f));
  delta = fminf(fmaxf(c->base_delta * scale, c->min_delta), c->max_delta);
  active = (uint32_t)fminf(fmaxf(o[1] + 0.5f, 1.0f), 8.0f);
}

// Schedule mode: the settings of phase `phase` from the table.
__device__ inline void sched_decide(const AddsPolicy* p, const AddsCtrl* c, uint32_t phase, float& delta,
                                    uint32_t& active, uint32_t& workers_now) {
  const float* e = p->sched + 3u * min(phase, p->sched_len - 1u);
  delta = fminf(fmaxf(c->base_delta * e[0], c->min_delta), c->max_delta);
  active = (uint32_t)fminf(fmaxf(e[1], 1.0f), 8.0f);
  const float share = fminf(fmaxf(e[2], 1.0f / 32.0f), 1.0f);
  workers_now = min(c->workers, (uint32_t)fmaxf((float)c->workers * share, 16.0f));
}

// MLP mode, before the first claim: the network also sets the first phase
// (the source alone in bucket 0, delta = delta0), so the run does not start
// with the heuristic width on every workgroup.  One wave.
__global__ void k_adds_first_phase(AddsArgs a) {
  const unsigned lane = threadIdx.x;
  AddsCtrl* c = a.c;
  float feat = 0.0f, raw = 0.0f;
  mlp_features(a.policy, c, lane == 0 ? 1u : 0u, 1u, c->delta, lane, &feat, &raw);
  float delta, light = 1.0f;
  uint32_t active, workers_now;
  if (a.policy->mode == kPolicySchedule) sched_decide(a.policy, c, 0u, delta, active, workers_now);
  else mlp_decide(a.policy, c, feat, lane, delta, active, workers_now, &light);
  if (lane == 0) {
    c->delta = delta;
    c->active = active;
    c->active_workers = workers_now;
    c->light_scale = light;
  }
}

// Wave-level phase advance; true when this wave moved the base.  P: pulls on
// (the boundary decides push or pull, and the next advance waits for both).
template <class W, bool P>
__device__ bool try_advance(const AddsArgs& a, unsigned long long bw) {
  AddsCtrl* c = a.c;
  const unsigned lane = lane_id();
  const uint32_t base = base_of(bw);
  const uint32_t s = base % kAddsSlots;
  bool ok = false;
  // The boundary work of `base` (policy, pull registration) is published and
  // the pull it registered, if any, is finished; the emptiness reads below
  // follow these loads.
  bool ready = false;
  if (lane == 0) {
    ready = true;
    if constexpr (P) {
      // one load when the boundary registered no pull
      const uint32_t dw = ld_rlx(&c->decided);
      ready = (dw >> 1) == base && (!(dw & 1u) || ld_rlx(&c->pull_done) >= ld_rlx(&c->pull_chunks));
    }
  }
  if (ready) {
    // Only holders of a claim (or hub chunk) of bucket `base` push into it,
    // and a claimer raises inflight before its head CAS.  head == tail (slot
    // empty at the tail read), then inflight == 0 (no holder), then an
    // unchanged tail (nothing pushed in between) leave no entry and no holder
    // that could still add one.
    // The heavy list of bucket base - 1 is drained the same way: its claims
    // hold inflight on slot base, and nothing is appended to it any more.
    const uint32_t hs = (base + kAddsSlots - 1) % kAddsSlots;
    const uint32_t h = __hip_atomic_load(&HEAD(c, s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t t = __hip_atomic_load(&TAIL(c, s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t hh = 0, ht = 0;
    if (base > 0) {
      hh = __hip_atomic_load(&HV_HEAD(c, hs), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ht = __hip_atomic_load(&HV_TAIL(c, hs), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const bool empty = h == t && hh == ht;
    const int inflight = empty ? __hip_atomic_load(&INFL(c, s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 1;
    bool same = false;
    if (inflight == 0)
      same = __hip_atomic_load(&TAIL(c, s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == t &&
             (base == 0 || __hip_atomic_load(&HV_TAIL(c, hs), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == ht);
    if (empty && inflight == 0 && same) {
      const float delta = ld_rlx(&c->delta);
      ok = atomicCAS(&c->base_word, bw, pack_base(base + 1, tbase_of(bw) + delta)) == bw;
    }
  }
  ok = __shfl(ok, 0, 64);
  if (!ok) return false;
  const long long t_adv = wall_clock64();
  // phase boundary: occupancy of the new window, then the policy
  const uint32_t nb = base + 1;
  uint32_t occ = 0;
  if (lane < (unsigned)kAddsSlots) {
    const uint32_t ss = (nb + lane) % kAddsSlots;
    occ = ld_rlx(&TAIL(c, ss)) - ld_rlx(&HEAD(c, ss));
    if ((int)occ < 0) occ = 0;
  }
  uint32_t window_items = occ;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) window_items += __shfl_xor(window_items, off, 64);
  // Heavy edges of the bucket just finished: pushed from its heavy list, or
  // pulled by the vertices they could still improve when the estimated push
  // work exceeds 2^margin x the estimated pull work.  Push work: the list's
  // length times the mean heavy degree of 256 entries spread over it.  Pull
  // work, over the pull range [0, pull_n): kPullUnreachedReads edge reads per
  // vertex not reached yet (its first four heavy in-edges, their labels and
  // the prefix they bound; pull_n / 256 times the unreached ones with edges
  // among 256 evenly spaced vertices), an edge read's worth per vertex (its
  // label and, at or above the bucket, a row search) and about two row reads
  // per queued vertex.
  // The bucket's claims are all finished (the base moved), so
  // its list is complete; nothing of it is claimed before `decided`.  Decided
  // first, before the policy: the heavy list waits for it.
  bool pulled = false;
  if constexpr (P) {
    const uint32_t fs = base % kAddsSlots;
    uint32_t lh = 0, lt = 0, mode_on = 0, key_min = kNoStamp;
    float margin = 0.0f;
    if (lane == 0) {
      lh = ld_rlx(&HV_HEAD(c, fs));
      lt = ld_rlx(&HV_TAIL(c, fs));
      mode_on = ld_rlx(&c->pull_mode);
      margin = exp2f(ld_rlx(&c->pull_margin));
      // the finished bucket's light passes are complete (each claim waited
      // for its updates before releasing inflight); the slot is next used by
      // bucket base + kAddsSlots, whose items are processed only once it is
      // active, long after this reset
      key_min = ld_rlx(&HV_MIN(c, fs));
      st_rlx(&HV_MIN(c, fs), kNoStamp);
    }
    lh = __shfl(lh, 0, 64);
    lt = __shfl(lt, 0, 64);
    mode_on = __shfl(mode_on, 0, 64);
    margin = __shfl(margin, 0, 64);
    key_min = __shfl(key_min, 0, 64);
    const uint32_t count = lt - lh;
    // the pull estimate is at least n: no sampling when even rows of the
    // largest degree could not outweigh it
    if (mode_on && count > 0 && a.pull_n > 0 && (float)count * (float)a.max_degree > margin * (float)a.pull_n) {
      const uint32_t ns = min(count, 256u);
      const int* ring = a.hring + (size_t)fs * (a.cap_mask + 1u);
      float heavy = 0.0f, unreached = 0.0f;
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t i = lane * 4u + j;
        if (i < ns) {
          const uint32_t pos = lh + (uint32_t)(((unsigned long long)i * count) / ns);
          const int u = ld_rlx(ring + (pos & a.cap_mask));
          if (u >= 0 && (uint32_t)u < a.n)
            heavy += (float)(a.rowptr[u + 1] - a.rowptr[u] - a.split[u]);
        }
        const uint32_t v = (uint32_t)((((unsigned long long)(2u * i + 1u)) * a.pull_n) / 512u);
        if ((uint32_t)(ld_rlx(a.labels + v) >> 32) == W::kInfKey && a.rowptr[v + 1] > a.rowptr[v]) unreached += 1.0f;
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        heavy += __shfl_xor(heavy, off, 64);
        unreached += __shfl_xor(unreached, off, 64);
      }
      const float push_work = (float)count * heavy / (float)ns;
      const float pull_work = kPullUnreachedReads * unreached * ((float)a.pull_n / 256.0f) + (float)a.pull_n +
                              2.0f * (float)window_items;
      pulled = push_work > margin * pull_work && key_min != kNoStamp;
    }
    if (lane == 0) {
      if (pulled) {
        // the pull holds the new phase open (pending, and pull_done before the
        // next advance); its chunks are offered once registered and claimed
        // once every list entry's key is folded into pull_key
        const uint32_t chunks = (a.pull_n + kPullChunk - 1) / kPullChunk;
        st_rlx(&c->pull_count, count);
        st_rlx(&c->pull_key, key_min);
        st_rlx(&c->pull_keyed, count);
        st_rlx(&c->pull_chunks, chunks);
        st_rlx(&c->pull_head, 0u);
        st_rlx(&c->pull_done, 0u);
        st_rlx(&c->pull_base, nb);
        atomicAdd(&c->pending, (unsigned long long)chunks);
        atomicAdd(&c->pulls, 1u);
        vm_wait();
        atomicAdd(&c->pull_avail, (int)chunks);
        vm_wait();
      }
      st_rlx(&c->decided, (nb << 1) | (pulled ? 1u : 0u));
    }
  }
  float delta = ld_rlx(&c->delta);
  uint32_t active = ld_rlx(&c->active);
  uint32_t workers_now = ld_rlx(&c->active_workers);
  const int mode = a.policy ? a.policy->mode : kPolicyFixed;
  float feat = 0.0f, raw = 0.0f;
  if (mode == kPolicyMlp || (a.policy && a.policy->trace))
    mlp_features(a.policy, c, occ, window_items, delta, lane, &feat, &raw);
  if (mode == kPolicyHeuristic) {
    const float capacity = (float)c->workers * (float)kAB;
    if ((float)window_items < c->util_low * capacity) delta = fminf(delta * c->growth, c->max_delta);
    else if ((float)window_items > c->util_high * capacity) delta = fmaxf(delta * c->shrink, c->min_delta);
  } else if (mode == kPolicyMlp) {
    mlp_decide(a.policy, c, feat, lane, delta, active, workers_now);
  } else if (mode == kPolicySchedule) {
    // only the wave that moved the base counts phases: the next phase's index
    sched_decide(a.policy, c, ld_rlx(&c->phases) + 1u, delta, active, workers_now);
  }
  // the record index comes from the phase counter itself: two waves that
  // advance one after the other never share a record
  const uint32_t ph = __shfl(lane == 0 ? atomicAdd(&c->phases, 1u) : 0u, 0, 64);
  if (a.policy && a.policy->trace && ph < a.policy->trace_cap) {
    float* t = a.policy->trace + (size_t)ph * kTraceStride;
    if (lane < (unsigned)kMlpIn) t[lane] = raw;
    if (lane == 0) {
      t[32] = (float)window_items;
      t[33] = delta;
      t[34] = (float)active;
      t[35] = (float)ld_rlx(&c->edges);
      t[36] = (float)ld_rlx(&c->pushes);
      t[37] = (float)workers_now;
      t[38] = (float)(t_adv - (long long)ld_rlx(&c->diag[15])) * 0.01f;  // 100 MHz clock -> us
      t[39] = pulled ? 1.0f : 0.0f;
    }
  }
  if (lane == 0) {
    __hip_atomic_store(&c->delta, delta, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&c->active, active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&c->active_workers, workers_now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicAdd(&c->diag[14], (unsigned long long)(wall_clock64() - t_adv));
  }
  return true;
}

// Relax the packed edges of up to kAB vertices held one per thread (deg = 0
// for threads without work) and stage the improved targets; u = the vertex's
// input id (the parent written into the labels), lo = producer bucket.  The edge list is packed across the whole workgroup and its waves
// take 128-edge steps of it in turn (two edges per lane), so one high-degree
// vertex does not keep the other waves waiting at the next barrier.  The
// arrays are kAB entries in LDS; s_wsum holds kAW totals.
template <class W, bool P>
__device__ inline void relax_block(const AddsArgs& a, PushStage& st, int* s_incl, int* s_base,
                                   uint32_t* s_key, int* s_u, uint32_t* s_lo, int* s_wsum, int u,
                                   int beg, int deg, uint32_t key, uint32_t lo, uint32_t base,
                                   float tbase, float inv, unsigned long long& edges,
                                   unsigned long long& pushed) {
  const unsigned lane = lane_id();
  const unsigned wv = threadIdx.x / kWave;
  int incl = deg;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int t = __shfl_up(incl, off, kWave);
    if ((int)lane >= off) incl += t;
  }
  if (lane == kWave - 1) s_wsum[wv] = incl;
  __syncthreads();
  int total = 0;
#pragma unroll
  for (int w = 0; w < kAW; ++w) {
    const int ws = s_wsum[w];
    if (w < (int)wv) incl += ws;
    total += ws;
  }
  s_incl[threadIdx.x] = incl;
  s_base[threadIdx.x] = beg - (incl - deg);
  s_key[threadIdx.x] = key;
  s_u[threadIdx.x] = u;
  s_lo[threadIdx.x] = lo;
  __syncthreads();
  if (a.live_edges && threadIdx.x == 0 && total > 0) atomicAdd(&a.c->edges, (unsigned long long)total);
  if (wv == 0) edges += (unsigned long long)total;
  // edge j belongs to the vertex p with s_incl[p - 1] <= j < s_incl[p];
  // a wave takes 256-edge steps, four edges per lane (j, j + 64, j + 128,
  // j + 192)
  static_assert((kAB & (kAB - 1)) == 0, "binary search over a power-of-two block");
  for (int t = (int)wv * 4 * kWave; t < total; t += 4 * kAB) {
    int pos[4];
    bool ok[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      pos[p] = 0;
      ok[p] = t + p * kWave + (int)lane < total;
    }
#pragma unroll
    for (int s2 = kAB / 2; s2 > 0; s2 >>= 1) {
#pragma unroll
      for (int p = 0; p < 4; ++p)
        if (s_incl[pos[p] + s2 - 1] <= t + p * kWave + (int)lane) pos[p] += s2;
    }
    uint2 e[4];
    // lanes past the end read edge 0 and discard it (no wait inside a branch)
#pragma unroll
    for (int p = 0; p < 4; ++p) e[p] = a.edges[ok[p] ? s_base[pos[p]] + t + p * kWave + (int)lane : 0];
    uint32_t v[4], ck[4], u[4], lo[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      v[p] = e[p].x;
      ck[p] = ok[p] ? W::relax(s_key[pos[p]], e[p].y) : W::kInfKey;
      u[p] = (uint32_t)s_u[pos[p]];
      lo[p] = s_lo[pos[p]];
    }
    relax_quad<W, P>(a, st, v, ck, u, lo, base, tbase, inv, pushed);
  }
  __builtin_amdgcn_wave_barrier();
}

// One pass of a pull chunk: thread t holds vertex v (has), the label it read
// (lab: a stale read is only larger, which widens the range below) and the
// in-edge range [beg, beg + len) whose edges can still lower or tie its label.
// The ranges are packed across the workgroup as in relax_block; every edge
// u -> v reads u's label with a device-coherent load (the candidate is built
// from it, so a stale cached copy could lose the improvement) and keeps the
// smallest (candidate key, input id of u) per vertex in LDS.  Afterwards each
// thread applies its vertex's best through the label atomicMin and queues an
// improved vertex like any push (producer bucket: the base).
template <class W, bool kDiag>
__device__ inline void pull_pass(const AddsArgs& a, PushStage& st, int* s_incl, int* s_base, uint32_t* s_key,
                                 unsigned long long* s_best, int* s_wsum, bool has, uint32_t v,
                                 unsigned long long lab, int beg, int len, uint32_t base, float tbase,
                                 float inv, unsigned long long& edges, unsigned long long& pushed) {
  const unsigned lane = lane_id();
  const unsigned wv = threadIdx.x / kWave;
  int incl = len;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int t = __shfl_up(incl, off, kWave);
    if ((int)lane >= off) incl += t;
  }
  if (lane == kWave - 1) s_wsum[wv] = incl;
  __syncthreads();
  int total = 0;
#pragma unroll
  for (int w = 0; w < kAW; ++w) {
    const int ws = s_wsum[w];
    if (w < (int)wv) incl += ws;
    total += ws;
  }
  s_incl[threadIdx.x] = incl;
  s_base[threadIdx.x] = beg - (incl - len);
  s_key[threadIdx.x] = (uint32_t)(lab >> 32);
  s_best[threadIdx.x] = lab;
  __syncthreads();
  if (a.live_edges && threadIdx.x == 0 && total > 0) atomicAdd(&a.c->edges, (unsigned long long)total);
  if (wv == 0) edges += (unsigned long long)total;
  if (kDiag && threadIdx.x == 0 && total > 0) atomicAdd(&a.c->pdiag[1], (unsigned long long)total);
  for (int t = (int)wv * 4 * kWave; t < total; t += 4 * kAB) {
    int pos[4];
    bool ok[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      pos[p] = 0;
      ok[p] = t + p * kWave + (int)lane < total;
    }
#pragma unroll
    for (int s2 = kAB / 2; s2 > 0; s2 >>= 1) {
#pragma unroll
      for (int p = 0; p < 4; ++p)
        if (s_incl[pos[p] + s2 - 1] <= t + p * kWave + (int)lane) pos[p] += s2;
    }
    uint2 e[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) e[p] = a.tedges[ok[p] ? s_base[pos[p]] + t + p * kWave + (int)lane : 0];
    unsigned long long lu[4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
      lu[p] = __hip_atomic_load(a.labels + (ok[p] ? e[p].x : 0u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const uint32_t ck = ok[p] ? W::relax((uint32_t)(lu[p] >> 32), e[p].y) : W::kInfKey;
      if (ck != W::kInfKey && ck <= s_key[pos[p]]) {
        const uint32_t pid = a.ids ? (uint32_t)a.ids[e[p].x] : e[p].x;
        const unsigned long long cand = ((unsigned long long)ck << 32) | pid;
        if (cand < s_best[pos[p]]) atomicMin(s_best + pos[p], cand);
      }
    }
  }
  __syncthreads();
  const unsigned long long best = s_best[threadIdx.x];
  bool imp = false;
  uint32_t bk = 0;
  if (has && best < lab) {
    const unsigned long long old = atomicMin(a.labels + v, best);
    const uint32_t bkey = (uint32_t)(best >> 32);
    imp = bkey < (uint32_t)(old >> 32);
    if (imp) bk = map_bucket(W::bucket_value(bkey), base, tbase, inv, base);
  }
  stage_wanted<true>(a, st, claim_stamp(a, imp, v, bk), v, bk, pushed);
  __syncthreads();  // the LDS arrays are refilled by the next pass
}

enum ClaimKind : uint32_t { kClaimItems = 0, kClaimHub = 1, kClaimHeavy = 2, kClaimPull = 3, kClaimHeavyDrop = 4 };

// Cut the row range [beg, end) of u into hub chunks (pending work held on
// inflight slot `slot`, pushes clamped to bucket >= lo).
__device__ inline void make_hub_chunks(const AddsArgs& a, uint32_t u, uint32_t key, int beg, int end,
                                       uint32_t slot, uint32_t lo) {
  AddsCtrl* c = a.c;
  const uint32_t nch = (uint32_t)((end - beg + kHubChunk - 1) / kHubChunk);
  atomicAdd(&c->pending, (unsigned long long)nch);
  atomicAdd(&INFL(c, slot), (int)nch);
  const uint32_t pos = atomicAdd(&c->hub_reserve, nch);
  if ((int)(pos + nch - ld_rlx(&c->hub_head)) > (int)(a.hub_mask + 1u)) {
    atomicOr(&c->err, kErrHub);
    atomicExch(&c->done, 1u);
  }
  for (uint32_t q = 0; q < nch; ++q) {
    HubChunk* hc = a.hub + ((pos + q) & a.hub_mask);
    const int e0 = beg + (int)q * kHubChunk;
    st_rlx(&hc->u, u);
    st_rlx(&hc->key, key);
    st_rlx(&hc->first, (uint32_t)e0);
    st_rlx(&hc->cnt, (uint32_t)(min(end, e0 + kHubChunk) - e0));
    st_rlx(&hc->slot, slot);
    st_rlx(&hc->bucket, lo);
    vm_wait();  // fields before the sequence number that publishes them
    st_rlx(&hc->seq, pos + q + 1u);
  }
  vm_wait();
  atomicAdd(&c->hub_avail, (int)nch);
  vm_wait();
}

// Claim from a (ring, head, avail) set through its counting semaphore while
// holding inflight on slot `pin`; returns the count and sets *begin.
__device__ inline uint32_t sem_claim(int* avail, uint32_t* head, int* infl, uint32_t* begin,
                                     int want = (int)kAB) {
  atomic_add_done(infl, 1);  // before the claim is visible
  const int old = atomic_add_done(avail, -want);
  const int k = min(old, want);
  if (k < want) atomic_add_done(avail, want - max(k, 0));  // give back the surplus
  if (k <= 0) {
    atomic_add_done(infl, -1);
    return 0;
  }
  *begin = atomicAdd(head, (uint32_t)k);
  return (uint32_t)k;
}

// kDiag: keep the per-workgroup counters and clocks of the MI355_ADDS_DEBUG
// report (a separate instantiation, so the default kernel carries none of
// their registers).
#define DIAG(...)            \
  do {                       \
    if constexpr (kDiag) {   \
      __VA_ARGS__;           \
    }                        \
  } while (0)

template <class W, bool kDiag, bool kPull>
__global__ __launch_bounds__(kAB) __attribute__((amdgpu_waves_per_eu(4))) void k_adds(AddsArgs a) {
  __shared__ uint32_t s_slot, s_qslot, s_begin, s_count, s_kind, s_done, s_base, s_active, s_hub_ok;
  __shared__ float s_tbase, s_inv;
  __shared__ int s_incl[kAW][kWave];
  __shared__ int s_beg[kAW][kWave];
  __shared__ uint32_t s_key[kAW][kWave];
  __shared__ int s_u[kAW][kWave];
  __shared__ uint32_t s_lo[kAW][kWave];
  __shared__ int s_wsum[kAW];
  __shared__ int s_pv[kAW][kPushStage];
  __shared__ uint32_t s_ps[kAW][kPushStage];
  __shared__ uint32_t s_ph[kAW][kAddsSlots];
  __shared__ int s_hv[kAW][kHeavyStage];
  // the local queue; a pull claim (never holding local entries) uses its
  // first kAB x 8 bytes for the per-vertex best candidates
  __shared__ __align__(16) int s_lq[kLocal];
  __shared__ uint32_t s_lfree, s_lhead, s_ltail, s_lnum, s_lbeg;
  static_assert(kLocal * sizeof(int) >= kAB * sizeof(unsigned long long), "pull candidates fit the local queue");
  AddsCtrl* c = a.c;
  const unsigned lane = threadIdx.x & 63;
  const unsigned wv = threadIdx.x >> 6;
  // run start for the trace's timestamps (block 0 starts first; bucket 0 is
  // not finished before it does)
  if (blockIdx.x == 0 && threadIdx.x == 0)
    __hip_atomic_store(&c->diag[15], (unsigned long long)wall_clock64(), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x == 0) s_lfree = s_lhead = s_ltail = 0;  // ordered by the first barrier below
  PushStage st{s_pv[wv], s_ps[wv], s_ph[wv], 0,
               LocalQueue{s_lq, &s_lfree, &s_lhead, &s_ltail, kNoLocal, min(a.local_keep, kLocal)}};
  uint32_t hv_n = 0;
  unsigned long long edges_local = 0, pushed_local = 0, items_local = 0;
  uint32_t idle = 0, backoff = 1;
  unsigned long long claims = 0;
  unsigned long long d_cas = 0, d_idle = 0, d_spin = 0, d_hub = 0, d_drop = 0, d_req = 0, d_heavy = 0;
  unsigned long long t_claim = 0, t_idle = 0, t_items = 0, t_hubs = 0, t_pub = 0, t_relax = 0;
  unsigned long long t_pull = 0, d_psearch = 0, d_punreached = 0;
  // A workgroup beyond the policy's worker share is dormant: it sleeps and
  // reads only the end flag and the share, so it adds no traffic on the
  // control words that the claimers update (wave 0's copy is the one used).
  bool dormant = false;

  for (;;) {
    long long t_top = 0;
    DIAG(t_top = wall_clock64());
    if (wv == 0) {
      uint32_t got = 0, slot = 0, qslot = 0, begin = 0, kind = kClaimItems, done = 0;
      if (dormant) {
        for (int k = 0; k < 4; ++k) __builtin_amdgcn_s_sleep(127);
        int pr = 0;
        if (lane == 0) pr = (int)ld_rlx(&c->done);
        else if (lane == 1) pr = (int)ld_rlx(&c->active_workers);
        done = (uint32_t)__shfl(pr, 0, 64);
        dormant = blockIdx.x >= max((uint32_t)__shfl(pr, 1, 64), 1u);
      }
      if (!dormant && !done) {
        // Probe the control words in two wave-wide rounds of loads (the base,
        // then done / hub / heavy-list / active counts and the window's
        // semaphores on separate lanes) instead of a chain of dependent loads.
        const unsigned long long bw0 = __shfl(lane == 0 ? ld_rlx(&c->base_word) : 0ull, 0, 64);
        const uint32_t base0 = base_of(bw0);
        int probe = 0;
        if (lane == 0) probe = (int)ld_rlx(&c->done);
        else if (lane == 1) probe = ld_rlx(&c->hub_avail);
        else if (lane == 2) probe = base0 > 0 ? ld_rlx(&HV_AVAIL(c, (base0 - 1) % kAddsSlots)) : 0;
        else if (lane == 3) probe = (int)ld_rlx(&c->active);
        else if (lane >= 4 && lane < 4u + kAddsSlots) probe = ld_rlx(&AVAIL(c, (base0 + lane - 4) % kAddsSlots));
        else if (lane == 4u + kAddsSlots) probe = (int)ld_rlx(&c->active_workers);
        else if (kPull && lane == 5u + kAddsSlots) probe = (int)ld_rlx(&c->decided);
        else if (kPull && lane == 6u + kAddsSlots) probe = ld_rlx(&c->pull_avail);
        else if (kPull && lane == 7u + kAddsSlots) probe = (int)ld_rlx(&c->pull_keyed);
        else if (kPull && lane == 8u + kAddsSlots) probe = (int)ld_rlx(&c->pull_count);
        done = (uint32_t)__shfl(probe, 0, 64);
        const int hub_n = __shfl(probe, 1, 64), hv_n0 = __shfl(probe, 2, 64);
        const uint32_t A = max(1u, min((uint32_t)__shfl(probe, 3, 64), (uint32_t)kAddsSlots));
        const unsigned long long ready = __ballot(lane >= 4 && lane < 4 + A && probe > 0) >> 4;
        // slots holding many items are claimed kClaim at a time (fewer
        // claims); otherwise kAB, so that scarce work spreads over more
        // workgroups
        const unsigned long long big = __ballot(lane >= 4 && lane < 4 + A && probe >= (int)kBigClaimAvail) >> 4;
        // Workgroups beyond the policy's worker share do not claim: fewer
        // claimers contend less on the control words when a phase holds little
        // work (high-diameter graphs).  They turn dormant until the share
        // includes them again or the run ends; the claimers advance the
        // phases and detect the end.
        const uint32_t share = (uint32_t)__shfl(probe, 4 + kAddsSlots, 64);
        // the heavy list of bucket base0 - 1 is claimed once the boundary
        // decided whether it is pushed or pulled (a read of an older word only
        // delays the claim, or pushes edges a pull also covers)
        const uint32_t dw = kPull ? (uint32_t)__shfl(probe, 5 + kAddsSlots, 64) : 0u;
        const bool decided = !kPull || (dw >> 1) == base0;
        const bool pulled = kPull && decided && (dw & 1u);
        // chunks once every entry of the dropped list has its key folded in
        const int pull_n = kPull && (uint32_t)__shfl(probe, 7 + kAddsSlots, 64) >= (uint32_t)__shfl(probe, 8 + kAddsSlots, 64)
                               ? __shfl(probe, 6 + kAddsSlots, 64) : 0;
        const bool claimer = blockIdx.x < max(share, 1u);
        dormant = !claimer;
        if (lane == 0 && !done && claimer) {
          // hub chunks first (they finish claimed work), then the heavy list of
          // the bucket just finished, then the active buckets.  Every claim is a
          // counting-semaphore decrement: no retry loop for hundreds of
          // workgroups to serialise on.
          if (hub_n > 0) {
            if (atomicSub(&c->hub_avail, 1) >= 1) {
              got = 1;
              kind = kClaimHub;
              begin = atomicAdd(&c->hub_head, 1u);
            } else {
              atomicAdd(&c->hub_avail, 1);
              DIAG(++d_cas);
            }
          }
          if (!got && base0 > 0 && hv_n0 > 0 && decided) {
            const uint32_t hs = (base0 - 1) % kAddsSlots;
            // pinned on slot base: bucket base cannot finish before the heavy
            // edges of bucket base - 1 are relaxed (or, pulled, dropped)
            const uint32_t ps = base0 % kAddsSlots;
            got = sem_claim(&HV_AVAIL(c, hs), &HV_HEAD(c, hs), &INFL(c, ps), &begin);
            if (got) {
              kind = pulled ? kClaimHeavyDrop : kClaimHeavy;
              slot = ps;
              qslot = hs;
            } else {
              DIAG(++d_cas);
            }
          }
          if (kPull && !got && pulled && pull_n > 0) {
            const uint32_t ps = base0 % kAddsSlots;
            got = sem_claim(&c->pull_avail, &c->pull_head, &INFL(c, ps), &begin, 1);
            if (got) {
              kind = kClaimPull;
              slot = ps;
              qslot = ps;
            }
          }
          for (unsigned long long m = ready; m && !got; m &= m - 1) {
            const uint32_t s = (base0 + (uint32_t)__ffsll((long long)m) - 1) % kAddsSlots;
            const int want = (big >> ((uint32_t)__ffsll((long long)m) - 1)) & 1ull ? (int)kClaim : (int)kAB;
            got = sem_claim(&AVAIL(c, s), &HEAD(c, s), &INFL(c, s), &begin, want);
            if (got) {
              slot = s;
              qslot = s;
            } else {
              DIAG(++d_cas);
            }
          }
        }
        got = __shfl(got, 0, 64);
        done = __shfl(done, 0, 64);
        if (got) {
          idle = 0;
          backoff = 1;
          if (++claims > a.claim_limit) {
            if (lane == 0) {
              atomicOr(&c->err, kErrStall);
              atomicExch(&c->done, 1u);
            }
            done = 1;
          }
        }
        if (!got && !done && claimer) {
          const unsigned long long pend = __shfl(lane == 0 ? ld_rlx(&c->pending) : 0ull, 0, 64);
          if (pend == 0) {
            if (lane == 0) atomicExch(&c->done, 1u);
            done = 1;
          } else {
            const unsigned long long bw = __shfl(lane == 0 ? ld_rlx(&c->base_word) : 0ull, 0, 64);
            if (try_advance<W, kPull>(a, bw)) {
              idle = 0;
              backoff = 1;
            } else {
              // exponential back-off keeps idle workgroups from flooding the
              // control words that the busy ones update
              DIAG(++d_idle);
              for (uint32_t k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(8);
              if (backoff < 32) backoff <<= 1;
              if (++idle > kIdleLimit) {  // no progress for seconds: abort, do not hang
                if (lane == 0) {
                  atomicOr(&c->err, kErrStall);
                  atomicExch(&c->done, 1u);
                }
                done = 1;
              }
            }
          }
        }
      }  // !dormant && !done
      if (lane == 0) {
        s_done = done;
        s_count = got;
        s_slot = slot;
        s_qslot = qslot;
        s_begin = begin;
        s_kind = kind;
        if (got) {  // the window a claim is mapped against
          const unsigned long long cur = ld_rlx(&c->base_word);
          s_base = base_of(cur);
          s_tbase = tbase_of(cur);
          s_inv = 1.0f / ld_rlx(&c->delta);
          s_active = max(1u, ld_rlx(&c->active));
        }
      }
    }
    __syncthreads();
    const uint32_t count = s_count;
    long long t_claimed = 0;
    DIAG(t_claimed = wall_clock64(); if (s_count) t_claim += t_claimed - t_top; else t_idle += t_claimed - t_top);
    if (s_done) break;
    if (count == 0) {
      __syncthreads();
      continue;
    }
    const uint32_t base = s_base;
    // absolute bucket of the slot this claim holds (the window starts at base)
    const uint32_t held_bucket = base + (s_slot + kAddsSlots - base % kAddsSlots) % kAddsSlots;
    const float tbase = s_tbase, inv = s_inv;
    const uint32_t kind = s_kind;

    if (kind == kClaimHub) {
      // one hub chunk; wait until its producer has published it
      HubChunk* hc = a.hub + (s_begin & a.hub_mask);
      if (threadIdx.x == 0) {
        uint32_t spins = 0;
        while (ld_rlx(&hc->seq) != s_begin + 1u && ++spins < kSpinLimit) __builtin_amdgcn_s_sleep(1);
        s_hub_ok = spins < kSpinLimit;
        if (!s_hub_ok) {
          atomicOr(&c->err, kErrHub);
          atomicExch(&c->done, 1u);
        }
      }
      __syncthreads();
      if (!s_hub_ok) break;
      // coherent reads: the record slot may hold an older chunk in a cache
      const uint32_t hu = ld_rlx(&hc->u), hkey = ld_rlx(&hc->key), hfirst = ld_rlx(&hc->first);
      const uint32_t hcnt = min(ld_rlx(&hc->cnt), (uint32_t)kHubChunk);
      const uint32_t hslot = ld_rlx(&hc->slot) % kAddsSlots, hbucket = ld_rlx(&hc->bucket);
      // wave wv takes edges [wv * 256, wv * 256 + 256) of every 4 * kAB, four per lane
      for (uint32_t t0 = wv * 4 * kWave; t0 < hcnt; t0 += 4 * kAB) {
        uint32_t v[4], ck[4], uu[4], lo[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const uint32_t i = t0 + p * kWave + lane;
          const uint2 e = a.edges[hfirst + (i < hcnt ? i : 0u)];
          v[p] = e.x;
          ck[p] = i < hcnt ? W::relax(hkey, e.y) : W::kInfKey;
          uu[p] = hu;
          lo[p] = hbucket;
        }
        relax_quad<W, kPull>(a, st, v, ck, uu, lo, base, tbase, inv, pushed_local);
      }
      stage_publish<kPull>(a, st);
      if (a.live_edges && wv == 0 && lane == 0) atomicAdd(&c->edges, (unsigned long long)hcnt);
      edges_local += hcnt;
      DIAG(d_hub += threadIdx.x == 0);
      vm_wait();  // every wave's publishes complete before the release
      __syncthreads();
      if (threadIdx.x == 0) {
        atomic_add_done(&INFL(c, hslot), -1);
        atomicAdd(&c->pending, ~0ull);
      }
      DIAG(t_hubs += wall_clock64() - t_claimed);
      continue;
    }

    if (kPull && kind == kClaimPull) {
      // One chunk of the pull of bucket base - 1.  Vertex v reads the in-edges
      // u -> v of its weight-sorted row with pull_key (+) w <= dist(v) (every
      // vertex of the pulled bucket has a key >= pull_key, and the rows are
      // ordered by weight, so the edges past that prefix can neither lower
      // nor tie v's label): all of them while v is unreached, none when
      // dist(v) < pull_key.  The prefix ends are found by binary search.
      const uint32_t pkey = ld_rlx(&c->pull_key);
      const uint32_t q = s_begin;
      unsigned long long* s_best = reinterpret_cast<unsigned long long*>(s_lq);
      if (q < ld_rlx(&c->pull_chunks)) {
        const uint32_t v0 = q * kPullChunk, v1 = min(a.pull_n, v0 + kPullChunk);
        for (uint32_t p0 = v0; p0 < v1; p0 += kAB) {
          const uint32_t v = p0 + threadIdx.x;
          const bool has = v < v1;
          const unsigned long long lab = has ? a.labels[v] : ~0ull;
          const uint32_t key_v = (uint32_t)(lab >> 32);
          int beg = 0, len = 0;
          if (has && key_v >= pkey) {
            const int rb = a.trow[v] + a.tsplit[v], re = a.trow[v + 1];
            int lo = rb;
            // the in-edges that can still lower or tie v: pkey (+) w <= bound,
            // bound = v's key.  An unreached vertex bounds its row by the best
            // candidate of its first four heavy in-edges instead: that
            // candidate is a real path length, and an edge with
            // pkey (+) w > bound gives a candidate above it, which can
            // neither lower nor tie the label the pass leaves (its first
            // edge lies inside the bound, so the pass applies the candidate).
            uint32_t bound = key_v;
            if (key_v == W::kInfKey && rb < re) {
              uint2 e4[4];
#pragma unroll
              for (int p = 0; p < 4; ++p) e4[p] = a.tedges[min(rb + p, re - 1)];
              unsigned long long l4[4];
#pragma unroll
              for (int p = 0; p < 4; ++p)
                l4[p] = __hip_atomic_load(a.labels + e4[p].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
              for (int p = 0; p < 4; ++p) {
                const uint32_t ku = (uint32_t)(l4[p] >> 32);
                if (rb + p < re && ku != W::kInfKey) bound = min(bound, W::relax(ku, e4[p].y));
              }
            }
            if (bound == W::kInfKey) {
              lo = re;
              DIAG(d_punreached += (unsigned long long)(re - rb));
            } else if (rb < re) {
              // the first four heavy in-edges in one round trip (most prefixes
              // end there; the weights ascend, so the reaching ones lead)
              uint32_t w4[4];
#pragma unroll
              for (int p = 0; p < 4; ++p) w4[p] = a.tedges[min(rb + p, re - 1)].y;
#pragma unroll
              for (int p = 0; p < 4; ++p) lo += (rb + p < re && W::relax(pkey, w4[p]) <= bound) ? 1 : 0;
              // all four reach: search the rest for the end
              int hi = lo == rb + 4 ? re : lo;
              while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (W::relax(pkey, a.tedges[mid].y) <= bound) lo = mid + 1;
                else hi = mid;
              }
              DIAG(if (key_v == W::kInfKey) d_punreached += (unsigned long long)(lo - rb));
            }
            beg = rb;
            len = lo - rb;
            DIAG(++d_psearch);
          }
          pull_pass<W, kDiag>(a, st, &s_incl[0][0], &s_beg[0][0], &s_key[0][0], s_best, s_wsum, has, v, lab, beg,
                       len, base, tbase, inv, edges_local, pushed_local);
        }
      }
      stage_publish<true>(a, st);
      vm_wait();  // every wave's publishes complete before the release
      __syncthreads();
      if (threadIdx.x == 0) {
        // the chunk's pushes have completed (each wave waited above)
        atomicAdd(&c->pull_done, 1u);
        atomic_add_done(&INFL(c, s_slot), -1);
        atomicAdd(&c->pending, ~0ull);
      }
      DIAG(t_pull += wall_clock64() - t_claimed);
      continue;
    }

    // items of a bucket ring or a heavy list, in passes of kAB: in pass p
    // wave w takes claim entries [p*kAB + w*64, p*kAB + w*64 + 64)
    const bool heavy = kind == kClaimHeavy || kind == kClaimHeavyDrop;
    // the bucket of a dropped heavy list was pulled: its entries are consumed
    // and their list stamps cleared, their edges not relaxed
    const bool drop = kPull && kind == kClaimHeavyDrop;
    const uint32_t slot = s_slot, begin = s_begin;
    // entries for other buckets are offered between passes once this many
    // are staged.  A light claim in a one-bucket window holds the base bucket:
    // nothing else is claimable until it finishes, so its other entries wait
    // for the end of the claim (or a full stage).
    const uint32_t publish_at = (!heavy && s_active == 1) ? kPushStage : kPublishAt;
    int* ring = (heavy ? a.hring : a.ring) + (size_t)s_qslot * (a.cap_mask + 1u);
    // improved vertices of the held bucket stay in this workgroup (item
    // claims), up to local_keep at a time; the rest go to the ring
    st.lq.slot = (!heavy && a.local_keep > 0) ? slot : kNoLocal;
    const uint32_t ring_passes = (count + kAB - 1) / kAB;
    for (uint32_t pass = 0;; ++pass) {
      // passes over the claimed ring entries, then over the local queue until
      // it is empty
      const bool from_ring = pass < ring_passes;
      uint32_t local_n = 0, local_beg = 0;
      if (!from_ring) {
        __syncthreads();  // the previous pass has read its entries and appended
        if (threadIdx.x == 0) {
          const uint32_t h = s_lhead;
          s_lfree = h;
          s_lnum = min(s_ltail - h, (uint32_t)kAB);
          s_lbeg = h;
          s_lhead = h + s_lnum;
        }
        __syncthreads();
        local_n = s_lnum;
        local_beg = s_lbeg;
        if (local_n == 0) break;
      }
      const uint32_t i = from_ring ? pass * kAB + wv * kWave + lane : wv * kWave + lane;
      const bool has = from_ring ? i < count : i < local_n;
      int u = 0, beg = 0, deg = 0;
      uint32_t key = W::kInfKey, bk = 0;
      bool entry_ok = true;
      bool requeue = false, to_heavy = false;
      uint32_t want = 0;
      if (has && !from_ring) {
        u = s_lq[(local_beg + i) & (kLocal - 1u)];
        entry_ok = u >= 0 && (uint32_t)u < a.n;
        if (!entry_ok) {
          atomicOr(&c->err, kErrRing);
          atomicExch(&c->done, 1u);
          u = 0;
        }
      } else if (has) {
        // The producer reserved this position before writing it: wait for the
        // write (bounded), then hand the entry back to the ring as empty.
        int* ep = ring + ((begin + i) & a.cap_mask);
        uint32_t spins = 0;
        while ((u = __hip_atomic_load(ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == kEmptyEntry &&
               ++spins < kSpinLimit)
          __builtin_amdgcn_s_sleep(1);
        DIAG(d_spin += spins);
        entry_ok = u != kEmptyEntry && u >= 0 && (uint32_t)u < a.n;
        if (!entry_ok) {
          atomicOr(&c->err, kErrRing);
          atomicExch(&c->done, 1u);
          u = 0;
        } else {
          __hip_atomic_store(ep, kEmptyEntry, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      // row bounds and light-prefix length of u depend only on u: issued now,
      // they arrive while the stamp exchange and the label read complete
      int r_beg = 0, r_end = 0, r_split = 0, pu = u;
      if (has && entry_ok) {
        r_beg = a.rowptr[u];
        r_end = a.rowptr[u + 1];
        r_split = a.split[u];
        if (a.ids) pu = a.ids[u];
      }
      if (has && entry_ok) {
        if (drop) {
          // a dropped entry keeps its list stamp (bucket base - 1, finished:
          // no later pass compares equal to it); the pull bound came from the
          // bucket's light passes
        } else if (heavy) {
          // Clearing the stamp first means a later light pass of u (after an
          // improvement) lists u again, or this pass reads the label it set.
          __hip_atomic_exchange(a.sstamp + u, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          vm_wait();  // the label is read after the stamp is cleared
          key = atomic_load_key(a.labels + u);
          bk = base;
          DIAG(++d_heavy);
        } else {
          // Whoever takes the stamp processes the vertex.  (Putting back a
          // stamp that belongs to an entry in another ring would race with that
          // entry's own pop and could lose the vertex.)  The bucket is the one
          // this claim holds; an entry whose stamp was already taken is stale.
          const uint32_t old = __hip_atomic_exchange(a.qstamp + u, kNoStamp, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
          if (old != kNoStamp) {
            bk = held_bucket;
            key = atomic_load_key(a.labels + u);
            want = map_bucket(W::bucket_value(key), base, tbase, inv, bk);
            if (want >= bk + s_active) {
              // clipped or re-mapped far item: queue it where it belongs now
              key = W::kInfKey;
              requeue = true;
              DIAG(++d_req);
            }
          } else {
            DIAG(++d_drop);
          }
        }
        if (key != W::kInfKey) {
          ++items_local;
          beg = r_beg;
          int end = r_end;
          const int split = beg + r_split;
          if (heavy) {
            beg = split;
          } else {
            // light prefix now; the heavy suffix once bucket bk is finished
            if (split < end)
              to_heavy = __hip_atomic_exchange(a.sstamp + u, bk + 1u, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT) != bk + 1u;
            end = split;
          }
          deg = end - beg;
          if (deg > kHubDegree) {
            make_hub_chunks(a, (uint32_t)pu, key, beg, end, slot, bk);
            deg = 0;
          }
        }
      }
      if constexpr (kPull) {
        // the bucket's pull bound: the smallest key with which a vertex that
        // has heavy edges is processed here (whether newly listed or already
        // on the list: a listed vertex improved later is processed again,
        // here or in a later bucket that lists it)
        if (!heavy) {
          uint32_t km = (key != W::kInfKey && r_split < r_end - r_beg) ? key : kNoStamp;
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) km = min(km, (uint32_t)__shfl_xor(km, off, 64));
          if (lane == 0 && km != kNoStamp) atomicMin(&HV_MIN(c, held_bucket % kAddsSlots), km);
        }
      }
      // heavy-list entries of this claim's bucket (list slot = bucket slot)
      {
        const unsigned long long hm = __ballot(to_heavy);
        if (hm) {
          const uint32_t k = (uint32_t)__popcll(hm);
          if (hv_n + k > kHeavyStage) heavy_publish(a, s_hv[wv], hv_n, slot);
          if (to_heavy) s_hv[wv][hv_n + prefix_count(hm)] = u;
          hv_n += k;
        }
      }
      // re-queued items go through the stage like any push
      push_improved<kPull>(a, st, requeue, (uint32_t)u, want, pushed_local);
      long long t_r0 = 0, t_r1 = 0;
      DIAG(t_r0 = wall_clock64());
      relax_block<W, kPull>(a, st, &s_incl[0][0], &s_beg[0][0], &s_key[0][0], &s_u[0][0], &s_lo[0][0], s_wsum,
                     pu, beg, deg, key, bk, base, tbase, inv, edges_local, pushed_local);
      DIAG(t_r1 = wall_clock64());
      // Between passes only the held bucket's entries move (to the local
      // queue: the next pass); entries for other buckets are offered once
      // publish_at of them are staged, heavy-list entries (claimable only
      // after the bucket) at the end of the claim.  A hop of a local chain
      // then costs its relaxation's round trips, not a publish's too.
      stage_keep_local(st);
      if (st.n >= publish_at) stage_publish<kPull>(a, st);
      DIAG(t_relax += t_r1 - t_r0; t_pub += wall_clock64() - t_r1);
    }
    // everything staged is offered to the rings before the claim is released
    // (the local queue is no longer read: entries of the held bucket too)
    st.lq.slot = kNoLocal;
    stage_publish<kPull>(a, st);
    heavy_publish(a, s_hv[wv], hv_n, slot);
    vm_wait();  // every wave's publishes complete before the release
    __syncthreads();
    if (threadIdx.x == 0) {
      atomic_add_done(&INFL(c, slot), -1);
      atomicAdd(&c->pending, 0ull - (unsigned long long)count);
    }
    DIAG(t_items += wall_clock64() - t_claimed);
  }
  if constexpr (kDiag) {
    if (threadIdx.x == 0) {
      atomicAdd(&c->diag[8], t_claim);
      atomicAdd(&c->diag[9], t_idle);
      atomicAdd(&c->diag[10], t_items);
      atomicAdd(&c->diag[11], t_hubs);
      atomicAdd(&c->diag[12], t_pub);
      atomicAdd(&c->diag[13], t_relax);
    }
    if (threadIdx.x == 0 && claims) atomicAdd(&c->diag[0], claims);
    if (d_cas) atomicAdd(&c->diag[1], d_cas);
    if (threadIdx.x == 0 && d_idle) atomicAdd(&c->diag[2], d_idle);
    if (d_spin) atomicAdd(&c->diag[3], d_spin);
    if (d_hub) atomicAdd(&c->diag[4], d_hub);
    if (d_drop) atomicAdd(&c->diag[5], d_drop);
    if (d_req) atomicAdd(&c->diag[6], d_req);
    if (d_heavy) atomicAdd(&c->diag[7], d_heavy);
    if (threadIdx.x == 0 && t_pull) atomicAdd(&c->pdiag[0], t_pull);
    if (d_psearch) atomicAdd(&c->pdiag[2], d_psearch);
    if (d_punreached) atomicAdd(&c->pdiag[3], d_punreached);
  }
  if (lane == 0) {
    if (!a.live_edges && edges_local) atomicAdd(&c->edges, edges_local);
    if (pushed_local) atomicAdd(&c->pushes, pushed_local);
  }
  if (items_local) atomicAdd(&c->items, items_local);
}

// The light width of a run under the delta network: its first bucket width
// (c->delta after k_adds_first_phase; `fallback` when not a positive finite
// number) times the caller's light_scale, or the network's light width when
// the caller set none (outside (0, 1]: 1).  sel[0] / sel[2]: the width each
// split (out-rows / transposed in-rows) holds; sel[1] / sel[3] become nonzero
// when that split must be recounted, and the held width is updated (the
// transposed one only when `transposed`).  The recount reads the width from
// sel[0] (sel[2]), which this kernel leaves equal to the new width.
__global__ void k_split_width(const AddsCtrl* __restrict__ c, float fallback, float req_light, int transposed,
                              uint32_t* __restrict__ sel) {
  float net_light = c->light_scale;
  if (!(net_light > 0.0f && net_light <= 1.0f)) net_light = 1.0f;
  const float first = c->delta;
  const float width = (first > 0.0f && isfinite(first) ? first : fallback) * (req_light > 0.0f ? req_light : net_light);
  const uint32_t bits = __float_as_uint(width);
  sel[1] = sel[0] != bits ? 1u : 0u;
  sel[0] = bits;
  if (transposed) {
    sel[3] = sel[2] != bits ? 1u : 0u;
    sel[2] = bits;
  }
}

template <class W>
__global__ void k_adds_init(AddsArgs a, int64_t n, uint32_t src, const uint32_t* __restrict__ init,
                            const AddsCtrl* __restrict__ init_ctrl) {
  const unsigned long long inf_label = ((unsigned long long)W::kInfKey << 32) | kNoPred;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t v = tid; v < n; v += stride) {
    unsigned long long l = inf_label;
    if (init) l = ((unsigned long long)W::output_to_key(init[a.ids ? a.ids[v] : v]) << 32) | kNoPred;
    a.labels[v] = v == (int64_t)src ? (((unsigned long long)W::kZeroKey << 32) | kNoPred) : l;
    a.qstamp[v] = v == (int64_t)src ? 0u : kNoStamp;
    a.sstamp[v] = 0u;
  }
  for (int64_t i = tid; i <= (int64_t)a.hub_mask; i += stride) a.hub[i].seq = 0u;
  // the header comes from the host's copy; the slot arrays start empty (slot
  // 0 holding the source: tail and avail 1) with no pull bound
  const uint32_t* src_words = reinterpret_cast<const uint32_t*>(init_ctrl);
  uint32_t* dst_words = reinterpret_cast<uint32_t*>(a.c);
  constexpr int64_t kHdr = (int64_t)(kCtrlHeader / 4);
  constexpr int64_t kTail0 = (int64_t)(offsetof(AddsCtrl, tail) / 4);
  constexpr int64_t kAvail0 = (int64_t)(offsetof(AddsCtrl, avail) / 4);
  constexpr int64_t kMin0 = (int64_t)(offsetof(AddsCtrl, hv_min) / 4);
  for (int64_t i = tid; i < (int64_t)(sizeof(AddsCtrl) / 4); i += stride) {
    uint32_t wv = 0u;
    if (i < kHdr) wv = src_words[i];
    else if (i == kTail0 || i == kAvail0) wv = 1u;
    else if (i >= kMin0 && (i - kMin0) % kPad == 0) wv = kNoStamp;
    dst_words[i] = wv;
  }
  if (tid == 0) a.ring[0] = (int)src;
}

template <class W>
void run_adds_typed(mi355_graph& g, uint32_t src, const RunRequest& req, RunResult& res,
                    hipStream_t s, const AddsPolicy* d_policy) {
  Workspace& ws = g.ws;
  const int64_t n = g.n;
  const float hdelta = heuristic_delta(g);
  float delta = req.delta > 0.0f ? req.delta : hdelta;
  if (!(delta > 0.0f) || !std::isfinite(delta)) fail(MI355_EINVAL, "delta must be a finite positive float");
  uint32_t cap = 1024;
  while ((int64_t)cap < n + 1024) cap <<= 1;
  uint32_t hub_cap = 1024;
  while ((int64_t)hub_cap < g.info.large_chunks * 2 + 1024) hub_cap <<= 1;
  ws.labels.ensure((size_t)n);
  if (ws.adds_ring.n < (size_t)2 * kAddsSlots * cap) {
    ws.adds_ring.reset((size_t)2 * kAddsSlots * cap);
    ws.adds_ring_clean = false;
  }
  ws.adds_qstamp.ensure((size_t)n * 2);
  ws.adds_hub.ensure(hub_cap);
  ws.adds_ctrl_mem.ensure(2 * sizeof(AddsCtrl));
  AddsCtrl* c = reinterpret_cast<AddsCtrl*>(ws.adds_ctrl_mem.p);
  AddsCtrl* c_init = c + 1;
  const DeviceProps p = device_props();
  const int workers = req.worker_blocks > 0 ? req.worker_blocks : p.cus * 2;
  // the host image of the control block: its header is filled here, uploaded,
  // and read back after the run
  std::vector<unsigned char>& hbuf = ws.adds_ctrl_host;
  if (hbuf.size() != sizeof(AddsCtrl)) hbuf.assign(sizeof(AddsCtrl), 0);
  else std::memset(hbuf.data(), 0, kCtrlHeader);
  AddsCtrl& h = *reinterpret_cast<AddsCtrl*>(hbuf.data());
  h.base_word = 0;  // bucket 0 starts at distance 0.0f
  h.delta = delta;
  // the policies scale the heuristic width: the MLP's output and feature 15
  // are log2(delta / delta0)
  h.base_delta = hdelta;
  h.min_delta = std::min(delta, hdelta) / 64.0f;
  h.max_delta = std::max(delta, hdelta) * 64.0f;
  h.util_low = 0.35f;
  h.util_high = 0.85f;
  h.growth = 1.5f;
  h.shrink = 0.75f;
  if (req.pull != 0 && req.pull != 1) fail(MI355_EINVAL, "pull must be 0 (push) or 1 (pull when cheaper)");
  if (!std::isfinite(req.pull_margin)) fail(MI355_EINVAL, "pull_margin must be finite");
  if (!(req.light_scale >= 0.0f) || !std::isfinite(req.light_scale))
    fail(MI355_EINVAL, "light_scale must be finite and nonnegative");
  h.decided = 0;
  h.pull_mode = (uint32_t)req.pull;
  h.pull_margin = req.pull_margin;
  h.pull_base = kNoStamp;
  h.light_scale = 1.0f;
  h.active = (uint32_t)std::min(std::max(req.active_buckets, 1), kAddsSlots);
  h.workers = (uint32_t)workers;
  h.active_workers = (uint32_t)workers;
  h.pending = 1;
  // the slot arrays (tail[0] = avail[0] = 1 for the source, hv_min empty) are
  // set by k_adds_init; only the header travels
  AddsArgs a{};
  a.rowptr = g.rowptr.p;
  a.edges = g.edges.p;
  a.labels = ws.labels.p;
  a.qstamp = ws.adds_qstamp.p;
  a.sstamp = ws.adds_qstamp.p + n;
  a.ring = ws.adds_ring.p;
  a.hring = ws.adds_ring.p + (size_t)kAddsSlots * cap;
  a.cap_mask = cap - 1;
  // a caller's per-bucket capacity counts every queued entry: with one, all
  // entries go through the rings
  a.ring_limit = req.bucket_capacity > 0 && req.bucket_capacity < (int64_t)cap ? (uint32_t)req.bucket_capacity : cap;
  a.local_keep = a.ring_limit < cap ? 0u : kLocalKeep;
  // light/heavy split at the starting bucket width (MLP mode: at the
  // network's first width, set below; one cached split per graph either way)
  // light edges: weight <= light_scale x the first bucket width
  const float lscale = req.light_scale > 0.0f ? req.light_scale : 1.0f;
  a.split = req.use_mlp ? nullptr : light_split(g, delta * lscale, s);
  float split_delta = delta * lscale;
  a.hub = ws.adds_hub.p;
  a.hub_mask = hub_cap - 1;
  a.c = c;
  a.policy = d_policy;
  a.ids = input_ids(g);
  a.trow = nullptr;
  a.tedges = nullptr;
  a.tsplit = nullptr;
  if (req.pull) {
    pull_rows(g, &a.trow, &a.tedges, s);
    a.pull_n = (uint32_t)g.pull_n;
  }
  a.max_degree = (uint32_t)std::max<int32_t>(g.info.max_degree, 0);
  a.n = (uint32_t)n;
  a.live_edges = req.trace ? 1 : 0;
  a.claim_limit = 16ull * (unsigned long long)(n + g.m) + (1ull << 20);
  MI_HIP(hipMemcpyAsync(c_init, hbuf.data(), kCtrlHeader, hipMemcpyHostToDevice, s));
  if (!ws.adds_ring_clean) {
    // every entry starts as the empty sentinel; a completed run leaves the
    // ring in that state again (all pushed entries are consumed)
    MI_HIP(hipMemsetAsync(ws.adds_ring.p, 0xFF, ws.adds_ring.n * sizeof(int), s));
    ws.adds_ring_clean = true;
  }

  for (hipEvent_t& e : ws.adds_ev)
    if (!e) MI_HIP(hipEventCreate(&e));
  hipEvent_t ev0 = ws.adds_ev[0], evk = ws.adds_ev[1], ev1 = ws.adds_ev[2];
  MI_HIP(hipEventRecord(ev0, s));
  k_adds_init<W><<<std::min<unsigned>(div_up((uint64_t)std::max<int64_t>(n, hub_cap), kBlock), 4096u),
                   kBlock, 0, s>>>(a, n, src, req.init_dist, c_init);
  MI_LAUNCH_CHECK();
  if (req.use_mlp) {
    k_adds_first_phase<<<1, kWave, 0, s>>>(a);
    MI_LAUNCH_CHECK();
    // The light prefix of each (weight-sorted) row follows the network's
    // first width, not the heuristic one: a split at a wider delta makes the
    // light pass relax edges of up to that width eagerly from vertices that
    // are not settled yet (on R-MAT scale 22, 1.4x the relaxations); the
    // network also scales it (its light width output), unless the caller
    // fixed light_scale.  The split is cached per width, so runs on one graph
    // compute it once.
    // The width is chosen on the device (k_split_width) and the splits are
    // recounted only when it changed, so the host does not wait for the
    // first phase.
    ws.mlp_split.ensure((size_t)n);
    ws.mlp_sel.ensure(4);
    if (!ws.mlp_sel_ready) {
      MI_HIP(hipMemsetAsync(ws.mlp_sel.p, 0xFF, 4 * sizeof(uint32_t), s));  // no width held yet
      ws.mlp_sel_ready = true;
    }
    const bool transposed = a.trow && g.pull_state == 2;
    if (transposed) ws.mlp_tsplit.ensure((size_t)n);
    k_split_width<<<1, 1, 0, s>>>(c, delta, req.light_scale, transposed ? 1 : 0, ws.mlp_sel.p);
    MI_LAUNCH_CHECK();
    const unsigned lgrid = std::min<unsigned>(div_up((uint64_t)n, kBlock), 2048u);
    k_light_count_sel<W><<<lgrid, kBlock, 0, s>>>(g.rowptr.p, g.edges.p, n, ws.mlp_sel.p, ws.mlp_split.p);
    MI_LAUNCH_CHECK();
    a.split = ws.mlp_split.p;
    if (a.trow) {
      // the pull reads the heavy in-edges at the same split width
      if (transposed) {
        k_light_count_sel<W><<<lgrid, kBlock, 0, s>>>(g.t_rowptr.p, g.t_edges.p, n, ws.mlp_sel.p + 2,
                                                       ws.mlp_tsplit.p);
        MI_LAUNCH_CHECK();
        a.tsplit = ws.mlp_tsplit.p;
      } else {
        a.tsplit = ws.mlp_split.p;
      }
    }
  } else if (a.trow) {
    // the pull reads the heavy in-edges at the same split width
    a.tsplit = pull_light_split(g, split_delta, s);
  }
  MI_HIP(hipEventRecord(evk, s));
  const bool dbg = std::getenv("MI355_ADDS_DEBUG") != nullptr;
  if (a.trow) {
    if (dbg) k_adds<W, true, true><<<workers, kAB, 0, s>>>(a);
    else k_adds<W, false, true><<<workers, kAB, 0, s>>>(a);
  } else {
    if (dbg) k_adds<W, true, false><<<workers, kAB, 0, s>>>(a);
    else k_adds<W, false, false><<<workers, kAB, 0, s>>>(a);
  }
  MI_LAUNCH_CHECK();
  MI_HIP(hipEventRecord(ev1, s));
  if (req.out_dist || req.out_pred || req.out_reach) {
    // the caller's output conversion runs right behind the kernel
    labels_to_output_internal(g, src, req.out_inf_mode, req.out_dist, req.out_pred, req.out_reach, s);
    res.converted = true;
  }
  MI_HIP(hipEventSynchronize(ev1));
  MI_HIP(hipEventElapsedTime(&res.device_ms, ev0, ev1));
  MI_HIP(hipEventElapsedTime(&res.kernel_ms, evk, ev1));
  MI_HIP(hipMemcpyAsync(hbuf.data(), c, kCtrlHeader, hipMemcpyDeviceToHost, s));
  MI_HIP(hipStreamSynchronize(s));
  if (h.err) ws.adds_ring_clean = false;
  if (std::getenv("MI355_ADDS_DEBUG"))
    std::fprintf(stderr,
                 "adds: %.3f ms phases %u pulls %u items %llu edges %llu pushes %llu | claims %llu cas_fail %llu "
                 "idle %llu ring_wait %llu hub %llu dropped %llu requeued %llu heavy %llu\n"
                 "  per-workgroup ms: claim %.2f idle %.2f items %.2f hubs %.2f pulls %.2f (publish %.2f relax %.2f)"
                 " | phase advances %.3f ms in total | pulled in-edges %llu (of unreached vertices %llu) searched vertices %llu\n",
                 res.device_ms, h.phases, h.pulls, h.items, h.edges, h.pushes, h.diag[0], h.diag[1], h.diag[2],
                 h.diag[3], h.diag[4], h.diag[5], h.diag[6], h.diag[7], h.diag[8] * 1e-5 / workers,
                 h.diag[9] * 1e-5 / workers, h.diag[10] * 1e-5 / workers, h.diag[11] * 1e-5 / workers,
                 h.pdiag[0] * 1e-5 / workers, h.diag[12] * 1e-5 / workers, h.diag[13] * 1e-5 / workers,
                 h.diag[14] * 1e-5, h.pdiag[1], h.pdiag[3], h.pdiag[2]);
  if (h.err & kErrRing)
    fail(MI355_ECAPACITY, a.ring_limit < cap ? "ADDS bucket overflow: a bucket exceeded the requested per-bucket capacity"
                                             : "ADDS bucket queue overflow");
  if (h.err & kErrHub) fail(MI355_ECAPACITY, "ADDS hub chunk queue overflow");
  if (h.err & kErrStall) fail(MI355_EDEVICE, "ADDS workers stalled (aborted by the spin bound)");
  res.iterations_used = h.phases;
  res.converged = h.done != 0 && h.pending == 0;
  res.rounds = h.phases;
  res.syncs = 1;
  res.edges = (int64_t)h.edges;
  res.vertices = (int64_t)h.items;
  res.initial_delta = delta;
  res.final_delta = h.delta;
  res.pulls = h.pulls;
}

}  // namespace

void run_adds(mi355_graph& g, uint32_t source, const RunRequest& req, RunResult& res,
              hipStream_t stream, const AddsPolicy* d_policy) {
  if (g.info.has_negative) fail(MI355_EINVAL, "ADDS requires nonnegative edge weights");
  if (g.weight_type == MI355_W_F32) run_adds_typed<WF32>(g, source, req, res, stream, d_policy);
  else run_adds_typed<WU32>(g, source, req, res, stream, d_policy);
}

}  // namespace mi355
