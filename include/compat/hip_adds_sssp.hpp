// Drop-in for HIP_kernel/faster_GPU_algo/jasper_oneshot_untested/
// hip_adds_sssp.hpp (:1-115) over the mi355sssp C ABI: same constants, views,
// options, statistics and Status codes; distances are written to the
// caller-owned device buffer (INF = +inf for float, 0xFFFFFFFF for uint32).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

#include "mi355sssp.h"

namespace hip_adds {

static constexpr uint32_t kBucketCount = 32u;
static constexpr uint32_t kPageWords = 1u << 16;
static constexpr uint32_t kSegmentWords = 64u;
static constexpr uint32_t kThreadsPerBlock = 256u;

template <typename Weight>
struct DeviceCsrGraph {
  uint32_t num_vertices = 0;
  uint32_t num_edges = 0;
  const uint32_t* row_offsets = nullptr;
  const uint32_t* col_indices = nullptr;
  const Weight* weights = nullptr;
};

struct Options {
  uint32_t worker_blocks = 0;
  uint32_t assignment_chunk = 256;
  uint32_t initial_active_buckets = 1;
  uint32_t max_active_buckets = kBucketCount;
  double initial_delta = 0.0;
  double initial_delta_scale = 16.0;
  double min_delta = 1.0e-3;
  double delta_growth = 1.5;
  double delta_shrink = 0.75;
  double clip_threshold = 0.65;
  double utilization_low = 0.35;
  double utilization_high = 0.85;
  uint32_t adjust_period_head_switches = 8;
  bool dynamic_delta = true;
  bool dynamic_active_buckets = true;
  size_t pool_words = 0;
  double memory_fraction = 0.45;
  bool initialize_distances = true;
};

struct RunStats {
  double initial_delta = 0.0;
  double final_delta = 0.0;
  double final_bucket_base = 0.0;
  uint32_t worker_blocks = 0;
  uint32_t final_active_buckets = 0;
  uint32_t pool_pages = 0;
  uint32_t device_error = 0;
  uint64_t assigned_items = 0;
  uint64_t processed_items = 0;
  uint64_t enqueued_items = 0;
  uint64_t edge_relaxations = 0;
  uint64_t successful_relaxations = 0;
  uint64_t head_switches = 0;
  uint64_t delta_updates = 0;
  uint64_t max_inflight_items = 0;
};

enum class Status : int {
  Success = 0,
  InvalidArgument,
  HipError,
  OutOfMemory,
  PoolExhausted,
  BucketOverflow,
  DeviceError
};

inline const char* status_string(Status status) {
  switch (status) {
    case Status::Success: return "success";
    case Status::InvalidArgument: return "invalid argument";
    case Status::HipError: return "HIP runtime error";
    case Status::OutOfMemory: return "device out of memory";
    case Status::PoolExhausted: return "ADDS worklist pool exhausted";
    case Status::BucketOverflow: return "ADDS bucket exceeded 32-bit/64K-page addressing capacity";
    case Status::DeviceError: return "ADDS device-side error";
    default: return "unknown";
  }
}

namespace detail {
inline mi355_adds_options to_c(const Options& o) {
  mi355_adds_options c{};
  c.worker_blocks = o.worker_blocks;
  c.assignment_chunk = o.assignment_chunk;
  c.initial_active_buckets = o.initial_active_buckets;
  c.max_active_buckets = o.max_active_buckets;
  c.initial_delta = o.initial_delta;
  c.initial_delta_scale = o.initial_delta_scale;
  c.min_delta = o.min_delta;
  c.delta_growth = o.delta_growth;
  c.delta_shrink = o.delta_shrink;
  c.clip_threshold = o.clip_threshold;
  c.utilization_low = o.utilization_low;
  c.utilization_high = o.utilization_high;
  c.adjust_period_head_switches = o.adjust_period_head_switches;
  c.dynamic_delta = o.dynamic_delta ? 1 : 0;
  c.dynamic_active_buckets = o.dynamic_active_buckets ? 1 : 0;
  c.initialize_distances = o.initialize_distances ? 1 : 0;
  c.pool_words = o.pool_words;
  c.memory_fraction = o.memory_fraction;
  return c;
}
inline void from_c(const mi355_adds_stats& c, RunStats* s) {
  if (!s) return;
  s->initial_delta = c.initial_delta;
  s->final_delta = c.final_delta;
  s->final_bucket_base = c.final_bucket_base;
  s->worker_blocks = c.worker_blocks;
  s->final_active_buckets = c.final_active_buckets;
  s->pool_pages = c.pool_pages;
  s->device_error = c.device_error;
  s->assigned_items = c.assigned_items;
  s->processed_items = c.processed_items;
  s->enqueued_items = c.enqueued_items;
  s->edge_relaxations = c.edge_relaxations;
  s->successful_relaxations = c.successful_relaxations;
  s->head_switches = c.head_switches;
  s->delta_updates = c.delta_updates;
  s->max_inflight_items = c.max_inflight_items;
}
template <typename W>
inline Status run(int wtype, const DeviceCsrGraph<W>& g, uint32_t source, W* d, RunStats* stats,
                  const Options& options, hipStream_t stream) {
  const mi355_adds_options o = to_c(options);
  mi355_adds_stats st{};
  const int rc = mi355_hip_adds_sssp(wtype, g.num_vertices, g.num_edges, g.row_offsets,
                                     g.col_indices, g.weights, source, d, stats ? &st : nullptr,
                                     &o, stream);
  if (rc == 0) from_c(st, stats);
  return static_cast<Status>(rc);
}
}  // namespace detail

inline Status sssp(const DeviceCsrGraph<float>& graph, uint32_t source, float* d_distances,
                   RunStats* stats = nullptr, const Options& options = Options{},
                   hipStream_t stream = nullptr) {
  return detail::run(MI355_W_F32, graph, source, d_distances, stats, options, stream);
}

inline Status sssp(const DeviceCsrGraph<uint32_t>& graph, uint32_t source, uint32_t* d_distances,
                   RunStats* stats = nullptr, const Options& options = Options{},
                   hipStream_t stream = nullptr) {
  return detail::run(MI355_W_U32, graph, source, d_distances, stats, options, stream);
}

}  // namespace hip_adds
