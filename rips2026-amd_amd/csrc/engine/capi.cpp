// C ABI of libmi355sssp.so (declared in include/mi355sssp.h).
//
// Every entry point runs inside guard(): C++ exceptions become status codes
// and a thread-local message, so nothing throws across the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "graph.hpp"
#include "sssp.hpp"

namespace {

thread_local std::string g_err;
thread_local int g_hip_err = 0;

template <typename F>
int guard(F&& f) {
  try {
    f();
    return MI355_OK;
  } catch (const mi355::Error& e) {
    g_err = e.what();
    g_hip_err = e.hip;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory";
    return MI355_ENOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return MI355_EINVAL;
  }
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

using mi355::fail;

struct GraphDeleter {
  void operator()(mi355_graph* g) const { delete g; }
};
using GraphPtr = std::unique_ptr<mi355_graph, GraphDeleter>;

// First round k >= 1 whose output frontier is empty (0 if none was observed).
int64_t first_empty_round(const std::vector<uint32_t>& log) {
  for (size_t k = 1; k < log.size(); ++k)
    if (log[k] == 0) return (int64_t)k;
  return 0;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void download_dist(mi355_graph& g, uint32_t source, int inf_mode, float* host_dist,
                   hipStream_t s) {
  mi355::DevBuf<uint32_t> d((size_t)g.n);
  mi355::labels_to_output(g, source, inf_mode, d.p, nullptr, nullptr, s);
  MI_HIP(hipMemcpyAsync(host_dist, d.p, (size_t)g.n * 4, hipMemcpyDeviceToHost, s));
  MI_HIP(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

const char* mi355_last_error(void) { return g_err.c_str(); }
int mi355_internal_set_error(int code, const char* msg, int hip) {
  g_err = msg ? msg : "";
  g_hip_err = hip;
  return code;
}
int mi355_last_hip_error(void) { return g_hip_err; }

// ---- device plumbing --------------------------------------------------------
int mi355_device_count(int* count) {
  return guard([&] {
    if (!count) fail(MI355_EINVAL, "null count");
    MI_HIP(hipGetDeviceCount(count));
  });
}
int mi355_set_device(int device) { return guard([&] { MI_HIP(hipSetDevice(device)); }); }
int mi355_device_malloc(void** ptr, size_t bytes) {
  return guard([&] {
    if (!ptr) fail(MI355_EINVAL, "null pointer");
    *ptr = nullptr;
    if (bytes) MI_HIP(hipMalloc(ptr, bytes));
  });
}
int mi355_device_free(void* ptr) { return guard([&] { MI_HIP(hipFree(ptr)); }); }
int mi355_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  return guard([&] {
    MI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    MI_HIP(hipStreamSynchronize(as_stream(stream)));
  });
}
int mi355_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  return guard([&] {
    MI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    MI_HIP(hipStreamSynchronize(as_stream(stream)));
  });
}
int mi355_device_gather_u64(void* dst, const void* src, const uint32_t* idx, int64_t count, void* stream) {
  return guard([&] { mi355::device_gather8(dst, src, idx, count, as_stream(stream)); });
}
int mi355_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  return guard([&] { mi355::device_copy(dst, src, bytes, as_stream(stream)); });
}
int mi355_memset_d8(void* dst, int value, size_t bytes, void* stream) {
  return guard([&] { MI_HIP(hipMemsetAsync(dst, value, bytes, as_stream(stream))); });
}
int mi355_stream_create(void** stream) {
  return guard([&] {
    if (!stream) fail(MI355_EINVAL, "null stream pointer");
    hipStream_t s;
    MI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
  });
}
int mi355_stream_destroy(void* stream) {
  return guard([&] { MI_HIP(hipStreamDestroy(as_stream(stream))); });
}
int mi355_stream_sync(void* stream) {
  return guard([&] { MI_HIP(hipStreamSynchronize(as_stream(stream))); });
}

// ---- graph handle -----------------------------------------------------------
int mi355_graph_create(const mi355_csr_desc* desc, int32_t allow_negative, void* stream,
                       mi355_graph** out) {
  return guard([&] {
    if (!desc || !out) fail(MI355_EINVAL, "null argument");
    *out = mi355::build_graph(*desc, allow_negative ? 1 : 0, as_stream(stream));
  });
}
int mi355_graph_create_ex(const mi355_csr_desc* desc, int32_t allow_negative, int32_t vertex_order,
                          void* stream, mi355_graph** out) {
  return guard([&] {
    if (!desc || !out) fail(MI355_EINVAL, "null argument");
    *out = mi355::build_graph(*desc, allow_negative ? 1 : 0, as_stream(stream), vertex_order);
  });
}
int mi355_graph_destroy(mi355_graph* g) {
  return guard([&] {
    if (g) {
      (void)hipSetDevice(g->device);
      delete g;
    }
  });
}
int mi355_graph_info_get(const mi355_graph* g, mi355_graph_info* info) {
  return guard([&] {
    if (!g || !info) fail(MI355_EINVAL, "null argument");
    *info = g->info;
  });
}
int mi355_graph_download(const mi355_graph* g, int64_t* rowptr, int32_t* col, uint32_t* w) {
  return guard([&] {
    if (!g || !rowptr || (g->m > 0 && (!col || !w))) fail(MI355_EINVAL, "null argument");
    MI_HIP(hipSetDevice(g->device));
    std::vector<int> rp((size_t)g->n + 1);
    std::vector<uint2> e((size_t)g->m);
    // the input-order arrays are a cache built on first use
    const int* drp = nullptr;
    const uint2* de = nullptr;
    mi355::input_order_arrays(*const_cast<mi355_graph*>(g), &drp, &de, nullptr);
    MI_HIP(hipMemcpy(rp.data(), drp, rp.size() * 4, hipMemcpyDeviceToHost));
    if (g->m) MI_HIP(hipMemcpy(e.data(), de, e.size() * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < rp.size(); ++i) rowptr[i] = rp[i];
    for (size_t i = 0; i < e.size(); ++i) {
      col[i] = (int32_t)e[i].x;
      w[i] = e[i].y;
    }
  });
}

// ---- SSSP on a handle -------------------------------------------------------
void mi355_sssp_options_default(mi355_sssp_options* opt) {
  if (!opt) return;
  *opt = mi355_sssp_options{};
  opt->algorithm = MI355_ALG_DELTA;
  opt->max_iters = -1;
  opt->check_interval = 1;
  opt->delta = 0.0f;
  opt->inf_mode = MI355_INF_IEEE;
  opt->dynamic_delta = 0;
}

static mi355::RunRequest make_request(const mi355_sssp_options* opt, bool want_work) {
  mi355::RunRequest r;
  if (opt) {
    r.algorithm = opt->algorithm;
    r.max_iters = opt->max_iters;
    r.check_interval = opt->check_interval;
    r.delta = opt->delta;
    r.rounds_per_sync = opt->rounds_per_sync;
    r.dynamic_delta = opt->dynamic_delta;
    r.use_mlp = opt->use_mlp;
    r.worker_blocks = opt->worker_blocks;
    r.active_buckets = opt->active_buckets;
    r.pull = opt->pull;
    r.pull_margin = opt->pull_margin;
    r.light_scale = opt->light_scale;
  } else {
    r.algorithm = MI355_ALG_DELTA;
  }
  r.want_work = want_work;
  return r;
}

static void fill_stats(const mi355::RunResult& res, mi355_sssp_stats* st) {
  if (!st) return;
  st->iterations_used = (int32_t)std::min<int64_t>(res.iterations_used, INT32_MAX);
  st->converged = res.converged ? 1 : 0;
  st->rounds = (int32_t)std::min<int64_t>(res.rounds, INT32_MAX);
  st->syncs = res.syncs;
  st->edges_relaxed = res.edges;
  st->vertices_processed = res.vertices;
  st->initial_delta = res.initial_delta;
  st->final_delta = res.final_delta;
  st->device_ms = res.device_ms;
  st->kernel_ms = res.kernel_ms;
  st->policy = res.policy;
  st->has_negative_cycle = res.negative_cycle ? 1 : 0;
  st->pulled_phases = (int32_t)std::min<int64_t>(res.pulls, INT32_MAX);
}

int mi355_sssp(mi355_graph* g, int32_t source, const mi355_sssp_options* opt, void* d_dist,
               int32_t* d_pred, mi355_sssp_stats* stats, void* stream) {
  return guard([&] {
    if (!g) fail(MI355_EINVAL, "null graph");
    if (source < 0 || source >= g->n) fail(MI355_ERANGE, "source vertex is outside the graph");
    MI_HIP(hipSetDevice(g->device));
    const hipStream_t s = as_stream(stream);
    mi355::RunResult res;
    unsigned long long* reach = nullptr;
    if (stats) {
      g->ws.reach.ensure(2);
      reach = g->ws.reach.p;
      MI_HIP(hipMemsetAsync(reach, 0, 16, s));
    }
    mi355::RunRequest req = make_request(opt, stats != nullptr);
    req.out_dist = d_dist;
    req.out_pred = d_pred;
    req.out_reach = reach;
    req.out_inf_mode = opt ? opt->inf_mode : MI355_INF_IEEE;
    mi355::run_sssp(*g, (uint32_t)source, req, res, s);
    if (!res.converted)
      mi355::labels_to_output(*g, (uint32_t)source, req.out_inf_mode, d_dist, d_pred, reach, s);
    unsigned long long hr[2] = {0, 0};
    if (stats) MI_HIP(hipMemcpyAsync(hr, reach, 16, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    if (stats) {
      *stats = mi355_sssp_stats{};
      fill_stats(res, stats);
      stats->reached = (int64_t)hr[0];
      stats->reached_edges = (int64_t)hr[1];
    }
  });
}

int mi355_sssp_labels(mi355_graph* g, int32_t source, const mi355_sssp_options* opt,
                      mi355_sssp_stats* stats, void* stream) {
  return guard([&] {
    if (!g) fail(MI355_EINVAL, "null graph");
    if (source < 0 || source >= g->n) fail(MI355_ERANGE, "source vertex is outside the graph");
    MI_HIP(hipSetDevice(g->device));
    mi355::RunResult res;
    mi355::run_sssp(*g, (uint32_t)source, make_request(opt, false), res, as_stream(stream));
    if (stats) {
      *stats = mi355_sssp_stats{};
      fill_stats(res, stats);
    }
  });
}

int mi355_sssp_labels_to_output(mi355_graph* g, int32_t source, int32_t inf_mode, void* d_dist,
                                int32_t* d_pred, void* stream) {
  return guard([&] {
    if (!g) fail(MI355_EINVAL, "null graph");
    if (!g->ws.labels.p) fail(MI355_EINVAL, "no labels: run mi355_sssp_labels first");
    mi355::labels_to_output(*g, (uint32_t)source, inf_mode, d_dist, d_pred, nullptr,
                            as_stream(stream));
    MI_HIP(hipStreamSynchronize(as_stream(stream)));
  });
}

int mi355_sssp_round_log(mi355_graph* g, int32_t* sizes, int32_t cap) {
  return guard([&] {
    if (!g) fail(MI355_EINVAL, "null graph");
    g->log_dst = sizes;
    g->log_cap = cap;
  });
}

// ---- reference entry points: CSR Bellman-Ford family -----------------------
int mi355_bf_csr_f32(const mi355_csr_desc* A, int source, int max_iters, int check_interval,
                     float* dist, mi355_bf_csr_stats* stats, mi355_bf_progress_cb cb,
                     void* user_data, void* stream) {
  return guard([&] {
    const auto t_total = std::chrono::steady_clock::now();
    if (!A || !dist) fail(MI355_EINVAL, "null argument");
    if (A->weight_type != MI355_W_F32) fail(MI355_EINVAL, "Bellman-Ford CSR expects fp32 weights");
    if (check_interval < 0) fail(MI355_EINVAL, "check_interval must be positive");
    if (A->rows <= 0) fail(MI355_EINVAL, "CSR graph must contain at least one vertex");
    if (source < 0 || source >= A->rows) fail(MI355_ERANGE, "source is outside graph");
    const hipStream_t s = as_stream(stream);
    const auto t_init = std::chrono::steady_clock::now();
    GraphPtr g(mi355::build_graph(*A, 1 | 2, s));
    const double init_ms = ms_since(t_init);
    const int64_t n = g->n;
    const int64_t mi = max_iters < 0 ? n - 1 : max_iters;
    mi355::RunRequest req;
    req.algorithm = MI355_ALG_BF_JACOBI;
    req.max_iters = mi;
    mi355::RunResult res;
    const auto t_loop = std::chrono::steady_clock::now();
    mi355::run_sssp(*g, (uint32_t)source, req, res, s);
    const double loop_ms = ms_since(t_loop);
    // r0: the first iteration whose merge changes nothing.
    const int64_t r0 = first_empty_round(res.frontier_log);
    int64_t iters = mi;
    bool converged = false;
    if (check_interval > 0 && r0 > 0) {
      const int64_t k = ((r0 + check_interval - 1) / check_interval) * check_interval;
      if (k <= mi) {
        iters = k;
        converged = true;
      }
    }
    const auto t_copy = std::chrono::steady_clock::now();
    download_dist(*g, (uint32_t)source, MI355_INF_IEEE, dist, s);
    const double copy_ms = ms_since(t_copy);
    if (cb) {
      for (int64_t k = 1; k <= iters; ++k) {
        mi355_bf_progress p;
        p.iteration = (int)k;
        p.max_iters = (int)mi;
        const bool checked = check_interval > 0 && k % check_interval == 0;
        const bool changed = r0 == 0 || k < r0;
        p.convergence_checked = checked ? 1 : 0;
        p.changed = checked && changed ? 1 : 0;
        cb(&p, user_data);
      }
    }
    if (stats) {
      *stats = mi355_bf_csr_stats{};
      stats->iterations_used = (int)iters;
      stats->converged = converged ? 1 : 0;
      stats->check_interval = check_interval;
      if (check_interval > 0) {
        stats->checks_performed = (int)(iters / check_interval);
        int changed_checks = 0;
        for (int64_t k = check_interval; k <= iters; k += check_interval)
          if (r0 == 0 || k < r0) ++changed_checks;
        stats->changed_checks = changed_checks;
        stats->unchanged_checks = stats->checks_performed - changed_checks;
      }
      stats->initialization_ms = init_ms;
      stats->bellman_ford_loop_ms = loop_ms;
      stats->relaxation_ms = res.device_ms;
      stats->final_copy_ms = copy_ms;
      stats->total_runtime_ms = ms_since(t_total);
    }
  });
}

int mi355_delta_stepping_csr_f32(const mi355_csr_desc* A, int source, float delta,
                                 int max_iters, float* dist, int* iterations_used,
                                 int* converged, mi355_bf_progress_cb cb, void* user_data,
                                 void* stream) {
  return guard([&] {
    if (!A || !dist) fail(MI355_EINVAL, "null argument");
    if (A->weight_type != MI355_W_F32) fail(MI355_EINVAL, "delta-stepping CSR expects fp32 weights");
    if (A->rows <= 0) fail(MI355_EINVAL, "CSR graph must contain at least one vertex");
    if (A->rows != A->cols) fail(MI355_EINVAL, "SSSP expects a square CSR adjacency matrix");
    if (A->nnz < 0) fail(MI355_EINVAL, "CSR nnz must be nonnegative");
    if (source < 0 || source >= A->rows) fail(MI355_ERANGE, "source vertex is outside CSR row range");
    if (A->rows > INT32_MAX) fail(MI355_EOVERFLOW, "rows must fit in int");
    if (!(delta > 0.0f) || !std::isfinite(delta)) fail(MI355_EINVAL, "delta must be a finite positive float");
    const hipStream_t s = as_stream(stream);
    GraphPtr g(mi355::build_graph(*A, 0, s));
    mi355::RunRequest req;
    req.algorithm = MI355_ALG_DELTA;
    req.max_iters = max_iters < 0 ? -1 : max_iters;
    req.delta = delta;
    mi355::RunResult res;
    mi355::run_sssp(*g, (uint32_t)source, req, res, s);
    download_dist(*g, (uint32_t)source, MI355_INF_IEEE, dist, s);
    if (iterations_used) *iterations_used = (int)res.iterations_used;
    if (converged) *converged = res.converged ? 1 : 0;
    if (cb) {
      const int64_t mi = max_iters < 0 ? (int64_t)INT32_MAX : max_iters;
      for (int64_t k = 1; k <= res.iterations_used; ++k) {
        mi355_bf_progress p;
        p.iteration = (int)k;
        p.max_iters = (int)mi;
        p.convergence_checked = 1;
        p.changed = (k == res.iterations_used && res.converged) ? 0 : 1;
        cb(&p, user_data);
      }
    }
  });
}

}  // extern "C"

// ---- min-plus GEMM / SpGEMM / dense Bellman-Ford ---------------------------
namespace mi355 {
void spgemm_minplus(int64_t a_rows, int64_t a_cols, int64_t a_nnz, const int64_t* arp,
                    const int32_t* acol, const float* aval, int64_t b_rows, int64_t b_cols,
                    int64_t b_nnz, const int64_t* brp, const int32_t* bcol, const float* bval,
                    int hash_capacity, int block_size, int64_t** c_rowptr, int32_t** c_col,
                    float** c_val, int64_t* c_nnz, hipStream_t s);
}

#include "minplus.hpp"

namespace {
// One device slot per host thread for single-GEMM descriptors (kept alive for
// the thread's lifetime so asynchronous launches never see a freed slot).
thread_local mi355::DevBuf<mi355::MinplusItem> t_item;
thread_local mi355::MinplusBatchCache t_batch;

// Upload a dense matrix if it lives on the host.
const float* stage_dense(const float* adj, int32_t memory, int n, mi355::DevBuf<float>& tmp,
                         hipStream_t s) {
  if (memory == MI355_DEVICE) return adj;
  if (memory != MI355_HOST) mi355::fail(MI355_EINVAL, "unknown memory kind");
  tmp.reset((size_t)n * (size_t)n);
  MI_HIP(hipMemcpyAsync(tmp.p, adj, (size_t)n * n * 4, hipMemcpyHostToDevice, s));
  return tmp.p;
}
}  // namespace

extern "C" {

int mi355_minplus_gemm_f32(const float* dA, const float* dB, float* dC, int M, int N, int K,
                           int lda, int ldb, int ldc, void* stream) {
  return guard([&] {
    if (!dA || !dB || !dC) fail(MI355_EINVAL, "null matrix pointer");
    if (M <= 0 || N <= 0 || K <= 0) fail(MI355_EINVAL, "matrix sizes must be positive");
    if (lda < K || ldb < N || ldc < N) fail(MI355_EINVAL, "leading dimension smaller than the matrix");
    // The descriptor slot is reused by the next call on this thread; the copy
    // is stream-ordered before the launch that reads it.
    t_item.ensure(1);
    mi355::minplus_gemm(dA, dB, dC, M, N, K, lda, ldb, ldc, as_stream(stream), t_item.p);
  });
}

int mi355_minplus_gemm_f32_batched(int count, const float* const* dA, const float* const* dB,
                                   float* const* dC, const int* M, const int* N, const int* K,
                                   const int* lda, const int* ldb, const int* ldc, void* stream) {
  return guard([&] {
    if (count < 0) fail(MI355_EINVAL, "negative batch count");
    if (count == 0) return;
    if (!dA || !dB || !dC || !M || !N || !K || !lda || !ldb || !ldc) fail(MI355_EINVAL, "null array");
    std::vector<mi355::MinplusItem> items((size_t)count);
    for (int i = 0; i < count; ++i) {
      if (!dA[i] || !dB[i] || !dC[i]) fail(MI355_EINVAL, "null matrix pointer in batch");
      if (M[i] <= 0 || N[i] <= 0 || K[i] <= 0) fail(MI355_EINVAL, "matrix sizes must be positive");
      if (lda[i] < K[i] || ldb[i] < N[i] || ldc[i] < N[i])
        fail(MI355_EINVAL, "leading dimension smaller than the matrix");
      items[(size_t)i] = mi355::MinplusItem{dA[i], dB[i], dC[i], M[i], N[i], K[i], lda[i], ldb[i], ldc[i], 0};
    }
    mi355::minplus_gemm_batched(items, as_stream(stream), t_batch);
  });
}

int mi355_minplus_gemm_f32_batched_noalloc(int count, const float* const* dA, const float* const* dB,
                                           float* const* dC, const int* M, const int* N, const int* K,
                                           const int* lda, const int* ldb, const int* ldc, void* stream) {
  return guard([&] {
    if (count < 0) fail(MI355_EINVAL, "negative batch count");
    if (count == 0) return;
    if (!dA || !dB || !dC || !M || !N || !K || !lda || !ldb || !ldc) fail(MI355_EINVAL, "null array");
    if ((size_t)count > mi355::minplus_noalloc_max_items())
      fail(MI355_ECAPACITY, "no-allocation min-plus batch holds at most 4096 products");
    thread_local std::vector<mi355::MinplusItem> items;
    items.resize((size_t)count);
    for (int i = 0; i < count; ++i) {
      if (!dA[i] || !dB[i] || !dC[i]) fail(MI355_EINVAL, "null matrix pointer in batch");
      if (M[i] <= 0 || N[i] <= 0 || K[i] <= 0) fail(MI355_EINVAL, "matrix sizes must be positive");
      if (lda[i] < K[i] || ldb[i] < N[i] || ldc[i] < N[i])
        fail(MI355_EINVAL, "leading dimension smaller than the matrix");
      items[(size_t)i] = mi355::MinplusItem{dA[i], dB[i], dC[i], M[i], N[i], K[i], lda[i], ldb[i], ldc[i], 0};
    }
    mi355::minplus_gemm_batched_noalloc(items, as_stream(stream));
  });
}

int mi355_minplus_tile_count(void) { return mi355::minplus_tile_count(); }
const char* mi355_minplus_tile_name(int index) { return mi355::minplus_tile_name(index); }

int mi355_minplus_autotune(const int* sizes, int count, double* timings_ms, void* stream) {
  return guard([&] {
    if (count < 0 || (count > 0 && !sizes)) fail(MI355_EINVAL, "invalid sizes");
    mi355::minplus_autotune(sizes, count, as_stream(stream), timings_ms);
  });
}
int mi355_minplus_tuning_get(int* table32) {
  return guard([&] {
    if (!table32) fail(MI355_EINVAL, "null table");
    mi355::minplus_tuning_get(table32);
  });
}
int mi355_minplus_tuning_set(const int* table32) {
  return guard([&] {
    if (!table32) fail(MI355_EINVAL, "null table");
    mi355::minplus_tuning_set(table32);
  });
}

int mi355_minplus_spgemm_csr_f32(int64_t a_rows, int64_t a_cols, int64_t a_nnz,
                                 const int64_t* a_rowptr, const int32_t* a_col,
                                 const float* a_val, int64_t b_rows, int64_t b_cols,
                                 int64_t b_nnz, const int64_t* b_rowptr, const int32_t* b_col,
                                 const float* b_val, int hash_capacity, int block_size,
                                 int64_t** c_rowptr, int32_t** c_col, float** c_val,
                                 int64_t* c_nnz, void* stream) {
  return guard([&] {
    mi355::spgemm_minplus(a_rows, a_cols, a_nnz, a_rowptr, a_col, a_val, b_rows, b_cols, b_nnz,
                          b_rowptr, b_col, b_val, hash_capacity, block_size, c_rowptr, c_col,
                          c_val, c_nnz, as_stream(stream));
  });
}

int mi355_bf_dense_f32(const float* adjacency, int32_t memory, int n, int source, int max_iters,
                       float* dist, int* iterations_used, int* converged, void* stream) {
  return guard([&] {
    if (!adjacency) fail(MI355_EINVAL, "d_adjacency must not be null");
    if (n <= 0) fail(MI355_EINVAL, "n must be positive");
    if (source < 0 || source >= n) fail(MI355_EINVAL, "source is outside graph");
    if (!dist) fail(MI355_EINVAL, "null dist");
    if (max_iters < 0) max_iters = n - 1;
    const hipStream_t s = as_stream(stream);
    mi355::DevBuf<float> tmp, d((size_t)n);
    const float* A = stage_dense(adjacency, memory, n, tmp, s);
    int conv = 0;
    const int used = mi355::dense_bf(A, n, source, max_iters, d.p, &conv, s);
    MI_HIP(hipMemcpyAsync(dist, d.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    if (iterations_used) *iterations_used = used;
    if (converged) *converged = conv;
  });
}

int mi355_bf_gauss_seidel_f32(const float* adjacency, int32_t memory, int n, int source,
                              int max_iters, float* dist, int* iterations_used, void* stream) {
  return guard([&] {
    if (!adjacency) fail(MI355_EINVAL, "d_adjacency must not be null");
    if (n <= 0) fail(MI355_EINVAL, "n must be positive");
    if (source < 0 || source >= n) fail(MI355_EINVAL, "source is outside graph");
    if (!dist) fail(MI355_EINVAL, "null dist");
    if (max_iters < 0) max_iters = n - 1;
    const hipStream_t s = as_stream(stream);
    mi355::DevBuf<float> tmp, d((size_t)n);
    const float* A = stage_dense(adjacency, memory, n, tmp, s);
    mi355::dense_gauss_seidel(A, n, source, max_iters, d.p, s);
    MI_HIP(hipMemcpyAsync(dist, d.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    if (iterations_used) *iterations_used = max_iters;
  });
}

// ---- ONE (adds_sssp_hip / bellman_ford_sssp_hip) ----------------------------
int mi355_one_sssp_f32(int32_t algorithm, int n, int m, const int* row_offsets,
                       const int* col_indices, const float* weights, int source,
                       const mi355_one_options* options, float* dist, mi355_one_result* result,
                       void* stream) {
  return guard([&] {
    if (n <= 0) fail(MI355_EINVAL, "graph must have at least one node");
    if (m < 0) fail(MI355_EINVAL, "num_edges must be nonnegative");
    if (source < 0 || source >= n) fail(MI355_EINVAL, "source out of range");
    if (!row_offsets || (m > 0 && (!col_indices || !weights)) || !dist)
      fail(MI355_EINVAL, "null graph array");
    mi355_one_options opt{};
    opt.max_iters = -1;
    opt.num_buckets = 32;
    opt.dynamic_delta = 1;
    opt.target_low_work = 8192;
    opt.target_high_work = 262144;
    if (options) opt = *options;
    const hipStream_t s = as_stream(stream);
    mi355_csr_desc d{};
    d.rows = d.cols = n;
    d.nnz = m;
    d.rowptr = row_offsets;
    d.col = col_indices;
    d.w = weights;
    d.offset_bits = 32;
    d.weight_type = MI355_W_F32;
    d.orientation = MI355_OUT;
    d.memory = MI355_HOST;
    mi355::RunRequest req;
    mi355::RunResult res;
    mi355_one_result out{};
    if (algorithm == MI355_ALG_BF) {
      // edge-parallel in-place Bellman-Ford (ONE/sssp_hip.cpp:350-426), with
      // its negative-cycle detection pass (:204-220)
      GraphPtr g(mi355::build_graph(d, 1, s));
      req.algorithm = MI355_ALG_BF;
      req.max_iters = opt.max_iters <= 0 ? std::max(n - 1, 0) : opt.max_iters;
      mi355::run_sssp(*g, (uint32_t)source, req, res, s);
      if (!res.converged && res.frontier_log.size() > 0 && res.frontier_log.back() > 0) {
        // one more round from the last frontier: any improvement means a
        // negative cycle is reachable from the source
        mi355::DevBuf<uint32_t> before((size_t)n);
        mi355::labels_to_output(*g, (uint32_t)source, MI355_INF_IEEE, before.p, nullptr, nullptr, s);
        std::vector<uint32_t> b0((size_t)n), b1((size_t)n);
        MI_HIP(hipMemcpyAsync(b0.data(), before.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
        mi355::RunRequest one = req;
        one.max_iters = 1;
        mi355::RunResult r1;
        mi355::run_bf_continue(*g, (uint32_t)source, one, r1, s);
        mi355::labels_to_output(*g, (uint32_t)source, MI355_INF_IEEE, before.p, nullptr, nullptr, s);
        MI_HIP(hipMemcpyAsync(b1.data(), before.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
        out.has_negative_cycle = b0 != b1 ? 1 : 0;
        // report the distances of the n-1 rounds, as the reference does
        std::memcpy(dist, b0.data(), (size_t)n * 4);
        out.iterations_used = (int)res.iterations_used;
        out.converged = 0;
        out.milliseconds = res.device_ms;
        out.work_count = (unsigned long long)res.edges;
        if (result) *result = out;
        return;
      }
      download_dist(*g, (uint32_t)source, MI355_INF_IEEE, dist, s);
      out.iterations_used = (int)res.iterations_used;
      out.converged = res.converged ? 1 : 0;
    } else if (algorithm == MI355_ALG_ADDS || algorithm == MI355_ALG_DELTA) {
      if (opt.num_buckets <= 1 || opt.num_buckets > 64)
        fail(MI355_EINVAL, "num_buckets must be in [2, 64]");
      GraphPtr g(mi355::build_graph(d, 0, s));
      req.algorithm = algorithm == MI355_ALG_DELTA ? MI355_ALG_DELTA : MI355_ALG_ADDS;
      req.delta = (opt.initial_delta > 0.0f && std::isfinite(opt.initial_delta))
                      ? opt.initial_delta : mi355::heuristic_delta(*g);
      req.max_iters = opt.max_iters > 0 ? opt.max_iters : -1;
      req.dynamic_delta = opt.dynamic_delta;
      req.want_work = true;
      // a caller's per-bucket capacity bounds every bucket's live entries;
      // exceeding it fails the call (ONE/sssp_hip.cpp:548-556)
      req.bucket_capacity = opt.bucket_capacity_per_bucket > 0 ? opt.bucket_capacity_per_bucket : 0;
      mi355::run_sssp(*g, (uint32_t)source, req, res, s);
      download_dist(*g, (uint32_t)source, MI355_INF_IEEE, dist, s);
      out.iterations_used = (int)res.iterations_used;
      out.converged = res.converged ? 1 : 0;
      out.final_delta = res.final_delta;
      // work_count counts processed vertices (ONE/sssp_hip.cpp:286)
      out.milliseconds = res.device_ms;
      out.work_count = (unsigned long long)res.vertices;
      if (result) *result = out;
      return;
    } else {
      fail(MI355_EINVAL, "algorithm must be MI355_ALG_BF or MI355_ALG_ADDS");
    }
    out.milliseconds = res.device_ms;
    out.work_count = (unsigned long long)res.edges;
    if (result) *result = out;
  });
}

// ---- hip_adds::sssp ----------------------------------------------------------
int mi355_hip_adds_sssp(int32_t weight_type, uint32_t num_vertices, uint32_t num_edges,
                        const uint32_t* d_row_offsets, const uint32_t* d_col_indices,
                        const void* d_weights, uint32_t source, void* d_distances,
                        mi355_adds_stats* stats, const mi355_adds_options* options,
                        void* stream) {
  // hip_adds::Status values
  enum { kSuccess = 0, kInvalidArgument = 1, kHipError = 2, kOutOfMemory = 3, kDeviceError = 6 };
  if (num_vertices == 0u || source >= num_vertices || !d_distances || !d_row_offsets ||
      !d_col_indices || !d_weights || (weight_type != MI355_W_F32 && weight_type != MI355_W_U32)) {
    g_err = "invalid argument";
    return kInvalidArgument;
  }
  const int rc = guard([&] {
    const hipStream_t s = as_stream(stream);
    mi355_csr_desc d{};
    d.rows = d.cols = num_vertices;
    d.nnz = num_edges;
    d.rowptr = d_row_offsets;
    d.col = d_col_indices;
    d.w = d_weights;
    d.offset_bits = 32;
    d.weight_type = weight_type;
    d.orientation = MI355_OUT;
    d.memory = MI355_DEVICE;
    GraphPtr g(mi355::build_graph(d, 0, s));
    mi355::RunRequest req;
    req.algorithm = MI355_ALG_ADDS;
    req.dynamic_delta = options ? options->dynamic_delta : 1;
    req.worker_blocks = options ? (int)options->worker_blocks : 0;
    req.active_buckets = options ? (int)std::min<uint32_t>(std::max<uint32_t>(options->initial_active_buckets, 1u), 32u) : 1;
    double delta = 0.0;
    const double scale = options ? options->initial_delta_scale : 16.0;
    if (options && options->initial_delta > 0.0) delta = options->initial_delta;
    else delta = scale * g->info.avg_weight / std::max(1.0, g->info.avg_degree);
    const double min_delta = options ? std::max(options->min_delta, 1.0e-12) : 1.0e-3;
    if (!std::isfinite(delta) || delta <= 0.0) delta = min_delta;
    delta = std::max(delta, min_delta);
    if (weight_type == MI355_W_U32) delta = std::max(delta, 1.0);
    req.delta = (float)delta;
    req.want_work = stats != nullptr;
    if (options && !options->initialize_distances)
      req.init_dist = static_cast<const uint32_t*>(d_distances);
    mi355::RunResult res;
    mi355::run_sssp(*g, source, req, res, s);
    mi355::labels_to_output(*g, source, MI355_INF_IEEE, d_distances, nullptr, nullptr, s);
    MI_HIP(hipStreamSynchronize(s));
    if (stats) {
      *stats = mi355_adds_stats{};
      stats->initial_delta = delta;
      stats->final_delta = res.final_delta;
      stats->edge_relaxations = (uint64_t)res.edges;
      stats->processed_items = (uint64_t)res.vertices;
      stats->head_switches = (uint64_t)res.iterations_used;
      stats->worker_blocks = (uint32_t)(req.worker_blocks > 0 ? req.worker_blocks : 0);
    }
  });
  if (rc == MI355_OK) return kSuccess;
  if (rc == MI355_EINVAL || rc == MI355_ERANGE || rc == MI355_EOVERFLOW) return kInvalidArgument;
  if (rc == MI355_ENOMEM) return kOutOfMemory;
  if (rc == MI355_EHIP) return kHipError;
  return kDeviceError;
}


// ---- learned delta policy ---------------------------------------------------
void mi355_mlp_train_options_default(mi355_mlp_train_options* o) {
  if (!o) return;
  *o = mi355_mlp_train_options{};
  o->epochs = 60;
  o->batch = 512;
  o->lr = 2.0e-3f;
  o->beta1 = 0.9f;
  o->beta2 = 0.999f;
  o->eps = 1.0e-8f;
  o->weight_decay = 0.0f;
  o->seed = 2026;
}

int mi355_mlp_forward(const mi355_mlp_params* p, const float* x, int64_t n, float* out) {
  return guard([&] { mi355::mlp_forward(p, x, n, out, nullptr); });
}

int mi355_mlp_loss_grad(const mi355_mlp_params* p, const float* x, const float* y, const float* w,
                        int64_t n, float* loss, mi355_mlp_params* grad) {
  return guard([&] { mi355::mlp_loss_grad(p, x, y, w, n, loss, grad, nullptr); });
}

int mi355_mlp_train(mi355_mlp_params* p, const float* x, const float* y, const float* w, int64_t n,
                    const mi355_mlp_train_options* o, float* epoch_loss) {
  return guard([&] {
    mi355_mlp_train_options d;
    mi355_mlp_train_options_default(&d);
    mi355::mlp_train(p, x, y, w, n, o ? *o : d, epoch_loss, nullptr);
  });
}

int mi355_adds_default_workers(int32_t* workers) {
  return guard([&] {
    if (!workers) fail(MI355_EINVAL, "null argument");
    *workers = mi355::device_props().cus * 2;
  });
}

int mi355_adds_mlp_set(const mi355_mlp_params* p) {
  return guard([&] { mi355::set_mlp_params(p); });
}

int mi355_adds_schedule_set(const float* table, int32_t count) {
  return guard([&] { mi355::set_schedule(table, count); });
}

int mi355_adds_mlp_clear(void) {
  return guard([&] { mi355::clear_mlp(); });
}

int mi355_adds_trace(mi355_graph* g, int32_t source, const mi355_sssp_options* opt, float* trace,
                     int32_t cap_phases, int32_t* phases, mi355_sssp_stats* stats, void* stream) {
  return guard([&] {
    if (!g) fail(MI355_EINVAL, "null graph");
    if (source < 0 || source >= g->n) fail(MI355_ERANGE, "source vertex is outside the graph");
    if (cap_phases < 0 || (cap_phases > 0 && !trace)) fail(MI355_EINVAL, "bad trace buffer");
    MI_HIP(hipSetDevice(g->device));
    const hipStream_t s = as_stream(stream);
    mi355::RunRequest req = make_request(opt, true);
    req.algorithm = MI355_ALG_ADDS;
    mi355::DevBuf<float> d_trace((size_t)std::max(cap_phases, 1) * MI355_ADDS_TRACE_STRIDE);
    MI_HIP(hipMemsetAsync(d_trace.p, 0, d_trace.n * sizeof(float), s));
    req.trace = d_trace.p;
    req.trace_cap = (uint32_t)std::max(cap_phases, 1);
    mi355::RunResult res;
    mi355::run_sssp(*g, (uint32_t)source, req, res, s);
    const int64_t recorded = std::min<int64_t>(res.iterations_used, cap_phases);
    if (recorded > 0)
      MI_HIP(hipMemcpyAsync(trace, d_trace.p, (size_t)recorded * MI355_ADDS_TRACE_STRIDE * sizeof(float),
                            hipMemcpyDeviceToHost, s));
    mi355::DevBuf<unsigned long long> reach(2);
    MI_HIP(hipMemsetAsync(reach.p, 0, 16, s));
    mi355::labels_to_output(*g, (uint32_t)source, MI355_INF_IEEE, nullptr, nullptr, reach.p, s);
    unsigned long long hr[2] = {0, 0};
    MI_HIP(hipMemcpyAsync(hr, reach.p, 16, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    if (phases) *phases = (int32_t)std::min<int64_t>(res.iterations_used, INT32_MAX);
    if (stats) {
      *stats = mi355_sssp_stats{};
      fill_stats(res, stats);
      stats->reached = (int64_t)hr[0];
      stats->reached_edges = (int64_t)hr[1];
    }
  });
}


// ---- batched multi-source queries -------------------------------------------
void mi355_batch_options_default(mi355_batch_options* o) {
  if (!o) return;
  *o = mi355_batch_options{};
  o->rounds_per_graph = 24;
  o->blocks_per_cu = 4;
  o->use_graph = 1;
  o->inf_mode = MI355_INF_IEEE;
}

int mi355_batch_layout(const mi355_graph* g, const mi355_query_batch* qb, int64_t* out_offsets) {
  return guard([&] {
    if (!g || !qb || !out_offsets) fail(MI355_EINVAL, "null argument");
    std::vector<int64_t> off;
    mi355::batch_layout(*qb, g->n, off);
    std::copy(off.begin(), off.end(), out_offsets);
  });
}

int mi355_batch_paths(mi355_graph* g, const mi355_query_batch* qb, const int64_t* out_offsets,
                      const void* d_dist, const int32_t* d_pred, const int32_t* query, const int32_t* target,
                      int64_t count, int64_t* path_offsets, uint32_t* target_dist, int32_t* path,
                      int64_t path_capacity, void* stream) {
  return guard([&] {
    if (!g || !qb || !out_offsets || !d_dist || !d_pred || !path_offsets) fail(MI355_EINVAL, "null argument");
    if (count > 0 && (!query || !target)) fail(MI355_EINVAL, "null target arrays");
    MI_HIP(hipSetDevice(g->device));
    mi355::batch_paths(*qb, out_offsets, d_pred, static_cast<const uint32_t*>(d_dist), query, target, count,
                       path_offsets, target_dist, path, path_capacity, as_stream(stream));
  });
}

int mi355_batch_autotune(mi355_graph* g, const mi355_query_batch* qb, int32_t sample_stride,
                         mi355_batch_options* best, double* timings_ms, void* stream) {
  return guard([&] {
    if (!g || !qb || !best) fail(MI355_EINVAL, "null argument");
    MI_HIP(hipSetDevice(g->device));
    mi355::batch_autotune(*g, *qb, sample_stride, *best, timings_ms, as_stream(stream));
  });
}

int mi355_sssp_batched(mi355_graph* g, const mi355_query_batch* qb, const mi355_batch_options* o,
                       void* d_dist, int32_t* d_pred, mi355_batch_stats* stats, void* stream) {
  return guard([&] {
    if (!g || !qb) fail(MI355_EINVAL, "null argument");
    if (!d_dist && !d_pred) fail(MI355_EINVAL, "no output requested");
    MI_HIP(hipSetDevice(g->device));
    mi355_batch_options d;
    mi355_batch_options_default(&d);
    mi355_batch_stats st{};
    mi355::run_batch(*g, *qb, o ? *o : d, nullptr, nullptr, d_dist, d_pred, st, as_stream(stream));
    if (stats) *stats = st;
  });
}

int mi355_sssp_batched_targets(mi355_graph* g, const mi355_query_batch* qb, const int32_t* target_offsets,
                               const int32_t* targets, const mi355_batch_options* o, void* d_dist,
                               int32_t* d_pred, mi355_batch_stats* stats, void* stream) {
  return guard([&] {
    if (!g || !qb || !target_offsets) fail(MI355_EINVAL, "null argument");
    if (!d_dist && !d_pred) fail(MI355_EINVAL, "no output requested");
    if (qb->count > 0 && target_offsets[qb->count] > 0 && !targets) fail(MI355_EINVAL, "null targets");
    MI_HIP(hipSetDevice(g->device));
    mi355_batch_options d;
    mi355_batch_options_default(&d);
    mi355_batch_stats st{};
    mi355::run_batch(*g, *qb, o ? *o : d, target_offsets, targets, d_dist, d_pred, st, as_stream(stream));
    if (stats) *stats = st;
  });
}

}  // extern "C"
