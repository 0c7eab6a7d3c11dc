// C ABI of libmi355sssp.so (declared in include/mi355sssp.h).
//
// Every entry point runs inside guard(): C++ exceptions become status codes
// and a thread-local message, so nothing throws across the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
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
    std::vector<int> rp((size_t)g->n + 1);
    std::vector<uint2> e((size_t)g->m);
    MI_HIP(hipMemcpy(rp.data(), g->rowptr.p, rp.size() * 4, hipMemcpyDeviceToHost));
    if (g->m) MI_HIP(hipMemcpy(e.data(), g->edges.p, e.size() * 8, hipMemcpyDeviceToHost));
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
  opt->dynamic_delta = 1;
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
  st->has_negative_cycle = res.negative_cycle ? 1 : 0;
}

int mi355_sssp(mi355_graph* g, int32_t source, const mi355_sssp_options* opt, void* d_dist,
               int32_t* d_pred, mi355_sssp_stats* stats, void* stream) {
  return guard([&] {
    if (!g) fail(MI355_EINVAL, "null graph");
    if (source < 0 || source >= g->n) fail(MI355_ERANGE, "source vertex is outside the graph");
    MI_HIP(hipSetDevice(g->device));
    const hipStream_t s = as_stream(stream);
    mi355::RunResult res;
    mi355::run_sssp(*g, (uint32_t)source, make_request(opt, stats != nullptr), res, s);
    mi355::DevBuf<unsigned long long> reach;
    if (stats) {
      reach.reset(2);
      MI_HIP(hipMemsetAsync(reach.p, 0, 16, s));
    }
    mi355::labels_to_output(*g, (uint32_t)source, opt ? opt->inf_mode : MI355_INF_IEEE, d_dist,
                            d_pred, stats ? reach.p : nullptr, s);
    unsigned long long hr[2] = {0, 0};
    if (stats) MI_HIP(hipMemcpyAsync(hr, reach.p, 16, hipMemcpyDeviceToHost, s));
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
