// Runtime check of the C++ compatibility layer (include/compat/*.hpp): calls
// the drop-in functions the way a caller of the reference would and writes
// what they return, so that tests/test_compat_layer.py can compare it with the
// oracle.  Modes:
//
//   host                         dense_to_csr_outgoing and its errors (no GPU)
//   detail  G SRC OUT            bf_csr_detail chain: validate, copy, source
//                                vector, sparse_minplus_relax + merge rounds
//                                until nothing changes, copy back
//   bfcsr   G SRC OUT            bellman_ford_minplus_hip_csr (host and device
//                                overloads, progress callback), no_checks,
//                                var_checks_with_stats; delta_stepping_minplus_hip_csr
//   dense   N MAT SRC OUT        adds_sssp_hip / bellman_ford_sssp_hip dense
//                                overloads (outgoing matrix)
//   errors                       each documented failure maps to the
//                                reference's exception type / error code
//
// G is a raw incoming CSR: int64 rows, int64 nnz, int64 rowptr[rows + 1],
// int32 colind[nnz], float values[nnz].  OUT receives float distances.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "bf_hip_CSR.hpp"
#include "bf_hip_CSR_device_utils.hpp"
#include "bf_hip_no_checks_CSR.hpp"
#include "bf_var_checks.hpp"
#include "delta_stepping_hip_CSR.hpp"
#include "hip_adds_sssp.hpp"
#include "minplus_hip.hpp"
#include "minplus_sparse_hip.hpp"
#include "sssp_hip.hpp"

namespace {

int g_failures = 0;

void expect(bool ok, const std::string& what) {
  std::printf("%s %s\n", ok ? "ok" : "FAIL", what.c_str());
  if (!ok) ++g_failures;
}

// Runs f and reports which of the standard exception types it threw.
std::string thrown(const std::function<void()>& f) {
  try {
    f();
  } catch (const std::out_of_range&) {
    return "out_of_range";
  } catch (const std::overflow_error&) {
    return "overflow_error";
  } catch (const std::invalid_argument&) {
    return "invalid_argument";
  } catch (const std::runtime_error&) {
    return "runtime_error";
  } catch (...) {
    return "other";
  }
  return "none";
}

HostCsrF32 read_graph(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  int64_t hdr[2];
  HostCsrF32 h;
  bool ok = std::fread(hdr, 8, 2, f) == 2 && hdr[0] > 0 && hdr[1] >= 0;
  if (ok) {
    h.rows = h.cols = hdr[0];
    h.nnz = hdr[1];
    h.rowptr.resize((size_t)h.rows + 1);
    h.colind.resize((size_t)h.nnz);
    h.values.resize((size_t)h.nnz);
    ok = std::fread(h.rowptr.data(), 8, h.rowptr.size(), f) == h.rowptr.size() &&
         std::fread(h.colind.data(), 4, h.colind.size(), f) == h.colind.size() &&
         std::fread(h.values.data(), 4, h.values.size(), f) == h.values.size();
  }
  std::fclose(f);
  if (!ok) throw std::runtime_error(std::string("malformed graph file ") + path);
  return h;
}

void write_floats(const char* path, const std::vector<float>& v) {
  FILE* f = std::fopen(path, "wb");
  if (!f || std::fwrite(v.data(), 4, v.size(), f) != v.size())
    throw std::runtime_error(std::string("cannot write ") + path);
  std::fclose(f);
}

bool same_bits(const std::vector<float>& a, const std::vector<float>& b) {
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * 4) == 0;
}

// ---- host: dense_to_csr_outgoing ------------------------------------------
int mode_host() {
  const float inf = std::numeric_limits<float>::infinity();
  const float nan = std::numeric_limits<float>::quiet_NaN();
  // row-major outgoing: diagonal, +inf and NaN entries are not edges
  const std::vector<float> adj = {5.0f, 1.5f, inf,  2.0f,
                                  inf,  0.0f, 3.0f, nan,
                                  0.25f, inf, 7.0f, 4.0f,
                                  inf,  inf,  inf,  1.0f};
  const CsrGraph g = dense_to_csr_outgoing(adj, 4);
  expect(g.num_nodes == 4 && g.num_edges == 5, "dense_to_csr_outgoing counts");
  expect(g.row_offsets == std::vector<int>({0, 2, 3, 5, 5}), "dense_to_csr_outgoing row offsets");
  expect(g.col_indices == std::vector<int>({1, 3, 2, 0, 3}), "dense_to_csr_outgoing columns");
  expect(g.weights == std::vector<float>({1.5f, 2.0f, 3.0f, 0.25f, 4.0f}),
         "dense_to_csr_outgoing weights");
  const CsrGraph e = dense_to_csr_outgoing(std::vector<float>(9, inf), 3);
  expect(e.num_edges == 0 && e.row_offsets == std::vector<int>(4, 0), "dense_to_csr_outgoing no edges");
  expect(thrown([] { dense_to_csr_outgoing({}, 0); }) == "invalid_argument", "n <= 0 rejected");
  expect(thrown([] { dense_to_csr_outgoing(std::vector<float>(8, 1.0f), 3); }) == "invalid_argument",
         "size != n*n rejected");
  expect(thrown([] { dense_to_csr_outgoing({0.0f, -1.0f, 2.0f, 0.0f}, 2); }) == "invalid_argument",
         "negative weight rejected");
  // a negative diagonal entry is skipped before the sign test
  expect(dense_to_csr_outgoing({-3.0f, 1.0f, 1.0f, -2.0f}, 2).num_edges == 2,
         "negative diagonal ignored");
  return g_failures;
}

// ---- detail: the sparse-vector Bellman-Ford of bf_hip_CSR.cpp:36-99 ---------
int mode_detail(const char* gpath, int src, const char* out) {
  using namespace bf_csr_detail;
  hipStream_t s = nullptr;
  check_hip(hipStreamCreate(&s), "create stream");
  const HostCsrF32 h = read_graph(gpath);
  validate_host_csr_adjacency(h, src);
  DeviceCsrOwner A = copy_host_csr_to_device(h, s);
  validate_device_csr_adjacency(A.view, src);
  DeviceCsrOwner d = make_device_single_source_vector(A.view.rows, src, s);
  int iters = 0;
  bool converged = false;
  const int max_iters = (int)h.rows - 1;
  while (iters < max_iters) {
    DeviceCsrOwner r = sparse_minplus_relax(A.view, d.view, s);
    int changed = 0;
    DeviceCsrOwner next = merge_distances_on_device(d.view, r.view, &changed, s);
    d = std::move(next);
    ++iters;
    if (!changed) {
      converged = true;
      break;
    }
  }
  const std::vector<float> dist = copy_dist_vector_to_host(d.view, s);
  write_floats(out, dist);
  std::printf("iterations %d converged %d nnz %lld\n", iters, converged ? 1 : 0, (long long)d.view.nnz);
  d.reset();
  A.reset();
  check_hip(hipStreamDestroy(s), "destroy stream");
  return 0;
}

// ---- bfcsr: the CSR Bellman-Ford and delta-stepping entry points ------------
struct ProgressLog {
  int calls = 0;
  int last_iteration = 0;  // iterations are numbered from 1 (bf_hip_CSR.cpp:79-84)
  int checked = 0;
  bool ordered = true;
};

void on_progress(const BellmanFordCsrProgress& p, void* user) {
  auto* log = static_cast<ProgressLog*>(user);
  if (p.iteration != log->last_iteration + 1) log->ordered = false;
  log->last_iteration = p.iteration;
  log->checked += p.convergence_checked ? 1 : 0;
  ++log->calls;
}

int mode_bfcsr(const char* gpath, int src, const char* out) {
  hipStream_t s = nullptr;
  bf_csr_detail::check_hip(hipStreamCreate(&s), "create stream");
  const HostCsrF32 h = read_graph(gpath);
  bf_csr_detail::DeviceCsrOwner A = bf_csr_detail::copy_host_csr_to_device(h, s);

  ProgressLog log;
  const BellmanFordCsrResult r1 = bellman_ford_minplus_hip_csr(A.view, src, -1, s, on_progress, &log);
  const BellmanFordCsrResult r2 = bellman_ford_minplus_hip_csr(h, src);
  expect(same_bits(r1.dist, r2.dist) && r1.iterations_used == r2.iterations_used,
         "host and device overloads agree");
  expect(log.calls == r1.iterations_used && log.ordered, "progress callback once per iteration, in order");
  expect(log.checked == log.calls, "every iteration of the checked entry point is checked");

  ProgressLog nlog;
  const BellmanFordCsrNoChecksResult nc =
      bellman_ford_minplus_hip_csr_no_checks(A.view, src, r1.iterations_used + 3, s, on_progress, &nlog);
  expect(same_bits(nc.dist, r1.dist) && nc.iterations_used == r1.iterations_used + 3,
         "no_checks runs exactly max_iters iterations");
  expect(nlog.calls == nc.iterations_used && nlog.checked == 0, "no_checks reports unchecked iterations");

  const BellmanFordCsrVarChecksResult vc =
      bellman_ford_minplus_hip_csr_var_checks_with_stats(A.view, src, -1, 3, s);
  expect(same_bits(vc.dist, r1.dist) && vc.converged, "var_checks distances");
  expect(vc.iterations_used >= r1.iterations_used && vc.iterations_used <= r1.iterations_used + 2 &&
             vc.iterations_used % 3 == 0,
         "var_checks stops at the first checked iteration after convergence");
  expect(vc.checks_performed == vc.iterations_used / 3, "var_checks counts its checks");

  const DeltaSteppingCsrResult ds = delta_stepping_minplus_hip_csr(A.view, src, 0.25f, -1, s);
  expect(same_bits(ds.dist, r1.dist) && ds.converged, "delta-stepping distances equal Bellman-Ford's");
  write_floats(out, r1.dist);
  std::printf("iterations %d converged %d ds_buckets %d\n", r1.iterations_used, r1.converged ? 1 : 0,
              ds.iterations_used);
  A.reset();
  bf_csr_detail::check_hip(hipStreamDestroy(s), "destroy stream");
  return g_failures;
}

// ---- dense: ONE's dense overloads ------------------------------------------
int mode_dense(int n, const char* mpath, int src, const char* out) {
  std::vector<float> adj((size_t)n * n);
  FILE* f = std::fopen(mpath, "rb");
  if (!f || std::fread(adj.data(), 4, adj.size(), f) != adj.size()) throw std::runtime_error("bad matrix");
  std::fclose(f);
  const SSSPResult a = adds_sssp_hip(adj, n, src);
  const SSSPResult b = bellman_ford_sssp_hip(adj, n, src);
  expect(a.converged && b.converged && !b.has_negative_cycle, "dense ADDS and Bellman-Ford converge");
  expect(same_bits(a.dist, b.dist), "dense ADDS and Bellman-Ford agree");
  const CsrGraph g = dense_to_csr_outgoing(adj, n);
  unsigned long long reached = 0;
  for (float x : a.dist) reached += std::isfinite(x) ? 1 : 0;
  // ADDS counts processed vertices: every reached vertex at least once
  expect(a.work_count >= reached, "ADDS work_count counts processed vertices");
  expect(b.work_count >= (unsigned long long)0 && b.iterations_used >= 1, "Bellman-Ford iterations");
  write_floats(out, a.dist);
  std::printf("edges %d adds_work %llu bf_work %llu\n", g.num_edges, a.work_count, b.work_count);
  return g_failures;
}

// ---- errors -------------------------------------------------------------------
int mode_errors() {
  hipStream_t s = nullptr;
  bf_csr_detail::check_hip(hipStreamCreate(&s), "create stream");
  // a 3-vertex incoming path 0 -> 1 -> 2
  HostCsrF32 h;
  h.rows = h.cols = 3;
  h.nnz = 2;
  h.rowptr = {0, 0, 1, 2};
  h.colind = {0, 1};
  h.values = {1.0f, 2.0f};
  bf_csr_detail::DeviceCsrOwner A = bf_csr_detail::copy_host_csr_to_device(h, s);
  const minplus_sparse::DeviceCsrF32 v = A.view;

  expect(thrown([&] { bellman_ford_minplus_hip_csr(v, 3, -1, s); }) == "invalid_argument",
         "BF CSR: source outside the graph -> invalid_argument");
  expect(thrown([&] { bellman_ford_minplus_hip_csr_var_checks(v, 0, -1, 0, s); }) == "invalid_argument",
         "BF var checks: check_interval 0 -> invalid_argument");
  expect(thrown([&] { delta_stepping_minplus_hip_csr(v, 7, 1.0f, -1, s); }) == "out_of_range",
         "delta-stepping: source outside the rows -> out_of_range");
  expect(thrown([&] { delta_stepping_minplus_hip_csr(v, 0, 0.0f, -1, s); }) == "invalid_argument",
         "delta-stepping: delta 0 -> invalid_argument");
  expect(thrown([&] { delta_stepping_minplus_hip_csr(v, 0, std::nanf(""), -1, s); }) == "invalid_argument",
         "delta-stepping: delta NaN -> invalid_argument");
  minplus_sparse::DeviceCsrF32 rect = v;
  rect.cols = 4;
  expect(thrown([&] { delta_stepping_minplus_hip_csr(rect, 0, 1.0f, -1, s); }) == "invalid_argument",
         "delta-stepping: non-square matrix -> invalid_argument");
  HostCsrF32 bad = h;
  bad.colind[1] = 9;
  expect(thrown([&] { bf_csr_detail::validate_host_csr_adjacency(bad, 0); }) == "invalid_argument",
         "bf_csr_detail: column out of range -> invalid_argument");
  expect(thrown([&] { bf_csr_detail::make_device_single_source_vector(3, 5, s); }) == "invalid_argument",
         "bf_csr_detail: source vector outside the rows -> invalid_argument");

  // ONE: bucket count outside [2, 64], and a bucket capacity too small for the run
  CsrGraph g;
  g.num_nodes = 3;
  g.num_edges = 2;
  g.row_offsets = {0, 1, 2, 2};
  g.col_indices = {1, 2};
  g.weights = {1.0f, 2.0f};
  SSSPOptions o;
  o.num_buckets = 1;
  expect(thrown([&] { adds_sssp_hip(g, 0, o); }) == "invalid_argument", "ONE ADDS: num_buckets 1 -> invalid_argument");
  expect(thrown([&] { adds_sssp_hip(g, 3); }) == "invalid_argument", "ONE ADDS: source outside -> invalid_argument");
  {
    // a star: the source's 64 neighbours all land in one bucket
    CsrGraph star;
    star.num_nodes = 65;
    star.row_offsets.assign(66, 64);
    star.row_offsets[0] = 0;
    for (int i = 1; i <= 64; ++i) {
      star.col_indices.push_back(i);
      star.weights.push_back(1.0f);
    }
    star.num_edges = 64;
    SSSPOptions small;
    small.bucket_capacity_per_bucket = 8;
    small.initial_delta = 100.0f;
    expect(thrown([&] { adds_sssp_hip(star, 0, small); }) == "runtime_error",
           "ONE ADDS: bucket capacity exceeded -> runtime_error");
    SSSPOptions roomy = small;
    roomy.bucket_capacity_per_bucket = 4096;
    expect(thrown([&] { adds_sssp_hip(star, 0, roomy); }) == "none", "ONE ADDS: enough capacity runs");
  }

  // minplus_gemm_f32 returns hipErrorInvalidValue on bad arguments, never throws
  expect(minplus_gemm_f32(nullptr, nullptr, nullptr, 4, 4, 4, 4, 4, 4, s) == hipErrorInvalidValue,
         "minplus_gemm_f32: null operands -> hipErrorInvalidValue");
  {
    float* d = nullptr;
    bf_csr_detail::check_hip(hipMalloc(reinterpret_cast<void**>(&d), 64 * sizeof(float)), "alloc");
    expect(minplus_gemm_f32(d, d, d, 4, 4, 4, 2, 4, 4, s) == hipErrorInvalidValue,
           "minplus_gemm_f32: lda < K -> hipErrorInvalidValue");
    (void)hipFree(d);
  }
  // SpGEMM: dimension mismatch -> hipErrorInvalidValue, outputs untouched
  {
    minplus_sparse::Offset* rp = nullptr;
    minplus_sparse::Index* ci = nullptr;
    float* val = nullptr;
    minplus_sparse::Offset nnz = -1;
    minplus_sparse::DeviceCsrF32 b = v;
    b.rows = 5;
    expect(minplus_sparse::minplus_spgemm_csr_f32(v, b, &rp, &ci, &val, &nnz, s) == hipErrorInvalidValue &&
               rp == nullptr,
           "SpGEMM: inner dimensions differ -> hipErrorInvalidValue");
  }
  // hip_adds: a source outside the graph is an InvalidArgument status
  {
    uint32_t* rp = nullptr;
    bf_csr_detail::check_hip(hipMalloc(reinterpret_cast<void**>(&rp), 16), "alloc");
    bf_csr_detail::check_hip(hipMemset(rp, 0, 16), "clear");
    float* dist = nullptr;
    bf_csr_detail::check_hip(hipMalloc(reinterpret_cast<void**>(&dist), 16), "alloc");
    hip_adds::DeviceCsrGraph<float> dg;
    dg.num_vertices = 3;
    dg.num_edges = 0;
    dg.row_offsets = rp;
    const hip_adds::Status st = hip_adds::sssp(dg, 9u, dist);
    expect(st == hip_adds::Status::InvalidArgument, "hip_adds::sssp: source outside -> InvalidArgument");
    expect(std::string(hip_adds::status_string(st)) == "invalid argument", "hip_adds::status_string");
    (void)hipFree(rp);
    (void)hipFree(dist);
  }
  A.reset();
  bf_csr_detail::check_hip(hipStreamDestroy(s), "destroy stream");
  return g_failures;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "";
    int rc = 2;
    if (mode == "host") rc = mode_host();
    else if (mode == "detail" && argc == 5) rc = mode_detail(argv[2], std::atoi(argv[3]), argv[4]);
    else if (mode == "bfcsr" && argc == 5) rc = mode_bfcsr(argv[2], std::atoi(argv[3]), argv[4]);
    else if (mode == "dense" && argc == 6) rc = mode_dense(std::atoi(argv[2]), argv[3], std::atoi(argv[4]), argv[5]);
    else if (mode == "errors") rc = mode_errors();
    else std::fprintf(stderr, "usage: compat_check host|detail|bfcsr|dense|errors ...\n");
    std::printf("failures %d\n", g_failures);
    return rc == 0 ? 0 : 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
