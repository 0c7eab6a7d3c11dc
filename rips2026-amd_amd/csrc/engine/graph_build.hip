// Graph ingest on the device: validation, offset narrowing, edge-record
// interleaving and a stable incoming->outgoing transpose.
//
// Replaces the per-call work of DS/src/delta_stepping_hip_CSR.cpp:258-324 and
// :546-605 (K7-K10: validate, count out-degrees, scan, atomic-cursor scatter)
// with one build per graph handle.  The transpose here is a stable radix sort
// of (source vertex, edge index) pairs, so each outgoing row lists its
// destinations in ascending order and the layout is deterministic.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "graph.hpp"

namespace mi355 {

DeviceProps device_props() {
  DeviceProps p;
  MI_HIP(hipGetDevice(&p.device));
  MI_HIP(hipDeviceGetAttribute(&p.cus, hipDeviceAttributeMultiprocessorCount, p.device));
  return p;
}

namespace {

unsigned grid_for(int64_t items) {
  const int64_t b = (items + kBlock - 1) / kBlock;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}


enum : uint32_t {
  kBadRowptr = 1u,
  kBadCol = 2u,
  kBadWeight = 4u,
  kNegWeight = 8u,
};

template <typename Off>
__global__ void k_validate(const Off* __restrict__ rowptr, const int32_t* __restrict__ col,
                           const uint32_t* __restrict__ w, int64_t rows, int64_t cols,
                           int64_t nnz, int32_t wtype, int32_t allow_negative,
                           uint32_t* __restrict__ flags) {
  uint32_t f = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t r = tid; r < rows; r += stride) {
    const int64_t a = (int64_t)rowptr[r], b = (int64_t)rowptr[r + 1];
    if (b < a || a < 0 || b > nnz) f |= kBadRowptr;
  }
  if (tid == 0 && ((int64_t)rowptr[0] != 0 || (int64_t)rowptr[rows] != nnz)) f |= kBadRowptr;
  for (int64_t e = tid; e < nnz; e += stride) {
    const int32_t c = col[e];
    if (c < 0 || (int64_t)c >= cols) f |= kBadCol;
    if (wtype == MI355_W_F32) {
      // allow_negative: bit 0 admits negative weights (incl. -inf), bit 1
      // admits +inf (an absent edge in the reference's sparse semantics).
      const float x = __uint_as_float(w[e]);
      if (isnan(x)) f |= kBadWeight;
      else if (x < 0.0f) f |= (allow_negative & 1) ? kNegWeight : kBadWeight;
      else if (isinf(x) && !(allow_negative & 2)) f |= kBadWeight;
    }
  }
  if (f) atomicOr(flags, f);
}

template <typename Off>
__global__ void k_narrow_offsets(const Off* __restrict__ in, int* __restrict__ out, int64_t count) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    out[i] = (int)in[i];
}

__global__ void k_interleave(const int32_t* __restrict__ col, const uint32_t* __restrict__ w,
                             uint2* __restrict__ edges, int64_t m) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride)
    edges[e] = make_uint2((uint32_t)col[e], w[e]);
}

// Row id of every edge of a CSR (thread per row).
__global__ void k_expand_rows(const int* __restrict__ rowptr, uint32_t* __restrict__ rowid,
                             int64_t rows) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += stride)
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) rowid[e] = (uint32_t)r;
}

__global__ void k_iota(uint32_t* __restrict__ v, int64_t m) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
    v[i] = (uint32_t)i;
}

// After sorting by source vertex: edges[i] = {row of the incoming edge, w}.
__global__ void k_gather_transposed(const uint32_t* __restrict__ order,
                                    const uint32_t* __restrict__ rowid,
                                    const uint32_t* __restrict__ w, uint2* __restrict__ edges,
                                    int64_t m) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    const uint32_t e = order[i];
    edges[i] = make_uint2(rowid[e], w[e]);
  }
}

// rowptr from sorted keys: rowptr[u] = first i with key[i] >= u (a thread per
// row and a binary search, so a long run of empty rows is not one thread's
// loop).
__global__ void k_offsets_from_sorted(const uint32_t* __restrict__ keys, int* __restrict__ rowptr,
                                      int64_t m, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u <= n; u += stride) {
    int64_t lo = 0, hi = m;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)keys[mid] < u) lo = mid + 1;
      else hi = mid;
    }
    rowptr[u] = (int)lo;
  }
}

// Row of every edge: the last u with rowptr[u] <= e (thread per edge, binary
// search, so hub rows do not serialise on one thread).
__global__ void k_rowid_search(const int* __restrict__ rowptr, uint32_t* __restrict__ rowid,
                               int64_t n, int64_t m) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
    int64_t lo = 0, hi = n;  // rowptr[lo] <= e < rowptr[hi]
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (rowptr[mid] <= e) lo = mid;
      else hi = mid;
    }
    rowid[e] = (uint32_t)lo;
  }
}

// Sort key of an edge: row in the high word, order-preserving weight key in
// the low word (f32: sign-flipped bits; u32: the value itself).
__global__ void k_row_weight_keys(const uint32_t* __restrict__ rowid, const uint2* __restrict__ edges,
                                  int32_t wtype, unsigned long long* __restrict__ keys,
                                  uint32_t* __restrict__ cols, int64_t m) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
    const uint2 ed = edges[e];
    const uint32_t wk = wtype == MI355_W_F32 ? f32_to_key(__uint_as_float(ed.y)) : ed.y;
    keys[e] = ((unsigned long long)rowid[e] << 32) | wk;
    cols[e] = ed.x;
  }
}

__global__ void k_row_weight_gather(const unsigned long long* __restrict__ keys,
                                    const uint32_t* __restrict__ cols, int32_t wtype,
                                    uint2* __restrict__ edges, int64_t m) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
    const uint32_t wk = (uint32_t)keys[e];
    edges[e] = make_uint2(cols[e], wtype == MI355_W_F32 ? __float_as_uint(key_to_f32(wk)) : wk);
  }
}

// Order every row's edges by weight (stable: equal weights keep their column
// order).  Distances and canonical predecessors do not depend on the order of
// a row; a weight-sorted row turns delta-stepping's light/heavy edge classes
// into a prefix and a suffix of the row.
void sort_rows_by_weight(mi355_graph& g, hipStream_t stream) {
  const int64_t n = g.n, m = g.m;
  if (m == 0) return;
  DevBuf<uint32_t> rowid((size_t)m), cols_in((size_t)m), cols_out((size_t)m);
  DevBuf<unsigned long long> keys_in((size_t)m), keys_out((size_t)m);
  k_rowid_search<<<grid_for(m), kBlock, 0, stream>>>(g.rowptr.p, rowid.p, n, m);
  k_row_weight_keys<<<grid_for(m), kBlock, 0, stream>>>(rowid.p, g.edges.p, g.weight_type, keys_in.p,
                                                        cols_in.p, m);
  MI_LAUNCH_CHECK();
  int bits = 1;
  while (bits < 32 && ((int64_t)1 << bits) < n) ++bits;
  size_t temp_bytes = 0;
  MI_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys_in.p, keys_out.p, cols_in.p, cols_out.p,
                                   (size_t)m, 0, 32 + bits, stream));
  DevBuf<unsigned char> temp(temp_bytes);
  MI_HIP(rocprim::radix_sort_pairs(temp.p, temp_bytes, keys_in.p, keys_out.p, cols_in.p, cols_out.p,
                                   (size_t)m, 0, 32 + bits, stream));
  k_row_weight_gather<<<grid_for(m), kBlock, 0, stream>>>(keys_out.p, cols_out.p, g.weight_type,
                                                          g.edges.p, m);
  MI_LAUNCH_CHECK();
  MI_HIP(hipStreamSynchronize(stream));
}

// Sort key of an edge u -> v for the transpose: target row in the high word,
// order-preserving weight key in the low word; the value is the source row.
__global__ void k_col_weight_keys(const uint2* __restrict__ edges, int32_t wtype,
                                  unsigned long long* __restrict__ keys, int64_t m) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
    const uint2 ed = edges[e];
    const uint32_t wk = wtype == MI355_W_F32 ? f32_to_key(__uint_as_float(ed.y)) : ed.y;
    keys[e] = ((unsigned long long)ed.x << 32) | wk;
  }
}

__global__ void k_key_rows(const unsigned long long* __restrict__ keys, uint32_t* __restrict__ rows,
                           int64_t m) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride)
    rows[e] = (uint32_t)(keys[e] >> 32);
}

// flag = 1 when the two CSRs differ anywhere (offsets or records, bit for bit).
__global__ void k_csr_differ(const int* __restrict__ rp_a, const uint2* __restrict__ e_a,
                             const int* __restrict__ rp_b, const uint2* __restrict__ e_b, int64_t n,
                             int64_t m, uint32_t* __restrict__ flag) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  bool differ = false;
  const int64_t total = n + 1 > m ? n + 1 : m;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    if (i <= n && rp_a[i] != rp_b[i]) differ = true;
    if (i < m) {
      const uint2 a = e_a[i], b = e_b[i];
      if (a.x != b.x || a.y != b.y) differ = true;
    }
  }
  if (differ) atomicOr(flag, 1u);
}

// Transpose of a CSR (n x n, m records): row v of the result lists {u, w} for
// every record u -> v, ordered by (weight, u) (a stable radix sort over a
// source-ordered input).
void transpose_sorted(const int* rowptr, const uint2* edges, int64_t n, int64_t m, int32_t wtype,
                      DevBuf<int>& out_rp, DevBuf<uint2>& out_e, hipStream_t stream) {
  out_rp.reset((size_t)n + 1);
  out_e.reset((size_t)std::max<int64_t>(m, 1));
  if (m == 0) {
    MI_HIP(hipMemsetAsync(out_rp.p, 0, (size_t)(n + 1) * 4, stream));
    return;
  }
  DevBuf<uint32_t> rowid((size_t)m), vals_out((size_t)m);
  DevBuf<unsigned long long> keys_in((size_t)m), keys_out((size_t)m);
  k_rowid_search<<<grid_for(m), kBlock, 0, stream>>>(rowptr, rowid.p, n, m);
  k_col_weight_keys<<<grid_for(m), kBlock, 0, stream>>>(edges, wtype, keys_in.p, m);
  MI_LAUNCH_CHECK();
  int bits = 1;
  while (bits < 32 && ((int64_t)1 << bits) < n) ++bits;
  size_t temp_bytes = 0;
  MI_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys_in.p, keys_out.p, rowid.p, vals_out.p,
                                   (size_t)m, 0, 32 + bits, stream));
  {
    DevBuf<unsigned char> temp(temp_bytes);
    MI_HIP(rocprim::radix_sort_pairs(temp.p, temp_bytes, keys_in.p, keys_out.p, rowid.p, vals_out.p,
                                     (size_t)m, 0, 32 + bits, stream));
    MI_HIP(hipStreamSynchronize(stream));
  }
  k_key_rows<<<grid_for(m), kBlock, 0, stream>>>(keys_out.p, rowid.p, m);
  k_offsets_from_sorted<<<grid_for(n + 1), kBlock, 0, stream>>>(rowid.p, out_rp.p, m, n);
  k_row_weight_gather<<<grid_for(m), kBlock, 0, stream>>>(keys_out.p, vals_out.p, wtype, out_e.p, m);
  MI_LAUNCH_CHECK();
  MI_HIP(hipStreamSynchronize(stream));
}

// Per-block partial statistics: [0] sum of positive finite weights (double),
// [1] their count, [2] max degree, [3] large vertices, [4] large chunks.
__global__ void k_stats(const int* __restrict__ rowptr, const uint2* __restrict__ edges, int64_t n,
                        int64_t m, int32_t wtype, double* __restrict__ part_sum,
                        unsigned long long* __restrict__ part_cnt) {
  __shared__ double s_sum[kBlock];
  __shared__ unsigned long long s_c[4][kBlock];
  double sum = 0.0;
  unsigned long long cnt = 0, maxdeg = 0, large = 0, chunks = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t e = tid; e < m; e += stride) {
    const uint32_t wb = edges[e].y;
    const double x = wtype == MI355_W_F32 ? (double)__uint_as_float(wb) : (double)wb;
    if (x > 0.0 && isfinite(x)) {
      sum += x;
      ++cnt;
    }
  }
  for (int64_t v = tid; v < n; v += stride) {
    const unsigned long long d = (unsigned long long)(rowptr[v + 1] - rowptr[v]);
    maxdeg = d > maxdeg ? d : maxdeg;
    if (d > (unsigned long long)kLargeDegree) {
      ++large;
      chunks += (d + kChunk - 1) / kChunk;
    }
  }
  s_sum[threadIdx.x] = sum;
  s_c[0][threadIdx.x] = cnt;
  s_c[1][threadIdx.x] = maxdeg;
  s_c[2][threadIdx.x] = large;
  s_c[3][threadIdx.x] = chunks;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
      s_c[0][threadIdx.x] += s_c[0][threadIdx.x + s];
      s_c[1][threadIdx.x] = max(s_c[1][threadIdx.x], s_c[1][threadIdx.x + s]);
      s_c[2][threadIdx.x] += s_c[2][threadIdx.x + s];
      s_c[3][threadIdx.x] += s_c[3][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_sum[blockIdx.x] = s_sum[0];
    for (int k = 0; k < 4; ++k) part_cnt[blockIdx.x * 4 + k] = s_c[k][0];
  }
}

// Sort key of a vertex for the degree order: ascending keys = descending
// degree (a stable sort keeps input order among equal degrees).
__global__ void k_degree_keys(const int* __restrict__ rowptr, uint32_t* __restrict__ keys,
                              uint32_t* __restrict__ ids, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    keys[v] = ~(uint32_t)(rowptr[v + 1] - rowptr[v]);
    ids[v] = (uint32_t)v;
  }
}

__global__ void k_inverse(const int* __restrict__ perm, int* __restrict__ inv, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    inv[perm[i]] = (int)i;
}

// Degree of output row i = input row rows[i].
__global__ void k_gathered_degrees(const int* __restrict__ rowptr, const int* __restrict__ rows,
                                   int* __restrict__ deg, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    deg[i] = rowptr[rows[i] + 1] - rowptr[rows[i]];
}

// Output edge j of row r (binary search over the output offsets) is edge
// j - out_rp[r] of input row rows[r], its destination renamed by ids[].
__global__ void k_permute_edges(const int* __restrict__ in_rp, const uint2* __restrict__ in_e,
                                const int* __restrict__ out_rp, const int* __restrict__ rows,
                                const int* __restrict__ ids, uint2* __restrict__ out_e, int64_t n,
                                int64_t m) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
    int64_t lo = 0, hi = n;  // out_rp[lo] <= j < out_rp[hi]
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (out_rp[mid] <= j) lo = mid;
      else hi = mid;
    }
    const uint2 e = in_e[in_rp[rows[lo]] + (j - out_rp[lo])];
    out_e[j] = make_uint2((uint32_t)ids[e.x], e.y);
  }
}

// Renumbered copy of a CSR: output row i is input row rows[i], and every
// destination x becomes ids[x].  Row contents keep their order.
void permute_csr(const int* in_rp, const uint2* in_e, const int* rows, const int* ids, int64_t n,
                 int64_t m, int* out_rp, uint2* out_e, hipStream_t stream) {
  DevBuf<int> deg((size_t)n);
  k_gathered_degrees<<<grid_for(n), kBlock, 0, stream>>>(in_rp, rows, deg.p, n);
  MI_LAUNCH_CHECK();
  MI_HIP(hipMemsetAsync(out_rp, 0, sizeof(int), stream));
  size_t temp_bytes = 0;
  MI_HIP(rocprim::inclusive_scan(nullptr, temp_bytes, deg.p, out_rp + 1, (size_t)n, rocprim::plus<int>(),
                                 stream));
  DevBuf<unsigned char> temp(std::max<size_t>(temp_bytes, 1));
  MI_HIP(rocprim::inclusive_scan(temp.p, temp_bytes, deg.p, out_rp + 1, (size_t)n, rocprim::plus<int>(),
                                 stream));
  if (m > 0) k_permute_edges<<<grid_for(m), kBlock, 0, stream>>>(in_rp, in_e, out_rp, rows, ids, out_e, n, m);
  MI_LAUNCH_CHECK();
  MI_HIP(hipStreamSynchronize(stream));
}

// Renumber g by descending degree (ties keep input order).
void renumber_by_degree(mi355_graph& g, hipStream_t stream) {
  const int64_t n = g.n, m = g.m;
  DevBuf<uint32_t> keys_in((size_t)n), keys_out((size_t)n), ids((size_t)n);
  g.perm.reset((size_t)n);
  g.inv.reset((size_t)n);
  k_degree_keys<<<grid_for(n), kBlock, 0, stream>>>(g.rowptr.p, keys_in.p, ids.p, n);
  MI_LAUNCH_CHECK();
  size_t temp_bytes = 0;
  uint32_t* perm = reinterpret_cast<uint32_t*>(g.perm.p);
  MI_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys_in.p, keys_out.p, ids.p, perm, (size_t)n, 0, 32,
                                   stream));
  DevBuf<unsigned char> temp(std::max<size_t>(temp_bytes, 1));
  MI_HIP(rocprim::radix_sort_pairs(temp.p, temp_bytes, keys_in.p, keys_out.p, ids.p, perm, (size_t)n, 0, 32,
                                   stream));
  k_inverse<<<grid_for(n), kBlock, 0, stream>>>(g.perm.p, g.inv.p, n);
  MI_LAUNCH_CHECK();
  DevBuf<int> rp((size_t)n + 1);
  DevBuf<uint2> e((size_t)std::max<int64_t>(m, 1));
  permute_csr(g.rowptr.p, g.edges.p, g.perm.p, g.inv.p, n, m, rp.p, e.p, stream);
  std::swap(g.rowptr, rp);
  std::swap(g.edges, e);
  g.inv_host.resize((size_t)n);
  MI_HIP(hipMemcpyAsync(g.inv_host.data(), g.inv.p, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
  MI_HIP(hipStreamSynchronize(stream));
  g.relabeled = true;
}

}  // namespace

void input_order_arrays(mi355_graph& g, const int** rowptr, const uint2** edges, hipStream_t stream) {
  if (!g.relabeled) {
    *rowptr = g.rowptr.p;
    *edges = g.edges.p;
    return;
  }
  if (!g.in_rowptr.p) {
    // input row x is internal row inv[x]; internal destinations become perm[]
    DevBuf<int> rp((size_t)g.n + 1);
    DevBuf<uint2> e((size_t)std::max<int64_t>(g.m, 1));
    permute_csr(g.rowptr.p, g.edges.p, g.inv.p, g.perm.p, g.n, g.m, rp.p, e.p, stream);
    std::swap(g.in_rowptr, rp);
    std::swap(g.in_edges, e);
  }
  *rowptr = g.in_rowptr.p;
  *edges = g.in_edges.p;
}

// The transpose T of the graph, ordered by (row, weight, column), and T's
// transpose, which is the graph itself in that order: the graph is symmetric
// exactly when the two are equal, and then its own rows are its in-edge rows
// (a pull reads a row by weight only, so the order of equal weights in it does
// not matter).  Otherwise T is kept.
// 1 + the last row with an entry (0 when every row is empty).
__global__ void k_last_nonempty_row(const int* __restrict__ rowptr, int64_t n, unsigned long long* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long best = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
    if (rowptr[v + 1] > rowptr[v]) best = (unsigned long long)(v + 1);
  if (best) atomicMax(out, best);
}

void pull_rows(mi355_graph& g, const int** rowptr, const uint2** edges, hipStream_t stream) {
  if (g.pull_state == 0) {
    DevBuf<int> t_rp, tt_rp;
    DevBuf<uint2> t_e, tt_e;
    transpose_sorted(g.rowptr.p, g.edges.p, g.n, g.m, g.weight_type, t_rp, t_e, stream);
    transpose_sorted(t_rp.p, t_e.p, g.n, g.m, g.weight_type, tt_rp, tt_e, stream);
    DevBuf<uint32_t> flag(1);
    MI_HIP(hipMemsetAsync(flag.p, 0, 4, stream));
    k_csr_differ<<<grid_for(std::max(g.n + 1, g.m)), kBlock, 0, stream>>>(t_rp.p, t_e.p, tt_rp.p, tt_e.p,
                                                                          g.n, g.m, flag.p);
    MI_LAUNCH_CHECK();
    uint32_t differ = 0;
    MI_HIP(hipMemcpyAsync(&differ, flag.p, 4, hipMemcpyDeviceToHost, stream));
    MI_HIP(hipStreamSynchronize(stream));
    if (differ) {
      g.t_rowptr = std::move(t_rp);
      g.t_edges = std::move(t_e);
      g.pull_state = 2;
    } else {
      g.pull_state = 1;
    }
    DevBuf<unsigned long long> last(1);
    MI_HIP(hipMemsetAsync(last.p, 0, 8, stream));
    k_last_nonempty_row<<<grid_for(g.n), kBlock, 0, stream>>>(g.pull_state == 2 ? g.t_rowptr.p : g.rowptr.p, g.n,
                                                              last.p);
    MI_LAUNCH_CHECK();
    unsigned long long h_last = 0;
    MI_HIP(hipMemcpyAsync(&h_last, last.p, 8, hipMemcpyDeviceToHost, stream));
    MI_HIP(hipStreamSynchronize(stream));
    g.pull_n = (int64_t)h_last;
  }
  *rowptr = g.pull_state == 2 ? g.t_rowptr.p : g.rowptr.p;
  *edges = g.pull_state == 2 ? g.t_edges.p : g.edges.p;
}

mi355_graph* build_graph(const mi355_csr_desc& d, int allow_negative, hipStream_t stream,
                         int32_t order) {
  if (order != kOrderAuto && order != kOrderInput && order != kOrderDegree)
    fail(MI355_EINVAL, "unknown vertex order");
  if (d.rows <= 0) fail(MI355_EINVAL, "CSR graph must contain at least one vertex");
  if (d.rows != d.cols) fail(MI355_EINVAL, "SSSP expects a square CSR adjacency matrix");
  if (d.nnz < 0) fail(MI355_EINVAL, "CSR nnz must be nonnegative");
  if (d.rows >= INT32_MAX) fail(MI355_EOVERFLOW, "rows must fit in int32");
  if (d.nnz >= INT32_MAX) fail(MI355_EOVERFLOW, "nnz must fit in int32 offsets");
  if (!d.rowptr || (d.nnz > 0 && (!d.col || !d.w))) fail(MI355_EINVAL, "null CSR array");
  if (d.offset_bits != 32 && d.offset_bits != 64) fail(MI355_EINVAL, "offset_bits must be 32 or 64");
  if (d.weight_type != MI355_W_F32 && d.weight_type != MI355_W_U32)
    fail(MI355_EINVAL, "unknown weight type");
  if (d.orientation != MI355_OUT && d.orientation != MI355_IN)
    fail(MI355_EINVAL, "unknown orientation");
  if (d.memory != MI355_HOST && d.memory != MI355_DEVICE) fail(MI355_EINVAL, "unknown memory kind");

  const int64_t n = d.rows, m = d.nnz;
  const size_t off_bytes = (size_t)(n + 1) * (d.offset_bits / 8);

  // Stage host input on the device so both paths share the kernels below.
  DevBuf<unsigned char> h_rowptr, h_col, h_w;
  const void* rp = d.rowptr;
  const int32_t* col = static_cast<const int32_t*>(d.col);
  const uint32_t* w = static_cast<const uint32_t*>(d.w);
  if (d.memory == MI355_HOST) {
    h_rowptr.reset(off_bytes);
    MI_HIP(hipMemcpyAsync(h_rowptr.p, d.rowptr, off_bytes, hipMemcpyHostToDevice, stream));
    rp = h_rowptr.p;
    if (m > 0) {
      h_col.reset((size_t)m * 4);
      h_w.reset((size_t)m * 4);
      MI_HIP(hipMemcpyAsync(h_col.p, d.col, (size_t)m * 4, hipMemcpyHostToDevice, stream));
      MI_HIP(hipMemcpyAsync(h_w.p, d.w, (size_t)m * 4, hipMemcpyHostToDevice, stream));
      col = reinterpret_cast<const int32_t*>(h_col.p);
      w = reinterpret_cast<const uint32_t*>(h_w.p);
    }
  }

  DevBuf<uint32_t> flags(1);
  MI_HIP(hipMemsetAsync(flags.p, 0, 4, stream));
  const unsigned gv = grid_for(std::max(n, m));
  if (d.offset_bits == 64)
    k_validate<int64_t><<<gv, kBlock, 0, stream>>>(static_cast<const int64_t*>(rp), col, w, n,
                                                  d.cols, m, d.weight_type, allow_negative, flags.p);
  else
    k_validate<uint32_t><<<gv, kBlock, 0, stream>>>(static_cast<const uint32_t*>(rp), col, w, n,
                                                   d.cols, m, d.weight_type, allow_negative,
                                                   flags.p);
  MI_LAUNCH_CHECK();
  uint32_t hflags = 0;
  MI_HIP(hipMemcpyAsync(&hflags, flags.p, 4, hipMemcpyDeviceToHost, stream));
  MI_HIP(hipStreamSynchronize(stream));
  if (hflags & kBadRowptr) fail(MI355_EINVAL, "CSR rowptr must be monotone and span [0, nnz]");
  if (hflags & kBadCol) fail(MI355_EINVAL, "CSR column index out of range");
  if (hflags & kBadWeight)
    fail(MI355_EINVAL, (allow_negative & 1) ? "edge weights must not be NaN"
                                            : "edge weights must be finite and nonnegative");

  auto g = new mi355_graph();
  try {
    g->device = device_props().device;
    g->n = n;
    g->m = m;
    g->weight_type = d.weight_type;
    g->rowptr.reset((size_t)n + 1);
    g->edges.reset((size_t)std::max<int64_t>(m, 1));
    DevBuf<int> in_rowptr;
    int* narrowed = g->rowptr.p;
    if (d.orientation == MI355_IN) {
      in_rowptr.reset((size_t)n + 1);
      narrowed = in_rowptr.p;
    }
    if (d.offset_bits == 64)
      k_narrow_offsets<int64_t><<<grid_for(n + 1), kBlock, 0, stream>>>(
          static_cast<const int64_t*>(rp), narrowed, n + 1);
    else
      k_narrow_offsets<uint32_t><<<grid_for(n + 1), kBlock, 0, stream>>>(
          static_cast<const uint32_t*>(rp), narrowed, n + 1);
    MI_LAUNCH_CHECK();

    if (d.orientation == MI355_OUT) {
      if (m > 0) {
        k_interleave<<<grid_for(m), kBlock, 0, stream>>>(col, w, g->edges.p, m);
        MI_LAUNCH_CHECK();
      }
    } else if (m == 0) {
      MI_HIP(hipMemsetAsync(g->rowptr.p, 0, (size_t)(n + 1) * 4, stream));
    } else {
      DevBuf<uint32_t> rowid((size_t)m), keys_out((size_t)m), order_in((size_t)m),
          order_out((size_t)m);
      k_expand_rows<<<grid_for(n), kBlock, 0, stream>>>(in_rowptr.p, rowid.p, n);
      k_iota<<<grid_for(m), kBlock, 0, stream>>>(order_in.p, m);
      MI_LAUNCH_CHECK();
      int bits = 1;
      while (bits < 32 && ((int64_t)1 << bits) < n) ++bits;
      size_t temp_bytes = 0;
      const uint32_t* keys_in = reinterpret_cast<const uint32_t*>(col);
      MI_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys_in, keys_out.p, order_in.p,
                                       order_out.p, (size_t)m, 0, bits, stream));
      DevBuf<unsigned char> temp(temp_bytes);
      MI_HIP(rocprim::radix_sort_pairs(temp.p, temp_bytes, keys_in, keys_out.p, order_in.p,
                                       order_out.p, (size_t)m, 0, bits, stream));
      k_gather_transposed<<<grid_for(m), kBlock, 0, stream>>>(order_out.p, rowid.p, w,
                                                              g->edges.p, m);
      k_offsets_from_sorted<<<grid_for(n + 1), kBlock, 0, stream>>>(keys_out.p, g->rowptr.p, m, n);
      MI_LAUNCH_CHECK();
      MI_HIP(hipStreamSynchronize(stream));
    }

    sort_rows_by_weight(*g, stream);

    // statistics (deterministic: fixed grid, fixed-order reductions)
    const unsigned sb = 512;
    DevBuf<double> psum(sb);
    DevBuf<unsigned long long> pcnt(sb * 4);
    k_stats<<<sb, kBlock, 0, stream>>>(g->rowptr.p, g->edges.p, n, m, d.weight_type, psum.p,
                                       pcnt.p);
    MI_LAUNCH_CHECK();
    std::vector<double> hs(sb);
    std::vector<unsigned long long> hc(sb * 4);
    MI_HIP(hipMemcpyAsync(hs.data(), psum.p, sb * sizeof(double), hipMemcpyDeviceToHost, stream));
    MI_HIP(hipMemcpyAsync(hc.data(), pcnt.p, sb * 4 * 8, hipMemcpyDeviceToHost, stream));
    MI_HIP(hipStreamSynchronize(stream));
    double sum = 0.0;
    unsigned long long cnt = 0, maxdeg = 0, large = 0, chunks = 0;
    for (unsigned b = 0; b < sb; ++b) {
      sum += hs[b];
      cnt += hc[b * 4 + 0];
      maxdeg = std::max(maxdeg, hc[b * 4 + 1]);
      large += hc[b * 4 + 2];
      chunks += hc[b * 4 + 3];
    }
    mi355_graph_info& inf = g->info;
    inf.n = n;
    inf.m = m;
    inf.weight_type = d.weight_type;
    inf.max_degree = (int32_t)maxdeg;
    inf.avg_degree = (double)m / (double)n;
    inf.avg_weight = cnt ? sum / (double)cnt : 1.0;
    inf.large_vertices = (int64_t)large;
    inf.large_chunks = (int64_t)chunks;
    inf.has_negative = (hflags & kNegWeight) ? 1 : 0;
    inf.device = g->device;
    // Degree order for skewed graphs: most relaxations then target the few
    // high-degree vertices, whose labels and stamps share cache lines.
    const bool skewed = n >= 4096 && (double)maxdeg >= 32.0 * inf.avg_degree;
    if (m > 0 && (order == kOrderDegree || (order == kOrderAuto && skewed))) {
      // the delta policy's features are sampled in the input numbering
      float feat[14];
      graph_features(*g, stream, feat);
      renumber_by_degree(*g, stream);
    }
    inf.vertex_order = g->relabeled ? kOrderDegree : kOrderInput;
  } catch (...) {
    delete g;
    throw;
  }
  return g;
}

}  // namespace mi355
