// Dense-matrix Bellman-Ford and Gauss-Seidel (the reference's dense paths).
//
// bellman_ford_minplus_hip (BF/src/bellman_ford_hip.cpp:64-264) runs
// d_k = min(d_{k-1}, A (x) d_{k-1}) with the product done as a 64x64-tiled
// GEMM with N = 1 (63/64 of every tile wasted) plus a merge kernel and a host
// round trip per iteration.  Here one wave owns one row v of the incoming
// matrix: lanes stream the row with coalesced loads, each lane folds
// min(A[v,u] + d[u]) over its columns, a 6-step butterfly reduces the wave,
// and lane 0 merges into d_next and flags a change.  The kernel is HBM-bound
// (n^2 * 4 bytes per iteration).  Iterations are enqueued in batches; an
// iteration whose predecessor changed nothing exits at once, and the host
// reads the per-iteration change flags once per batch.
//
// bellman_ford_minplus_hip_gauss_seidel (BF/src/bf_gauss_seidel.cpp:37-62)
// is inherently sequential in v; its sweep runs in one 1024-thread workgroup
// that reduces each row in parallel instead of one thread doing n^2 work.
#include <algorithm>
#include <vector>

#include "minplus.hpp"

namespace mi355 {

namespace {

__device__ inline float wave_min(float x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fminf(x, __shfl_xor(x, off, 64));
  return x;
}

__global__ __launch_bounds__(256) void k_dense_bf_iter(const float* __restrict__ A, int n,
                                                       const float* __restrict__ d_prev,
                                                       float* __restrict__ d_next,
                                                       int* __restrict__ changed, int iter) {
  if (iter > 0 && changed[iter - 1] == 0) return;  // converged: nothing to do
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int v = wave; v < n; v += nwaves) {
    const float* row = A + (int64_t)v * n;
    float acc = __builtin_huge_valf();
    for (int u = lane; u < n; u += 64) acc = fminf(acc, row[u] + d_prev[u]);
    acc = wave_min(acc);
    if (lane == 0) {
      // relax_dist_kernel (BF/src/bellman_ford_hip.cpp:43-59)
      const float old = d_prev[v];
      const float nv = acc < old ? acc : old;
      d_next[v] = nv;
      if (nv < old) changed[iter] = 1;
    }
  }
}

__global__ void k_dense_init(float* __restrict__ d, int n, int src, int* __restrict__ changed,
                             int nflags) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    d[i] = i == src ? 0.0f : __builtin_huge_valf();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nflags; i += gridDim.x * blockDim.x)
    changed[i] = 0;
}

// All sweeps in one workgroup: for v in order, best = min(d[v], min_{u != v}
// d[u] + A[v,u]) with d updated in place (u < v already swept).
__global__ __launch_bounds__(1024) void k_gauss_seidel(const float* __restrict__ A, int n,
                                                       float* __restrict__ d, int sweeps) {
  __shared__ float s_part[16];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  for (int it = 0; it < sweeps; ++it) {
    for (int v = 0; v < n; ++v) {
      const float* row = A + (int64_t)v * n;
      float acc = __builtin_huge_valf();
      for (int u = tid; u < n; u += 1024)
        if (u != v) acc = fminf(acc, d[u] + row[u]);
      acc = wave_min(acc);
      if (lane == 0) s_part[wave] = acc;
      __syncthreads();
      if (wave == 0) {
        float x = lane < 16 ? s_part[lane] : __builtin_huge_valf();
        x = wave_min(x);
        if (lane == 0) {
          const float best = d[v];
          d[v] = x < best ? x : best;
        }
      }
      __syncthreads();
    }
  }
}

}  // namespace

int dense_bf(const float* A, int n, int source, int max_iters, float* d_dist, int* converged,
             hipStream_t s) {
  DevBuf<float> other((size_t)n);
  DevBuf<int> changed((size_t)std::max(max_iters, 1));
  k_dense_init<<<std::min<unsigned>(div_up((uint64_t)std::max(n, max_iters), 256), 1024u), 256, 0, s>>>(
      d_dist, n, source, changed.p, std::max(max_iters, 1));
  MI_LAUNCH_CHECK();
  float* buf[2] = {d_dist, other.p};
  const unsigned grid = std::min<unsigned>(div_up((uint64_t)n, 4), 8192u);
  int it = 0, batch = 1, used = 0;
  *converged = 0;
  std::vector<int> flags;
  while (it < max_iters) {
    const int k = std::min(batch, max_iters - it);
    for (int i = 0; i < k; ++i, ++it)
      k_dense_bf_iter<<<grid, 256, 0, s>>>(A, n, buf[it & 1], buf[(it + 1) & 1], changed.p, it);
    MI_LAUNCH_CHECK();
    flags.resize((size_t)it);
    MI_HIP(hipMemcpyAsync(flags.data(), changed.p, (size_t)it * 4, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    bool done = false;
    for (int i = 0; i < it; ++i)
      if (flags[i] == 0) {
        used = i + 1;
        *converged = 1;
        done = true;
        break;
      }
    if (done) break;
    used = it;
    batch = std::min(batch * 2, 32);
  }
  // Iteration i (0-based) writes buf[(i+1)&1]; the result of the last
  // executed iteration (index used-1) holds the distances.
  if (used > 0 && (used & 1)) MI_HIP(hipMemcpyAsync(d_dist, other.p, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  MI_HIP(hipStreamSynchronize(s));
  return used;
}

void dense_gauss_seidel(const float* A, int n, int source, int max_iters, float* d_dist,
                        hipStream_t s) {
  DevBuf<int> dummy(1);
  k_dense_init<<<std::min<unsigned>(div_up((uint64_t)n, 256), 1024u), 256, 0, s>>>(d_dist, n, source,
                                                                                  dummy.p, 1);
  MI_LAUNCH_CHECK();
  if (max_iters > 0) {
    k_gauss_seidel<<<1, 1024, 0, s>>>(A, n, d_dist, max_iters);
    MI_LAUNCH_CHECK();
  }
  MI_HIP(hipStreamSynchronize(s));
}

}  // namespace mi355
