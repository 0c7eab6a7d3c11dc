// Min-plus GEMM (minplus.hip) and dense Bellman-Ford / Gauss-Seidel (dense.hip).
#pragma once

#include <vector>

#include "common.hpp"

namespace mi355 {

struct MinplusItem {
  const float* A;
  const float* B;
  float* C;
  int M, N, K;
  int lda, ldb, ldc;
  int pad;
};

int minplus_tile_count();
const char* minplus_tile_name(int i);
// Asynchronous on stream s; scratch_item is a device slot for the descriptor.
void minplus_gemm(const float* A, const float* B, float* C, int M, int N, int K, int lda, int ldb,
                  int ldc, hipStream_t s, MinplusItem* scratch_item, int force_tile = -1);
// A batch's device descriptors and (product, tile) work list, kept for the
// next call: a call with the same products (pointers and sizes) on the same
// device and tuning launches again without rebuilding or uploading anything.
// Kernels only read these buffers, so reuse needs no synchronisation; a
// rebuild first waits for the event recorded after the last launches that
// read them (whatever stream those ran on).
struct MinplusBatchCache {
  DevBuf<MinplusItem> table;
  DevBuf<int2> work;
  std::vector<MinplusItem> items;
  std::vector<std::pair<int, uint32_t>> launches;  // (tile, workgroups), in work-list order
  uint64_t tune_gen = 0;
  int device = -1;
  hipEvent_t done = nullptr;  // recorded after the last launches, on the device of `device`
  bool valid = false;
  MinplusBatchCache() = default;
  MinplusBatchCache(const MinplusBatchCache&) = delete;
  MinplusBatchCache& operator=(const MinplusBatchCache&) = delete;
  ~MinplusBatchCache() {
    if (done) (void)hipEventDestroy(done);
  }
};

void minplus_gemm_batched(const std::vector<MinplusItem>& items, hipStream_t s, MinplusBatchCache& cache);
// The (product, tile) work list of a batch and its launches (tile shape,
// workgroups), in work-list order (host only).
void minplus_plan(const std::vector<MinplusItem>& items, std::vector<int2>& work,
                  std::vector<std::pair<int, uint32_t>>& launches);
// Launch a plan whose descriptors and work list are in device memory.
void minplus_launch_plan(const std::vector<std::pair<int, uint32_t>>& launches, const MinplusItem* d_table,
                         const int2* d_work, hipStream_t s);
// The batch with its plan in the library's fixed per-GPU array
// (minplus_noalloc.hip): no allocation; at most minplus_noalloc_max_items()
// products and minplus_noalloc_max_tiles() work entries.
void minplus_gemm_batched_noalloc(const std::vector<MinplusItem>& items, hipStream_t s);
size_t minplus_noalloc_max_items();
size_t minplus_noalloc_max_tiles();
void minplus_autotune(const int* sizes, int count, hipStream_t s, double* report);
void minplus_tuning_get(int* table32);
void minplus_tuning_set(const int* table32);

// Dense incoming matrix A[v*n+u] = w(u->v); dist is device memory (n floats).
// Returns iterations_used; converged via *converged.
int dense_bf(const float* A, int n, int source, int max_iters, float* d_dist, int* converged,
             hipStream_t s);
void dense_gauss_seidel(const float* A, int n, int source, int max_iters, float* d_dist,
                        hipStream_t s);

}  // namespace mi355
