// Dense min-plus ("tropical") GEMM on the VALU: C[i,j] = min_k A[i,k] + B[k,j].
//
// Replaces minplus_gemm_f32 / minplus_kernel_f32 (MP/src/minplus_hip.cpp:7-173).
// (min,+) is not a ring contraction, so MFMA cannot compute it; the kernel is
// a register-blocked VALU tile:
//   * a 256-thread workgroup owns a BM x BN tile of C; each thread a TM x TN
//     micro-tile held in registers (TM*TN accumulators);
//   * A and B k-slices of depth BK are staged through LDS, both stored k-major
//     (A transposed on the way in) so a thread fetches its TM A values and TN
//     B values of one k with 128-bit LDS reads;
//   * LDS is double-buffered: the next slice is loaded from HBM into registers
//     while the current one is consumed, then written to the other buffer, so
//     each k-slice costs one barrier instead of two;
//   * two consecutive k feed one v_min3_f32 (acc = min(acc, a0+b0, a1+b1)):
//     3 VALU instructions per 2 (add, min) pairs.  Packed fp32 adds are not
//     used (the build disables them);
//   * the k loop within a slice is unrolled UK pairs at a time.  Fully
//     unrolled, the compiler hoists every LDS fragment of the slice into
//     registers (193-256 VGPRs, one or two waves per SIMD); UK = 2 or 4 keeps
//     64x64 tiles at 80-120 VGPRs and four to six waves per SIMD, which is
//     what a launch with many workgroups needs.  A launch with fewer
//     workgroups than CUs runs faster fully unrolled (more independent work
//     per wave).
// Out-of-range elements are padded with +INF, as the reference does (:49-73),
// which leaves every real output unchanged.  NaN candidates never win (the
// reference's select `cand < acc ? cand : acc`, :96-97, behaves the same);
// the sign of a zero result may differ from the reference when -0.0 inputs
// tie with +0.0, the only divergence from its k-ordered select.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <cmath>
#include <mutex>
#include <vector>

#include "minplus.hpp"
#include "xcd_deal.hpp"

namespace mi355 {

namespace {

constexpr float kInf = __builtin_huge_valf();

template <int BM, int BN, int BK, int TM, int TN>
struct TileCfg {
  static constexpr int kThreads = (BM / TM) * (BN / TN);
  static_assert(kThreads == 256, "tile must map to 256 threads");
  static_assert(TM % 4 == 0 && TN % 4 == 0 && BK % 2 == 0, "128-bit LDS fragments, k pairs");
  static constexpr int kTX = BN / TN;  // threads along N
  static constexpr int kTY = BM / TM;  // threads along M
  static constexpr int kALoads = BM * BK / 256;  // A elements per thread per slice
  static constexpr int kBLoads = BK * BN / 256;
  static constexpr int kAPad = 4;  // keeps rows 16-byte aligned, spreads banks
  static constexpr int kBPad = 4;
};

// One workgroup per C tile.  Without a work list the grid is (tiles along N,
// tiles along M) over items[0].  A batch passes `work`: entry blockIdx.x is
// {item, tile index within the item (row-major over its tile grid)}, item -1
// marks padding.
template <int BM, int BN, int BK, int TM, int TN, int UK>
__global__ __launch_bounds__(256) void k_minplus_tile(const MinplusItem* __restrict__ items,
                                                      const int2* __restrict__ work) {
  using Cfg = TileCfg<BM, BN, BK, TM, TN>;
  int item = 0, tile_m = blockIdx.y, tile_n = blockIdx.x;
  if (work) {
    const int2 w = work[blockIdx.x];
    if (w.x < 0) return;
    item = w.x;
    const int tiles_n = (items[item].N + BN - 1) / BN;
    tile_m = w.y / tiles_n;
    tile_n = w.y - tile_m * tiles_n;
  }
  const MinplusItem it = items[item];
  const int row0 = tile_m * BM;
  const int col0 = tile_n * BN;
  if (row0 >= it.M || col0 >= it.N) return;

  __shared__ float As[2][BK][BM + Cfg::kAPad];
  __shared__ float Bs[2][BK][BN + Cfg::kBPad];

  const int tid = threadIdx.x;
  const int tx = tid % Cfg::kTX;
  const int ty = tid / Cfg::kTX;

  float acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = kInf;

  float ra[Cfg::kALoads];
  float rb[Cfg::kBLoads];

  // Global -> register staging of one k-slice.  A: element e of the slice is
  // (r, k) = (e / BK, e % BK) -- consecutive threads read consecutive k of a
  // row.  B: (k, c) = (e / BN, e % BN) -- consecutive threads read a row.
  auto load_slice = [&](int k0) {
#pragma unroll
    for (int l = 0; l < Cfg::kALoads; ++l) {
      const int e = tid + l * 256;
      const int r = e / BK, k = e % BK;
      const int gr = row0 + r, gk = k0 + k;
      ra[l] = (gr < it.M && gk < it.K) ? it.A[(int64_t)gr * it.lda + gk] : kInf;
    }
#pragma unroll
    for (int l = 0; l < Cfg::kBLoads; ++l) {
      const int e = tid + l * 256;
      const int k = e / BN, c = e % BN;
      const int gk = k0 + k, gc = col0 + c;
      rb[l] = (gk < it.K && gc < it.N) ? it.B[(int64_t)gk * it.ldb + gc] : kInf;
    }
  };
  auto store_slice = [&](int buf) {
#pragma unroll
    for (int l = 0; l < Cfg::kALoads; ++l) {
      const int e = tid + l * 256;
      As[buf][e % BK][e / BK] = ra[l];
    }
#pragma unroll
    for (int l = 0; l < Cfg::kBLoads; ++l) {
      const int e = tid + l * 256;
      Bs[buf][e / BN][e % BN] = rb[l];
    }
  };

  const int nslices = (it.K + BK - 1) / BK;
  load_slice(0);
  store_slice(0);
  __syncthreads();
  for (int s = 0; s < nslices; ++s) {
    const int cur = s & 1;
    if (s + 1 < nslices) load_slice((s + 1) * BK);
#pragma unroll UK
    for (int k = 0; k < BK; k += 2) {
      float a0[TM], a1[TM], b0[TN], b1[TN];
#pragma unroll
      for (int i = 0; i < TM; i += 4) {
        const float4 x = *reinterpret_cast<const float4*>(&As[cur][k][ty * TM + i]);
        const float4 y = *reinterpret_cast<const float4*>(&As[cur][k + 1][ty * TM + i]);
        a0[i] = x.x; a0[i + 1] = x.y; a0[i + 2] = x.z; a0[i + 3] = x.w;
        a1[i] = y.x; a1[i + 1] = y.y; a1[i + 2] = y.z; a1[i + 3] = y.w;
      }
#pragma unroll
      for (int j = 0; j < TN; j += 4) {
        const float4 x = *reinterpret_cast<const float4*>(&Bs[cur][k][tx * TN + j]);
        const float4 y = *reinterpret_cast<const float4*>(&Bs[cur][k + 1][tx * TN + j]);
        b0[j] = x.x; b0[j + 1] = x.y; b0[j + 2] = x.z; b0[j + 3] = x.w;
        b1[j] = y.x; b1[j + 1] = y.y; b1[j + 2] = y.z; b1[j + 3] = y.w;
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_fminf(__builtin_fminf(acc[i][j], a0[i] + b0[j]), a1[i] + b1[j]);
    }
    if (s + 1 < nslices) store_slice(cur ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int r = row0 + ty * TM + i;
    if (r >= it.M) continue;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int c = col0 + tx * TN + j;
      if (c < it.N) it.C[(int64_t)r * it.ldc + c] = acc[i][j];
    }
  }
}

using TileKernel = void (*)(const MinplusItem*, const int2*);

struct TileEntry {
  int bm, bn, bk, tm, tn;
  TileKernel fn;
  const char* name;
};

// Index 2 is the default for launches of many workgroups, index 5 for
// launches of few.
constexpr int kTileMany = 2;
constexpr int kTileFew = 5;
const TileEntry kTiles[] = {
    {128, 128, 16, 8, 8, k_minplus_tile<128, 128, 16, 8, 8, 2>, "128x128x16/8x8/u2"},
    {128, 64, 16, 8, 4, k_minplus_tile<128, 64, 16, 8, 4, 2>, "128x64x16/8x4/u2"},
    {64, 64, 16, 4, 4, k_minplus_tile<64, 64, 16, 4, 4, 4>, "64x64x16/4x4/u4"},
    {64, 64, 32, 4, 4, k_minplus_tile<64, 64, 32, 4, 4, 4>, "64x64x32/4x4/u4"},
    {64, 128, 16, 4, 8, k_minplus_tile<64, 128, 16, 4, 8, 2>, "64x128x16/4x8/u2"},
    {64, 64, 16, 4, 4, k_minplus_tile<64, 64, 16, 4, 4, 8>, "64x64x16/4x4/full"},
};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

// Autotuned choice per size class (class = ceil(log2(max(M, N)))); -1 = use
// the static rule.  Guarded by a mutex: tuning may run from several threads.
std::mutex g_tune_mu;
int g_tuned[32];
bool g_tuned_init = false;
uint64_t g_tune_gen = 0;  // bumped on every change of g_tuned

// Static rule by the launch's 64x64 tile count: below 512 (two workgroups
// per CU) the fully unrolled 64x64 tile, else the UK = 4 one.  Measured on
// MI355X, square products: n = 512 single 6.2 vs 5.4 T ops/s, n = 1024 single
// 24.5 vs 19.6, n = 2048 single 45.4 vs 52.9, 512 products of n in [32, 512]
// 28.9 vs 40.5.
constexpr uint64_t kFewTiles = 512;

uint64_t tiles64(int M, int N) { return (uint64_t)div_up((uint64_t)M, 64) * div_up((uint64_t)N, 64); }

int static_choice(uint64_t launch_tiles64) { return launch_tiles64 < kFewTiles ? kTileFew : kTileMany; }

int size_class(int M, int N) {
  int cls = 0;
  while ((1 << cls) < std::max(M, N) && cls < 31) ++cls;
  return cls;
}

// launch_tiles64: 64x64 tiles of the whole launch the product is part of.
int choose_tile(int M, int N, uint64_t launch_tiles64) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (!g_tuned_init) {
    for (int& x : g_tuned) x = -1;
    g_tuned_init = true;
  }
  const int cls = size_class(M, N);
  return g_tuned[cls] >= 0 ? g_tuned[cls] : static_choice(launch_tiles64);
}

void launch_single(int tile, const MinplusItem* d_item, int M, int N, hipStream_t s) {
  const TileEntry& t = kTiles[tile];
  dim3 grid(div_up((uint64_t)N, t.bn), div_up((uint64_t)M, t.bm), 1);
  hipLaunchKernelGGL(t.fn, grid, dim3(256), 0, s, d_item, static_cast<const int2*>(nullptr));
}

}  // namespace

int minplus_tile_count() { return kNumTiles; }
const char* minplus_tile_name(int i) { return (i >= 0 && i < kNumTiles) ? kTiles[i].name : ""; }

void minplus_gemm(const float* A, const float* B, float* C, int M, int N, int K, int lda,
                  int ldb, int ldc, hipStream_t s, MinplusItem* scratch_item, int force_tile) {
  const MinplusItem it{A, B, C, M, N, K, lda, ldb, ldc, 0};
  MI_HIP(hipMemcpyAsync(scratch_item, &it, sizeof(it), hipMemcpyHostToDevice, s));
  const int tile = force_tile >= 0 ? force_tile : choose_tile(M, N, tiles64(M, N));
  launch_single(tile, scratch_item, M, N, s);
  MI_LAUNCH_CHECK();
}

// Work list of a batch: one launch per tile shape (each product takes its
// size class's tuned tile, else the static rule over the whole batch's tile
// count) over a flat list of (product, tile) entries, so every workgroup owns
// a real tile whatever the spread of sizes.  512 products of n in [32, 512]
// on the 64x64 full-unroll tile: one grid sized for the largest product,
// 0.93 ms; one grid per size class, 0.75 ms; the work list, 0.54 ms.
void minplus_plan(const std::vector<MinplusItem>& items, std::vector<int2>& flat,
                  std::vector<std::pair<int, uint32_t>>& launches) {
  uint64_t all_tiles64 = 0;
  for (const MinplusItem& it : items) all_tiles64 += tiles64(it.M, it.N);
  std::vector<std::vector<TileRun>> lists(kNumTiles);
  std::vector<uint64_t> list_tiles(kNumTiles, 0);
  for (size_t i = 0; i < items.size(); ++i) {
    const MinplusItem& it = items[i];
    const int t = choose_tile(it.M, it.N, all_tiles64);
    const uint64_t tiles = (uint64_t)div_up((uint64_t)it.M, kTiles[t].bm) * div_up((uint64_t)it.N, kTiles[t].bn);
    if (tiles > (uint64_t)INT32_MAX) fail(MI355_EINVAL, "min-plus product too large for one batch item");
    lists[t].push_back(TileRun{(int)i, 0, (int)tiles, it.K});
    list_tiles[t] += tiles;
  }
  std::vector<std::vector<int2>> ordered(kNumTiles);
  size_t total = 0;
  for (int t = 0; t < kNumTiles; ++t) {
    if (lists[t].empty()) continue;
    std::vector<TileSlot> dealt;
    xcd_deal(lists[t], list_tiles[t], dealt);
    ordered[t].resize(dealt.size());
    for (size_t j = 0; j < dealt.size(); ++j) ordered[t][j] = int2{dealt[j].item, dealt[j].tile};
    // a launch holds at most 2^32 - 1 threads
    if (ordered[t].size() > (size_t)UINT32_MAX / 256) fail(MI355_EINVAL, "min-plus batch too large");
    total += ordered[t].size();
  }
  flat.clear();
  flat.reserve(total);
  for (auto& o : ordered) flat.insert(flat.end(), o.begin(), o.end());
  launches.clear();
  for (int t = 0; t < kNumTiles; ++t)
    if (!ordered[t].empty()) launches.emplace_back(t, (uint32_t)ordered[t].size());
}

// The launches of a plan whose descriptors and work list are on the device.
void minplus_launch_plan(const std::vector<std::pair<int, uint32_t>>& launches, const MinplusItem* d_table,
                         const int2* d_work, hipStream_t s) {
  size_t off = 0;
  for (const auto& l : launches) {
    hipLaunchKernelGGL(kTiles[l.first].fn, dim3(l.second), dim3(256), 0, s, d_table, d_work + off);
    MI_LAUNCH_CHECK();
    off += l.second;
  }
}

// Descriptors and work list of a batch, uploaded into the cache's buffers.
static void build_batch(const std::vector<MinplusItem>& items, hipStream_t s, MinplusBatchCache& cache) {
  DevBuf<MinplusItem>& table = cache.table;
  DevBuf<int2>& work = cache.work;
  table.ensure(items.size());
  MI_HIP(hipMemcpyAsync(table.p, items.data(), items.size() * sizeof(MinplusItem), hipMemcpyHostToDevice, s));
  std::vector<int2> flat;
  minplus_plan(items, flat, cache.launches);
  work.ensure(flat.size());
  MI_HIP(hipMemcpyAsync(work.p, flat.data(), flat.size() * sizeof(int2), hipMemcpyHostToDevice, s));
}

void minplus_gemm_batched(const std::vector<MinplusItem>& items, hipStream_t s, MinplusBatchCache& cache) {
  if (items.empty()) return;
  int dev = 0;
  MI_HIP(hipGetDevice(&dev));
  uint64_t gen;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    gen = g_tune_gen;
  }
  const bool same = cache.valid && cache.device == dev && cache.tune_gen == gen && cache.items.size() == items.size() &&
                    std::memcmp(cache.items.data(), items.data(), items.size() * sizeof(MinplusItem)) == 0;
  if (!same) {
    // the previous call's kernels may still read the buffers being replaced
    if (cache.done) MI_HIP(hipEventSynchronize(cache.done));
    cache.valid = false;
    if (cache.device != dev) {
      if (cache.done) (void)hipEventDestroy(cache.done);
      cache.done = nullptr;
      cache.table.release();
      cache.work.release();
      MI_HIP(hipEventCreateWithFlags(&cache.done, hipEventDisableTiming));
      cache.device = dev;
    }
    build_batch(items, s, cache);
    cache.items = items;
    cache.tune_gen = gen;
    cache.valid = true;
  }
  minplus_launch_plan(cache.launches, cache.table.p, cache.work.p, s);
  MI_HIP(hipEventRecord(cache.done, s));
}

// Times every tile variant on a sizes x sizes product per size class and keeps
// the fastest.  Returns per-class timings through `report` (may be null).
void minplus_autotune(const int* sizes, int count, hipStream_t s, double* report) {
  for (int q = 0; q < count; ++q) {
    const int n = sizes[q];
    if (n <= 0) fail(MI355_EINVAL, "autotune sizes must be positive");
    DevBuf<float> A((size_t)n * n), B((size_t)n * n), C((size_t)n * n);
    MI_HIP(hipMemsetAsync(A.p, 0, (size_t)n * n * 4, s));
    MI_HIP(hipMemsetAsync(B.p, 0, (size_t)n * n * 4, s));
    DevBuf<MinplusItem> item(1);
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    int best = 0;
    float best_ms = 1e30f;
    for (int t = 0; t < kNumTiles; ++t) {
      minplus_gemm(A.p, B.p, C.p, n, n, n, n, n, n, s, item.p, t);  // warm-up
      const int reps = n <= 128 ? 50 : 10;
      MI_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; ++r) {
        const TileEntry& te = kTiles[t];
        dim3 grid(div_up((uint64_t)n, te.bn), div_up((uint64_t)n, te.bm), 1);
        hipLaunchKernelGGL(te.fn, grid, dim3(256), 0, s, static_cast<const MinplusItem*>(item.p),
                           static_cast<const int2*>(nullptr));
      }
      MI_HIP(hipEventRecord(e1, s));
      MI_HIP(hipEventSynchronize(e1));
      float ms = 0;
      MI_HIP(hipEventElapsedTime(&ms, e0, e1));
      ms /= (float)reps;
      if (report) report[q * kNumTiles + t] = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int cls = 0;
    while ((1 << cls) < n && cls < 31) ++cls;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tuned_init) {
      for (int& x : g_tuned) x = -1;
      g_tuned_init = true;
    }
    g_tuned[cls] = best;
    ++g_tune_gen;
  }
}

void minplus_tuning_get(int* table32) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int i = 0; i < 32; ++i) table32[i] = g_tuned_init ? g_tuned[i] : -1;
}

void minplus_tuning_set(const int* table32) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int i = 0; i < 32; ++i)
    g_tuned[i] = (table32[i] >= 0 && table32[i] < kNumTiles) ? table32[i] : -1;
  g_tuned_init = true;
  ++g_tune_gen;
}

}  // namespace mi355
