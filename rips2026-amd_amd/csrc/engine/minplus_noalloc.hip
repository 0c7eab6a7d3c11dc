// Batched min-plus GEMM that allocates nothing: mi355_minplus_gemm_f32_batched
// with the batch's product descriptors and (product, tile) work list in a
// fixed device array instead of buffers the library allocates.
//
// For callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops).  The array `scratch` below is one per GPU, held until the
// process ends; every call on a GPU overwrites it, so two calls that overlap
// on one GPU (on different streams, or from different threads) race and must
// be ordered by the caller.  Nothing in it needs to start at zero: a call
// writes the descriptors and the work list before its kernels read them.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "minplus.hpp"

namespace mi355 {

namespace {

constexpr size_t kScratchBytes = (size_t)4 << 20;  // 4 MiB per GPU
constexpr size_t kMaxItems = 4096;                  // product descriptors (48 B each)
constexpr size_t kTableBytes = kMaxItems * sizeof(MinplusItem);
constexpr size_t kMaxWork = (kScratchBytes - kTableBytes) / sizeof(int2);
static_assert(kTableBytes % 256 == 0, "work list starts on a 256-byte boundary");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices];  // device address of `scratch` on each GPU (set up once)

// The device address of this GPU's copy of `scratch`.  In host code the
// array's name is only a stand-in for the symbol: the address comes from the
// runtime, once per device, with that device current.
unsigned char* scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) fail(MI355_EDEVICE, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) throw Error(MI355_EHIP, "hipGetSymbolAddress(scratch) failed", (int)e);
    g_scratch[dev] = p;
  }
  return static_cast<unsigned char*>(g_scratch[dev]);
}

}  // namespace

size_t minplus_noalloc_max_items() { return kMaxItems; }
size_t minplus_noalloc_max_tiles() { return kMaxWork; }

void minplus_gemm_batched_noalloc(const std::vector<MinplusItem>& items, hipStream_t s) {
  if (items.empty()) return;
  if (items.size() > kMaxItems) fail(MI355_ECAPACITY, "no-allocation min-plus batch holds at most 4096 products");
  // host plan buffers reused per thread (their capacity survives the call)
  thread_local std::vector<int2> work;
  thread_local std::vector<std::pair<int, uint32_t>> launches;
  minplus_plan(items, work, launches);
  if (work.size() > kMaxWork)
    fail(MI355_ECAPACITY, "no-allocation min-plus batch: more output tiles than its 4 MiB plan area holds");
  int dev = 0;
  MI_HIP(hipGetDevice(&dev));
  unsigned char* base = scratch_on(dev);
  MinplusItem* d_table = reinterpret_cast<MinplusItem*>(base);
  int2* d_work = reinterpret_cast<int2*>(base + kTableBytes);
  // pageable host sources: the runtime has taken the bytes when the call returns
  MI_HIP(hipMemcpyAsync(d_table, items.data(), items.size() * sizeof(MinplusItem), hipMemcpyHostToDevice, s));
  MI_HIP(hipMemcpyAsync(d_work, work.data(), work.size() * sizeof(int2), hipMemcpyHostToDevice, s));
  minplus_launch_plan(launches, d_table, d_work, s);
}

}  // namespace mi355
