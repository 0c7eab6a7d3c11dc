// ADDS phase policy set-up: graph features for the learned delta model, the
// process-wide MLP weights, and the device policy object of a run.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "sssp.hpp"

namespace mi355 {

namespace {

std::mutex g_mlp_mu;
bool g_mlp_loaded = false;
AddsPolicy g_mlp{};  // weights + feature normalisation (mode/trace unused)
float g_feat_lo[14], g_feat_hi[14];  // training range of the graph features
std::vector<float> g_sched;           // explicit phase schedule, 3 floats per phase

// True when every bounded graph feature lies in the model's training range
// (widened by 1 % of the range on each side).
bool in_training_range(const float* f, const float* lo, const float* hi) {
  for (int i = 0; i < 14; ++i) {
    if (!(hi[i] > lo[i])) continue;  // not bounded
    const float tol = 0.01f * (hi[i] - lo[i]);
    if (!(f[i] >= lo[i] - tol && f[i] <= hi[i] + tol)) return false;
  }
  return true;
}

constexpr int kSamples = 4096;

// Deterministic strided sample of out-degrees and edge weights.
__global__ void k_sample(const int* __restrict__ rowptr, const uint2* __restrict__ edges, int64_t n,
                         int64_t m, int32_t wtype, float* __restrict__ deg, float* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kSamples) return;
  const int64_t v = (int64_t)((double)i * (double)n / kSamples);
  deg[i] = (float)(rowptr[v + 1] - rowptr[v]);
  if (m > 0) {
    const int64_t e = (int64_t)((double)i * (double)m / kSamples);
    const uint32_t b = edges[e].y;
    w[i] = wtype == MI355_W_F32 ? __uint_as_float(b) : (float)b;
  } else {
    w[i] = 1.0f;
  }
}

float quantile(std::vector<float>& v, double q) {
  const size_t k = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
  std::nth_element(v.begin(), v.begin() + (long)k, v.end());
  return v[k];
}

}  // namespace

// 14 graph features: log2(1 + degree quantile) at 5/25/50/75/95/99 % and
// log2(mean degree); weight quantiles at the same points over the mean weight,
// and log2(n) / 8.
void graph_features(mi355_graph& g, hipStream_t s, float* out14) {
  if (!g.graph_feat_ready) {
    DevBuf<float> d(kSamples), w(kSamples);
    k_sample<<<kSamples / 256, 256, 0, s>>>(g.rowptr.p, g.edges.p, g.n, g.m, g.weight_type, d.p, w.p);
    MI_LAUNCH_CHECK();
    std::vector<float> hd(kSamples), hw(kSamples);
    MI_HIP(hipMemcpyAsync(hd.data(), d.p, kSamples * 4, hipMemcpyDeviceToHost, s));
    MI_HIP(hipMemcpyAsync(hw.data(), w.p, kSamples * 4, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    const double qs[6] = {0.05, 0.25, 0.5, 0.75, 0.95, 0.99};
    const double wbar = g.info.avg_weight > 0 ? g.info.avg_weight : 1.0;
    for (int k = 0; k < 6; ++k) {
      g.graph_feat[k] = std::log2(1.0f + quantile(hd, qs[k]));
      g.graph_feat[7 + k] = (float)(quantile(hw, qs[k]) / wbar);
    }
    g.graph_feat[6] = (float)std::log2(std::max(1e-9, g.info.avg_degree));
    g.graph_feat[13] = (float)(std::log2((double)std::max<int64_t>(g.n, 1)) / 8.0);
    g.graph_feat_ready = true;
  }
  std::copy(g.graph_feat, g.graph_feat + 14, out14);
}

void set_schedule(const float* table, int32_t count) {
  if (count < 0 || count > kSchedMax) fail(MI355_EINVAL, "a phase schedule holds 1..64 phases (0 clears it)");
  if (count > 0 && !table) fail(MI355_EINVAL, "null schedule");
  for (int32_t i = 0; i < count; ++i) {
    const float sc = table[3 * i], ac = table[3 * i + 1], sh = table[3 * i + 2];
    if (!(sc > 0.0f) || !std::isfinite(sc) || !(ac >= 1.0f && ac <= 8.0f) || !(sh > 0.0f && sh <= 1.0f))
      fail(MI355_EINVAL, "schedule entry: width scale > 0, active buckets in [1, 8], worker share in (0, 1]");
  }
  std::lock_guard<std::mutex> lk(g_mlp_mu);
  g_sched.assign(table, table + 3 * (size_t)count);
}

void set_mlp(const AddsPolicy& p, const float* feat_lo, const float* feat_hi) {
  std::lock_guard<std::mutex> lk(g_mlp_mu);
  g_mlp = p;
  std::copy(feat_lo, feat_lo + 14, g_feat_lo);
  std::copy(feat_hi, feat_hi + 14, g_feat_hi);
  g_mlp_loaded = true;
}

void clear_mlp() {
  std::lock_guard<std::mutex> lk(g_mlp_mu);
  g_mlp_loaded = false;
}

bool get_mlp(AddsPolicy* p, float* feat_lo, float* feat_hi) {
  std::lock_guard<std::mutex> lk(g_mlp_mu);
  if (g_mlp_loaded && p) *p = g_mlp;
  if (g_mlp_loaded && feat_lo) std::copy(g_feat_lo, g_feat_lo + 14, feat_lo);
  if (g_mlp_loaded && feat_hi) std::copy(g_feat_hi, g_feat_hi + 14, feat_hi);
  return g_mlp_loaded;
}

const AddsPolicy* prepare_adds_policy(mi355_graph& g, const RunRequest& req, hipStream_t s,
                                      int* mode_out) {
  AddsPolicy h{};
  float lo[14], hi[14];
  int mode = req.dynamic_delta ? kPolicyHeuristic : kPolicyFixed;
  if (req.use_mlp == 2) {
    std::lock_guard<std::mutex> lk(g_mlp_mu);
    if (g_sched.empty()) fail(MI355_EINVAL, "no phase schedule set (mi355_adds_schedule_set)");
    h.sched_len = (uint32_t)(g_sched.size() / 3);
    std::copy(g_sched.begin(), g_sched.end(), h.sched);
    mode = kPolicySchedule;
  } else if (req.use_mlp) {
    if (!get_mlp(&h, lo, hi)) fail(MI355_EINVAL, "no delta model loaded (mi355_adds_mlp_set)");
    mode = kPolicyMlp;
  }
  if (mode == kPolicyMlp || req.trace) graph_features(g, s, h.graph_feat);
  // a graph unlike every training graph runs with the fixed heuristic delta:
  // the network's output there is an extrapolation
  if (mode == kPolicyMlp && !in_training_range(h.graph_feat, lo, hi)) mode = kPolicyFixed;
  if (mode != kPolicyMlp) {
    for (int i = 0; i < kMlpIn; ++i) {
      h.feat_mean[i] = 0.0f;
      h.feat_inv_std[i] = 1.0f;
    }
  }
  h.mode = mode;
  h.trace = req.trace;
  h.trace_cap = req.trace ? req.trace_cap : 0u;
  *mode_out = mode;
  if (mode == kPolicyFixed && !h.trace) return nullptr;
  // the same policy as the last upload (same model, graph and options): the
  // device copy is current (the kernel only writes the trace buffer it points to)
  std::vector<unsigned char>& last = g.ws.adds_policy_host;
  const unsigned char* hb = reinterpret_cast<const unsigned char*>(&h);
  if (g.ws.adds_policy.p && last.size() == sizeof(AddsPolicy) && std::memcmp(last.data(), hb, sizeof(AddsPolicy)) == 0)
    return g.ws.adds_policy.p;
  g.ws.adds_policy.ensure(1);
  MI_HIP(hipMemcpyAsync(g.ws.adds_policy.p, &h, sizeof(AddsPolicy), hipMemcpyHostToDevice, s));
  last.assign(hb, hb + sizeof(AddsPolicy));
  return g.ws.adds_policy.p;
}

}  // namespace mi355
