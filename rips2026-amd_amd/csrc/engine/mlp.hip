// The learned delta policy's network, trained and evaluated on the device.
//
// 32 phase features -> 64 hidden (ReLU) -> 4 outputs (log2 of the bucket-width
// scale, active buckets, log2 of the claiming worker share, log2 of the light
// width relative to the first bucket width).  Weights are kept as an fp32 master copy for
// training; every evaluation (here and inside the ADDS kernel) uses the
// weights rounded to bf16 (round to nearest even) with fp32 accumulation, in
// the same operation order, so a trained model behaves on the worklist
// exactly as it did on its training data.  Gradients flow through the rounding
// unchanged (straight-through).
//
// One wave evaluates one sample at a time with lane j owning hidden unit j:
// the inputs are broadcast with shuffles, the two outputs reduced with an xor
// butterfly.  Training: per-wave gradient partials, a fixed-order reduction
// (deterministic), then Adam on the fp32 master weights.  Loss:
// (1/N) sum_s w_s sum_k (o_k - y_k)^2.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "adds.hpp"
#include "relax.hpp"
#include "sssp.hpp"

namespace mi355 {

namespace {

constexpr int kP = kMlpHidden * kMlpIn + kMlpHidden + kMlpOut * kMlpHidden + kMlpOut;  // 2372
constexpr int kOffB1 = kMlpHidden * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpOut * kMlpHidden;
constexpr int kOffMean = kP;
constexpr int kOffInv = kP + kMlpIn;
constexpr int kParamsWithNorm = kP + 2 * kMlpIn;
constexpr int kRec = kP + 1;  // gradient partial record: P gradients + loss
constexpr int kGraphFeat = 14;  // graph features: inputs 0..13, range-checked per graph
static_assert(sizeof(mi355_mlp_params) == (kParamsWithNorm + 2 * kGraphFeat) * sizeof(float),
              "parameter layout");

__device__ inline float bf16_round(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xFFFF0000u);
}

// Wave-resident copy of the model: lane j holds row j of w1 and column j of
// w2 (bf16-rounded).
struct WaveModel {
  float w1[kMlpIn];
  float b1;
  float w2[kMlpOut], b2[kMlpOut];
  float mean, inv;  // normalisation of feature `lane` (lanes < 32)
};

__device__ inline void load_model(const float* __restrict__ p, unsigned lane, WaveModel& m) {
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) m.w1[i] = bf16_round(p[lane * kMlpIn + i]);
  m.b1 = p[kOffB1 + lane];
#pragma unroll
  for (int k = 0; k < kMlpOut; ++k) {
    m.w2[k] = bf16_round(p[kOffW2 + k * kMlpHidden + lane]);
    m.b2[k] = p[kOffB2 + k];
  }
  m.mean = lane < (unsigned)kMlpIn ? p[kOffMean + lane] : 0.0f;
  m.inv = lane < (unsigned)kMlpIn ? p[kOffInv + lane] : 0.0f;
}

// Same arithmetic as mlp_forward_wave() in the ADDS kernel.
__device__ inline void forward(const WaveModel& m, float xn, float& h, float (&o)[kMlpOut]) {
  h = m.b1;
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) h += m.w1[i] * __shfl(xn, i, 64);
  h = fmaxf(h, 0.0f);
#pragma unroll
  for (int k = 0; k < kMlpOut; ++k) {
    float v = m.w2[k] * h;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    o[k] = v + m.b2[k];
  }
}

__global__ __launch_bounds__(kBlock) void k_mlp_forward(const float* __restrict__ p,
                                                        const float* __restrict__ x, int64_t n,
                                                        float* __restrict__ out) {
  const unsigned lane = lane_id();
  WaveModel m;
  load_model(p, lane, m);
  const int64_t waves = (int64_t)gridDim.x * kWavesPerBlock;
  for (int64_t s = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; s < n; s += waves) {
    const float xn = lane < (unsigned)kMlpIn ? (x[s * kMlpIn + lane] - m.mean) * m.inv : 0.0f;
    float h, o[kMlpOut];
    forward(m, xn, h, o);
    if (lane == 0)
      for (int k = 0; k < kMlpOut; ++k) out[s * kMlpOut + k] = o[k];
  }
}

// Gradient partials of samples order[first .. first + count) (order may be
// null: identity), normalised by `norm` (the batch size); one record per wave.
__global__ __launch_bounds__(kBlock) void k_mlp_grad(const float* __restrict__ p,
                                                     const float* __restrict__ x,
                                                     const float* __restrict__ y,
                                                     const float* __restrict__ sw,
                                                     const int* __restrict__ order, int64_t first,
                                                     int64_t count, float norm,
                                                     float* __restrict__ partial) {
  const unsigned lane = lane_id();
  WaveModel m;
  load_model(p, lane, m);
  float g1[kMlpIn];
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) g1[i] = 0.0f;
  float gb1 = 0.0f, loss = 0.0f, gw2[kMlpOut], gb2[kMlpOut];
#pragma unroll
  for (int k = 0; k < kMlpOut; ++k) gw2[k] = gb2[k] = 0.0f;
  const int64_t wave = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  const int64_t waves = (int64_t)gridDim.x * kWavesPerBlock;
  const float inv_n = 1.0f / norm;
  for (int64_t s = wave; s < count; s += waves) {
    const int64_t idx = order ? (int64_t)order[first + s] : first + s;
    const float xn = lane < (unsigned)kMlpIn ? (x[idx * kMlpIn + lane] - m.mean) * m.inv : 0.0f;
    float h, o[kMlpOut];
    forward(m, xn, h, o);
    const float wt = sw ? sw[idx] : 1.0f;
    float dh = 0.0f;
#pragma unroll
    for (int k = 0; k < kMlpOut; ++k) {
      const float e = o[k] - y[idx * kMlpOut + k];
      const float d = 2.0f * wt * e * inv_n;
      loss += wt * e * e * inv_n;
      gb2[k] += d;
      gw2[k] += d * h;
      dh += d * m.w2[k];
    }
    dh = h > 0.0f ? dh : 0.0f;
    gb1 += dh;
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) g1[i] += dh * __shfl(xn, i, 64);
  }
  float* rec = partial + wave * kRec;
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) rec[lane * kMlpIn + i] = g1[i];
  rec[kOffB1 + lane] = gb1;
#pragma unroll
  for (int k = 0; k < kMlpOut; ++k) rec[kOffW2 + k * kMlpHidden + lane] = gw2[k];
  if (lane == 0) {
    for (int k = 0; k < kMlpOut; ++k) rec[kOffB2 + k] = gb2[k];
    rec[kP] = loss;
  }
}

// out[t] = sum over waves of partial[w][t], w in order (deterministic).
__global__ void k_mlp_reduce(const float* __restrict__ partial, int waves, float* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kRec) return;
  float acc = 0.0f;
  for (int w = 0; w < waves; ++w) acc += partial[(int64_t)w * kRec + t];
  out[t] = acc;
}

__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ mom, float* __restrict__ vel,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps,
                           float wd, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kP) return;
  const float gt = g[t] + wd * p[t];
  const float m = b1 * mom[t] + (1.0f - b1) * gt;
  const float v = b2 * vel[t] + (1.0f - b2) * gt * gt;
  mom[t] = m;
  vel[t] = v;
  p[t] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

constexpr int kGradBlocks = 64;  // 256 waves of partials
constexpr int kGradWaves = kGradBlocks * kWavesPerBlock;

struct DevData {
  DevBuf<float> p, x, y, sw, partial, grad, mom, vel;
  DevBuf<int> order;
};

void upload(DevBuf<float>& d, const float* h, size_t n, hipStream_t s) {
  d.ensure(std::max<size_t>(n, 1));
  if (n) MI_HIP(hipMemcpyAsync(d.p, h, n * sizeof(float), hipMemcpyHostToDevice, s));
}

void check_params(const mi355_mlp_params* prm) {
  if (!prm) fail(MI355_EINVAL, "null model parameters");
  const float* f = reinterpret_cast<const float*>(prm);
  for (int i = 0; i < kParamsWithNorm; ++i)
    if (!std::isfinite(f[i])) fail(MI355_EINVAL, "model parameters must be finite");
}

void check_data(const float* x, const float* y, int64_t n, bool need_y) {
  if (n <= 0) fail(MI355_EINVAL, "need at least one sample");
  if (!x || (need_y && !y)) fail(MI355_EINVAL, "null sample array");
}

}  // namespace

void mlp_forward(const mi355_mlp_params* prm, const float* h_x, int64_t n, float* h_out,
                 hipStream_t s) {
  check_params(prm);
  check_data(h_x, nullptr, n, false);
  if (!h_out) fail(MI355_EINVAL, "null output array");
  DevData d;
  upload(d.p, reinterpret_cast<const float*>(prm), kParamsWithNorm, s);
  upload(d.x, h_x, (size_t)n * kMlpIn, s);
  DevBuf<float> out((size_t)n * kMlpOut);
  const unsigned grid = (unsigned)std::min<int64_t>(div_up((uint64_t)n, kWavesPerBlock), 4096);
  k_mlp_forward<<<grid, kBlock, 0, s>>>(d.p.p, d.x.p, n, out.p);
  MI_LAUNCH_CHECK();
  MI_HIP(hipMemcpyAsync(h_out, out.p, (size_t)n * kMlpOut * sizeof(float), hipMemcpyDeviceToHost, s));
  MI_HIP(hipStreamSynchronize(s));
}

void mlp_loss_grad(const mi355_mlp_params* prm, const float* h_x, const float* h_y,
                   const float* h_w, int64_t n, float* loss, mi355_mlp_params* grad,
                   hipStream_t s) {
  check_params(prm);
  check_data(h_x, h_y, n, true);
  DevData d;
  upload(d.p, reinterpret_cast<const float*>(prm), kParamsWithNorm, s);
  upload(d.x, h_x, (size_t)n * kMlpIn, s);
  upload(d.y, h_y, (size_t)n * kMlpOut, s);
  if (h_w) upload(d.sw, h_w, (size_t)n, s);
  d.partial.ensure((size_t)kGradWaves * kRec);
  d.grad.ensure(kRec);
  k_mlp_grad<<<kGradBlocks, kBlock, 0, s>>>(d.p.p, d.x.p, d.y.p, h_w ? d.sw.p : nullptr, nullptr, 0,
                                            n, (float)n, d.partial.p);
  k_mlp_reduce<<<div_up((uint64_t)kRec, 256), 256, 0, s>>>(d.partial.p, kGradWaves, d.grad.p);
  MI_LAUNCH_CHECK();
  std::vector<float> g(kRec);
  MI_HIP(hipMemcpyAsync(g.data(), d.grad.p, kRec * sizeof(float), hipMemcpyDeviceToHost, s));
  MI_HIP(hipStreamSynchronize(s));
  if (loss) *loss = g[kP];
  if (grad) {
    float* out = reinterpret_cast<float*>(grad);
    std::memcpy(out, g.data(), kP * sizeof(float));
    // normalisation and the feature range are not trained
    std::fill(out + kP, out + kParamsWithNorm + 2 * kGraphFeat, 0.0f);
  }
}

void mlp_train(mi355_mlp_params* prm, const float* h_x, const float* h_y, const float* h_w,
               int64_t n, const mi355_mlp_train_options& o, float* h_epoch_loss, hipStream_t s) {
  check_params(prm);
  check_data(h_x, h_y, n, true);
  if (o.epochs <= 0 || o.batch <= 0) fail(MI355_EINVAL, "epochs and batch must be positive");
  if (!(o.lr > 0.0f) || !(o.beta1 >= 0.0f && o.beta1 < 1.0f) || !(o.beta2 >= 0.0f && o.beta2 < 1.0f) ||
      !(o.eps > 0.0f) || !(o.weight_decay >= 0.0f))
    fail(MI355_EINVAL, "invalid optimiser settings");
  DevData d;
  upload(d.p, reinterpret_cast<const float*>(prm), kParamsWithNorm, s);
  upload(d.x, h_x, (size_t)n * kMlpIn, s);
  upload(d.y, h_y, (size_t)n * kMlpOut, s);
  if (h_w) upload(d.sw, h_w, (size_t)n, s);
  d.partial.ensure((size_t)kGradWaves * kRec);
  d.grad.ensure(kRec);
  d.mom.ensure(kP);
  d.vel.ensure(kP);
  d.order.ensure((size_t)n);
  MI_HIP(hipMemsetAsync(d.mom.p, 0, kP * sizeof(float), s));
  MI_HIP(hipMemsetAsync(d.vel.p, 0, kP * sizeof(float), s));
  std::vector<int> order((size_t)n);
  std::vector<float> rec(kRec);
  uint64_t rng = o.seed;
  auto next = [&rng]() {  // SplitMix64
    uint64_t z = (rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const int64_t batch = std::min<int64_t>(o.batch, n);
  const int64_t steps = (n + batch - 1) / batch;
  int64_t t = 0;
  std::vector<float> losses((size_t)o.epochs, 0.0f);
  for (int e = 0; e < o.epochs; ++e) {
    for (int64_t i = 0; i < n; ++i) order[(size_t)i] = (int)i;
    for (int64_t i = n - 1; i > 0; --i) std::swap(order[(size_t)i], order[(size_t)(next() % (uint64_t)(i + 1))]);
    MI_HIP(hipMemcpyAsync(d.order.p, order.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
    double eloss = 0.0;
    for (int64_t b = 0; b < steps; ++b) {
      const int64_t first = b * batch, cnt = std::min(batch, n - first);
      k_mlp_grad<<<kGradBlocks, kBlock, 0, s>>>(d.p.p, d.x.p, d.y.p, h_w ? d.sw.p : nullptr, d.order.p,
                                                first, cnt, (float)cnt, d.partial.p);
      k_mlp_reduce<<<div_up((uint64_t)kRec, 256), 256, 0, s>>>(d.partial.p, kGradWaves, d.grad.p);
      ++t;
      const float c1 = 1.0f - std::pow(o.beta1, (float)t), c2 = 1.0f - std::pow(o.beta2, (float)t);
      k_mlp_adam<<<div_up((uint64_t)kP, 256), 256, 0, s>>>(d.p.p, d.mom.p, d.vel.p, d.grad.p, o.lr,
                                                           o.beta1, o.beta2, o.eps, o.weight_decay,
                                                           c1, c2);
      MI_LAUNCH_CHECK();
      MI_HIP(hipMemcpyAsync(rec.data() + kP, d.grad.p + kP, sizeof(float), hipMemcpyDeviceToHost, s));
      MI_HIP(hipStreamSynchronize(s));
      eloss += (double)rec[kP] * (double)cnt;
    }
    losses[(size_t)e] = (float)(eloss / (double)n);
  }
  MI_HIP(hipMemcpyAsync(reinterpret_cast<float*>(prm), d.p.p, kP * sizeof(float), hipMemcpyDeviceToHost, s));
  MI_HIP(hipStreamSynchronize(s));
  if (h_epoch_loss) std::copy(losses.begin(), losses.end(), h_epoch_loss);
}

// Engine copy of a model: bf16 weights and fp32 biases / normalisation.
void set_mlp_params(const mi355_mlp_params* prm) {
  check_params(prm);
  AddsPolicy h{};
  auto to_bf16 = [](float x) -> uint16_t {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  };
  for (int i = 0; i < kMlpHidden * kMlpIn; ++i) h.w1[i] = to_bf16(prm->w1[i]);
  for (int i = 0; i < kMlpHidden; ++i) h.b1[i] = prm->b1[i];
  for (int i = 0; i < kMlpOut * kMlpHidden; ++i) h.w2[i] = to_bf16(prm->w2[i]);
  for (int i = 0; i < kMlpOut; ++i) h.b2[i] = prm->b2[i];
  for (int i = 0; i < kMlpIn; ++i) {
    h.feat_mean[i] = prm->feat_mean[i];
    h.feat_inv_std[i] = prm->feat_inv_std[i];
  }
  set_mlp(h, prm->graph_feat_lo, prm->graph_feat_hi);
}

}  // namespace mi355
