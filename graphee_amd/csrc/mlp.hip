// The scheduling MLP: 16 features -> 64 hidden (ReLU) -> 72 outputs = 8 bucket-width logits
// (delta in {8, 16, ..., 1024}) + 64 shard-activity logits, FP32, trained and run on the GPU by
// the kernels below.  An engine holds two networks of this shape -- the delta-stepping schedule
// (bucket widths; its shard outputs predict the degree bands an epoch touches) and the
// semi-external window predictor (shard outputs: the 64ths of the row range whose windows the
// sweep after next runs, for prefetch and order) -- and GMP_OPT_MLP_NET selects the one the
// gmp_mlp_* calls act on.  Neither ever changes which edges are relaxed or gathered, so results
// are identical with or without them.
//
// Parameters, 5768 floats: W1[16][64] (input-major), b1[64], W2[64][72], b2[72].
// Loss per sample: cross-entropy of softmax(y[0:8]) against the best width class (none for a
// row labelled -1), plus the mean binary cross-entropy of sigmoid(y[8:72]) against the 64-bit
// shard-activity mask; averaged over the batch.  Every reduction runs in a fixed order (bit-reproducible).
#include <cmath>
#include <cstring>
#include <vector>

#include "gmp_internal.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void mlp_forward_kernel(const float *P, const float *X, uint32_t n, float *H,
                                                             float *Y) {
  __shared__ float w[kMlpParams];
  for (uint32_t i = threadIdx.x; i < kMlpParams; i += blockDim.x) w[i] = P[i];
  __syncthreads();
  const float *W1 = w, *b1 = w + kMlpIn * kMlpHid, *W2 = b1 + kMlpHid, *b2 = W2 + kMlpHid * kMlpOut;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    float x[kMlpIn];
    for (int i = 0; i < kMlpIn; i++) x[i] = X[(size_t)s * kMlpIn + i];
    float h[kMlpHid];
    for (int j = 0; j < kMlpHid; j++) {
      float a = b1[j];
      for (int i = 0; i < kMlpIn; i++) a = fmaf(x[i], W1[i * kMlpHid + j], a);
      h[j] = a > 0.0f ? a : 0.0f;
      if (H) H[(size_t)s * kMlpHid + j] = h[j];
    }
    for (int k = 0; k < kMlpOut; k++) {
      float a = b2[k];
      for (int j = 0; j < kMlpHid; j++) a = fmaf(h[j], W2[j * kMlpOut + k], a);
      Y[(size_t)s * kMlpOut + k] = a;
    }
  }
}

// dY = dL/dY per sample (already divided by n), and the per-sample loss.
__global__ __launch_bounds__(kBlock) void mlp_output_grad_kernel(const float *Y, const int32_t *label,
                                                                 const uint64_t *mask, uint32_t n, float *dY,
                                                                 float *loss) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const float *y = Y + (size_t)s * kMlpOut;
    float *d = dY + (size_t)s * kMlpOut;
    float m = y[0];
    for (int c = 1; c < kMlpClasses; c++) m = fmaxf(m, y[c]);
    float z = 0;
    for (int c = 0; c < kMlpClasses; c++) z += expf(y[c] - m);
    const int lab = label[s];
    const float inv_n = 1.0f / (float)n;
    // a row without a class (label < 0) trains the shard outputs only
    float l = lab >= 0 ? logf(z) + m - y[lab] : 0.0f;
    for (int c = 0; c < kMlpClasses; c++) d[c] = lab >= 0 ? (expf(y[c] - m) / z - (c == lab ? 1.0f : 0.0f)) * inv_n : 0.0f;
    const uint64_t msk = mask[s];
    float b = 0;
    for (int j = 0; j < kMlpShards; j++) {
      const float t = (msk >> j) & 1ull ? 1.0f : 0.0f;
      const float v = y[kMlpClasses + j];
      const float p = 1.0f / (1.0f + expf(-v));
      // log(1 + e^v) - t v, stable form
      b += fmaxf(v, 0.0f) - t * v + log1pf(expf(-fabsf(v)));
      d[kMlpClasses + j] = (p - t) * inv_n / (float)kMlpShards;
    }
    loss[s] = l + b / (float)kMlpShards;
  }
}

// dW2[j][k] = sum_s H[s][j] dY[s][k]; db2[k] = sum_s dY[s][k]  (thread per output, sample order)
__global__ __launch_bounds__(kBlock) void mlp_grad_w2_kernel(const float *H, const float *dY, uint32_t n, float *G) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  float *dW2 = G + kMlpIn * kMlpHid + kMlpHid, *db2 = dW2 + kMlpHid * kMlpOut;
  if (t < (uint32_t)(kMlpHid * kMlpOut)) {
    const uint32_t j = t / kMlpOut, k = t % kMlpOut;
    float a = 0;
    for (uint32_t s = 0; s < n; s++) a = fmaf(H[(size_t)s * kMlpHid + j], dY[(size_t)s * kMlpOut + k], a);
    dW2[t] = a;
  } else if (t < (uint32_t)(kMlpHid * kMlpOut + kMlpOut)) {
    const uint32_t k = t - kMlpHid * kMlpOut;
    float a = 0;
    for (uint32_t s = 0; s < n; s++) a += dY[(size_t)s * kMlpOut + k];
    db2[k] = a;
  }
}

// dH[s][j] = (sum_k dY[s][k] W2[j][k]) * [H[s][j] > 0]
__global__ __launch_bounds__(kBlock) void mlp_grad_hidden_kernel(const float *P, const float *H, const float *dY,
                                                                 uint32_t n, float *dH) {
  const float *W2 = P + kMlpIn * kMlpHid + kMlpHid;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (uint64_t)n * kMlpHid;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = t / kMlpHid;
    const uint32_t j = (uint32_t)(t % kMlpHid);
    float a = 0;
    if (H[t] > 0.0f)
      for (int k = 0; k < kMlpOut; k++) a = fmaf(dY[s * kMlpOut + k], W2[j * kMlpOut + k], a);
    dH[t] = a;
  }
}

// dW1[i][j] = sum_s X[s][i] dH[s][j]; db1[j] = sum_s dH[s][j]
__global__ __launch_bounds__(kBlock) void mlp_grad_w1_kernel(const float *X, const float *dH, uint32_t n, float *G) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  float *dW1 = G, *db1 = G + kMlpIn * kMlpHid;
  if (t < (uint32_t)(kMlpIn * kMlpHid)) {
    const uint32_t i = t / kMlpHid, j = t % kMlpHid;
    float a = 0;
    for (uint32_t s = 0; s < n; s++) a = fmaf(X[(size_t)s * kMlpIn + i], dH[(size_t)s * kMlpHid + j], a);
    dW1[t] = a;
  } else if (t < (uint32_t)(kMlpIn * kMlpHid + kMlpHid)) {
    const uint32_t j = t - kMlpIn * kMlpHid;
    float a = 0;
    for (uint32_t s = 0; s < n; s++) a += dH[(size_t)s * kMlpHid + j];
    db1[j] = a;
  }
}

__global__ void adam_kernel(float *P, const float *G, float *m, float *v, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint32_t)kMlpParams) return;
  const float g = G[t];
  const float mt = b1 * m[t] + (1.0f - b1) * g;
  const float vt = b2 * v[t] + (1.0f - b2) * g * g;
  m[t] = mt;
  v[t] = vt;
  P[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

__global__ void sum_loss_kernel(const float *loss, uint32_t n, float *out) {
  __shared__ float sh[kBlock];
  float a = 0;
  for (uint32_t s = threadIdx.x; s < n; s += kBlock) a += loss[s];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0] / (float)n;
}

inline uint32_t grid_for(uint64_t n) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 4096));
}

struct MlpBatch {
  DevBuf X, y, mask, H, Y, dY, dH, loss, lsum;
  void alloc(uint32_t n) {
    X.alloc((size_t)n * kMlpIn * 4);
    y.alloc((size_t)n * 4);
    mask.alloc((size_t)n * 8);
    H.alloc((size_t)n * kMlpHid * 4);
    Y.alloc((size_t)n * kMlpOut * 4);
    dY.alloc((size_t)n * kMlpOut * 4);
    dH.alloc((size_t)n * kMlpHid * 4);
    loss.alloc((size_t)n * 4);
    lsum.alloc(4);
  }
};

// Full-batch loss and gradient into the selected network's gradient; returns the mean loss.
float loss_grad(Engine &e, MlpBatch &b, uint32_t n) {
  Mlp &m = e.net();
  mlp_forward_kernel<<<grid_for(n), kBlock, 0, e.stream>>>(m.params.as<float>(), b.X.as<float>(), n, b.H.as<float>(),
                                                         b.Y.as<float>());
  mlp_output_grad_kernel<<<grid_for(n), kBlock, 0, e.stream>>>(b.Y.as<float>(), b.y.as<int32_t>(),
                                                             b.mask.as<uint64_t>(), n, b.dY.as<float>(),
                                                             b.loss.as<float>());
  mlp_grad_w2_kernel<<<grid_for(kMlpHid * kMlpOut + kMlpOut), kBlock, 0, e.stream>>>(b.H.as<float>(), b.dY.as<float>(),
                                                                                    n, m.grad.as<float>());
  mlp_grad_hidden_kernel<<<grid_for((uint64_t)n * kMlpHid), kBlock, 0, e.stream>>>(
      m.params.as<float>(), b.H.as<float>(), b.dY.as<float>(), n, b.dH.as<float>());
  mlp_grad_w1_kernel<<<grid_for(kMlpIn * kMlpHid + kMlpHid), kBlock, 0, e.stream>>>(b.X.as<float>(), b.dH.as<float>(),
                                                                                  n, m.grad.as<float>());
  sum_loss_kernel<<<1, kBlock, 0, e.stream>>>(b.loss.as<float>(), n, b.lsum.as<float>());
  GMP_CHECK_LAUNCH();
  float l = 0;
  GMP_HIP(hipMemcpyAsync(&l, b.lsum.p, 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  return l;
}

void upload_batch(Engine &e, MlpBatch &b, const float *X, const int32_t *y, const uint64_t *mask, uint32_t n) {
  for (uint32_t s = 0; s < n; s++)
    if (y[s] < -1 || y[s] >= kMlpClasses) fail(GMP_EINVAL, "mlp: label out of range");
  b.alloc(n);
  GMP_HIP(hipMemcpyAsync(b.X.p, X, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(b.y.p, y, (size_t)n * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(b.mask.p, mask, (size_t)n * 8, hipMemcpyHostToDevice, e.stream));
}

}  // namespace

void mlp_ensure(Engine &e) {
  Mlp &m = e.net();
  if (m.params.p) return;
  m.params.alloc(kMlpParams * 4);
  m.grad.alloc(kMlpParams * 4);
  m.adam_m.alloc(kMlpParams * 4);
  m.adam_v.alloc(kMlpParams * 4);
  mlp_init(e, 1);
  m.trained = false;
}

// He-style uniform initialisation from the SplitMix64 stream of `seed`.
void mlp_init(Engine &e, uint64_t seed) {
  Mlp &m = e.net();
  if (!m.params.p) {
    m.params.alloc(kMlpParams * 4);
    m.grad.alloc(kMlpParams * 4);
    m.adam_m.alloc(kMlpParams * 4);
    m.adam_v.alloc(kMlpParams * 4);
  }
  std::vector<float> p(kMlpParams, 0.0f);
  uint64_t s = seed;
  auto uni = [&]() {
    s += kGamma;
    return (double)(mix64(s) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  const double a1 = std::sqrt(6.0 / kMlpIn), a2 = std::sqrt(6.0 / kMlpHid);
  for (int i = 0; i < kMlpIn * kMlpHid; i++) p[i] = (float)(a1 * uni());
  float *w2 = p.data() + kMlpIn * kMlpHid + kMlpHid;
  for (int i = 0; i < kMlpHid * kMlpOut; i++) w2[i] = (float)(a2 * uni());
  mlp_set(e, p.data());
  m.trained = false;
}

void mlp_set(Engine &e, const float *params) {
  Mlp &m = e.net();
  if (!m.params.p) mlp_ensure(e);
  GMP_HIP(hipMemcpy(m.params.p, params, kMlpParams * 4, hipMemcpyHostToDevice));
  GMP_HIP(hipMemset(m.adam_m.p, 0, kMlpParams * 4));
  GMP_HIP(hipMemset(m.adam_v.p, 0, kMlpParams * 4));
  m.step = 0;
  m.trained = true;
}

void mlp_get(Engine &e, float *params) {
  mlp_ensure(e);
  GMP_HIP(hipMemcpy(params, e.net().params.p, kMlpParams * 4, hipMemcpyDeviceToHost));
}

void mlp_forward(Engine &e, const float *X, uint32_t n, float *Y) {
  mlp_ensure(e);
  if (n == 0) return;
  DevBuf dX, dY;
  dX.alloc((size_t)n * kMlpIn * 4);
  dY.alloc((size_t)n * kMlpOut * 4);
  GMP_HIP(hipMemcpyAsync(dX.p, X, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, e.stream));
  mlp_forward_kernel<<<grid_for(n), kBlock, 0, e.stream>>>(e.net().params.as<float>(), dX.as<float>(), n, nullptr,
                                                         dY.as<float>());
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipMemcpyAsync(Y, dY.p, (size_t)n * kMlpOut * 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
}

float mlp_loss_grad(Engine &e, const float *X, const int32_t *y, const uint64_t *mask, uint32_t n, float *grad) {
  mlp_ensure(e);
  if (n == 0) fail(GMP_EINVAL, "mlp: empty batch");
  MlpBatch b;
  upload_batch(e, b, X, y, mask, n);
  const float l = loss_grad(e, b, n);
  if (grad) GMP_HIP(hipMemcpy(grad, e.net().grad.p, kMlpParams * 4, hipMemcpyDeviceToHost));
  return l;
}

float mlp_train(Engine &e, const float *X, const int32_t *y, const uint64_t *mask, uint32_t n, int32_t epochs,
                float lr) {
  mlp_ensure(e);
  if (n == 0) fail(GMP_EINVAL, "mlp: empty batch");
  MlpBatch b;
  upload_batch(e, b, X, y, mask, n);
  Mlp &m = e.net();
  float l = 0;
  const float b1 = 0.9f, b2 = 0.999f;
  for (int32_t ep = 0; ep < epochs; ep++) {
    l = loss_grad(e, b, n);
    m.step++;
    const float c1 = 1.0f - std::pow(b1, (float)m.step), c2 = 1.0f - std::pow(b2, (float)m.step);
    adam_kernel<<<grid_for(kMlpParams), kBlock, 0, e.stream>>>(m.params.as<float>(), m.grad.as<float>(),
                                                             m.adam_m.as<float>(), m.adam_v.as<float>(), lr, b1, b2,
                                                             1e-7f, c1, c2);
    GMP_CHECK_LAUNCH();
  }
  GMP_HIP(hipStreamSynchronize(e.stream));
  m.trained = true;
  return l;
}

}  // namespace gmp

extern "C" {

int gmp_sssp_actions(gmp_engine *h, uint32_t *counts, uint32_t n) {
  GMP_API_BEGIN
  if (!h || (n && !counts)) gmp::fail(GMP_EINVAL, "gmp_sssp_actions: null argument");
  if (n > (uint32_t)gmp::kMlpClasses) gmp::fail(GMP_EINVAL, "gmp_sssp_actions: n above the class count");
  const gmp::Engine &e = *reinterpret_cast<gmp::Engine *>(h);
  for (uint32_t c = 0; c < n; c++) counts[c] = e.last_actions[c];
  GMP_API_END
}

int gmp_mlp_init(gmp_engine *h, uint64_t seed) {
  GMP_API_BEGIN
  if (!h) gmp::fail(GMP_EINVAL, "null engine handle");
  gmp::mlp_init(*reinterpret_cast<gmp::Engine *>(h), seed);
  GMP_API_END
}

int gmp_mlp_set(gmp_engine *h, const float *params) {
  GMP_API_BEGIN
  if (!h || !params) gmp::fail(GMP_EINVAL, "gmp_mlp_set: null argument");
  gmp::mlp_set(*reinterpret_cast<gmp::Engine *>(h), params);
  GMP_API_END
}

int gmp_mlp_get(gmp_engine *h, float *params) {
  GMP_API_BEGIN
  if (!h || !params) gmp::fail(GMP_EINVAL, "gmp_mlp_get: null argument");
  gmp::mlp_get(*reinterpret_cast<gmp::Engine *>(h), params);
  GMP_API_END
}

int gmp_mlp_forward(gmp_engine *h, const float *X, uint32_t n, float *Y) {
  GMP_API_BEGIN
  if (!h || (n && (!X || !Y))) gmp::fail(GMP_EINVAL, "gmp_mlp_forward: null argument");
  gmp::mlp_forward(*reinterpret_cast<gmp::Engine *>(h), X, n, Y);
  GMP_API_END
}

int gmp_mlp_loss_grad(gmp_engine *h, const float *X, const int32_t *labels, const uint64_t *masks, uint32_t n,
                      float *loss, float *grad) {
  GMP_API_BEGIN
  if (!h || !X || !labels || !masks) gmp::fail(GMP_EINVAL, "gmp_mlp_loss_grad: null argument");
  const float l = gmp::mlp_loss_grad(*reinterpret_cast<gmp::Engine *>(h), X, labels, masks, n, grad);
  if (loss) *loss = l;
  GMP_API_END
}

int gmp_mlp_train(gmp_engine *h, const float *X, const int32_t *labels, const uint64_t *masks, uint32_t n,
                  int32_t epochs, float lr, float *loss) {
  GMP_API_BEGIN
  if (!h || !X || !labels || !masks) gmp::fail(GMP_EINVAL, "gmp_mlp_train: null argument");
  const float l = gmp::mlp_train(*reinterpret_cast<gmp::Engine *>(h), X, labels, masks, n, epochs, lr);
  if (loss) *loss = l;
  GMP_API_END
}

}  // extern "C"
