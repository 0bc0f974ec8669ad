// mlp.hip -- the cost model of include/lnt_dispatch.h: a 3-layer perceptron
// (F=LNT_NUM_FEATURES -> H=64 -> H -> S=LNT_NUM_STRATEGIES, ReLU, FP32)
// trained and evaluated by the kernels below.  No counterpart in the
// reference, whose dispatch is a fixed rule (sparse_times_polyalgorithm.m).
//
// Parameters (one float array): W1[H][F], b1[H], W2[H][H], b2[H], W3[S][H],
// b3[S], then the feature standardisation mean[F], inv_std[F].
// Loss over a set: L = (1/m) sum over measured (sample, strategy) of
// (y - t)^2, m = number of measured entries; t = log2(ms), NaN = unmeasured.
//
// Kernels: forward (one thread per sample, weights in LDS, activations kept
// feature-major [unit][sample] for coalesced reuse), backward (per-sample
// deltas), gradient (one thread per parameter, a fixed-order sum over the
// samples, so gradients are deterministic), Adam, minibatch gather.
#include <math.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/lnt_dispatch.h"
#include "lnt_internal.h"

namespace lnt {
namespace {

constexpr int F = LNT_NUM_FEATURES;
constexpr int H = LNT_MODEL_HIDDEN;
constexpr int S = LNT_NUM_STRATEGIES;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + H * F;
constexpr int OFF_W2 = OFF_B1 + H;
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;
constexpr int OFF_B3 = OFF_W3 + S * H;
constexpr int NW = OFF_B3 + S;  // trainable parameters
constexpr int OFF_MEAN = NW;
constexpr int OFF_ISTD = NW + F;
constexpr int NP = NW + 2 * F;
constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void mlp_forward_kernel(const float *__restrict__ P, const float *__restrict__ X,
                                                               int n, float *__restrict__ Z1, float *__restrict__ Z2,
                                                               float *__restrict__ Y) {
  __shared__ float w[NP];
  for (int i = threadIdx.x; i < NP; i += kThreads) w[i] = P[i];
  __syncthreads();
  const int s = blockIdx.x * kThreads + threadIdx.x;
  if (s >= n) return;
  float x[F];
#pragma unroll
  for (int k = 0; k < F; ++k) x[k] = (X[(size_t)s * F + k] - w[OFF_MEAN + k]) * w[OFF_ISTD + k];
  float h1[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float z = w[OFF_B1 + j];
#pragma unroll
    for (int k = 0; k < F; ++k) z = fmaf(w[OFF_W1 + j * F + k], x[k], z);
    if (Z1) Z1[(size_t)j * n + s] = z;
    h1[j] = z > 0.f ? z : 0.f;
  }
  float h2[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float z = w[OFF_B2 + j];
#pragma unroll
    for (int i = 0; i < H; ++i) z = fmaf(w[OFF_W2 + j * H + i], h1[i], z);
    if (Z2) Z2[(size_t)j * n + s] = z;
    h2[j] = z > 0.f ? z : 0.f;
  }
#pragma unroll
  for (int o = 0; o < S; ++o) {
    float y = w[OFF_B3 + o];
#pragma unroll
    for (int j = 0; j < H; ++j) y = fmaf(w[OFF_W3 + o * H + j], h2[j], y);
    Y[(size_t)s * S + o] = y;
  }
}

// dL/dz for every layer of every sample; L[s] = this sample's share of the loss.
__global__ __launch_bounds__(kThreads) void mlp_backward_kernel(
    const float *__restrict__ P, const float *__restrict__ T, const float *__restrict__ Y,
    const float *__restrict__ Z1, const float *__restrict__ Z2, int n, float scale, float *__restrict__ D1,
    float *__restrict__ D2, float *__restrict__ D3, float *__restrict__ L) {
  __shared__ float w[NW];
  for (int i = threadIdx.x; i < NW; i += kThreads) w[i] = P[i];
  __syncthreads();
  const int s = blockIdx.x * kThreads + threadIdx.x;
  if (s >= n) return;
  float d3[S];
  float loss = 0.f;
#pragma unroll
  for (int o = 0; o < S; ++o) {
    const float t = T[(size_t)s * S + o];
    float d = 0.f;
    if (t == t) {  // measured
      const float e = Y[(size_t)s * S + o] - t;
      d = 2.f * scale * e;
      loss += scale * e * e;
    }
    d3[o] = d;
    D3[(size_t)o * n + s] = d;
  }
  L[s] = loss;
  float dz2[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float g = 0.f;
#pragma unroll
    for (int o = 0; o < S; ++o) g = fmaf(w[OFF_W3 + o * H + j], d3[o], g);
    dz2[j] = Z2[(size_t)j * n + s] > 0.f ? g : 0.f;
    D2[(size_t)j * n + s] = dz2[j];
  }
#pragma unroll 4
  for (int i = 0; i < H; ++i) {
    float g = 0.f;
#pragma unroll
    for (int j = 0; j < H; ++j) g = fmaf(w[OFF_W2 + j * H + i], dz2[j], g);
    D1[(size_t)i * n + s] = Z1[(size_t)i * n + s] > 0.f ? g : 0.f;
  }
}

// One thread per trainable parameter: the sum over samples in sample order.
__global__ __launch_bounds__(kThreads) void mlp_grad_kernel(const float *__restrict__ P, const float *__restrict__ X,
                                                            const float *__restrict__ Z1, const float *__restrict__ Z2,
                                                            const float *__restrict__ D1, const float *__restrict__ D2,
                                                            const float *__restrict__ D3, int n,
                                                            float *__restrict__ G) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= NW) return;
  float acc = 0.f;
  if (p < OFF_B1) {
    const int j = p / F, k = p % F;
    const float mean = P[OFF_MEAN + k], istd = P[OFF_ISTD + k];
    for (int s = 0; s < n; ++s) acc = fmaf(D1[(size_t)j * n + s], (X[(size_t)s * F + k] - mean) * istd, acc);
  } else if (p < OFF_W2) {
    const int j = p - OFF_B1;
    for (int s = 0; s < n; ++s) acc += D1[(size_t)j * n + s];
  } else if (p < OFF_B2) {
    const int q = p - OFF_W2, j = q / H, i = q % H;
    for (int s = 0; s < n; ++s) {
      const float z = Z1[(size_t)i * n + s];
      acc = fmaf(D2[(size_t)j * n + s], z > 0.f ? z : 0.f, acc);
    }
  } else if (p < OFF_W3) {
    const int j = p - OFF_B2;
    for (int s = 0; s < n; ++s) acc += D2[(size_t)j * n + s];
  } else if (p < OFF_B3) {
    const int q = p - OFF_W3, o = q / H, j = q % H;
    for (int s = 0; s < n; ++s) {
      const float z = Z2[(size_t)j * n + s];
      acc = fmaf(D3[(size_t)o * n + s], z > 0.f ? z : 0.f, acc);
    }
  } else {
    const int o = p - OFF_B3;
    for (int s = 0; s < n; ++s) acc += D3[(size_t)o * n + s];
  }
  G[p] = acc;
}

// Adam (Kingma & Ba, 2015) with bias correction c1 = 1-beta1^t, c2 = 1-beta2^t.
__global__ void mlp_adam_kernel(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M,
                                float *__restrict__ V, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NW) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.f - b1) * g;
  const float v = b2 * V[p] + (1.f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void mlp_gather_kernel(const float *__restrict__ X, const float *__restrict__ T,
                                  const int32_t *__restrict__ idx, int nb, float *__restrict__ Xb,
                                  float *__restrict__ Tb) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb) return;
  const size_t s = (size_t)idx[r];
#pragma unroll
  for (int k = 0; k < F; ++k) Xb[(size_t)r * F + k] = X[s * F + k];
#pragma unroll
  for (int o = 0; o < S; ++o) Tb[(size_t)r * S + o] = T[s * S + o];
}

uint64_t splitmix(uint64_t &state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double uniform01(uint64_t &state) { return (double)(splitmix(state) >> 11) * 0x1p-53; }

}  // namespace
}  // namespace lnt

struct lnt_model_s {
  lnt_context ctx = nullptr;
  std::vector<float> host;  // parameters (kept equal to dev)
  lnt::DeviceBuffer dev;    // parameters
  lnt::DeviceBuffer opt;    // Adam moments M, V and the gradient G
  lnt::DeviceBuffer data;   // features, targets, indices
  lnt::DeviceBuffer work;   // activations and deltas
};

namespace lnt {
namespace {

struct Work {
  float *Xb, *Tb, *Z1, *Z2, *Y, *D1, *D2, *D3, *L;
};

// Work arrays for up to `n` samples.
int work_for(lnt_model m, int64_t n, Work *w) {
  const size_t per = (size_t)F + S + 2 * H + S + 2 * H + S + 1;
  LNT_TRY(ensure(m->work, per * (size_t)n * sizeof(float) + 64 * 10));
  float *p = (float *)m->work.ptr;
  auto take = [&](size_t count) {
    float *r = p;
    p += (count + 15) & ~(size_t)15;
    return r;
  };
  w->Xb = take((size_t)n * F);
  w->Tb = take((size_t)n * S);
  w->Z1 = take((size_t)n * H);
  w->Z2 = take((size_t)n * H);
  w->Y = take((size_t)n * S);
  w->D1 = take((size_t)n * H);
  w->D2 = take((size_t)n * H);
  w->D3 = take((size_t)n * S);
  w->L = take((size_t)n);
  return LNT_OK;
}

inline uint32_t blocks(int64_t n) { return (uint32_t)((n + kThreads - 1) / kThreads); }

int upload_params(lnt_model m) {
  LNT_HIP(hipMemcpyAsync(m->dev.ptr, m->host.data(), NP * sizeof(float), hipMemcpyHostToDevice, m->ctx->stream));
  return LNT_OK;
}

int download_params(lnt_model m) {
  LNT_HIP(hipMemcpyAsync(m->host.data(), m->dev.ptr, NP * sizeof(float), hipMemcpyDeviceToHost, m->ctx->stream));
  LNT_HIP(hipStreamSynchronize(m->ctx->stream));
  return LNT_OK;
}

// Forward + backward + gradient over Xd/Td (device, n samples); the
// gradient lands in G.  `scale` = 1 / measured entries.
int step_grad(lnt_model m, const float *Xd, const float *Td, int n, float scale, const Work &w, float *G) {
  hipStream_t st = m->ctx->stream;
  const float *P = (const float *)m->dev.ptr;
  mlp_forward_kernel<<<blocks(n), kThreads, 0, st>>>(P, Xd, n, w.Z1, w.Z2, w.Y);
  mlp_backward_kernel<<<blocks(n), kThreads, 0, st>>>(P, Td, w.Y, w.Z1, w.Z2, n, scale, w.D1, w.D2, w.D3, w.L);
  mlp_grad_kernel<<<blocks(NW), kThreads, 0, st>>>(P, Xd, w.Z1, w.Z2, w.D1, w.D2, w.D3, n, G);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

int64_t measured_count(const float *T, int64_t n) {
  int64_t c = 0;
  for (int64_t i = 0; i < n * S; ++i) c += T[i] == T[i];
  return c;
}

// Uploads features and targets; returns device pointers.
int upload_set(lnt_model m, const float *X, const float *T, int64_t n, float **Xd, float **Td, int32_t **idx) {
  const size_t bx = (size_t)n * F * 4, bt = (size_t)n * S * 4, bi = (size_t)n * 4;
  LNT_TRY(ensure(m->data, bx + bt + bi + 512));
  char *p = (char *)m->data.ptr;
  *Xd = (float *)p;
  *Td = (float *)(p + ((bx + 255) & ~(size_t)255));
  *idx = (int32_t *)(p + ((bx + 255) & ~(size_t)255) + ((bt + 255) & ~(size_t)255));
  LNT_HIP(hipMemcpyAsync(*Xd, X, bx, hipMemcpyHostToDevice, m->ctx->stream));
  if (T) LNT_HIP(hipMemcpyAsync(*Td, T, bt, hipMemcpyHostToDevice, m->ctx->stream));
  return LNT_OK;
}

double full_loss(lnt_model m, const float *Xd, const float *Td, int64_t n, float scale, const Work &w, float *G) {
  if (step_grad(m, Xd, Td, (int)n, scale, w, G) != LNT_OK) return NAN;
  std::vector<float> L(n);
  if (hipMemcpyAsync(L.data(), w.L, n * sizeof(float), hipMemcpyDeviceToHost, m->ctx->stream) != hipSuccess ||
      hipStreamSynchronize(m->ctx->stream) != hipSuccess)
    return NAN;
  double s = 0;
  for (float x : L) s += x;
  return s;
}

}  // namespace
}  // namespace lnt

using namespace lnt;

namespace lnt {
// One sample's prediction on the host, from the parameters the model keeps
// equal to its device copy: the same fused multiply-adds in the same order
// as mlp_forward_kernel, so the same floats.  A per-call dispatch decision
// then needs no device round trip (a launch and a synchronisation cost more
// than the strategy choice saves on small contractions).
int model_predict_one_host(lnt_model m, const float *features, float *y) {
  if (!m || m->host.size() != (size_t)NP) return fail(LNT_ERR_INVALID_ARGUMENT, "model has no parameters");
  const float *w = m->host.data();
  float x[F], h1[H], h2[H];
  for (int k = 0; k < F; ++k) x[k] = (features[k] - w[OFF_MEAN + k]) * w[OFF_ISTD + k];
  for (int j = 0; j < H; ++j) {
    float z = w[OFF_B1 + j];
    for (int k = 0; k < F; ++k) z = fmaf(w[OFF_W1 + j * F + k], x[k], z);
    h1[j] = z > 0.f ? z : 0.f;
  }
  for (int j = 0; j < H; ++j) {
    float z = w[OFF_B2 + j];
    for (int i = 0; i < H; ++i) z = fmaf(w[OFF_W2 + j * H + i], h1[i], z);
    h2[j] = z > 0.f ? z : 0.f;
  }
  for (int o = 0; o < S; ++o) {
    float v = w[OFF_B3 + o];
    for (int j = 0; j < H; ++j) v = fmaf(w[OFF_W3 + o * H + j], h2[j], v);
    y[o] = v;
  }
  return LNT_OK;
}
}  // namespace lnt

extern "C" {

int64_t lnt_model_param_count(void) { return NP; }

int lnt_model_create(lnt_context ctx, uint64_t seed, lnt_model *out) {
  if (!ctx || !out) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  LNT_HIP(hipSetDevice(ctx->device));
  lnt_model m = new lnt_model_s();
  m->ctx = ctx;
  m->host.assign(NP, 0.f);
  // He-uniform weights (bound sqrt(6 / fan_in)), zero biases, identity
  // standardisation.
  uint64_t state = seed;
  auto fill = [&](int off, int rows, int cols) {
    const double bound = sqrt(6.0 / cols);
    for (int i = 0; i < rows * cols; ++i) m->host[off + i] = (float)((2.0 * uniform01(state) - 1.0) * bound);
  };
  fill(OFF_W1, H, F);
  fill(OFF_W2, H, H);
  fill(OFF_W3, S, H);
  for (int k = 0; k < F; ++k) m->host[OFF_ISTD + k] = 1.f;
  int rc = ensure(m->dev, NP * sizeof(float));
  if (rc == LNT_OK) rc = ensure(m->opt, 3 * NW * sizeof(float) + 256);
  if (rc == LNT_OK) rc = upload_params(m);
  if (rc == LNT_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(LNT_ERR_DEVICE, "upload failed");
  if (rc != LNT_OK) {
    lnt_model_destroy(m);
    return rc;
  }
  *out = m;
  return LNT_OK;
}

int lnt_model_destroy(lnt_model m) {
  if (!m) return LNT_OK;
  if (m->ctx) {
    (void)hipSetDevice(m->ctx->device);
    if (m->ctx->model == m) m->ctx->model = nullptr;
  }
  release(m->dev);
  release(m->opt);
  release(m->data);
  release(m->work);
  delete m;
  return LNT_OK;
}

int lnt_model_get_params(lnt_model m, float *params) {
  if (!m || !params) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  memcpy(params, m->host.data(), NP * sizeof(float));
  return LNT_OK;
}

int lnt_model_set_params(lnt_model m, const float *params) {
  if (!m || !params) return fail(LNT_ERR_INVALID_ARGUMENT, "null argument");
  for (int i = 0; i < NP; ++i)
    if (!(params[i] == params[i]) || isinf(params[i])) return fail(LNT_ERR_INVALID_ARGUMENT, "non-finite parameter");
  LNT_HIP(hipSetDevice(m->ctx->device));
  memcpy(m->host.data(), params, NP * sizeof(float));
  LNT_TRY(upload_params(m));
  LNT_HIP(hipStreamSynchronize(m->ctx->stream));
  return LNT_OK;
}


int lnt_model_predict(lnt_model m, const float *features, int64_t n, float *out) {
  if (!m || (n > 0 && (!features || !out)) || n < 0) return fail(LNT_ERR_INVALID_ARGUMENT, "bad arguments");
  if (n == 0) return LNT_OK;
  if (n > (int64_t(1) << 26)) return fail(LNT_ERR_TOO_LARGE, "too many samples");
  LNT_HIP(hipSetDevice(m->ctx->device));
  float *Xd, *Td;
  int32_t *idx;
  LNT_TRY(upload_set(m, features, nullptr, n, &Xd, &Td, &idx));
  Work w;
  LNT_TRY(work_for(m, n, &w));
  mlp_forward_kernel<<<blocks(n), kThreads, 0, m->ctx->stream>>>((const float *)m->dev.ptr, Xd, (int)n, nullptr,
                                                                   nullptr, w.Y);
  LNT_HIP(hipGetLastError());
  LNT_HIP(hipMemcpyAsync(out, w.Y, (size_t)n * S * sizeof(float), hipMemcpyDeviceToHost, m->ctx->stream));
  LNT_HIP(hipStreamSynchronize(m->ctx->stream));
  return LNT_OK;
}

int lnt_model_gradient(lnt_model m, const float *features, const float *targets, int64_t n, float *grad,
                       float *loss) {
  if (!m || !features || !targets || !grad || n < 1) return fail(LNT_ERR_INVALID_ARGUMENT, "bad arguments");
  if (n > (int64_t(1) << 26)) return fail(LNT_ERR_TOO_LARGE, "too many samples");
  const int64_t cnt = measured_count(targets, n);
  if (cnt == 0) return fail(LNT_ERR_INVALID_ARGUMENT, "no measured target");
  LNT_HIP(hipSetDevice(m->ctx->device));
  float *Xd, *Td;
  int32_t *idx;
  LNT_TRY(upload_set(m, features, targets, n, &Xd, &Td, &idx));
  Work w;
  LNT_TRY(work_for(m, n, &w));
  float *G = (float *)m->opt.ptr + 2 * NW;
  const double l = full_loss(m, Xd, Td, n, (float)(1.0 / cnt), w, G);
  if (l != l) return fail(LNT_ERR_DEVICE, "gradient evaluation failed");
  LNT_HIP(hipMemcpyAsync(grad, G, NW * sizeof(float), hipMemcpyDeviceToHost, m->ctx->stream));
  LNT_HIP(hipStreamSynchronize(m->ctx->stream));
  if (loss) *loss = (float)l;
  return LNT_OK;
}

int lnt_model_train(lnt_model m, const float *features, const float *targets, int64_t n, int epochs, int batch,
                    float learning_rate, uint64_t seed, float *loss) {
  if (!m || !features || !targets || n < 1 || epochs < 0 || batch < 1 || !(learning_rate > 0))
    return fail(LNT_ERR_INVALID_ARGUMENT, "bad arguments");
  if (n > (int64_t(1) << 26)) return fail(LNT_ERR_TOO_LARGE, "too many samples");
  LNT_HIP(hipSetDevice(m->ctx->device));
  hipStream_t st = m->ctx->stream;
  // Standardisation from the training features.
  for (int k = 0; k < F; ++k) {
    double s = 0, s2 = 0;
    for (int64_t i = 0; i < n; ++i) s += features[i * F + k];
    const double mean = s / n;
    for (int64_t i = 0; i < n; ++i) s2 += (features[i * F + k] - mean) * (features[i * F + k] - mean);
    const double sd = sqrt(s2 / n);
    m->host[OFF_MEAN + k] = (float)mean;
    m->host[OFF_ISTD + k] = sd > 1e-9 ? (float)(1.0 / sd) : 1.f;
  }
  LNT_TRY(upload_params(m));
  std::vector<int32_t> per(n);
  for (int64_t i = 0; i < n; ++i) {
    int c = 0;
    for (int o = 0; o < S; ++o) c += targets[i * S + o] == targets[i * S + o];
    per[i] = c;
  }
  float *Xd, *Td;
  int32_t *idx;
  LNT_TRY(upload_set(m, features, targets, n, &Xd, &Td, &idx));
  Work w;
  LNT_TRY(work_for(m, std::max<int64_t>(n, batch), &w));
  float *M = (float *)m->opt.ptr, *V = M + NW, *G = V + NW;
  LNT_HIP(hipMemsetAsync(M, 0, 2 * NW * sizeof(float), st));
  std::vector<int32_t> order(n);
  for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
  uint64_t state = seed;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double p1 = 1.0, p2 = 1.0;
  for (int e = 0; e < epochs; ++e) {
    for (int64_t i = n - 1; i > 0; --i) {  // Fisher-Yates
      const int64_t j = (int64_t)(splitmix(state) % (uint64_t)(i + 1));
      std::swap(order[i], order[j]);
    }
    LNT_HIP(hipMemcpyAsync(idx, order.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    LNT_HIP(hipStreamSynchronize(st));  // `order` is rewritten next epoch
    for (int64_t r0 = 0; r0 < n; r0 += batch) {
      const int nb = (int)std::min<int64_t>(batch, n - r0);
      int64_t cnt = 0;
      for (int r = 0; r < nb; ++r) cnt += per[order[r0 + r]];
      if (cnt == 0) continue;
      mlp_gather_kernel<<<blocks(nb), kThreads, 0, st>>>(Xd, Td, idx + r0, nb, w.Xb, w.Tb);
      LNT_TRY(step_grad(m, w.Xb, w.Tb, nb, (float)(1.0 / cnt), w, G));
      p1 *= b1;
      p2 *= b2;
      mlp_adam_kernel<<<blocks(NW), kThreads, 0, st>>>((float *)m->dev.ptr, G, M, V, learning_rate, b1, b2, eps,
                                                         (float)(1.0 - p1), (float)(1.0 - p2));
      LNT_HIP(hipGetLastError());
    }
  }
  if (loss) {
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; ++i) cnt += per[i];
    const double l = cnt ? full_loss(m, Xd, Td, n, (float)(1.0 / cnt), w, G) : 0.0;
    *loss = (float)l;
  }
  return download_params(m);
}

int lnt_context_set_model(lnt_context ctx, lnt_model model) {
  if (!ctx) return fail(LNT_ERR_INVALID_ARGUMENT, "null context");
  if (model && model->ctx != ctx) return fail(LNT_ERR_INVALID_ARGUMENT, "the model belongs to another context");
  ctx->model = model;
  return LNT_OK;
}

}  // extern "C"
