// mlp.hip -- the cost model of the learned tuner: a small fully connected
// network (inputs -> hidden x layers, ReLU -> 1) in FP32, with forward,
// backward (mean squared error) and Adam kernels, all on the device.
//
// Parameters are one flat device array: for each layer l with fan-in a and
// fan-out b, W_l (a x b, row-major: W_l[i * b + j] connects input i to unit
// j) followed by its bias (b).  Gradients use the same layout.
//
// The reference has no learned tuner: its tuner (src/DedispersionTuning.cpp)
// is exhaustive.  This network ranks configurations so that only the most
// promising ones are timed; the exhaustive sweep stays the fallback.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "dedispersion_amd.h"
#include "internal.hpp"

namespace {

using ddamd::set_error;

constexpr uint32_t kMaxWidth = 256;

// Y[r][j] = act(b[j] + sum_i X[r][i] W[i][j]) for r < n, j < dout.
__global__ void dense_forward(const float *__restrict__ X, uint32_t n, uint32_t din, const float *__restrict__ W,
                              const float *__restrict__ b, uint32_t dout, int relu, float *__restrict__ Y) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (uint64_t)n * dout) return;
  const uint32_t r = (uint32_t)(idx / dout), j = (uint32_t)(idx % dout);
  const float *x = X + (uint64_t)r * din;
  float acc = b[j];
  for (uint32_t i = 0; i < din; ++i) acc += x[i] * W[(uint64_t)i * dout + j];
  Y[idx] = relu ? fmaxf(acc, 0.0f) : acc;
}

// Output layer error of L = (1/n) sum_r (y_r - t_r)^2: G[r] = 2 (y_r - t_r) / n,
// and the per-row squared error into E.
__global__ void mse_error(const float *__restrict__ y, const float *__restrict__ t, uint32_t n, float *__restrict__ G,
                          float *__restrict__ E) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const float d = y[r] - t[r];
  G[r] = 2.0f * d / (float)n;
  E[r] = d * d;
}

// loss = (1/n) sum E, one workgroup of 256 threads.
__global__ void mean_of(const float *__restrict__ E, uint32_t n, float *__restrict__ out) {
  __shared__ float part[256];
  float s = 0.0f;
  for (uint32_t r = threadIdx.x; r < n; r += blockDim.x) s += E[r];
  part[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t w = blockDim.x / 2; w > 0; w /= 2) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0] / (float)n;
}

// dW[i][j] = sum_r X[r][i] G[r][j]; db[j] = sum_r G[r][j].  One thread per
// weight (and per bias, the extra row i = din).
__global__ void dense_weight_grad(const float *__restrict__ X, const float *__restrict__ G, uint32_t n, uint32_t din,
                                  uint32_t dout, float *__restrict__ dW, float *__restrict__ db) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (din + 1) * dout) return;
  const uint32_t i = idx / dout, j = idx % dout;
  float acc = 0.0f;
  if (i < din) {
    for (uint32_t r = 0; r < n; ++r) acc += X[(uint64_t)r * din + i] * G[(uint64_t)r * dout + j];
    dW[(uint64_t)i * dout + j] = acc;
  } else {
    for (uint32_t r = 0; r < n; ++r) acc += G[(uint64_t)r * dout + j];
    db[j] = acc;
  }
}

// Error at the layer input, through the ReLU of the layer below:
// GX[r][i] = (H[r][i] > 0) * sum_j G[r][j] W[i][j].
__global__ void dense_input_grad(const float *__restrict__ G, const float *__restrict__ W, const float *__restrict__ H,
                                 uint32_t n, uint32_t din, uint32_t dout, float *__restrict__ GX) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (uint64_t)n * din) return;
  const uint32_t r = (uint32_t)(idx / din), i = (uint32_t)(idx % din);
  if (H[idx] <= 0.0f) {
    GX[idx] = 0.0f;
    return;
  }
  const float *g = G + (uint64_t)r * dout;
  const float *w = W + (uint64_t)i * dout;
  float acc = 0.0f;
  for (uint32_t j = 0; j < dout; ++j) acc += g[j] * w[j];
  GX[idx] = acc;
}

// Adam (Kingma and Ba, 2015) with bias correction factors c1 = 1 - beta1^t,
// c2 = 1 - beta2^t computed on the host.
__global__ void adam_update(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, uint32_t count, float lr, float beta1, float beta2, float eps,
                            float c1, float c2) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const float gk = g[k];
  const float mk = beta1 * m[k] + (1.0f - beta1) * gk;
  const float vk = beta2 * v[k] + (1.0f - beta2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

inline uint32_t blocks_for(uint64_t n, uint32_t threads) { return (uint32_t)((n + threads - 1) / threads); }

} // namespace

struct dd_mlp {
  uint32_t inputs = 0, hidden = 0, layers = 0;  // hidden layers
  std::vector<uint32_t> fan_in, fan_out;         // per dense layer (layers + 1)
  std::vector<size_t> w_off, b_off;              // offsets into the flat parameter array
  size_t count = 0;
  float *params = nullptr, *m = nullptr, *v = nullptr;
  uint32_t step = 0;
  // activations / errors for the last n rows
  uint32_t cap = 0;
  std::vector<float *> act;   // act[0] unused (the input), act[l] = output of layer l-1 (n x fan_out)
  float *err_a = nullptr, *err_b = nullptr, *sq = nullptr, *grad_tmp = nullptr, *loss_tmp = nullptr;

  ~dd_mlp() {
    (void)hipFree(params);
    (void)hipFree(m);
    (void)hipFree(v);
    release_rows();
    (void)hipFree(grad_tmp);
    (void)hipFree(loss_tmp);
  }
  void release_rows() {
    for (float *p : act) (void)hipFree(p);
    act.clear();
    (void)hipFree(err_a);
    (void)hipFree(err_b);
    (void)hipFree(sq);
    err_a = err_b = sq = nullptr;
    cap = 0;
  }
};

namespace {

dd_status fail_mlp(dd_status s, const std::string &msg) {
  set_error(msg);
  return s;
}

#define MLP_HIP(call)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) return fail_mlp(DD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

dd_status reserve_rows(dd_mlp *net, uint32_t n) {
  if (n <= net->cap) return DD_OK;
  net->release_rows();
  net->act.assign(net->layers + 2, nullptr);
  for (uint32_t l = 0; l <= net->layers; ++l)
    MLP_HIP(hipMalloc(reinterpret_cast<void **>(&net->act[l + 1]), (size_t)n * net->fan_out[l] * sizeof(float)));
  const uint32_t widest = std::max(net->hidden, net->inputs);
  MLP_HIP(hipMalloc(reinterpret_cast<void **>(&net->err_a), (size_t)n * widest * sizeof(float)));
  MLP_HIP(hipMalloc(reinterpret_cast<void **>(&net->err_b), (size_t)n * widest * sizeof(float)));
  MLP_HIP(hipMalloc(reinterpret_cast<void **>(&net->sq), (size_t)n * sizeof(float)));
  net->cap = n;
  return DD_OK;
}

dd_status forward(dd_mlp *net, const float *d_x, uint32_t n, hipStream_t s) {
  const float *in = d_x;
  for (uint32_t l = 0; l <= net->layers; ++l) {
    const uint32_t a = net->fan_in[l], b = net->fan_out[l];
    const int relu = l < net->layers ? 1 : 0;
    hipLaunchKernelGGL(dense_forward, dim3(blocks_for((uint64_t)n * b, 256)), dim3(256), 0, s, in, n, a,
                       net->params + net->w_off[l], net->params + net->b_off[l], b, relu, net->act[l + 1]);
    MLP_HIP(hipGetLastError());
    in = net->act[l + 1];
  }
  return DD_OK;
}

dd_status loss_grad(dd_mlp *net, const float *d_x, const float *d_t, uint32_t n, float *d_loss, float *d_grad,
                    hipStream_t s) {
  dd_status st = forward(net, d_x, n, s);
  if (st != DD_OK) return st;
  const uint32_t L = net->layers;
  hipLaunchKernelGGL(mse_error, dim3(blocks_for(n, 256)), dim3(256), 0, s, net->act[L + 1], d_t, n, net->err_a,
                     net->sq);
  MLP_HIP(hipGetLastError());
  hipLaunchKernelGGL(mean_of, dim3(1), dim3(256), 0, s, net->sq, n, d_loss);
  MLP_HIP(hipGetLastError());
  float *G = net->err_a, *GX = net->err_b;
  for (int l = (int)L; l >= 0; --l) {
    const uint32_t a = net->fan_in[l], b = net->fan_out[l];
    const float *X = l == 0 ? d_x : net->act[l];
    hipLaunchKernelGGL(dense_weight_grad, dim3(blocks_for((uint64_t)(a + 1) * b, 256)), dim3(256), 0, s, X, G, n, a,
                       b, d_grad + net->w_off[l], d_grad + net->b_off[l]);
    MLP_HIP(hipGetLastError());
    if (l > 0) {
      hipLaunchKernelGGL(dense_input_grad, dim3(blocks_for((uint64_t)n * a, 256)), dim3(256), 0, s, G,
                         net->params + net->w_off[l], net->act[l], n, a, b, GX);
      MLP_HIP(hipGetLastError());
      std::swap(G, GX);
    }
  }
  return DD_OK;
}

dd_status adam(dd_mlp *net, const float *d_grad, float lr, hipStream_t s) {
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  net->step += 1;
  const float c1 = 1.0f - (float)std::pow((double)beta1, (double)net->step);
  const float c2 = 1.0f - (float)std::pow((double)beta2, (double)net->step);
  hipLaunchKernelGGL(adam_update, dim3(blocks_for(net->count, 256)), dim3(256), 0, s, net->params, d_grad, net->m,
                     net->v, (uint32_t)net->count, lr, beta1, beta2, eps, c1, c2);
  MLP_HIP(hipGetLastError());
  return DD_OK;
}

} // namespace

extern "C" {

dd_status dd_mlp_create(dd_mlp **out, uint32_t inputs, uint32_t hidden, uint32_t hidden_layers) {
  if (!out) return fail_mlp(DD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (inputs == 0 || inputs > kMaxWidth || hidden == 0 || hidden > kMaxWidth || hidden_layers == 0 ||
      hidden_layers > 8)
    return fail_mlp(DD_ERR_INVALID_ARGUMENT, "network shape out of range");
  dd_mlp *net = new dd_mlp();
  net->inputs = inputs;
  net->hidden = hidden;
  net->layers = hidden_layers;
  size_t off = 0;
  for (uint32_t l = 0; l <= hidden_layers; ++l) {
    const uint32_t a = l == 0 ? inputs : hidden, b = l == hidden_layers ? 1 : hidden;
    net->fan_in.push_back(a);
    net->fan_out.push_back(b);
    net->w_off.push_back(off);
    off += (size_t)a * b;
    net->b_off.push_back(off);
    off += b;
  }
  net->count = off;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&net->params), off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&net->m), off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&net->v), off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&net->grad_tmp), off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&net->loss_tmp), sizeof(float));
  if (e == hipSuccess) e = hipMemset(net->params, 0, off * sizeof(float));
  if (e == hipSuccess) e = hipMemset(net->m, 0, off * sizeof(float));
  if (e == hipSuccess) e = hipMemset(net->v, 0, off * sizeof(float));
  if (e != hipSuccess) {
    delete net;
    return fail_mlp(DD_ERR_HIP, std::string("allocating the network: ") + hipGetErrorString(e));
  }
  *out = net;
  return DD_OK;
}

dd_status dd_mlp_destroy(dd_mlp *net) {
  delete net;
  return DD_OK;
}

dd_status dd_mlp_param_count(const dd_mlp *net, uint32_t *count) {
  if (!net || !count) return fail_mlp(DD_ERR_INVALID_ARGUMENT, "null argument");
  *count = (uint32_t)net->count;
  return DD_OK;
}

dd_status dd_mlp_set_params(dd_mlp *net, const float *params) {
  if (!net || !params) return fail_mlp(DD_ERR_INVALID_ARGUMENT, "null argument");
  MLP_HIP(hipMemcpy(net->params, params, net->count * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(hipMemset(net->m, 0, net->count * sizeof(float)));
  MLP_HIP(hipMemset(net->v, 0, net->count * sizeof(float)));
  net->step = 0;
  return DD_OK;
}

dd_status dd_mlp_get_params(const dd_mlp *net, float *params) {
  if (!net || !params) return fail_mlp(DD_ERR_INVALID_ARGUMENT, "null argument");
  MLP_HIP(hipMemcpy(params, net->params, net->count * sizeof(float), hipMemcpyDeviceToHost));
  return DD_OK;
}

dd_status dd_mlp_init(dd_mlp *net, uint64_t seed) {
  if (!net) return fail_mlp(DD_ERR_INVALID_ARGUMENT, "null argument");
  // He-uniform weights (bound sqrt(6 / fan_in)), zero biases.
  std::mt19937_64 rng(seed);
  std::vector<float> p(net->count, 0.0f);
  for (uint32_t l = 0; l <= net->layers; ++l) {
    const float bound = std::sqrt(6.0f / (float)net->fan_in[l]);
    std::uniform_real_distribution<float> u(-bound, bound);
    for (size_t k = 0; k < (size_t)net->fan_in[l] * net->fan_out[l]; ++k) p[net->w_off[l] + k] = u(rng);
  }
  return dd_mlp_set_params(net, p.data());
}

dd_status dd_mlp_forward(dd_mlp *net, const float *d_x, uint32_t n, float *d_y, void *stream) {
  if (!net || !d_x || !d_y) return fail_mlp(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0) return DD_OK;
  dd_status s = reserve_rows(net, n);
  if (s == DD_OK) s = forward(net, d_x, n, (hipStream_t)stream);
  if (s != DD_OK) return s;
  MLP_HIP(hipMemcpyAsync(d_y, net->act[net->layers + 1], (size_t)n * sizeof(float), hipMemcpyDeviceToDevice,
                         (hipStream_t)stream));
  return DD_OK;
}

dd_status dd_mlp_loss_grad(dd_mlp *net, const float *d_x, const float *d_target, uint32_t n, float *d_loss,
                           float *d_grad, void *stream) {
  if (!net || !d_x || !d_target || !d_loss || !d_grad || n == 0)
    return fail_mlp(DD_ERR_INVALID_ARGUMENT, "invalid argument");
  dd_status s = reserve_rows(net, n);
  return s != DD_OK ? s : loss_grad(net, d_x, d_target, n, d_loss, d_grad, (hipStream_t)stream);
}

dd_status dd_mlp_adam_step(dd_mlp *net, const float *d_grad, float lr, void *stream) {
  if (!net || !d_grad) return fail_mlp(DD_ERR_INVALID_ARGUMENT, "null argument");
  return adam(net, d_grad, lr, (hipStream_t)stream);
}

dd_status dd_mlp_train(dd_mlp *net, const float *d_x, const float *d_target, uint32_t n, uint32_t steps, float lr,
                       float *losses, void *stream) {
  if (!net || !d_x || !d_target || n == 0) return fail_mlp(DD_ERR_INVALID_ARGUMENT, "invalid argument");
  dd_status s = reserve_rows(net, n);
  if (s != DD_OK) return s;
  hipStream_t st = (hipStream_t)stream;
  for (uint32_t k = 0; k < steps; ++k) {
    s = loss_grad(net, d_x, d_target, n, net->loss_tmp, net->grad_tmp, st);
    if (s == DD_OK) s = adam(net, net->grad_tmp, lr, st);
    if (s != DD_OK) return s;
    if (losses) MLP_HIP(hipMemcpyAsync(losses + k, net->loss_tmp, sizeof(float), hipMemcpyDeviceToHost, st));
  }
  MLP_HIP(hipStreamSynchronize(st));
  return DD_OK;
}

} // extern "C"
