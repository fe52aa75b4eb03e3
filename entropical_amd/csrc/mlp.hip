// mlp.hip -- the learned bandwidth model: a small MLP (n_in -> H -> H -> H -> 1,
// tanh hidden units) trained and evaluated on the GPU with these kernels.
//
// Weights live in FP32 (master copy, updated by Adam) and are rounded to
// BF16 (round-to-nearest-even) before every forward pass; forward and
// backward read the BF16 weights and accumulate in FP32.  The layer products
// (forward, and both products of the backward pass) run on the matrix cores
// (FP32 MFMA, 16 x 16 tiles, one wave each).
//
// Layer l maps a_{l} (B x n_l, FP32) to z_{l+1} = a_l W_l + b_l, with
// a_{l+1} = tanh(z_{l+1}) for hidden layers and the identity for the output.
// Loss: mean squared error over the batch.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/entropical_kde.h"
#include "ctx.hpp"
#include "engine.hpp"

namespace {

using ent::set_error;

constexpr int kMaxLayers = 4;

__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan pass through
  u += 0x7fffu + ((u >> 16) & 1u);                                     // round to nearest even
  return (uint16_t)(u >> 16);
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// One wave's 16 x 16 tile of C = A B (M x N, inner size K) on the matrix
// cores, FP32 in and out (v_mfma_f32_16x16x4_f32): per step of 4 along K,
// lane l supplies A(m0 + l % 16, k0 + l / 16) and B(k0 + l / 16, n0 + l % 16);
// element i of lane l's accumulator is C(m0 + 4 (l / 16) + i, n0 + l % 16).
// Operands outside the matrices are zero; epi(m, n, c) receives each
// element of the tile inside C.  The BF16 weights enter as their exact FP32
// values, so the products are those of the FP32 model with BF16 weights.
template <class FA, class FB, class FE>
__device__ __forceinline__ void mfma_tile(int M, int N, int K, int m0, int n0, FA a, FB b, FE epi) {
  const int l = (int)__lane_id();
  const int r = l % 16, g = l / 16;
  floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + g;
    const float av = (m0 + r < M && k < K) ? a(m0 + r, k) : 0.0f;
    const float bv = (k < K && n0 + r < N) ? b(k, n0 + r) : 0.0f;
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + 4 * g + i, n = n0 + r;
    if (m < M && n < N) epi(m, n, acc[i]);
  }
}

// Y[b][o] = act(sum_i A[b][i] * W[i][o] + bias[o]); W is BF16 [n_in][n_out].
// Grid: (ceil(n_out / 16), ceil(B / 16)) waves.
__global__ __launch_bounds__(64) void k_linear_fwd(const float* __restrict__ A, const uint16_t* __restrict__ W,
                                                   const float* __restrict__ bias, float* __restrict__ Y, int B,
                                                   int n_in, int n_out, int use_tanh) {
  mfma_tile(
      B, n_out, n_in, (int)blockIdx.y * 16, (int)blockIdx.x * 16,
      [&](int b, int i) { return A[(size_t)b * n_in + i]; },
      [&](int i, int o) { return bf16_to_f(W[(size_t)i * n_out + o]); },
      [&](int b, int o, float c) {
        const float z = c + bias[o];
        Y[(size_t)b * n_out + o] = use_tanh ? tanhf(z) : z;
      });
}

// dZ of the output layer for MSE: dZ[b] = 2 (y[b] - t[b]) / B; loss partial.
__global__ void k_mse_grad(const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ dz,
                           float* __restrict__ loss_part, int B) {
  __shared__ float s[256];
  float acc = 0.0f;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    const float r = y[b] - t[b];
    dz[b] = 2.0f * r / (float)B;
    acc += r * r;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss_part = s[0] / (float)B;
}

// dA[b][i] = (sum_o dZ[b][o] * W[i][o]) * (1 - A[b][i]^2)   (tanh input layer)
// Grid: (ceil(n_in / 16), ceil(B / 16)) waves.
__global__ __launch_bounds__(64) void k_linear_bwd_data(const float* __restrict__ dZ,
                                                        const uint16_t* __restrict__ W, const float* __restrict__ A,
                                                        float* __restrict__ dA, int B, int n_in, int n_out) {
  mfma_tile(
      B, n_in, n_out, (int)blockIdx.y * 16, (int)blockIdx.x * 16,
      [&](int b, int o) { return dZ[(size_t)b * n_out + o]; },
      [&](int o, int i) { return bf16_to_f(W[(size_t)i * n_out + o]); },
      [&](int b, int i, float c) {
        const float a = A[(size_t)b * n_in + i];
        dA[(size_t)b * n_in + i] = c * (1.0f - a * a);
      });
}

// dW[i][o] = sum_b A[b][i] dZ[b][o];  db[o] = sum_b dZ[b][o] (row i = n_in
// of the same product, with A's extra column all ones).
// Grid: (ceil(n_out / 16), ceil((n_in + 1) / 16)) waves.
__global__ __launch_bounds__(64) void k_linear_bwd_weight(const float* __restrict__ A, const float* __restrict__ dZ,
                                                          float* __restrict__ dW, float* __restrict__ db, int B,
                                                          int n_in, int n_out) {
  mfma_tile(
      n_in + 1, n_out, B, (int)blockIdx.y * 16, (int)blockIdx.x * 16,
      [&](int i, int b) { return i < n_in ? A[(size_t)b * n_in + i] : 1.0f; },
      [&](int b, int o) { return dZ[(size_t)b * n_out + o]; },
      [&](int i, int o, float c) {
        if (i < n_in) dW[(size_t)i * n_out + o] = c;
        else db[o] = c;
      });
}

// Adam on FP32 master weights; refreshes the BF16 copy.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, uint16_t* __restrict__ w16, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1.0f - b1) * gk;
  const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
  w[k] = wk;
  if (w16) w16[k] = f_to_bf16(wk);
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ w16, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) w16[k] = f_to_bf16(w[k]);
}

// rows of X selected by idx -> contiguous batch
__global__ void k_gather_rows(const float* __restrict__ X, const float* __restrict__ T,
                              const uint32_t* __restrict__ idx, float* __restrict__ xb, float* __restrict__ tb,
                              int B, int n_in) {
  const int b = blockIdx.x;
  if (b >= B) return;
  const uint32_t r = idx[b];
  for (int i = threadIdx.x; i < n_in; i += blockDim.x) xb[(size_t)b * n_in + i] = X[(size_t)r * n_in + i];
  if (threadIdx.x == 0 && T) tb[b] = T[r];
}

}  // namespace

// ---------------------------------------------------------------- host side
struct ent_mlp {
  int n_in = 0, hidden = 0, n_layers = 0;  // n_layers linear maps
  int dims[kMaxLayers + 1] = {};
  size_t w_off[kMaxLayers] = {}, b_off[kMaxLayers] = {};
  size_t n_params = 0;
  // device: params (FP32), grads, adam m/v, bf16 weights (same offsets)
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* p16 = nullptr;
  // scratch for activations: per layer B x dims[l]
  float* act = nullptr;
  float* dact = nullptr;
  size_t act_cap = 0;
  float* xb = nullptr;  // batch inputs / targets / dz
  float* tb = nullptr;
  float* loss = nullptr;
  uint32_t* idx = nullptr;
  size_t batch_cap = 0;
  int step = 0;
  int device = 0;
};

namespace {

#define MLP_HIP(call)                                                                \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call);  \
      return e_ == hipErrorOutOfMemory ? ENT_ERR_OOM : ENT_ERR_HIP;                  \
    }                                                                                \
  } while (0)

// splitmix64 -> uniform / normal for initialisation
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

int ensure_batch(ent_mlp* M, size_t B) {
  if (B <= M->batch_cap) return ENT_OK;
  // the old buffers are dropped before the larger ones are allocated
  (void)hipFree(M->act);
  (void)hipFree(M->dact);
  (void)hipFree(M->xb);
  (void)hipFree(M->tb);
  (void)hipFree(M->idx);
  size_t per = 0;
  for (int l = 0; l <= M->n_layers; ++l) per += (size_t)M->dims[l];
  MLP_HIP(hipMalloc(&M->act, sizeof(float) * B * per));
  MLP_HIP(hipMalloc(&M->dact, sizeof(float) * B * per));
  MLP_HIP(hipMalloc(&M->xb, sizeof(float) * B * M->n_in));
  MLP_HIP(hipMalloc(&M->tb, sizeof(float) * B));
  MLP_HIP(hipMalloc(&M->idx, sizeof(uint32_t) * B));
  M->batch_cap = B;
  return ENT_OK;
}

float* act_ptr(ent_mlp* M, float* base, size_t B, int l) {
  size_t off = 0;
  for (int k = 0; k < l; ++k) off += (size_t)M->dims[k];
  return base + off * B;
}

// forward from M->act layer 0 (already filled with the batch inputs)
void forward(ent_mlp* M, hipStream_t s, int B) {
  for (int l = 0; l < M->n_layers; ++l) {
    const int ni = M->dims[l], no = M->dims[l + 1];
    dim3 grid((no + 15) / 16, (B + 15) / 16);
    hipLaunchKernelGGL(k_linear_fwd, grid, dim3(64), 0, s, act_ptr(M, M->act, B, l), M->p16 + M->w_off[l],
                       M->p + M->b_off[l], act_ptr(M, M->act, B, l + 1), B, ni, no, l + 1 < M->n_layers ? 1 : 0);
  }
}

// backward: gradients into M->g; loss partial into M->loss[0]
void backward(ent_mlp* M, hipStream_t s, int B, const float* targets) {
  const int L = M->n_layers;
  float* dz = act_ptr(M, M->dact, B, L);
  hipLaunchKernelGGL(k_mse_grad, dim3(1), dim3(256), 0, s, act_ptr(M, M->act, B, L), targets, dz, M->loss, B);
  for (int l = L - 1; l >= 0; --l) {
    const int ni = M->dims[l], no = M->dims[l + 1];
    float* dZl = act_ptr(M, M->dact, B, l + 1);
    hipLaunchKernelGGL(k_linear_bwd_weight, dim3((no + 15) / 16, (ni + 1 + 15) / 16), dim3(64), 0, s,
                       act_ptr(M, M->act, B, l), dZl, M->g + M->w_off[l], M->g + M->b_off[l], B, ni, no);
    if (l > 0)
      hipLaunchKernelGGL(k_linear_bwd_data, dim3((ni + 15) / 16, (B + 15) / 16), dim3(64), 0, s, dZl,
                         M->p16 + M->w_off[l], act_ptr(M, M->act, B, l), act_ptr(M, M->dact, B, l), B, ni, no);
  }
}

}  // namespace

extern "C" {

int ent_mlp_create(ent_ctx* ctx, int n_in, int hidden, int n_hidden_layers, uint64_t seed, ent_mlp** out) {
  if (!ctx || !ctx->eng || !out) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  if (n_in < 1 || n_in > 4096 || hidden < 1 || hidden > 4096 || n_hidden_layers < 1 ||
      n_hidden_layers > kMaxLayers - 1) {
    set_error("invalid MLP shape (1..4096 inputs/hidden units, 1..3 hidden layers)");
    return ENT_ERR_ARG;
  }
  MLP_HIP(hipSetDevice(ctx->eng->device()));
  ent_mlp* M = new ent_mlp;
  M->device = ctx->eng->device();
  M->n_in = n_in;
  M->hidden = hidden;
  M->n_layers = n_hidden_layers + 1;
  M->dims[0] = n_in;
  for (int l = 1; l <= n_hidden_layers; ++l) M->dims[l] = hidden;
  M->dims[M->n_layers] = 1;
  size_t off = 0;
  for (int l = 0; l < M->n_layers; ++l) {
    M->w_off[l] = off;
    off += (size_t)M->dims[l] * M->dims[l + 1];
    M->b_off[l] = off;
    off += (size_t)M->dims[l + 1];
  }
  M->n_params = off;
  std::vector<float> init(off, 0.0f);
  Rng r{seed};
  for (int l = 0; l < M->n_layers; ++l) {
    const double sd = std::sqrt(1.0 / M->dims[l]);  // LeCun normal (tanh)
    for (size_t k = 0; k < (size_t)M->dims[l] * M->dims[l + 1]; ++k) init[M->w_off[l] + k] = (float)(sd * r.normal());
  }
  hipError_t e = hipMalloc(&M->p, sizeof(float) * off);
  if (e == hipSuccess) e = hipMalloc(&M->g, sizeof(float) * off);
  if (e == hipSuccess) e = hipMalloc(&M->m, sizeof(float) * off);
  if (e == hipSuccess) e = hipMalloc(&M->v, sizeof(float) * off);
  if (e == hipSuccess) e = hipMalloc(&M->p16, sizeof(uint16_t) * off);
  if (e == hipSuccess) e = hipMalloc(&M->loss, sizeof(float) * 4);
  if (e != hipSuccess) {
    set_error(std::string("HIP error: ") + hipGetErrorString(e));
    delete M;
    return ENT_ERR_OOM;
  }
  hipStream_t s = ctx->eng->stream();
  MLP_HIP(hipMemcpyAsync(M->p, init.data(), sizeof(float) * off, hipMemcpyHostToDevice, s));
  std::vector<float> zeros(off, 0.0f);
  MLP_HIP(hipMemcpyAsync(M->m, zeros.data(), sizeof(float) * off, hipMemcpyHostToDevice, s));
  MLP_HIP(hipMemcpyAsync(M->v, zeros.data(), sizeof(float) * off, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((off + 255) / 256)), dim3(256), 0, s, M->p, M->p16, (int)off);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipStreamSynchronize(s));
  *out = M;
  return ENT_OK;
}

int ent_mlp_destroy(ent_mlp* M) {
  if (!M) return ENT_OK;
  (void)hipSetDevice(M->device);
  (void)hipDeviceSynchronize();
  for (void* p : {(void*)M->p, (void*)M->g, (void*)M->m, (void*)M->v, (void*)M->p16, (void*)M->act,
                  (void*)M->dact, (void*)M->xb, (void*)M->tb, (void*)M->loss, (void*)M->idx})
    if (p) (void)hipFree(p);
  delete M;
  return ENT_OK;
}

int ent_mlp_num_params(ent_mlp* M, size_t* n) {
  if (!M || !n) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  *n = M->n_params;
  return ENT_OK;
}

int ent_mlp_get_params(ent_ctx* ctx, ent_mlp* M, float* out) {
  if (!ctx || !M || !out) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  hipStream_t s = ctx->eng->stream();
  MLP_HIP(hipMemcpyAsync(out, M->p, sizeof(float) * M->n_params, hipMemcpyDeviceToHost, s));
  MLP_HIP(hipStreamSynchronize(s));
  return ENT_OK;
}

int ent_mlp_set_params(ent_ctx* ctx, ent_mlp* M, const float* in) {
  if (!ctx || !M || !in) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  hipStream_t s = ctx->eng->stream();
  MLP_HIP(hipMemcpyAsync(M->p, in, sizeof(float) * M->n_params, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((M->n_params + 255) / 256)), dim3(256), 0, s, M->p, M->p16,
                     (int)M->n_params);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipStreamSynchronize(s));
  return ENT_OK;
}

int ent_mlp_predict(ent_ctx* ctx, ent_mlp* M, const float* X, size_t n, float* y) {
  if (!ctx || !M || !X || !y) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  hipStream_t s = ctx->eng->stream();
  const size_t B = 4096;
  if (ensure_batch(M, std::min(B, std::max<size_t>(n, 1))) != ENT_OK) return ENT_ERR_OOM;
  for (size_t s0 = 0; s0 < n; s0 += M->batch_cap) {
    const int b = (int)std::min(M->batch_cap, n - s0);
    MLP_HIP(hipMemcpyAsync(M->act, X + s0 * M->n_in, sizeof(float) * b * M->n_in, hipMemcpyHostToDevice, s));
    forward(M, s, b);
    MLP_HIP(hipMemcpyAsync(y + s0, act_ptr(M, M->act, b, M->n_layers), sizeof(float) * b, hipMemcpyDeviceToHost, s));
  }
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipStreamSynchronize(s));
  return ENT_OK;
}

// Gradient of the mean squared error over (X, t) at the current parameters
// (no update): grads has n_params floats, *loss the batch MSE.
int ent_mlp_gradient(ent_ctx* ctx, ent_mlp* M, const float* X, const float* t, size_t n, float* grads,
                     float* loss) {
  if (!ctx || !M || !X || !t || !grads || n == 0 || n > (1u << 20)) {
    set_error("null argument or batch size outside 1..2^20");
    return ENT_ERR_ARG;
  }
  hipStream_t s = ctx->eng->stream();
  if (ensure_batch(M, n) != ENT_OK) return ENT_ERR_OOM;
  const int B = (int)n;
  MLP_HIP(hipMemcpyAsync(M->act, X, sizeof(float) * n * M->n_in, hipMemcpyHostToDevice, s));
  MLP_HIP(hipMemcpyAsync(M->tb, t, sizeof(float) * n, hipMemcpyHostToDevice, s));
  forward(M, s, B);
  backward(M, s, B, M->tb);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipMemcpyAsync(grads, M->g, sizeof(float) * M->n_params, hipMemcpyDeviceToHost, s));
  float l = 0.0f;
  MLP_HIP(hipMemcpyAsync(&l, M->loss, sizeof(float), hipMemcpyDeviceToHost, s));
  MLP_HIP(hipStreamSynchronize(s));
  if (loss) *loss = l;
  return ENT_OK;
}

// Minibatch Adam training on (X [n x n_in], t [n]) resident on the device for
// the call.  order: host array of epochs*n row indices (the caller's
// shuffles).  losses: per-epoch mean training loss (host, epochs floats).
int ent_mlp_train(ent_ctx* ctx, ent_mlp* M, const float* X, const float* t, size_t n, const uint32_t* order,
                  int epochs, int batch, float lr, float* losses) {
  if (!ctx || !M || !X || !t || !order || n == 0 || epochs < 1 || batch < 1) {
    set_error("null argument or empty training set");
    return ENT_ERR_ARG;
  }
  hipStream_t s = ctx->eng->stream();
  ent::Engine& E = *ctx->eng;
  if (ensure_batch(M, (size_t)batch) != ENT_OK) return ENT_ERR_OOM;
  // device copies of X, t, order
  size_t bx = sizeof(float) * n * M->n_in, bt = sizeof(float) * n, bo = sizeof(uint32_t) * n * epochs;
  if (E.ensure(ctx->aux, bx + bt + bo) != ent::kOk) return ENT_ERR_OOM;
  char* base = (char*)ctx->aux.p;
  float* dX = (float*)base;
  float* dT = (float*)(base + bx);
  uint32_t* dO = (uint32_t*)(base + bx + bt);
  MLP_HIP(hipMemcpyAsync(dX, X, bx, hipMemcpyHostToDevice, s));
  MLP_HIP(hipMemcpyAsync(dT, t, bt, hipMemcpyHostToDevice, s));
  MLP_HIP(hipMemcpyAsync(dO, order, bo, hipMemcpyHostToDevice, s));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  std::vector<float> step_loss;
  std::vector<float> hl(1);
  for (int ep = 0; ep < epochs; ++ep) {
    double ep_loss = 0.0;
    int nb = 0;
    for (size_t s0 = 0; s0 + (size_t)batch <= n; s0 += batch) {
      const int B = batch;
      hipLaunchKernelGGL(k_gather_rows, dim3(B), dim3(64), 0, s, dX, dT, dO + (size_t)ep * n + s0, M->act,
                         M->tb, B, M->n_in);
      forward(M, s, B);
      backward(M, s, B, M->tb);
      ++M->step;
      const float c1 = 1.0f - std::pow(b1, (float)M->step), c2 = 1.0f - std::pow(b2, (float)M->step);
      hipLaunchKernelGGL(k_adam, dim3((unsigned)((M->n_params + 255) / 256)), dim3(256), 0, s, M->p, M->m, M->v,
                         M->g, M->p16, (int)M->n_params, lr, b1, b2, eps, c1, c2);
      MLP_HIP(hipMemcpyAsync(hl.data(), M->loss, sizeof(float), hipMemcpyDeviceToHost, s));
      MLP_HIP(hipStreamSynchronize(s));
      ep_loss += hl[0];
      ++nb;
    }
    MLP_HIP(hipGetLastError());
    if (losses) losses[ep] = nb ? (float)(ep_loss / nb) : 0.0f;
  }
  return ENT_OK;
}

}  // extern "C"
