// reductions.hip -- FP64 reductions of density arrays into per-result scalars.
//
// One "item" reduces one formula over k in [0, n) of up to four density
// arrays (device doubles).  Two deterministic stages: fixed-size partial sums
// per (item, block), then one ordered sum per item.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include "reductions.hpp"

namespace ent {
namespace {

constexpr int kRBlock = 256;

__device__ __forceinline__ double inv0(double x) { return x == 0.0 ? x : 1.0 / x; }  // tools.hxx:45-53

// One term of formula OP (FP64 throughout; log of 0 -> 0 as select_log
// does, tools.cpp:18-37).
template <uint32_t OP>
__device__ __forceinline__ double term(const RedItem& it, uint64_t k, const double* __restrict__ D) {
  const double a = D[it.off[0] + k];
  if constexpr (OP == kRedMI) {  // mi.cpp:53-56: p_ij * log(p_ij / p_i / p_j), one division
    const double r = a / (D[it.off[1] + k] * D[it.off[2] + k]);
    return r == 0.0 ? 0.0 : a * log(r) * it.scale;
  } else if constexpr (OP == kRedLogMean) {  // sum of ln p (entropy = -mean)
    return log(a);
  } else if constexpr (OP == kRedTE) {  // transs.cpp:77-100: p3*pref*log(p3*p1*inv(p2)*inv(ptau))
    if (!(a > 0.0)) return 0.0;
    // inv(p2) * inv(ptau) as one division (0 when either is 0, as inv does)
    const double den = D[it.off[2] + k] * D[it.off[3] + k];
    const double r = den == 0.0 ? 0.0 : a * D[it.off[1] + k] / den;
    if (r == 0.0) return 0.0;
    const double te = a * it.pref * log(r) * it.scale;
    return (isfinite(te) && a <= 1.0) ? te : 0.0;
  } else if constexpr (OP == kRedAutoEnt) {  // autoent.cpp:52-56: p2*h*h*log(p1*inv(p2*h)), p2 = a
    if (!(a > 0.0)) return 0.0;
    const double r = D[it.off[1] + k] * inv0(a * it.pref);
    return r == 0.0 ? 0.0 : a * it.pref * it.pref * log(r) * it.scale;
  } else if constexpr (OP == kRedLoo) {  // p = density incl. self pair, divided by n (n_rows == n)
    const double nd = (double)it.n;
    const double loo = (nd * a - it.pref) / (nd - 1.0);
    return log(loo > it.scale ? loo : it.scale);
  } else if constexpr (OP == kRedSum) {
    return a * it.scale;
  } else if constexpr (OP == kRedKL) {  // select_log: log of a zero ratio is 0
    const double qa = a * it.pref;
    const double r = qa / (D[it.off[1] + k] * it.pref2);
    return r == 0.0 ? 0.0 : qa * log(r) * it.scale;
  } else {  // kRedNegs
    const double q = a * it.pref;
    return q == 0.0 ? 0.0 : log(q) * q * it.scale;
  }
}

// Terms k0 + t, k0 + t + kRBlock, ... below k1 of thread t, summed in that
// order; four terms at a time are computed before they are added, so their
// loads are in flight together.
template <uint32_t OP>
__device__ __forceinline__ double sum_terms(const RedItem& it, const double* __restrict__ D, uint64_t k0,
                                            uint64_t k1) {
  double acc = 0.0;
  uint64_t k = k0 + threadIdx.x;
  for (; k + 3 * (uint64_t)kRBlock < k1; k += 4 * (uint64_t)kRBlock) {
    const double t0 = term<OP>(it, k, D);
    const double t1 = term<OP>(it, k + kRBlock, D);
    const double t2 = term<OP>(it, k + 2 * kRBlock, D);
    const double t3 = term<OP>(it, k + 3 * kRBlock, D);
    acc += t0;
    acc += t1;
    acc += t2;
    acc += t3;
  }
  for (; k < k1; k += kRBlock) acc += term<OP>(it, k, D);
  return acc;
}

__global__ __launch_bounds__(kRBlock) void k_reduce_partial(const RedItem* __restrict__ items,
                                                            const double* __restrict__ D,
                                                            double* __restrict__ partial,
                                                            uint32_t blocks_per_item) {
  __shared__ double s[kRBlock];
  const uint32_t item = blockIdx.y;
  const RedItem it = items[item];
  const uint64_t n = it.n;
  const uint64_t per = (n + blocks_per_item - 1) / blocks_per_item;
  const uint64_t k0 = (uint64_t)blockIdx.x * per;
  const uint64_t k1 = k0 + per < n ? k0 + per : n;
  double acc = 0.0;
  switch (it.op) {  // uniform over the workgroup
    case kRedMI: acc = sum_terms<kRedMI>(it, D, k0, k1); break;
    case kRedLogMean: acc = sum_terms<kRedLogMean>(it, D, k0, k1); break;
    case kRedTE: acc = sum_terms<kRedTE>(it, D, k0, k1); break;
    case kRedAutoEnt: acc = sum_terms<kRedAutoEnt>(it, D, k0, k1); break;
    case kRedLoo: acc = sum_terms<kRedLoo>(it, D, k0, k1); break;
    case kRedSum: acc = sum_terms<kRedSum>(it, D, k0, k1); break;
    case kRedKL: acc = sum_terms<kRedKL>(it, D, k0, k1); break;
    case kRedNegs: acc = sum_terms<kRedNegs>(it, D, k0, k1); break;
    default: break;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kRBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(uint64_t)item * blocks_per_item + blockIdx.x] = s[0];
}

__global__ void k_reduce_final(const double* __restrict__ partial, uint32_t blocks_per_item,
                               uint32_t n_items, double* __restrict__ out) {
  const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= n_items) return;
  double acc = 0.0;
  for (uint32_t b = 0; b < blocks_per_item; ++b) acc += partial[(uint64_t)item * blocks_per_item + b];
  out[item] = acc;
}

// Column sums (mean == nullptr) or centred square sums of X, one partial per
// (column, block); blockIdx.y = column.
__global__ __launch_bounds__(kRBlock) void k_col_partial(const float* __restrict__ X, uint64_t n_rows,
                                                         const double* __restrict__ mean,
                                                         double* __restrict__ partial,
                                                         uint32_t blocks_per_col) {
  __shared__ double s[kRBlock];
  const uint32_t col = blockIdx.y;
  const float* x = X + (uint64_t)col * n_rows;
  const uint64_t per = (n_rows + blocks_per_col - 1) / blocks_per_col;
  const uint64_t k0 = (uint64_t)blockIdx.x * per;
  const uint64_t k1 = k0 + per < n_rows ? k0 + per : n_rows;
  const double m = mean ? mean[col] : 0.0;
  double acc = 0.0;
  for (uint64_t k = k0 + threadIdx.x; k < k1; k += kRBlock) {
    const double v = (double)x[k] - m;
    acc += mean ? v * v : v;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kRBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(uint64_t)col * blocks_per_col + blockIdx.x] = s[0];
}

__global__ void k_col_final(const double* __restrict__ partial, uint32_t blocks_per_col, uint32_t n_cols,
                            double inv_n, double* __restrict__ out) {
  const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n_cols) return;
  double acc = 0.0;
  for (uint32_t b = 0; b < blocks_per_col; ++b) acc += partial[(uint64_t)col * blocks_per_col + b];
  out[col] = acc * inv_n;
}

}  // namespace

void launch_col_moments(hipStream_t s, const float* X, uint64_t n_rows, uint32_t n_cols, double* partial,
                        uint32_t blocks_per_col, double* mean, double* var) {
  if (n_cols == 0 || n_rows == 0) return;
  const double inv_n = 1.0 / (double)n_rows;
  const dim3 grid(blocks_per_col, n_cols);
  const dim3 fin((n_cols + 255) / 256);
  hipLaunchKernelGGL(k_col_partial, grid, dim3(kRBlock), 0, s, X, n_rows, (const double*)nullptr, partial,
                     blocks_per_col);
  hipLaunchKernelGGL(k_col_final, fin, dim3(256), 0, s, partial, blocks_per_col, n_cols, inv_n, mean);
  hipLaunchKernelGGL(k_col_partial, grid, dim3(kRBlock), 0, s, X, n_rows, (const double*)mean, partial,
                     blocks_per_col);
  hipLaunchKernelGGL(k_col_final, fin, dim3(256), 0, s, partial, blocks_per_col, n_cols, inv_n, var);
}

void launch_reduce(hipStream_t s, const RedItem* items, uint32_t n_items, const double* D,
                   double* partial, uint32_t blocks_per_item, double* out) {
  if (n_items == 0) return;
  hipLaunchKernelGGL(k_reduce_partial, dim3(blocks_per_item, n_items), dim3(kRBlock), 0, s, items, D,
                     partial, blocks_per_item);
  hipLaunchKernelGGL(k_reduce_final, dim3((n_items + 255) / 256), dim3(256), 0, s, partial,
                     blocks_per_item, n_items, out);
}

}  // namespace ent
