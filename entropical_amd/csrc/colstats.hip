// colstats.hip -- per-column sample statistics of the learned bandwidth's
// features, computed on the device for every column of a column-major
// float32 array X[col * n_rows + row]:
//   mean and population variance (FP64, two deterministic passes),
//   the order statistics at 18 given ranks (exact, by a 4-pass radix select
//   over the columns' float bit patterns: per pass, a 256-bin histogram of
//   the next 8 bits among the values that share the bits found so far),
//   and, with z = (x - mean) / sd: the sums of z^3, z^4 and |z|, the count of
//   |z| > 3 and the 32-bin histogram of z over [-4, 4] (bin edges -4 + k/4,
//   the last bin closed on the right, values outside not counted).
// Integer counts use atomics (exact); FP64 sums are per-block partials added
// in a fixed order, so results are deterministic.
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
#include "kde_kernels.hpp"
#include "reductions.hpp"

namespace {

using ent::set_error;

constexpr int kBlk = 256;
constexpr int kRanks = ENT_COLSTATS_RANKS;  // order statistics per column
constexpr int kHist = 32;                   // z histogram bins

// float -> uint32 whose unsigned order is the float order
__device__ __forceinline__ uint32_t f2o(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) {
  const uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __uint_as_float(u);
}

// Digit histogram of one radix-select pass.  sel[c * R + t] = (prefix, k):
// the bits found so far of rank t's key and the rank within that prefix.
// hist[(c * R + t) * 256 + digit] (cleared by the caller).
__global__ __launch_bounds__(kBlk) void k_select_hist(const float* __restrict__ X, uint64_t n_rows,
                                                      const uint2* __restrict__ sel, int shift,
                                                      uint32_t* __restrict__ hist, uint32_t per_block) {
  __shared__ uint32_t h[kRanks][256];
  __shared__ uint32_t pre[kRanks];
  const uint32_t c = blockIdx.y;
  for (int i = threadIdx.x; i < kRanks * 256; i += kBlk) (&h[0][0])[i] = 0u;
  if (threadIdx.x < kRanks) pre[threadIdx.x] = sel[c * kRanks + threadIdx.x].x;
  __syncthreads();
  const float* x = X + (uint64_t)c * n_rows;
  const uint64_t k0 = (uint64_t)blockIdx.x * per_block;
  const uint64_t k1 = k0 + per_block < n_rows ? k0 + per_block : n_rows;
  // the bits above this pass's digit must equal the rank's prefix
  const uint32_t hi_mask = shift >= 24 ? 0u : ~((1u << (shift + 8)) - 1u);
  for (uint64_t k = k0 + threadIdx.x; k < k1; k += kBlk) {
    const uint32_t key = f2o(x[k]);
    const uint32_t dig = (key >> shift) & 255u;
#pragma unroll 6
    for (int t = 0; t < kRanks; ++t)
      if (((key ^ pre[t]) & hi_mask) == 0u) atomicAdd(&h[t][dig], 1u);
  }
  __syncthreads();
  uint32_t* hc = hist + (uint64_t)c * kRanks * 256;
  for (int i = threadIdx.x; i < kRanks * 256; i += kBlk) {
    const uint32_t v = (&h[0][0])[i];
    if (v) atomicAdd(hc + i, v);
  }
}

// One thread per (column, rank): the digit whose cumulative count passes the
// rank's remaining rank k joins its prefix; k becomes the rank within it.
__global__ void k_select_step(const uint32_t* __restrict__ hist, uint2* __restrict__ sel, int shift,
                              uint32_t n_items) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_items) return;
  const uint32_t* h = hist + (uint64_t)i * 256;
  uint2 s = sel[i];
  uint32_t cum = 0, d = 0;
  for (; d < 255u; ++d) {
    if (cum + h[d] > s.y) break;
    cum += h[d];
  }
  s.x |= d << shift;
  s.y -= cum;
  sel[i] = s;
}

// z sums, tail count and histogram per (column, block).  part: 3 doubles
// per (column, block); counts: per column [tail, hist[32]] (atomics).
__global__ __launch_bounds__(kBlk) void k_z_stats(const float* __restrict__ X, uint64_t n_rows,
                                                  const double* __restrict__ mean, const double* __restrict__ var,
                                                  double* __restrict__ part, uint32_t* __restrict__ counts,
                                                  uint32_t per_block) {
  __shared__ double s[3][kBlk];
  __shared__ uint32_t hc[kHist + 1];
  const uint32_t c = blockIdx.y;
  if (threadIdx.x <= kHist) hc[threadIdx.x] = 0u;
  __syncthreads();
  const float* x = X + (uint64_t)c * n_rows;
  const double mu = mean[c];
  double sd = sqrt(var[c]);
  if (sd == 0.0) sd = 1.0;
  const uint64_t k0 = (uint64_t)blockIdx.x * per_block;
  const uint64_t k1 = k0 + per_block < n_rows ? k0 + per_block : n_rows;
  double a3 = 0.0, a4 = 0.0, aa = 0.0;
  uint32_t tail = 0;
  for (uint64_t k = k0 + threadIdx.x; k < k1; k += kBlk) {
    const double z = ((double)x[k] - mu) / sd;
    const double z2 = z * z;
    a3 += z2 * z;
    a4 += z2 * z2;
    aa += fabs(z);
    tail += fabs(z) > 3.0 ? 1u : 0u;
    if (z >= -4.0 && z <= 4.0) {  // bins [-4 + b/4, -4 + (b+1)/4), the last closed
      int b = (int)((z + 4.0) * 4.0);
      if (b > kHist - 1) b = kHist - 1;
      if (z < -4.0 + 0.25 * b) --b;
      else if (b < kHist - 1 && z >= -4.0 + 0.25 * (b + 1)) ++b;
      atomicAdd(&hc[b], 1u);
    }
  }
  s[0][threadIdx.x] = a3;
  s[1][threadIdx.x] = a4;
  s[2][threadIdx.x] = aa;
  if (tail) atomicAdd(&hc[kHist], tail);
  __syncthreads();
  for (int o = kBlk / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int q = 0; q < 3; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[((uint64_t)c * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = s[threadIdx.x][0];
  if (threadIdx.x <= kHist && hc[threadIdx.x]) {
    // counts[c][0] = tail, counts[c][1 + b] = bin b
    const uint32_t slot = threadIdx.x == kHist ? 0u : threadIdx.x + 1u;
    atomicAdd(counts + (uint64_t)c * (kHist + 1) + slot, hc[threadIdx.x]);
  }
}

// Per column: the z sums in block order, and the order statistics' values.
__global__ void k_colstats_final(const double* __restrict__ part, uint32_t blocks, const uint2* __restrict__ sel,
                                 const uint32_t* __restrict__ counts, const double* __restrict__ mean,
                                 const double* __restrict__ var, uint32_t n_cols, double* __restrict__ out) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  double* o = out + (uint64_t)c * ENT_COLSTATS_WIDTH;
  o[0] = mean[c];
  o[1] = var[c];
  for (int t = 0; t < kRanks; ++t) o[2 + t] = (double)o2f(sel[c * kRanks + t].x);
  double a3 = 0.0, a4 = 0.0, aa = 0.0;
  for (uint32_t b = 0; b < blocks; ++b) {
    const double* p = part + ((uint64_t)c * blocks + b) * 3;
    a3 += p[0];
    a4 += p[1];
    aa += p[2];
  }
  o[2 + kRanks] = a3;
  o[3 + kRanks] = a4;
  o[4 + kRanks] = aa;
  const uint32_t* cc = counts + (uint64_t)c * (kHist + 1);
  for (int b = 0; b <= kHist; ++b) o[5 + kRanks + b] = (double)cc[b];
}

}  // namespace

extern "C" {

int ent_column_stats_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const uint32_t* ranks,
                         double* out) {
  if (!ctx || !ctx->eng || !X_dev || !ranks || !out || n_rows == 0 || n_cols == 0) {
    set_error("null argument or empty input");
    return ENT_ERR_ARG;
  }
  if (n_rows >= (1ull << 32) || n_cols > 65535) {
    set_error("column statistics: at most 2^32 - 1 rows and 65535 columns");
    return ENT_ERR_UNSUPPORTED;
  }
  for (int t = 0; t < kRanks; ++t)
    if (ranks[t] >= n_rows) {
      set_error("column statistics: rank beyond the column length");
      return ENT_ERR_ARG;
    }
  ent::Engine& E = *ctx->eng;
  hipStream_t s = E.stream();
  const uint32_t C = (uint32_t)n_cols;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(64, (n_rows + 8191) / 8192);
  const uint32_t per = (uint32_t)((n_rows + blocks - 1) / blocks);
  // device layout: sel (C*R uint2) | hist (C*R*256 u32) | counts (C*33 u32) |
  //                mean, var (C doubles each) | part (C*blocks*3 doubles,
  //                also the moments' partials) | out (C*WIDTH doubles)
  const size_t n_sel = (size_t)C * kRanks;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_sel = take(sizeof(uint2) * n_sel), o_hist = take(sizeof(uint32_t) * n_sel * 256),
               o_cnt = take(sizeof(uint32_t) * C * (kHist + 1)), o_mean = take(sizeof(double) * C),
               o_var = take(sizeof(double) * C), o_part = take(sizeof(double) * C * blocks * 3),
               o_out = take(sizeof(double) * C * ENT_COLSTATS_WIDTH);
  if (E.ensure(ctx->aux, off) != ent::kOk) return ENT_ERR_OOM;
  char* base = (char*)ctx->aux.p;
  uint2* sel = (uint2*)(base + o_sel);
  uint32_t* hist = (uint32_t*)(base + o_hist);
  uint32_t* cnt = (uint32_t*)(base + o_cnt);
  double* mean = (double*)(base + o_mean);
  double* var = (double*)(base + o_var);
  double* part = (double*)(base + o_part);
  double* dout = (double*)(base + o_out);
  // initial selection state: prefix 0, rank k
  std::vector<uint2> h_sel(n_sel);
  for (uint32_t c = 0; c < C; ++c)
    for (int t = 0; t < kRanks; ++t) h_sel[(size_t)c * kRanks + t] = make_uint2(0u, ranks[t]);
  if (E.ensure_host(sizeof(uint2) * n_sel) != ent::kOk) return ENT_ERR_OOM;
  std::memcpy(E.host_stage(), h_sel.data(), sizeof(uint2) * n_sel);
#define CS_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call);    \
      return e_ == hipErrorOutOfMemory ? ENT_ERR_OOM : ENT_ERR_HIP;                    \
    }                                                                                  \
  } while (0)
  CS_HIP(hipSetDevice(E.device()));
  CS_HIP(hipMemcpyAsync(sel, E.host_stage(), sizeof(uint2) * n_sel, hipMemcpyHostToDevice, s));
  ent::launch_col_moments(s, X_dev, n_rows, C, part, blocks, mean, var);
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 24 - 8 * pass;
    ent::launch_fill_u32(s, hist, 0u, n_sel * 256);
    hipLaunchKernelGGL(k_select_hist, dim3(blocks, C), dim3(kBlk), 0, s, X_dev, (uint64_t)n_rows, sel, shift, hist,
                       per);
    hipLaunchKernelGGL(k_select_step, dim3((unsigned)((n_sel + 255) / 256)), dim3(256), 0, s, hist, sel, shift,
                       (uint32_t)n_sel);
  }
  ent::launch_fill_u32(s, cnt, 0u, (uint64_t)C * (kHist + 1));
  hipLaunchKernelGGL(k_z_stats, dim3(blocks, C), dim3(kBlk), 0, s, X_dev, (uint64_t)n_rows, mean, var, part, cnt,
                     per);
  hipLaunchKernelGGL(k_colstats_final, dim3((C + 255) / 256), dim3(256), 0, s, part, blocks, sel, cnt, mean, var, C,
                     dout);
  CS_HIP(hipGetLastError());
  CS_HIP(hipMemcpyAsync(out, dout, sizeof(double) * C * ENT_COLSTATS_WIDTH, hipMemcpyDeviceToHost, s));
  CS_HIP(hipStreamSynchronize(s));
#undef CS_HIP
  return ENT_OK;
}

}  // extern "C"
