// drivers.hip -- batched MI / transfer-entropy / entropy drivers (C ABI).
//
// Each driver builds the full job table of its reference driver loop, runs it
// as ONE engine batch (one persistent density launch) and reduces the
// density arrays on the device in FP64; only the per-result scalars return
// to the host.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/entropical_kde.h"
#include "ctx.hpp"
#include "engine.hpp"
#include "reductions.hpp"

namespace {

using ent::set_error;

#define DRV_HIP(call)                                                                \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call);  \
      return e_ == hipErrorOutOfMemory ? ENT_ERR_OOM : ENT_ERR_HIP;                  \
    }                                                                                \
  } while (0)

#define DRV_TRY(expr)            \
  do {                           \
    int r_ = (expr);             \
    if (r_ != ENT_OK) return r_; \
  } while (0)

int check_common(ent_ctx* ctx, const float* X, size_t n_rows, size_t n_cols, const float* h) {
  if (!ctx || !ctx->eng) {
    set_error("null context");
    return ENT_ERR_ARG;
  }
  if (!X || !h || n_rows == 0 || n_cols == 0) {
    set_error("null argument or empty input");
    return ENT_ERR_ARG;
  }
  if (n_rows >= (1ull << 32)) {
    set_error("more than 2^32 rows");
    return ENT_ERR_UNSUPPORTED;
  }
  for (size_t c = 0; c < n_cols; ++c)
    if (!(h[c] > 0.0f) || !std::isfinite(h[c])) {
      set_error("error: bandwidths must be finite and positive");
      return ENT_ERR_ARG;
    }
  return ENT_OK;
}

ent::JobSpec make_job(size_t n_rows, std::initializer_list<uint32_t> cols, std::initializer_list<float> hs,
                      std::initializer_list<uint32_t> taus, uint64_t out_off) {
  ent::JobSpec j;
  j.d = (uint32_t)cols.size();
  uint32_t tmax = 0;
  for (uint32_t t : taus) tmax = t > tmax ? t : tmax;
  j.n = (uint32_t)(n_rows - tmax);
  j.n_rows = (uint32_t)n_rows;
  j.out_off = out_off;
  auto ci = cols.begin();
  auto hi = hs.begin();
  auto ti = taus.begin();
  for (uint32_t e = 0; e < j.d; ++e, ++ci, ++hi, ++ti) {
    j.off[e] = (uint64_t)(*ci) * n_rows + *ti;
    j.h[e] = *hi;
  }
  return j;
}

// Reduction items over the densities already in ctx->out; results (one
// double per item) are copied to host `res`.
int reduce_only(ent_ctx* ctx, const std::vector<ent::RedItem>& items, std::vector<double>& res) {
  ent::Engine& E = *ctx->eng;
  const uint32_t NI = (uint32_t)items.size();
  uint64_t maxn = 1;
  for (const auto& it : items) maxn = std::max<uint64_t>(maxn, it.n);
  // blocks per item: one per 16384 terms, and for batches of few items up to
  // 1024 blocks in all (at least 1024 terms each), so a single long item
  // does not run on a handful of workgroups; at most 64
  uint64_t want = std::max<uint64_t>(maxn / 16384, (1024 + NI - 1) / std::max<uint32_t>(NI, 1));
  want = std::min<uint64_t>(want, std::max<uint64_t>(1, maxn / 1024));
  uint32_t bpi = (uint32_t)std::min<uint64_t>(64, std::max<uint64_t>(1, want));
  size_t items_bytes = (sizeof(ent::RedItem) * NI + 255) / 256 * 256;
  size_t part_bytes = sizeof(double) * (size_t)NI * bpi;
  DRV_TRY(E.ensure(ctx->red, items_bytes + part_bytes + sizeof(double) * NI));
  char* base = (char*)ctx->red.p;
  ent::RedItem* d_items = (ent::RedItem*)base;
  double* d_part = (double*)(base + items_bytes);
  double* d_out = (double*)(base + items_bytes + part_bytes);
  DRV_TRY(E.ensure_host(sizeof(ent::RedItem) * NI + sizeof(double) * NI));
  char* hs = (char*)E.host_stage();
  std::memcpy(hs, items.data(), sizeof(ent::RedItem) * NI);
  double* h_out = (double*)(hs + sizeof(ent::RedItem) * NI);
  DRV_HIP(hipMemcpyAsync(d_items, hs, sizeof(ent::RedItem) * NI, hipMemcpyHostToDevice, E.stream()));
  ent::launch_reduce(E.stream(), d_items, NI, (const double*)ctx->out.p, d_part, bpi, d_out);
  DRV_HIP(hipGetLastError());
  DRV_HIP(hipMemcpyAsync(h_out, d_out, sizeof(double) * NI, hipMemcpyDeviceToHost, E.stream()));
  DRV_HIP(hipStreamSynchronize(E.stream()));
  res.assign(h_out, h_out + NI);
  return ENT_OK;
}

// Densities of all jobs into ctx->out, then the reduction items.
int run_and_reduce(ent_ctx* ctx, const float* X, const std::vector<ent::JobSpec>& jobs, int method,
                   const std::vector<ent::RedItem>& items, std::vector<double>& res) {
  ent::Engine& E = *ctx->eng;
  uint64_t total = 0;
  for (const auto& j : jobs) total = std::max<uint64_t>(total, j.out_off + j.n);
  DRV_TRY(E.ensure(ctx->out, sizeof(double) * total));
  DRV_TRY(E.densities(X, jobs, (double*)ctx->out.p, method));
  return reduce_only(ctx, items, res);
}

ent::RedItem item(uint32_t op, std::initializer_list<uint64_t> offs, uint64_t n, double pref, double scale) {
  ent::RedItem it;
  std::memset(&it, 0, sizeof(it));
  it.op = op;
  int e = 0;
  for (uint64_t o : offs) it.off[e++] = o;
  it.n = n;
  it.pref = pref;
  it.scale = scale;
  return it;
}

// mi.cpp:37-52 job set: 2-D joints first (the larger ones), then the 1-D
// marginals once per column (mi.cpp:43-46 recomputes p_j for every pair).
void build_mi(size_t N, size_t C, const float* h, int bits, std::vector<ent::JobSpec>& jobs,
              std::vector<ent::RedItem>& items) {
  const double sc = bits ? 1.0 / std::log(2.0) : 1.0;
  uint64_t slot = C;
  for (uint32_t i = 0; i < C; ++i)
    for (uint32_t j = i + 1; j < C; ++j) {
      jobs.push_back(make_job(N, {i, j}, {h[i], h[j]}, {0, 0}, slot * N));
      items.push_back(item(ent::kRedMI, {slot * N, (uint64_t)i * N, (uint64_t)j * N}, N, 0.0, sc));
      ++slot;
    }
  for (uint32_t c = 0; c < C; ++c) jobs.push_back(make_job(N, {c}, {h[c]}, {0}, (uint64_t)c * N));
}

int check_taus(size_t n_rows, const uint32_t* taus, size_t n_taus) {
  for (size_t t = 0; t < n_taus; ++t)
    if (taus[t] >= n_rows) {
      set_error("error: lag time (tau) must be smaller than the number of rows");
      return ENT_ERR_ARG;
    }
  return ENT_OK;
}

// transs.cpp:42-74 job set for every lag (3-D jobs first), and the
// reduction items of transs.cpp:75-113 (both directions per pair).
void build_te(size_t N, size_t C, const float* h, const uint32_t* taus, size_t n_taus, int bits,
              std::vector<ent::JobSpec>& jobs, std::vector<ent::RedItem>& items) {
  const double sc = bits ? 1.0 / std::log(2.0) : 1.0;
  uint64_t next = 0;
  auto alloc = [&](uint64_t n) {
    uint64_t o = next;
    next += n;
    return o;
  };
  std::vector<uint64_t> off_p(C), off_pij(C * C, 0);
  std::vector<std::vector<uint64_t>> off_ptau(n_taus, std::vector<uint64_t>(C));
  std::vector<std::vector<uint64_t>> off_itau(n_taus, std::vector<uint64_t>(C * C)), off_jtau = off_itau;
  for (size_t t = 0; t < n_taus; ++t) {
    const uint32_t tau = taus[t];
    for (uint32_t i = 0; i < C; ++i)
      for (uint32_t j = i + 1; j < C; ++j) {
        off_itau[t][i * C + j] = alloc(N - tau);
        jobs.push_back(make_job(N, {i, i, j}, {h[i], h[i], h[j]}, {0, tau, 0}, off_itau[t][i * C + j]));
        // p(x_t, y_t, y_t+tau) with its dims listed as (y_t, y_t+tau, x_t): the
        // same density, and its dims 0 and 1 then match those of the other
        // jobs of series j and lag tau, so they share one two-dim sort
        off_jtau[t][i * C + j] = alloc(N - tau);
        jobs.push_back(make_job(N, {j, j, i}, {h[j], h[j], h[i]}, {0, tau, 0}, off_jtau[t][i * C + j]));
      }
  }
  for (uint32_t i = 0; i < C; ++i)
    for (uint32_t j = i + 1; j < C; ++j) {
      off_pij[i * C + j] = alloc(N);
      jobs.push_back(make_job(N, {i, j}, {h[i], h[j]}, {0, 0}, off_pij[i * C + j]));
    }
  for (size_t t = 0; t < n_taus; ++t)
    for (uint32_t i = 0; i < C; ++i) {
      off_ptau[t][i] = alloc(N - taus[t]);
      jobs.push_back(make_job(N, {i, i}, {h[i], h[i]}, {0, taus[t]}, off_ptau[t][i]));
    }
  for (uint32_t i = 0; i < C; ++i) {
    off_p[i] = alloc(N);
    jobs.push_back(make_job(N, {i}, {h[i]}, {0}, off_p[i]));
  }
  for (size_t t = 0; t < n_taus; ++t) {
    const uint64_t n = N - taus[t];
    for (uint32_t i = 0; i < C; ++i)
      for (uint32_t j = i + 1; j < C; ++j) {
        const double hi = h[i], hj = h[j];
        items.push_back(item(ent::kRedTE, {off_itau[t][i * C + j], off_p[i], off_pij[i * C + j], off_ptau[t][i]},
                             n, hi * hi * hj, sc));
        items.push_back(item(ent::kRedTE, {off_jtau[t][i * C + j], off_p[j], off_pij[i * C + j], off_ptau[t][j]},
                             n, hi * hj * hi, sc));
      }
  }
}

// Runs the engine's autotuner on a job set, with ctx->out as scratch output.
int tune_jobs(ent_ctx* ctx, const float* X, const std::vector<ent::JobSpec>& jobs, int* out5, double* best_ms) {
  ent::Engine& E = *ctx->eng;
  uint64_t total = 0;
  for (const auto& j : jobs) total = std::max<uint64_t>(total, j.out_off + j.n);
  DRV_TRY(E.ensure(ctx->out, sizeof(double) * total));
  ent::Tuning best;
  double ms = 0.0;
  DRV_TRY(E.autotune(X, jobs, (double*)ctx->out.p, &best, &ms));
  if (out5) {
    out5[0] = best.q;
    out5[1] = best.m0;
    out5[2] = best.mr;
    out5[3] = best.group_h;
    out5[4] = best.sub;
  }
  if (best_ms) *best_ms = ms;
  return ENT_OK;
}

}  // namespace

extern "C" {

int ent_mi_autotune_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h,
                        int* tuning5) {
  DRV_TRY(check_common(ctx, X_dev, n_rows, n_cols, h));
  std::vector<ent::JobSpec> jobs;
  std::vector<ent::RedItem> items;
  build_mi(n_rows, n_cols, h, 0, jobs, items);
  return tune_jobs(ctx, X_dev, jobs, tuning5, nullptr);
}

int ent_te_autotune_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h,
                        const uint32_t* taus, size_t n_taus, int* tuning5) {
  DRV_TRY(check_common(ctx, X_dev, n_rows, n_cols, h));
  if (!taus || n_taus == 0) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  DRV_TRY(check_taus(n_rows, taus, n_taus));
  std::vector<ent::JobSpec> jobs;
  std::vector<ent::RedItem> items;
  build_te(n_rows, n_cols, h, taus, n_taus, 0, jobs, items);
  return tune_jobs(ctx, X_dev, jobs, tuning5, nullptr);
}

int ent_autotune_dev(ent_ctx* ctx, const float* X_dev, const ent_job* jobs, size_t n_jobs, int* tuning5,
                     double* best_ms) {
  if (!ctx || !ctx->eng || !X_dev || !jobs || n_jobs == 0) {
    set_error("null argument or empty job list");
    return ENT_ERR_ARG;
  }
  std::vector<ent::JobSpec> js(n_jobs);
  uint64_t next = 0;
  for (size_t j = 0; j < n_jobs; ++j) {
    const ent_job& a = jobs[j];
    if (a.n_dim < 1 || a.n_dim > 3 || a.n == 0 || a.n_rows == 0) {
      set_error("invalid job: n_dim must be 1..3 and n, n_rows > 0");
      return ENT_ERR_ARG;
    }
    js[j].d = a.n_dim;
    js[j].n = a.n;
    js[j].n_rows = a.n_rows;
    js[j].out_off = next;  // tuning output is scratch: pack the jobs
    next += a.n;
    for (uint32_t e = 0; e < 3; ++e) {
      js[j].off[e] = a.off[e];
      js[j].h[e] = e < a.n_dim ? a.h[e] : 1.0f;
    }
  }
  return tune_jobs(ctx, X_dev, js, tuning5, best_ms);
}

int ent_mi_matrix_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h,
                      int bits, int method, double* mi) {
  DRV_TRY(check_common(ctx, X_dev, n_rows, n_cols, h));
  if (!mi) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  auto t0 = std::chrono::steady_clock::now();
  const size_t C = n_cols;
  std::vector<ent::JobSpec> jobs;
  std::vector<ent::RedItem> items;
  build_mi(n_rows, n_cols, h, bits, jobs, items);
  std::vector<double> res;
  DRV_TRY(run_and_reduce(ctx, X_dev, jobs, method, items, res));
  size_t p = 0;
  for (size_t i = 0; i < C; ++i) mi[i * C + i] = 0.0;
  for (size_t i = 0; i < C; ++i)
    for (size_t j = i + 1; j < C; ++j, ++p) {
      mi[i * C + j] = res[p];
      mi[j * C + i] = res[p];
    }
  ctx->driver_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ENT_OK;
}

int ent_te_matrix_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h,
                      const uint32_t* taus, size_t n_taus, int bits, int method, double* te) {
  DRV_TRY(check_common(ctx, X_dev, n_rows, n_cols, h));
  if (!te || !taus || n_taus == 0) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  DRV_TRY(check_taus(n_rows, taus, n_taus));
  auto t0 = std::chrono::steady_clock::now();
  const size_t C = n_cols;
  std::vector<ent::JobSpec> jobs;
  std::vector<ent::RedItem> items;
  build_te(n_rows, n_cols, h, taus, n_taus, bits, jobs, items);
  std::vector<double> res;
  DRV_TRY(run_and_reduce(ctx, X_dev, jobs, method, items, res));
  size_t p = 0;
  for (size_t t = 0; t < n_taus; ++t) {
    double* M = te + t * C * C;
    for (size_t k = 0; k < C * C; ++k) M[k] = 0.0;
    for (size_t i = 0; i < C; ++i)
      for (size_t j = i + 1; j < C; ++j) {
        M[i * C + j] = res[p++];
        M[j * C + i] = res[p++];
      }
  }
  ctx->driver_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ENT_OK;
}

int ent_batch_entropy_dev(ent_ctx* ctx, const float* X_dev, const ent_job* jobs, size_t n_jobs, int method,
                          double* H) {
  if (!ctx || !ctx->eng || !X_dev || !jobs || !H) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  auto t0 = std::chrono::steady_clock::now();
  std::vector<ent::JobSpec> js(n_jobs);
  std::vector<ent::RedItem> items(n_jobs);
  uint64_t next = 0;
  for (size_t j = 0; j < n_jobs; ++j) {
    const ent_job& a = jobs[j];
    if (a.n_dim < 1 || a.n_dim > 3 || a.n == 0 || a.n_rows == 0) {
      set_error("invalid job: n_dim must be 1..3 and n, n_rows > 0");
      return ENT_ERR_ARG;
    }
    js[j].d = a.n_dim;
    js[j].n = a.n;
    js[j].n_rows = a.n_rows;
    js[j].out_off = next;
    for (uint32_t e = 0; e < 3; ++e) {
      js[j].off[e] = a.off[e];
      js[j].h[e] = e < a.n_dim ? a.h[e] : 1.0f;
    }
    items[j] = item(ent::kRedLogMean, {next}, a.n, 0.0, 1.0);
    next += a.n;
  }
  std::vector<double> res;
  DRV_TRY(run_and_reduce(ctx, X_dev, js, method, items, res));
  for (size_t j = 0; j < n_jobs; ++j) H[j] = -res[j] / (double)jobs[j].n;
  ctx->driver_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ENT_OK;
}

int ent_batch_loo_dev(ent_ctx* ctx, const float* X_dev, const ent_job* jobs, size_t n_jobs, double floor_rel,
                      int method, double* loglik) {
  if (!ctx || !ctx->eng || !X_dev || !jobs || !loglik) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  if (!(floor_rel > 0.0)) {
    set_error("floor_rel must be > 0");
    return ENT_ERR_ARG;
  }
  std::vector<ent::JobSpec> js(n_jobs);
  std::vector<ent::RedItem> items(n_jobs);
  uint64_t next = 0;
  for (size_t j = 0; j < n_jobs; ++j) {
    const ent_job& a = jobs[j];
    if (a.n_dim < 1 || a.n_dim > 3 || a.n < 2) {
      set_error("invalid job: n_dim must be 1..3 and n >= 2");
      return ENT_ERR_ARG;
    }
    js[j].d = a.n_dim;
    js[j].n = a.n;
    js[j].n_rows = a.n;  // p = S * K0 / n, so n p - K0 = (S - 1) K0
    js[j].out_off = next;
    double k0 = 1.0;
    for (uint32_t e = 0; e < 3; ++e) {
      js[j].off[e] = a.off[e];
      js[j].h[e] = e < a.n_dim ? a.h[e] : 1.0f;
      if (e < a.n_dim) k0 *= 0.75 / (double)a.h[e];
    }
    items[j] = item(ent::kRedLoo, {next}, a.n, k0, floor_rel * k0 / (double)(a.n - 1));
    next += a.n;
  }
  std::vector<double> res;
  DRV_TRY(run_and_reduce(ctx, X_dev, js, method, items, res));
  for (size_t j = 0; j < n_jobs; ++j) loglik[j] = res[j];
  return ENT_OK;
}

// AutoEnt::main (autoent.cpp:36-55): per column i, the lagged 1-D density
// p(x_{t+tau}) and the 2-D density p(x_t, x_{t+tau}), both over the
// n_rows - tau lagged rows; one batch for all columns.
int ent_autoent_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h,
                    uint32_t tau, int bits, int method, double* out) {
  DRV_TRY(check_common(ctx, X_dev, n_rows, n_cols, h));
  if (!out) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  DRV_TRY(check_taus(n_rows, &tau, 1));
  auto t0 = std::chrono::steady_clock::now();
  const double sc = bits ? 1.0 / std::log(2.0) : 1.0;
  const uint64_t m = n_rows - tau;
  std::vector<ent::JobSpec> jobs;
  std::vector<ent::RedItem> items;
  for (uint32_t i = 0; i < n_cols; ++i) {
    const uint64_t o2 = 2 * (uint64_t)i * m, o1 = o2 + m;
    jobs.push_back(make_job(n_rows, {i, i}, {h[i], h[i]}, {0, tau}, o2));
    jobs.push_back(make_job(n_rows, {i}, {h[i]}, {tau}, o1));
    items.push_back(item(ent::kRedAutoEnt, {o2, o1}, m, (double)h[i], sc));
  }
  std::vector<double> res;
  DRV_TRY(run_and_reduce(ctx, X_dev, jobs, method, items, res));
  for (size_t i = 0; i < n_cols; ++i) out[i] = res[i];
  ctx->driver_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ENT_OK;
}

namespace {

// 1-D densities of every column into ctx->out (column c at c*n_rows) and
// the reciprocal of each one's sum (Tools::sum1_normalized, tools.hxx:33-43).
int column_densities_normalised(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols,
                                const float* h, int method, std::vector<double>& inv_sum) {
  std::vector<ent::JobSpec> jobs;
  std::vector<ent::RedItem> items;
  for (uint32_t c = 0; c < n_cols; ++c) {
    jobs.push_back(make_job(n_rows, {c}, {h[c]}, {0}, (uint64_t)c * n_rows));
    items.push_back(item(ent::kRedSum, {(uint64_t)c * n_rows}, n_rows, 0.0, 1.0));
  }
  std::vector<double> sums;
  DRV_TRY(run_and_reduce(ctx, X_dev, jobs, method, items, sums));
  inv_sum.resize(n_cols);
  for (size_t c = 0; c < n_cols; ++c) inv_sum[c] = sums[c] != 0.0 ? 1.0 / sums[c] : 1.0;
  return ENT_OK;
}

}  // namespace

// KLDiv::main (kldiv.cpp:34-50): for every ordered column pair,
// sum_k p_i log(p_i / p_j) over sum1-normalised 1-D densities.  The
// densities are computed once per column (the reference recomputes p_j for
// every pair, kldiv.cpp:42-46).
int ent_kldiv_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h, int bits,
                  int method, double* out) {
  DRV_TRY(check_common(ctx, X_dev, n_rows, n_cols, h));
  if (!out) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  auto t0 = std::chrono::steady_clock::now();
  const double sc = bits ? 1.0 / std::log(2.0) : 1.0;
  std::vector<double> inv;
  DRV_TRY(column_densities_normalised(ctx, X_dev, n_rows, n_cols, h, method, inv));
  std::vector<ent::RedItem> items;
  for (uint64_t i = 0; i < n_cols; ++i)
    for (uint64_t j = 0; j < n_cols; ++j) {
      if (i == j) continue;
      ent::RedItem it = item(ent::kRedKL, {i * n_rows, j * n_rows}, n_rows, inv[i], sc);
      it.pref2 = inv[j];
      items.push_back(it);
    }
  std::vector<double> res;
  DRV_TRY(reduce_only(ctx, items, res));
  size_t p = 0;
  for (size_t i = 0; i < n_cols; ++i)
    for (size_t j = 0; j < n_cols; ++j) out[i * n_cols + j] = (i == j) ? 0.0 : res[p++];
  ctx->driver_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ENT_OK;
}

// Negs::main (negs.cpp:34-46): per column, -sum_k p log p over the
// sum1-normalised 1-D density.
int ent_negs_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h, int bits,
                 int method, double* out) {
  DRV_TRY(check_common(ctx, X_dev, n_rows, n_cols, h));
  if (!out) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  auto t0 = std::chrono::steady_clock::now();
  const double sc = bits ? 1.0 / std::log(2.0) : 1.0;
  std::vector<double> inv;
  DRV_TRY(column_densities_normalised(ctx, X_dev, n_rows, n_cols, h, method, inv));
  std::vector<ent::RedItem> items;
  for (uint64_t c = 0; c < n_cols; ++c) items.push_back(item(ent::kRedNegs, {c * n_rows}, n_rows, inv[c], sc));
  std::vector<double> res;
  DRV_TRY(reduce_only(ctx, items, res));
  for (size_t c = 0; c < n_cols; ++c) out[c] = -res[c];
  ctx->driver_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ENT_OK;
}

namespace {

// hestimate.cpp:42-53 with the self-convolution value supplied: float
// intermediates exactly where the reference rounds to float.
float amise_of(float conv, float h, size_t n) {
  float a = (float)(0.04 * 1.0 / std::pow((double)h, 5.0) * (double)conv);
  a = (float)std::pow((double)a, 0.2);
  return (float)((double)a * 1.25 * std::pow(0.6, 0.8) / std::pow((double)(float)n, 0.8));
}

// One step of AmiseMin::Kernel::_interval_optimized_h (hestimate.cpp:55-86)
// for one column, unrolled from the reference's recursion into a loop.
struct AmiseSearch {
  float prec, lo, hi, pl, ph, h;
  unsigned iter;
  bool done;
  // evaluates the termination test; when not done, sets the two pivots
  void step() {
    const float delta = hi - lo;
    const float pivot = (float)((double)lo + 0.5 * (double)delta);
    if (iter == 0 || delta <= prec) {
      h = pivot;
      done = true;
      return;
    }
    pl = (float)((double)pivot - 0.25 * (double)delta);
    ph = (float)((double)pivot + 0.25 * (double)delta);
  }
  void update(float al, float ah) {
    const float delta = hi - lo;
    const float pivot = (float)((double)lo + 0.5 * (double)delta);
    if (al < ah) hi = pivot;
    else lo = pivot;
    --iter;
  }
};

}  // namespace

int ent_hestimate_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, int kind, float* h_out) {
  if (!ctx || !ctx->eng || !X_dev || !h_out || n_rows == 0 || n_cols == 0) {
    set_error("null argument or empty input");
    return ENT_ERR_ARG;
  }
  if (kind != ENT_HEST_THUMB && kind != ENT_HEST_AMISE) {
    set_error("unknown bandwidth estimator");
    return ENT_ERR_ARG;
  }
  if (n_rows >= (1ull << 32)) {
    set_error("more than 2^32 rows");
    return ENT_ERR_UNSUPPORTED;
  }
  auto t0 = std::chrono::steady_clock::now();
  ent::Engine& E = *ctx->eng;
  DRV_HIP(hipSetDevice(E.device()));
  // Thumb::Kernel (hestimate.cpp:19-30): N^(-1/7) * sqrt(population variance);
  // the variance in FP64 on the device (the reference accumulates in float)
  const uint32_t C = (uint32_t)n_cols;
  const uint32_t bpc = (uint32_t)std::min<uint64_t>(64, std::max<uint64_t>(1, n_rows / 16384));
  DRV_TRY(E.ensure(ctx->red, sizeof(double) * ((size_t)C * bpc + 2 * (size_t)C)));
  double* d_part = (double*)ctx->red.p;
  double* d_mean = d_part + (size_t)C * bpc;
  double* d_var = d_mean + C;
  ent::launch_col_moments(E.stream(), X_dev, n_rows, C, d_part, bpc, d_mean, d_var);
  DRV_HIP(hipGetLastError());
  std::vector<double> var(C);
  DRV_HIP(hipMemcpyAsync(var.data(), d_var, sizeof(double) * C, hipMemcpyDeviceToHost, E.stream()));
  DRV_HIP(hipStreamSynchronize(E.stream()));
  std::vector<float> h0(C);
  for (uint32_t c = 0; c < C; ++c) {
    h0[c] = (float)(std::pow((double)n_rows, -1.0 / 7.0) * std::sqrt(var[c]));
    if (!std::isfinite(h0[c])) {
      set_error("non-finite coordinate in bandwidth estimation input");
      return ENT_ERR_NONFINITE;
    }
  }
  if (kind == ENT_HEST_THUMB) {
    std::memcpy(h_out, h0.data(), sizeof(float) * C);
    ctx->driver_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return ENT_OK;
  }
  for (uint32_t c = 0; c < C; ++c)
    if (!(h0[c] > 0.0f)) {
      set_error("error: column with zero variance has no AMISE bandwidth");
      return ENT_ERR_ARG;
    }
  // AmiseMin::Kernel (hestimate.cpp:88-104): every column's interval search in
  // lockstep, both pivots of every unfinished column in one batch per step;
  // the self-convolution is the engine's triangular-kernel job, as in
  // ent_epa_convolution: (1/N^2) sum_ij 4.5 max(0, 1 - |x_i - x_j| / (2h)).
  std::vector<AmiseSearch> S(C);
  for (uint32_t c = 0; c < C; ++c) {
    S[c].prec = (float)(0.01 * (double)h0[c]);
    S[c].lo = (float)(0.1 * (double)h0[c]);
    S[c].hi = 2.0f * h0[c];
    S[c].iter = 20;
    S[c].done = false;
  }
  const double norm = 4.5 / ((double)n_rows * (double)n_rows);
  for (;;) {
    std::vector<uint32_t> act;
    for (uint32_t c = 0; c < C; ++c)
      if (!S[c].done) {
        S[c].step();
        if (!S[c].done) act.push_back(c);
      }
    if (act.empty()) break;
    std::vector<ent::JobSpec> js;
    std::vector<ent::RedItem> items;
    uint64_t next = 0;
    for (uint32_t c : act)
      for (int side = 0; side < 2; ++side) {
        ent::JobSpec j;
        j.d = 1;
        j.n = (uint32_t)n_rows;
        j.n_rows = (uint32_t)n_rows;
        j.off[0] = (uint64_t)c * n_rows;
        j.h[0] = 2.0f * (side ? S[c].ph : S[c].pl);
        j.flags = ent::kJobTri;
        j.norm = norm;
        j.out_off = next;
        js.push_back(j);
        items.push_back(item(ent::kRedSum, {next}, n_rows, 0.0, 1.0));
        next += n_rows;
      }
    std::vector<double> conv;
    DRV_TRY(run_and_reduce(ctx, X_dev, js, ent::kMethodCell, items, conv));
    for (size_t a = 0; a < act.size(); ++a) {
      AmiseSearch& s = S[act[a]];
      const float al = amise_of((float)conv[2 * a], s.pl, n_rows);
      const float ah = amise_of((float)conv[2 * a + 1], s.ph, n_rows);
      s.update(al, ah);
    }
  }
  for (uint32_t c = 0; c < C; ++c) h_out[c] = S[c].h;
  ctx->driver_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ENT_OK;
}

int ent_ctx_last_driver_ms(ent_ctx* ctx, double* ms) {
  if (!ctx || !ms) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  *ms = ctx->driver_ms;
  return ENT_OK;
}

}  // extern "C"
