// capi.hip -- extern "C" boundary (include/entropical_kde.h) over the engine.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include "../../include/entropical_kde.h"
#include "ctx.hpp"
#include "engine.hpp"
#include "kde_kernels.hpp"

namespace {

using ent::set_error;

int check_ctx(ent_ctx* ctx) {
  if (!ctx || !ctx->eng) {
    set_error("null context");
    return ENT_ERR_ARG;
  }
  return ENT_OK;
}

#define CAPI_HIP(call)                                                               \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call);  \
      return e_ == hipErrorOutOfMemory ? ENT_ERR_OOM : ENT_ERR_HIP;                  \
    }                                                                                \
  } while (0)

#define CAPI_TRY(expr)          \
  do {                          \
    int r_ = (expr);            \
    if (r_ != ENT_OK) return r_; \
  } while (0)

}  // namespace

// Argument checks of prob_dens_coord_prep (densities_common.hxx:15-33) plus
// the conditions the reference leaves undefined (tau_max >= n_rows, column
// index out of range, non-positive bandwidth).
int ent_validate_density_args(size_t n_rows, size_t n_cols, const uint32_t* i_cols, const float* h,
                          const uint32_t* tau, uint32_t n_dim, uint32_t* tau_max_out) {
  if (n_dim < 1 || n_dim > 3) {
    set_error("error: can only compute combined probabilities in 1, 2 or 3 dimensions.");
    return ENT_ERR_ARG;
  }
  if (!i_cols || !h) {
    set_error("error: number of bandwidth parameters does not match number of selected columns.");
    return ENT_ERR_ARG;
  }
  uint32_t tau_max = 0;
  for (uint32_t e = 0; e < n_dim; ++e) {
    if (i_cols[e] >= n_cols) {
      set_error("error: given column index is not compatible with available number of columns");
      return ENT_ERR_ARG;
    }
    if (!(h[e] > 0.0f) || !std::isfinite(h[e])) {
      set_error("error: bandwidths must be finite and positive");
      return ENT_ERR_ARG;
    }
    if (tau) tau_max = std::max(tau_max, tau[e]);
  }
  if (n_rows == 0 || (size_t)tau_max >= n_rows) {
    set_error("error: lag time (tau) must be smaller than the number of rows");
    return ENT_ERR_ARG;
  }
  if (n_rows - tau_max >= (1ull << 32)) {
    set_error("error: more than 2^32 rows");
    return ENT_ERR_UNSUPPORTED;
  }
  *tau_max_out = tau_max;
  return ENT_OK;
}

namespace {
int validate_density_args(size_t n_rows, size_t n_cols, const uint32_t* i_cols, const float* h,
                          const uint32_t* tau, uint32_t n_dim, uint32_t* tau_max_out) {
  return ent_validate_density_args(n_rows, n_cols, i_cols, h, tau, n_dim, tau_max_out);
}
}  // namespace

extern "C" {

int ent_version(void) { return 1; }

const char* ent_last_error(void) { return ent::last_error(); }

int ent_device_count(int* n) {
  if (!n) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *n = c;
  return ENT_OK;
}

int ent_ctx_create(int device, ent_ctx** out) {
  if (!out) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  *out = nullptr;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c == 0) {
    set_error("error: no HIP-capable GPU found");
    return ENT_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= c) {
    set_error("error: device index out of range");
    return ENT_ERR_ARG;
  }
  ent_ctx* ctx = new (std::nothrow) ent_ctx;
  if (!ctx) {
    set_error("out of host memory");
    return ENT_ERR_OOM;
  }
  ctx->eng = new (std::nothrow) ent::Engine(device);
  if (!ctx->eng || !ctx->eng->stream()) {
    delete ctx->eng;
    delete ctx;
    set_error("error: could not initialise the HIP device");
    return ENT_ERR_NO_DEVICE;
  }
  // set-up of the no-allocation entry point: the device's scratch array and
  // an engine that carves its workspace from it (host staging allocated now)
  void* sp = nullptr;
  size_t sb = 0;
  int rc = ent_noalloc_scratch(device, &sp, &sb);
  if (rc == ENT_OK) {
    ctx->eng_na = new (std::nothrow) ent::Engine(device);
    if (!ctx->eng_na || !ctx->eng_na->stream()) {
      set_error("error: could not initialise the HIP device");
      rc = ENT_ERR_NO_DEVICE;
    } else {
      rc = ctx->eng_na->use_arena(sp, sb, 1u << 20);
    }
  }
  if (rc == ENT_OK) ctx->na_jobs.reserve(1);
  if (rc != ENT_OK) {
    delete ctx->eng_na;
    delete ctx->eng;
    delete ctx;
    return rc;
  }
  *out = ctx;
  return ENT_OK;
}

int ent_ctx_destroy(ent_ctx* ctx) {
  if (!ctx) return ENT_OK;
  if (ctx->eng) {
    // teardown: errors here are not actionable
    (void)hipSetDevice(ctx->eng->device());
    (void)hipStreamSynchronize(ctx->eng->stream());
    if (ctx->upload.p) (void)hipFree(ctx->upload.p);
    if (ctx->out.p) (void)hipFree(ctx->out.p);
    if (ctx->aux.p) (void)hipFree(ctx->aux.p);
    if (ctx->red.p) (void)hipFree(ctx->red.p);
  }
  delete ctx->eng_na;
  delete ctx->eng;
  delete ctx;
  return ENT_OK;
}

int ent_ctx_set_tuning(ent_ctx* ctx, int q, int m0, int mr, int group_h, int sub) {
  CAPI_TRY(check_ctx(ctx));
  if (!(q == 1 || q == 2 || q == 4) || m0 < 1 || m0 > 64 || mr < 1 || mr > 8 || group_h < 1 || group_h > 64 ||
      !(sub == 0 || sub == 16 || sub == 32) || (sub != 0 && q != 1)) {
    set_error("invalid tuning (q in {1,2,4}, 1<=m0<=64, 1<=mr<=8, 1<=group_h<=64, sub in {0,16,32}; "
              "sub-tiles with q = 1)");
    return ENT_ERR_ARG;
  }
  for (ent::Engine* e : {ctx->eng, ctx->eng_na}) {
    if (!e) continue;
    e->tuning.q = q;
    e->tuning.m0 = m0;
    e->tuning.mr = mr;
    e->tuning.group_h = group_h;
    e->tuning.sub = sub;
  }
  return ENT_OK;
}

int ent_ctx_get_tuning(ent_ctx* ctx, int* tuning5) {
  CAPI_TRY(check_ctx(ctx));
  if (!tuning5) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  const ent::Tuning& t = ctx->eng->tuning;
  tuning5[0] = t.q;
  tuning5[1] = t.m0;
  tuning5[2] = t.mr;
  tuning5[3] = t.group_h;
  tuning5[4] = t.sub;
  return ENT_OK;
}

int ent_ctx_last_stats(ent_ctx* ctx, double* density_ms, double* total_ms, uint64_t* pairs,
                       uint64_t* tiles) {
  CAPI_TRY(check_ctx(ctx));
  const ent::RunStats& s = ctx->eng->last;
  if (density_ms) *density_ms = s.density_ms;
  if (total_ms) *total_ms = s.total_ms;
  if (pairs) *pairs = s.pairs;
  if (tiles) *tiles = s.tiles;
  return ENT_OK;
}

int ent_ctx_set_graph(ent_ctx* ctx, int enable) {
  CAPI_TRY(check_ctx(ctx));
  ctx->eng->use_graph = enable != 0;
  return ENT_OK;
}

int ent_ctx_last_sort(ent_ctx* ctx, uint32_t* out2) {
  CAPI_TRY(check_ctx(ctx));
  if (!out2) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  out2[0] = ctx->eng->last.radix_passes;
  out2[1] = ctx->eng->last.sort_bases;
  return ENT_OK;
}

int ent_ctx_last_pairs(ent_ctx* ctx, uint64_t* out3) {  // out3: 4 entries
  CAPI_TRY(check_ctx(ctx));
  if (!out3) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  for (int d = 0; d < 3; ++d) out3[d] = ctx->eng->last.pairs_by_d[d];
  out3[3] = ctx->eng->last.lane_pairs;
  return ENT_OK;
}

int ent_ctx_last_moments(ent_ctx* ctx, uint64_t* out4) {
  CAPI_TRY(check_ctx(ctx));
  if (!out4) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  for (int d = 0; d < 2; ++d) {
    out4[d] = ctx->eng->last.moment_pairs[d];
    out4[2 + d] = ctx->eng->last.moment_terms[d];
  }
  return ENT_OK;
}

int ent_ctx_last_tiles(ent_ctx* ctx, uint64_t* out2) {
  CAPI_TRY(check_ctx(ctx));
  if (!out2) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  out2[0] = ctx->eng->last.tiles;
  out2[1] = ctx->eng->last.raw_tiles;
  return ENT_OK;
}

int ent_ctx_last_graph_replays(ent_ctx* ctx, uint32_t* out) {
  CAPI_TRY(check_ctx(ctx));
  if (!out) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  *out = ctx->eng->last.graph_replays;
  return ENT_OK;
}

int ent_dev_alloc(ent_ctx* ctx, size_t bytes, void** ptr) {
  CAPI_TRY(check_ctx(ctx));
  if (!ptr) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  CAPI_HIP(hipSetDevice(ctx->eng->device()));
  CAPI_HIP(hipMalloc(ptr, bytes ? bytes : 4));
  return ENT_OK;
}

int ent_dev_free(ent_ctx* ctx, void* ptr) {
  CAPI_TRY(check_ctx(ctx));
  CAPI_HIP(hipSetDevice(ctx->eng->device()));
  CAPI_HIP(hipStreamSynchronize(ctx->eng->stream()));
  if (ptr) CAPI_HIP(hipFree(ptr));
  return ENT_OK;
}

int ent_h2d(ent_ctx* ctx, void* dst, const void* src, size_t bytes) {
  CAPI_TRY(check_ctx(ctx));
  CAPI_HIP(hipSetDevice(ctx->eng->device()));
  CAPI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->eng->stream()));
  CAPI_HIP(hipStreamSynchronize(ctx->eng->stream()));
  return ENT_OK;
}

int ent_d2h(ent_ctx* ctx, void* dst, const void* src, size_t bytes) {
  CAPI_TRY(check_ctx(ctx));
  CAPI_HIP(hipSetDevice(ctx->eng->device()));
  CAPI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->eng->stream()));
  CAPI_HIP(hipStreamSynchronize(ctx->eng->stream()));
  return ENT_OK;
}

int ent_ctx_sync(ent_ctx* ctx) {
  CAPI_TRY(check_ctx(ctx));
  CAPI_HIP(hipSetDevice(ctx->eng->device()));
  CAPI_HIP(hipStreamSynchronize(ctx->eng->stream()));
  return ENT_OK;
}

int ent_coord_prep(const float* coords, size_t n_rows, const uint32_t* i_cols, const float* h,
                   const uint32_t* tau, uint32_t n_dim, int row_major, float* out, size_t* n_sel) {
  if (!coords || !out) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  uint32_t tau_max = 0;
  // n_cols is not part of the reference signature; column bounds are the
  // caller's responsibility there as here.
  CAPI_TRY(validate_density_args(n_rows, (size_t)-1, i_cols, h, tau, n_dim, &tau_max));
  const size_t ns = n_rows - tau_max;
  for (size_t i = 0; i < ns; ++i)
    for (uint32_t j = 0; j < n_dim; ++j) {
      float v = coords[(size_t)i_cols[j] * n_rows + i + (tau ? tau[j] : 0u)];
      if (row_major) out[i * n_dim + j] = v;
      else out[(size_t)j * ns + i] = v;
    }
  if (n_sel) *n_sel = ns;
  return ENT_OK;
}

int ent_combined_densities_dev(ent_ctx* ctx, const float* coords_dev, size_t n_rows, size_t n_cols,
                               const uint32_t* i_cols, const float* h, const uint32_t* tau,
                               uint32_t n_dim, int method, double* out_dev) {
  CAPI_TRY(check_ctx(ctx));
  if (!coords_dev || !out_dev) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  if (method != ENT_METHOD_CELL && method != ENT_METHOD_BRUTE) {
    set_error("unknown method");
    return ENT_ERR_ARG;
  }
  uint32_t tau_max = 0;
  CAPI_TRY(validate_density_args(n_rows, n_cols, i_cols, h, tau, n_dim, &tau_max));
  ent::JobSpec js;
  js.d = n_dim;
  js.n = (uint32_t)(n_rows - tau_max);
  js.n_rows = (uint32_t)n_rows;
  js.out_off = 0;
  for (uint32_t e = 0; e < n_dim; ++e) {
    js.off[e] = (uint64_t)i_cols[e] * n_rows + (tau ? tau[e] : 0u);
    js.h[e] = h[e];
  }
  std::vector<ent::JobSpec> jobs(1, js);
  return ctx->eng->densities(coords_dev, jobs, out_dev, method);
}

int ent_combined_densities(ent_ctx* ctx, const float* coords, size_t n_rows, size_t n_cols,
                           const uint32_t* i_cols, const float* h, const uint32_t* tau,
                           uint32_t n_dim, int method, double* out) {
  CAPI_TRY(check_ctx(ctx));
  if (!coords || !out) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  uint32_t tau_max = 0;
  CAPI_TRY(validate_density_args(n_rows, n_cols, i_cols, h, tau, n_dim, &tau_max));
  ent::Engine& E = *ctx->eng;
  CAPI_HIP(hipSetDevice(E.device()));
  // upload the distinct selected columns only
  uint32_t slot_of[3];
  std::vector<uint32_t> cols;
  for (uint32_t e = 0; e < n_dim; ++e) {
    auto it = std::find(cols.begin(), cols.end(), i_cols[e]);
    if (it == cols.end()) {
      slot_of[e] = (uint32_t)cols.size();
      cols.push_back(i_cols[e]);
    } else {
      slot_of[e] = (uint32_t)(it - cols.begin());
    }
  }
  const size_t ns = n_rows - tau_max;
  CAPI_TRY(E.ensure(ctx->upload, sizeof(float) * n_rows * cols.size()));
  CAPI_TRY(E.ensure(ctx->out, sizeof(double) * ns));
  for (size_t c = 0; c < cols.size(); ++c)
    CAPI_HIP(hipMemcpyAsync((float*)ctx->upload.p + c * n_rows, coords + (size_t)cols[c] * n_rows,
                            sizeof(float) * n_rows, hipMemcpyHostToDevice, E.stream()));
  std::vector<uint32_t> local_cols(n_dim);
  for (uint32_t e = 0; e < n_dim; ++e) local_cols[e] = slot_of[e];
  int rc = ent_combined_densities_dev(ctx, (const float*)ctx->upload.p, n_rows, cols.size(),
                                      local_cols.data(), h, tau, n_dim, method, (double*)ctx->out.p);
  if (rc != ENT_OK) return rc;
  CAPI_HIP(hipMemcpyAsync(out, ctx->out.p, sizeof(double) * ns, hipMemcpyDeviceToHost, E.stream()));
  CAPI_HIP(hipStreamSynchronize(E.stream()));
  return ENT_OK;
}

int ent_batch_densities_dev(ent_ctx* ctx, const float* X_dev, const ent_job* jobs, size_t n_jobs,
                            int method, double* dens_dev) {
  CAPI_TRY(check_ctx(ctx));
  if (!X_dev || !dens_dev || (!jobs && n_jobs)) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  if (method != ENT_METHOD_CELL && method != ENT_METHOD_BRUTE) {
    set_error("unknown method");
    return ENT_ERR_ARG;
  }
  std::vector<ent::JobSpec> js(n_jobs);
  for (size_t j = 0; j < n_jobs; ++j) {
    const ent_job& a = jobs[j];
    if (a.n_dim < 1 || a.n_dim > 3 || a.n == 0 || a.n_rows == 0) {
      set_error("invalid job: n_dim must be 1..3 and n, n_rows > 0");
      return ENT_ERR_ARG;
    }
    js[j].d = a.n_dim;
    js[j].n = a.n;
    js[j].n_rows = a.n_rows;
    js[j].out_off = a.out_off;
    for (uint32_t e = 0; e < 3; ++e) {
      js[j].off[e] = a.off[e];
      js[j].h[e] = e < a.n_dim ? a.h[e] : 1.0f;
    }
  }
  return ctx->eng->densities(X_dev, js, dens_dev, method);
}

int ent_epa_convolution(ent_ctx* ctx, const float* x, size_t n, float h, float* out) {
  CAPI_TRY(check_ctx(ctx));
  if (!x || !out || n == 0) {
    set_error("null argument or empty input");
    return ENT_ERR_ARG;
  }
  if (!(h > 0.0f) || !std::isfinite(h)) {
    set_error("error: bandwidth must be finite and positive");
    return ENT_ERR_ARG;
  }
  if (n >= (1ull << 32)) {
    set_error("error: more than 2^32 points");
    return ENT_ERR_UNSUPPORTED;
  }
  ent::Engine& E = *ctx->eng;
  CAPI_HIP(hipSetDevice(E.device()));
  CAPI_TRY(E.ensure(ctx->upload, sizeof(float) * n));
  CAPI_TRY(E.ensure(ctx->out, sizeof(double) * n));
  CAPI_HIP(hipMemcpyAsync(ctx->upload.p, x, sizeof(float) * n, hipMemcpyHostToDevice, E.stream()));
  // (1/N^2) sum_i sum_j 4.5 * max(0, 1 - |x_i - x_j| / (2h))  (densities_omp.cpp:259-276)
  ent::JobSpec js;
  js.d = 1;
  js.n = (uint32_t)n;
  js.n_rows = (uint32_t)n;
  js.h[0] = 2.0f * h;
  js.flags = ent::kJobTri;
  js.norm = 4.5 / ((double)n * (double)n);
  std::vector<ent::JobSpec> jobs(1, js);
  CAPI_TRY(E.densities((const float*)ctx->upload.p, jobs, (double*)ctx->out.p, ent::kMethodCell));
  std::vector<double> rows(n);
  CAPI_HIP(hipMemcpyAsync(rows.data(), ctx->out.p, sizeof(double) * n, hipMemcpyDeviceToHost, E.stream()));
  CAPI_HIP(hipStreamSynchronize(E.stream()));
  double s = 0.0, c = 0.0;  // Neumaier
  for (size_t i = 0; i < n; ++i) {
    double t = s + rows[i];
    c += (std::fabs(s) >= std::fabs(rows[i])) ? (s - t) + rows[i] : (rows[i] - t) + s;
    s = t;
  }
  *out = (float)(s + c);
  return ENT_OK;
}

int ent_bin(ent_ctx* ctx, const float* x, size_t n, float box_size, int32_t* box, uint32_t* n_boxes,
            float* mn, float* mx) {
  CAPI_TRY(check_ctx(ctx));
  if (!x || !box || n == 0) {
    set_error("null argument or empty input");
    return ENT_ERR_ARG;
  }
  if (!(box_size > 0.0f) || !std::isfinite(box_size)) {
    set_error("error: box size must be finite and positive");
    return ENT_ERR_ARG;
  }
  if (n >= (1ull << 32)) {
    set_error("error: more than 2^32 points");
    return ENT_ERR_UNSUPPORTED;
  }
  ent::Engine& E = *ctx->eng;
  CAPI_HIP(hipSetDevice(E.device()));
  CAPI_TRY(E.ensure(ctx->upload, sizeof(float) * n));
  CAPI_TRY(E.ensure(ctx->aux, sizeof(int32_t) * n));
  CAPI_HIP(hipMemcpyAsync(ctx->upload.p, x, sizeof(float) * n, hipMemcpyHostToDevice, E.stream()));
  ent::JobSpec js;
  js.d = 1;
  js.n = (uint32_t)n;
  js.n_rows = (uint32_t)n;
  js.h[0] = box_size;  // with one box per bandwidth the box size is exactly box_size
  std::vector<ent::JobSpec> jobs(1, js);
  ent::Tuning saved = E.tuning;
  E.tuning.m0 = 1;
  int rc = E.densities((const float*)ctx->upload.p, jobs, nullptr, ent::kMethodCell, ent::kStopAfterPlan);
  E.tuning = saved;
  if (rc != ENT_OK) return rc;
  ent::launch_bin_ids(E.stream(), (const ent::DJob*)E.d_jobs.p, (const float*)ctx->upload.p,
                      (int32_t*)ctx->aux.p, (uint32_t)n);
  CAPI_HIP(hipGetLastError());
  CAPI_HIP(hipMemcpyAsync(box, ctx->aux.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, E.stream()));
  CAPI_HIP(hipStreamSynchronize(E.stream()));
  const ent::DJob& D = E.h_jobs[0];
  if (n_boxes) *n_boxes = D.nb[0];
  if (mn) *mn = D.mn[0];
  if (mx) {
    // max is not kept in the plan; recompute from the box geometry is not
    // exact, so take it from the input (same value the GPU reduced).
    float m = x[0];
    for (size_t i = 1; i < n; ++i) m = std::max(m, x[i]);
    *mx = m;
  }
  return ENT_OK;
}

int ent_cell_sort(ent_ctx* ctx, const float* coords, size_t n_rows, size_t n_cols,
                  const uint32_t* i_cols, const float* h, const uint32_t* tau, uint32_t n_dim,
                  uint64_t* keys_sorted, uint32_t* perm, uint32_t* nb, float* mn, float* bs) {
  CAPI_TRY(check_ctx(ctx));
  if (!coords || !keys_sorted || !perm) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  uint32_t tau_max = 0;
  CAPI_TRY(validate_density_args(n_rows, n_cols, i_cols, h, tau, n_dim, &tau_max));
  ent::Engine& E = *ctx->eng;
  CAPI_HIP(hipSetDevice(E.device()));
  CAPI_TRY(E.ensure(ctx->upload, sizeof(float) * n_rows * n_cols));
  CAPI_HIP(hipMemcpyAsync(ctx->upload.p, coords, sizeof(float) * n_rows * n_cols, hipMemcpyHostToDevice,
                          E.stream()));
  ent::JobSpec js;
  js.d = n_dim;
  js.n = (uint32_t)(n_rows - tau_max);
  js.n_rows = (uint32_t)n_rows;
  for (uint32_t e = 0; e < n_dim; ++e) {
    js.off[e] = (uint64_t)i_cols[e] * n_rows + (tau ? tau[e] : 0u);
    js.h[e] = h[e];
  }
  std::vector<ent::JobSpec> jobs(1, js);
  CAPI_TRY(E.densities((const float*)ctx->upload.p, jobs, nullptr, ent::kMethodCell, ent::kStopAfterSort));
  const size_t ns = js.n;
  std::vector<uint32_t> narrow;
  if (E.keys_wide) {
    CAPI_HIP(hipMemcpyAsync(keys_sorted, E.d_keys[E.sorted_buf].p, sizeof(uint64_t) * ns,
                            hipMemcpyDeviceToHost, E.stream()));
  } else {
    narrow.resize(ns);
    CAPI_HIP(hipMemcpyAsync(narrow.data(), E.d_keys[E.sorted_buf].p, sizeof(uint32_t) * ns,
                            hipMemcpyDeviceToHost, E.stream()));
  }
  CAPI_HIP(hipMemcpyAsync(perm, E.d_idx[E.sorted_buf].p, sizeof(uint32_t) * ns, hipMemcpyDeviceToHost,
                          E.stream()));
  CAPI_HIP(hipStreamSynchronize(E.stream()));
  for (size_t i = 0; i < narrow.size(); ++i) keys_sorted[i] = narrow[i];
  const ent::DJob& D = E.h_jobs[0];
  for (uint32_t e = 0; e < n_dim; ++e) {
    if (nb) nb[e] = D.nb[e];
    if (mn) mn[e] = D.mn[e];
    if (bs) bs[e] = D.bs[e];
  }
  return ENT_OK;
}

}  // extern "C"
