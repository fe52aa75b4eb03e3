// noalloc.hip -- the density entry point that allocates nothing.
//
// Its workspace (cell keys, permutations, sorted coordinates, tile records,
// cell-start and moment tables, claim and pair counters, host-planned job
// tables) lives in `scratch`, a fixed device array of ENT_NOALLOC_SCRATCH_BYTES
// that every GPU the library runs on holds until the process ends.  The array
// is located once per device, with that device current, when a context is
// created (ent_ctx_create); from then on only that device address is used.
// All contexts on one device share the array: calls that overlap on one GPU
// race and must be serialised by the caller.  Every buffer that a run needs
// cleared (claim and pair counters, min/max slots, flags) is cleared by a
// fill kernel at the start of the run, so the contents left by the previous
// call do not matter.
#include <hip/hip_runtime.h>
#include <map>
#include <mutex>
#include <string>
#include "../../include/entropical_kde.h"
#include "ctx.hpp"
#include "engine.hpp"

namespace {

static __device__ __align__(256) unsigned char scratch[ENT_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;  // device -> address of its copy of `scratch`

}  // namespace

int ent_noalloc_scratch(int device, void** ptr, size_t* bytes) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it == g_scratch_ptr.end()) {
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) prev = device;
    hipError_t e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess || !p) {
      ent::set_error(std::string("HIP error: ") + hipGetErrorString(e) + " locating the no-allocation scratch array");
      return ENT_ERR_HIP;
    }
    it = g_scratch_ptr.emplace(device, p).first;
  }
  *ptr = it->second;
  *bytes = ENT_NOALLOC_SCRATCH_BYTES;
  return ENT_OK;
}

extern "C" {

int ent_combined_densities_dev_noalloc(ent_ctx* ctx, const float* coords_dev, size_t n_rows, size_t n_cols,
                                       const uint32_t* i_cols, const float* h, const uint32_t* tau,
                                       uint32_t n_dim, int method, double* out_dev) {
  if (!ctx || !ctx->eng_na) {
    ent::set_error("null context");
    return ENT_ERR_ARG;
  }
  if (!coords_dev || !out_dev) {
    ent::set_error("null argument");
    return ENT_ERR_ARG;
  }
  if (method != ENT_METHOD_CELL && method != ENT_METHOD_BRUTE) {
    ent::set_error("unknown method");
    return ENT_ERR_ARG;
  }
  uint32_t tau_max = 0;
  int rc = ent_validate_density_args(n_rows, n_cols, i_cols, h, tau, n_dim, &tau_max);
  if (rc != ENT_OK) return rc;
  const size_t n = n_rows - tau_max;
  if (n > ENT_NOALLOC_MAX_POINTS) {
    ent::set_error("no-allocation density: " + std::to_string(n) + " points exceed the maximum of " +
                   std::to_string(ENT_NOALLOC_MAX_POINTS) + " that the fixed scratch array holds");
    return ENT_ERR_UNSUPPORTED;
  }
  ent::JobSpec js;
  js.d = n_dim;
  js.n = (uint32_t)n;
  js.n_rows = (uint32_t)n_rows;
  js.out_off = 0;
  for (uint32_t e = 0; e < n_dim; ++e) {
    js.off[e] = (uint64_t)i_cols[e] * n_rows + (tau ? tau[e] : 0u);
    js.h[e] = h[e];
  }
  ctx->na_jobs.assign(1, js);  // capacity reserved at set-up
  return ctx->eng_na->densities(coords_dev, ctx->na_jobs, out_dev, method);
}

}  // extern "C"
