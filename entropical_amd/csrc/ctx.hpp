// ctx.hpp -- definition of the opaque ent_ctx handle of the C ABI.
#pragma once
#include <vector>
#include "engine.hpp"

// Sets *ptr / *bytes to the calling device's fixed scratch array (looked up
// once per device, with that device current).  noalloc.hip.
int ent_noalloc_scratch(int device, void** ptr, size_t* bytes);

// Argument checks shared by the density entry points (capi.hip).
int ent_validate_density_args(size_t n_rows, size_t n_cols, const uint32_t* i_cols, const float* h,
                              const uint32_t* tau, uint32_t n_dim, uint32_t* tau_max_out);

struct ent_ctx {
  ent::Engine* eng = nullptr;
  // Engine of the no-allocation entry point: its workspace is the device's
  // fixed scratch array (noalloc.hip), shared by every context on the device.
  ent::Engine* eng_na = nullptr;
  std::vector<ent::JobSpec> na_jobs;  // its one-job list (capacity kept)
  ent::DevBuf upload;   // staged coordinates for host-pointer entry points
  ent::DevBuf out;      // staged densities / batch density arrays
  ent::DevBuf aux;      // small per-call buffers
  ent::DevBuf red;      // reduction items, partials and results
  double driver_ms = 0.0;  // wall time of the last driver call (HIP events)
};
