// ctx.hpp -- definition of the opaque ent_ctx handle of the C ABI.
#pragma once
#include "engine.hpp"

struct ent_ctx {
  ent::Engine* eng = nullptr;
  ent::DevBuf upload;   // staged coordinates for host-pointer entry points
  ent::DevBuf out;      // staged densities / batch density arrays
  ent::DevBuf aux;      // small per-call buffers
  ent::DevBuf red;      // reduction items, partials and results
  double driver_ms = 0.0;  // wall time of the last driver call (HIP events)
};
