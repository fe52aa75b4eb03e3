// kde_device.hpp -- device helpers shared by the kernel files of the KDE
// engine (wave reductions and scans on the DPP network, XCD ordering, the
// box arithmetic).  Each translation unit gets its own copy (internal
// linkage).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "kde_engine.hpp"

namespace ent {
namespace {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Monotone float <-> uint32 map (order preserving) for atomic min/max.
__device__ __forceinline__ uint32_t f2o(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Cross-lane steps on the DPP network (no LDS round trip).  Controls of the
// gfx9 encoding: row_shr:n = 0x110 + n, row_ror:n = 0x120 + n,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROW_MASK, 0xF, false);
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dppf(float old, float v) {
  return __uint_as_float(dpp<CTRL, ROW_MASK>(__float_as_uint(old), __float_as_uint(v)));
}

// Wave minimum / maximum, returned wave-uniform: rotations within each row
// of 16 lanes (every lane has a source), then rows combined through lanes
// 15 and 31; lane 63 ends with the result.
template <bool MAX>
__device__ __forceinline__ float wave_extreme(float v) {
  auto op = [](float a, float b) { return MAX ? fmaxf(a, b) : fminf(a, b); };
  v = op(v, dppf<0x121>(v, v));
  v = op(v, dppf<0x122>(v, v));
  v = op(v, dppf<0x124>(v, v));
  v = op(v, dppf<0x128>(v, v));
  v = op(v, dppf<0x142, 0xA>(v, v));
  v = op(v, dppf<0x143, 0xC>(v, v));
  return __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), kWave - 1));
}
__device__ __forceinline__ float wave_min(float v) { return wave_extreme<false>(v); }
__device__ __forceinline__ float wave_max(float v) { return wave_extreme<true>(v); }

// Inclusive wave scan of uint32: shifted adds within rows of 16 (lanes
// shifted in from outside the row add 0), then row totals carried through
// lanes 15 and 31.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += dpp<0x111>(0u, v);
  v += dpp<0x112>(0u, v);
  v += dpp<0x114>(0u, v);
  v += dpp<0x118>(0u, v);
  v += dpp<0x142, 0xA>(0u, v);
  v += dpp<0x143, 0xC>(0u, v);
  return v;
}

// Reductions and scans over groups of W = 16 or 32 consecutive lanes (the
// sub-tiles of a packed tile); every lane of a group gets its group's
// result.  Rotations within rows of 16, then for W = 32 the row pair is
// combined across lane distance 16.
template <int W, typename Op>
__device__ __forceinline__ uint32_t sub_reduce(uint32_t v, Op op) {
  static_assert(W == 16 || W == 32, "sub-group width");
  v = op(v, dpp<0x121>(v, v));
  v = op(v, dpp<0x122>(v, v));
  v = op(v, dpp<0x124>(v, v));
  v = op(v, dpp<0x128>(v, v));
  if constexpr (W == 32) v = op(v, (uint32_t)__shfl_xor((int)v, 16, kWave));
  return v;
}
template <int W>
__device__ __forceinline__ float sub_min(float v) {
  return __uint_as_float(sub_reduce<W>(__float_as_uint(v), [](uint32_t a, uint32_t b) {
    return __float_as_uint(fminf(__uint_as_float(a), __uint_as_float(b)));
  }));
}
template <int W>
__device__ __forceinline__ float sub_max(float v) {
  return __uint_as_float(sub_reduce<W>(__float_as_uint(v), [](uint32_t a, uint32_t b) {
    return __float_as_uint(fmaxf(__uint_as_float(a), __uint_as_float(b)));
  }));
}
template <int W>
__device__ __forceinline__ uint32_t sub_sum(uint32_t v) {
  return sub_reduce<W>(v, [](uint32_t a, uint32_t b) { return a + b; });
}
// inclusive scan within each group of W lanes
template <int W>
__device__ __forceinline__ uint32_t sub_incl_scan(uint32_t v) {
  v += dpp<0x111>(0u, v);
  v += dpp<0x112>(0u, v);
  v += dpp<0x114>(0u, v);
  v += dpp<0x118>(0u, v);
  if constexpr (W == 32) v += dpp<0x142, 0xA>(0u, v);
  return v;
}
// wave maximum of uint32, wave-uniform
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  auto op = [](uint32_t a, uint32_t b) { return a > b ? a : b; };
  v = op(v, dpp<0x121>(v, v));
  v = op(v, dpp<0x122>(v, v));
  v = op(v, dpp<0x124>(v, v));
  v = op(v, dpp<0x128>(v, v));
  v = op(v, dpp<0x142, 0xA>(v, v));
  v = op(v, dpp<0x143, 0xC>(v, v));
  return __builtin_amdgcn_readlane(v, kWave - 1);
}

// Workgroups are dispatched round-robin over the XCDs (workgroup b runs on
// XCD b % 8).  This bijection hands each XCD a contiguous share of the
// workgroup-indexed work list, so the few jobs one XCD works on at a time
// keep their coordinate columns in that XCD's own L2 instead of every XCD
// touching every active job.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t G) {
  const uint32_t q = G / kXcds, r = G % kXcds, x = b % kXcds;
  return x * q + (x < r ? x : r) + b / kXcds;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Box of x along one dim (boxedsearch.cpp:30-31 arithmetic): floor of the
// correctly rounded fl(fl(x - mn) / bs), clamped to [0, nb-1].
__device__ __forceinline__ uint32_t box_of(float x, float mn, float bs, uint32_t nb) {
  float q = __fdiv_rn(__fsub_rn(x, mn), bs);
  float f = floorf(q);
  if (!(f > 0.0f)) return 0u;  // also maps NaN to 0
  if (f >= (float)nb) return nb - 1u;
  uint32_t b = (uint32_t)f;
  return b > nb - 1u ? nb - 1u : b;
}

__device__ __forceinline__ uint64_t cell_key(const DJob& J, const float* v) {
  uint64_t k = box_of(v[0], J.mn[0], J.bs[0], J.nb[0]);
  if (J.d >= 2) {
    uint64_t r = box_of(v[1], J.mn[1], J.bs[1], J.nb[1]);
    if (J.d == 3) r += (uint64_t)box_of(v[2], J.mn[2], J.bs[2], J.nb[2]) * J.nb[1];
    k += r * J.nb[0];
  }
  return k;
}

}  // namespace

}  // namespace ent
