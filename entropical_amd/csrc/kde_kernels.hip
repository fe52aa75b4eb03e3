// kde_kernels.hip -- gfx950 kernels of the KDE engine.
//
// Every chunked pass gives one wavefront one chunk of kChunk consecutive
// points of one job; a 256-thread workgroup carries four independent chunks.
// The chunk table (job id, job-local start) is built on the host.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>
#include "kde_engine.hpp"
#include "kde_kernels.hpp"

// Register budget of the density kernel: at least this many waves per SIMD
// (HIP's second launch-bounds argument), i.e. at most 512/N VGPRs per lane.
// 6 keeps the two-queries-per-lane variant at 6 waves without spills (its
// free allocation took 90 VGPRs, 5 waves); one query per lane is bound at 7
// waves by its scalar registers either way, and four per lane spills (the
// autotuner times every setting, so a slower one is not picked).
#ifndef ENT_DENSITY_MIN_WAVES
#define ENT_DENSITY_MIN_WAVES 6
#endif


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

// ------------------------------------------------------------------ init
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// minmax slots: [job][dim][0=min,1=max] as ordered uint32
__global__ void k_init_minmax(uint32_t* mm, uint32_t n_jobs) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_jobs * 6u) mm[i] = (i & 1u) ? 0u : 0xFFFFFFFFu;
}

// ---------------------------------------------------------------- minmax
__global__ __launch_bounds__(kBlock) void k_minmax(const DJob* __restrict__ jobs,
                                                   const uint2* __restrict__ chunks,
                                                   uint32_t n_chunks, const float* __restrict__ X,
                                                   uint32_t* __restrict__ mm,
                                                   uint32_t* __restrict__ bad) {
  const uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  const int l = lane_id();
  for (uint32_t e = 0; e < J.d; ++e) {
    float mn = __builtin_inff(), mx = -__builtin_inff();
    uint32_t nonfinite = 0;
    const float* src = X + J.off[e];
#pragma unroll 4
    for (int i = 0; i < kChunkItems; ++i) {
      uint32_t k = ch.y + i * kWave + l;
      if (k < J.n) {
        float v = src[k];
        nonfinite |= !__builtin_isfinite(v);
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
      }
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    uint64_t anybad = __ballot(nonfinite != 0);
    if (l == 0) {
      atomicMin(&mm[ch.x * 6 + e * 2 + 0], f2o(mn));
      atomicMax(&mm[ch.x * 6 + e * 2 + 1], f2o(mx));
      if (anybad) atomicOr(bad, 1u);
    }
  }
}

// ------------------------------------------------------------- cell keys
// KeyT is uint32_t when every key of the batch fits 32 bits, else uint64_t.
template <typename KeyT>
__global__ __launch_bounds__(kBlock) void k_cell_keys(const DJob* __restrict__ jobs,
                                                      const uint2* __restrict__ chunks,
                                                      uint32_t n_chunks, const float* __restrict__ X,
                                                      KeyT* __restrict__ keys,
                                                      uint32_t* __restrict__ idx) {
  const uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  const int l = lane_id();
  const uint32_t d = __builtin_amdgcn_readfirstlane(J.d);
  const uint32_t last = J.n - 1;
  // coordinates of the chunk loaded in batches of 8 per lane before use
  for (int i0 = 0; i0 < kChunkItems; i0 += 8) {
    float v[8][3];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint32_t k = ch.y + (uint32_t)(i0 + i) * kWave + l;
      const uint32_t kc = k < last ? k : last;
      v[i][0] = X[J.off[0] + kc];
      v[i][1] = d >= 2 ? X[J.off[1] + kc] : 0.f;
      v[i][2] = d >= 3 ? X[J.off[2] + kc] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint32_t k = ch.y + (uint32_t)(i0 + i) * kWave + l;
      if (k < J.n) {
        keys[J.pbase + k] = (KeyT)cell_key(J, v[i]);
        idx[J.pbase + k] = k;
      }
    }
  }
}

// Bin ids of one column (test/bit-exactness hook): box[k] for every point of
// job 0 dim 0.
__global__ void k_bin_ids(const DJob* __restrict__ jobs, const float* __restrict__ X,
                          int32_t* __restrict__ box) {
  const DJob& J = jobs[0];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < J.n; k += gridDim.x * blockDim.x)
    box[k] = (int32_t)box_of(X[J.off[0] + k], J.mn[0], J.bs[0], J.nb[0]);
}

// ------------------------------------------------------ LSD radix passes
// Histogram layout per job: hist[256*chunk0 + digit*nch + (c - chunk0)], so
// one global exclusive scan over the whole array yields, for every
// (job, digit, chunk), the global destination of that chunk's first element
// with that digit (jobs are laid out back to back at pbase).
template <typename KeyT>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const DJob* __restrict__ jobs,
                                                       const uint2* __restrict__ chunks,
                                                       const uint32_t* __restrict__ chunk_first,
                                                       uint32_t n_chunks,
                                                       const KeyT* __restrict__ keys, int shift,
                                                       uint32_t* __restrict__ hist) {
  __shared__ uint32_t bins[kBlock / kWave][kRadixBins];
  const int w = threadIdx.x / kWave;
  const int l = lane_id();
  const uint32_t c = blockIdx.x * (kBlock / kWave) + w;
  for (int b = l; b < kRadixBins; b += kWave) bins[w][b] = 0u;
  wave_sync();
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  // all of the chunk's keys are loaded before the first use
  KeyT kr[kChunkItems];
  const uint32_t last = J.n - 1;
#pragma unroll
  for (int i = 0; i < kChunkItems; ++i) {
    const uint32_t k = ch.y + i * kWave + l;
    kr[i] = keys[J.pbase + (k < last ? k : last)];
  }
#pragma unroll
  for (int i = 0; i < kChunkItems; ++i) {
    const uint32_t k = ch.y + i * kWave + l;
    if (k < J.n) atomicAdd(&bins[w][(uint32_t)(kr[i] >> shift) & (kRadixBins - 1)], 1u);
  }
  wave_sync();
  const uint32_t c0 = chunk_first[ch.x];
  const uint32_t nch = chunk_first[ch.x + 1] - c0;
  for (int b = l; b < kRadixBins; b += kWave)
    hist[(uint64_t)kRadixBins * c0 + (uint64_t)b * nch + (c - c0)] = bins[w][b];
}

// Stable scatter: a wave walks its chunk in order, 64 elements at a time;
// equal digits inside a step are ranked by lane via 8 ballots.  With GATHER
// (last pass) the point's coordinates are written to the sorted SoA array at
// the same destination, so no separate gather pass re-reads the permutation.
template <typename KeyT, bool GATHER>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const DJob* __restrict__ jobs, const uint2* __restrict__ chunks,
    const uint32_t* __restrict__ chunk_first, uint32_t n_chunks,
    const KeyT* __restrict__ keys_in, const uint32_t* __restrict__ idx_in,
    KeyT* __restrict__ keys_out, uint32_t* __restrict__ idx_out, int shift,
    const uint32_t* __restrict__ offs, const float* __restrict__ X, float* __restrict__ scoord) {
  __shared__ uint32_t run[kBlock / kWave][kRadixBins];
  const int w = threadIdx.x / kWave;
  const int l = lane_id();
  const uint32_t c = blockIdx.x * (kBlock / kWave) + w;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  const uint32_t c0 = chunk_first[ch.x];
  const uint32_t nch = chunk_first[ch.x + 1] - c0;
  for (int b = l; b < kRadixBins; b += kWave)
    run[w][b] = offs[(uint64_t)kRadixBins * c0 + (uint64_t)b * nch + (c - c0)];
  wave_sync();
  const uint64_t lt = lanemask_lt();
  for (int i = 0; i < kChunkItems; ++i) {
    uint32_t k = ch.y + i * kWave + l;
    bool valid = k < J.n;
    KeyT key = 0;
    uint32_t id = 0, dg = 0;
    if (valid) {
      key = keys_in[J.pbase + k];
      id = idx_in[J.pbase + k];
      dg = (uint32_t)(key >> shift) & (kRadixBins - 1);
    }
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      bool bit = (dg >> b) & 1u;
      uint64_t bal = __ballot(bit);
      match &= bit ? bal : ~bal;
    }
    uint32_t rank = __popcll(match & lt);
    uint32_t base = valid ? run[w][dg] : 0u;
    wave_sync();
    if (valid && rank == 0) run[w][dg] = base + (uint32_t)__popcll(match);
    wave_sync();
    if (valid) {
      uint32_t dst = base + rank;
      idx_out[dst] = id;
      if constexpr (GATHER) {
        const uint64_t p = dst - J.pbase;
        float v[3] = {0.f, 0.f, 0.f};
        for (uint32_t e = 0; e < J.d; ++e) {
          v[e] = X[J.off[e] + id];
          scoord[3 * J.pbase + (uint64_t)e * J.n + p] = v[e];
        }
        keys_out[dst] = (KeyT)cell_key(J, v);
      } else {
        keys_out[dst] = key;
      }
    }
  }
}

// LDS-staged stable scatter, one wavefront per block and chunk.  Phase 1
// ranks the chunk's elements by digit (stable, as above) into LDS, so the
// chunk sits there sorted by digit; phase 2 streams it out in that order:
// consecutive lanes then write consecutive global slots within each digit's
// run, instead of 64 unrelated addresses per store.  The chunk's per-digit
// counts come from the scanned offsets: offs[i + 1] - offs[i] (the scan is
// contiguous over all jobs and ends with the grand total at offs[m]).
// With GATHER (last pass) the full cell key is recomputed from the gathered
// coordinates, so a pass over a partial key (the two-stage sort) also ends
// with full keys.
// Output phase of a gathering scatter pass for jobs of dimension D: staged
// points leave in digit order, kB per lane at a time (positions from LDS,
// then the coordinate gathers, then the stores).
template <int D, typename KeyT>
__device__ __forceinline__ void scatter_gather_out(const DJob& J, const KeyT* skey, const uint32_t* sid,
                                                   const uint32_t* delta, uint32_t nv, int shift,
                                                   const float* __restrict__ X, KeyT* __restrict__ keys_out,
                                                   uint32_t* __restrict__ idx_out, float* __restrict__ scoord,
                                                   int l) {
  constexpr int kB = 8;
  for (uint32_t j0 = 0; j0 < nv; j0 += kB * kWave) {
    uint32_t id[kB], dst[kB];
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const uint32_t j = j0 + (uint32_t)i * kWave + l;
      const uint32_t jj = j < nv ? j : 0u;
      id[i] = sid[jj];
      dst[i] = delta[(uint32_t)(skey[jj] >> shift) & (kRadixBins - 1)] + jj;
    }
    float v[kB][3];
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      v[i][0] = v[i][1] = v[i][2] = 0.0f;
#pragma unroll
      for (int e = 0; e < D; ++e) v[i][e] = X[J.off[e] + id[i]];
    }
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      if (j0 + (uint32_t)i * kWave + l >= nv) continue;
      idx_out[dst[i]] = id[i];
      const uint64_t p = dst[i] - J.pbase;
#pragma unroll
      for (int e = 0; e < D; ++e) scoord[3 * J.pbase + (uint64_t)e * J.n + p] = v[i][e];
      keys_out[dst[i]] = (KeyT)cell_key(J, v[i]);
    }
  }
}
template <typename KeyT, bool GATHER>
__global__ __launch_bounds__(kWave) void k_radix_scatter_lds(
    const DJob* __restrict__ jobs, const uint2* __restrict__ chunks,
    const uint32_t* __restrict__ chunk_first, uint32_t n_chunks,
    const KeyT* __restrict__ keys_in, const uint32_t* __restrict__ idx_in,
    KeyT* __restrict__ keys_out, uint32_t* __restrict__ idx_out, int shift,
    const uint32_t* __restrict__ offs, const float* __restrict__ X, float* __restrict__ scoord) {
  __shared__ uint32_t lrun[kRadixBins];
  __shared__ uint32_t delta[kRadixBins];
  __shared__ KeyT skey[kChunk];
  __shared__ uint32_t sid[kChunk];
  const int l = lane_id();
  const uint32_t c = GATHER ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  const uint32_t c0 = chunk_first[ch.x];
  const uint32_t nch = chunk_first[ch.x + 1] - c0;
  // digits 4l .. 4l+3: global start, chunk count, chunk-local start
  {
    constexpr int kPer = kRadixBins / kWave;
    uint32_t g[kPer], cnt[kPer], s = 0;
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
      const uint64_t i = (uint64_t)kRadixBins * c0 + (uint64_t)(kPer * l + t) * nch + (c - c0);
      g[t] = offs[i];
      cnt[t] = offs[i + 1] - g[t];
      s += cnt[t];
    }
    uint32_t ls = wave_incl_scan(s) - s;
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
      lrun[kPer * l + t] = ls;
      delta[kPer * l + t] = g[t] - ls;
      ls += cnt[t];
    }
  }
  const uint32_t nv = min((uint32_t)kChunk, J.n - ch.y);
  KeyT kr[kChunkItems];
  uint32_t ir[kChunkItems];
#pragma unroll
  for (int i = 0; i < kChunkItems; ++i) {
    const uint32_t k = ch.y + i * kWave + l;
    kr[i] = 0;
    ir[i] = 0;
    if (k < J.n) {
      kr[i] = keys_in[J.pbase + k];
      ir[i] = idx_in[J.pbase + k];
    }
  }
  wave_sync();
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int i = 0; i < kChunkItems; ++i) {
    const bool valid = (uint32_t)(i * kWave + l) < nv;
    const uint32_t dg = (uint32_t)(kr[i] >> shift) & (kRadixBins - 1);
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const bool bit = (dg >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      match &= bit ? bal : ~bal;
    }
    const uint32_t rank = __popcll(match & lt);
    const uint32_t base = valid ? lrun[dg] : 0u;
    wave_sync();
    if (valid && rank == 0) lrun[dg] = base + (uint32_t)__popcll(match);
    if (valid) {
      skey[base + rank] = kr[i];
      sid[base + rank] = ir[i];
    }
    wave_sync();
  }
  if constexpr (GATHER) {
    const uint32_t d = __builtin_amdgcn_readfirstlane(J.d);
    if (d == 1) scatter_gather_out<1>(J, skey, sid, delta, nv, shift, X, keys_out, idx_out, scoord, l);
    else if (d == 2) scatter_gather_out<2>(J, skey, sid, delta, nv, shift, X, keys_out, idx_out, scoord, l);
    else scatter_gather_out<3>(J, skey, sid, delta, nv, shift, X, keys_out, idx_out, scoord, l);
  } else {
    for (uint32_t j = l; j < nv; j += kWave) {
      const KeyT key = skey[j];
      const uint32_t dst = delta[(uint32_t)(key >> shift) & (kRadixBins - 1)] + j;
      idx_out[dst] = sid[j];
      keys_out[dst] = key;
    }
  }
}

// ------------------------------------------------ two-stage sort, stage 2
// Jobs that share their dim-0 window and bandwidth share one sort of that
// window by dim-0 box (the base, stable from the identity order).  Each job
// starts from its base's order and is then sorted stably by the key of its
// remaining dims, r = b1 + b2*nb1; stability keeps the dim-0 order inside
// equal r, so the result is the order by full key b0 + r*nb0 with ties in
// point order -- the same permutation as one sort by the full key.
// Writes (r, point) at the job's own positions.  With scoord (no pass over r
// follows: every r is 0) it writes the full key and the sorted coordinates.
// One chunk of k_major_keys for jobs of dimension D, in batches of kB
// points: base positions first, then the coordinate gathers, then the
// stores, so a batch's loads are in flight together.
template <int D, bool FINAL, typename KeyT>
__device__ __forceinline__ void major_keys_chunk(const DJob& J, uint32_t y, const uint32_t* __restrict__ bi,
                                                 const float* __restrict__ X, KeyT* __restrict__ keys,
                                                 uint32_t* __restrict__ idx, float* __restrict__ scoord, int l) {
  constexpr int kB = 16;
  constexpr int E0 = FINAL ? 0 : 1;  // without coordinates only the dims above 0 are needed
  const uint32_t last = J.n - 1;
#pragma unroll
  for (int i0 = 0; i0 < kChunkItems; i0 += kB) {
    uint32_t id[kB];
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const uint32_t k = y + (uint32_t)(i0 + i) * kWave + l;
      id[i] = bi[k < last ? k : last];
    }
    float v[kB][3];
    // over a base sorted by dims 0 and 1, only dim 2 is needed
    const int e0 = (!FINAL && D == 3 && (J.flags & kJobBase2)) ? 2 : E0;
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      v[i][0] = v[i][1] = v[i][2] = 0.0f;
#pragma unroll
      for (int e = E0; e < D; ++e)
        if (e >= e0) v[i][e] = X[J.off[e] + id[i]];
    }
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const uint32_t k = y + (uint32_t)(i0 + i) * kWave + l;
      if (k >= J.n) continue;
      idx[J.pbase + k] = id[i];
      if constexpr (FINAL) {
#pragma unroll
        for (int e = 0; e < D; ++e) scoord[3 * J.pbase + (uint64_t)e * J.n + k] = v[i][e];
        keys[J.pbase + k] = (KeyT)cell_key(J, v[i]);
      } else {
        // key of the dims the base does not order: 1.. (or 2 over a base
        // sorted by dims 0 and 1)
        uint64_t r = 0;
        if constexpr (D == 3) {
          r = box_of(v[i][2], J.mn[2], J.bs[2], J.nb[2]);
          if (!(J.flags & kJobBase2)) r = r * J.nb[1] + box_of(v[i][1], J.mn[1], J.bs[1], J.nb[1]);
        } else if constexpr (D == 2) {
          r = box_of(v[i][1], J.mn[1], J.bs[1], J.nb[1]);
        }
        keys[J.pbase + k] = (KeyT)r;
      }
    }
  }
}

template <typename KeyT>
__global__ __launch_bounds__(kBlock) void k_major_keys(const DJob* __restrict__ jobs,
                                                       const uint2* __restrict__ chunks,
                                                       uint32_t n_chunks,
                                                       const uint64_t* __restrict__ bbase,
                                                       const uint32_t* __restrict__ bidx,
                                                       const float* __restrict__ X,
                                                       KeyT* __restrict__ keys,
                                                       uint32_t* __restrict__ idx,
                                                       float* __restrict__ scoord) {
  const uint32_t c = xcd_order(blockIdx.x, gridDim.x) * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob J = jobs[ch.x];
  const uint32_t* bi = bidx + bbase[ch.x];
  const int l = lane_id();
  const uint32_t d = __builtin_amdgcn_readfirstlane(J.d);
  if (scoord) {
    if (d == 1) major_keys_chunk<1, true>(J, ch.y, bi, X, keys, idx, scoord, l);
    else if (d == 2) major_keys_chunk<2, true>(J, ch.y, bi, X, keys, idx, scoord, l);
    else major_keys_chunk<3, true>(J, ch.y, bi, X, keys, idx, scoord, l);
  } else {
    if (d == 1) major_keys_chunk<1, false>(J, ch.y, bi, X, keys, idx, scoord, l);
    else if (d == 2) major_keys_chunk<2, false>(J, ch.y, bi, X, keys, idx, scoord, l);
    else major_keys_chunk<3, false>(J, ch.y, bi, X, keys, idx, scoord, l);
  }
}

// --------------------------------------------------------- exclusive scan
// Three-phase device-wide exclusive scan of uint32 (in place allowed).
// kScanItems elements per thread, kBlock threads per block.
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

__global__ __launch_bounds__(kBlock) void k_scan_blocks(const uint32_t* in, uint32_t* out,
                                                        uint64_t m, uint32_t* block_sums) {
  __shared__ uint32_t s[kBlock];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t tsum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < m) ? in[base + i] : 0u;
    tsum += v[i];
  }
  s[threadIdx.x] = tsum;
  __syncthreads();
  for (int o = 1; o < kBlock; o <<= 1) {
    uint32_t t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - tsum;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < m) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == kBlock - 1) block_sums[blockIdx.x] = s[kBlock - 1];
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* out, uint64_t m,
                                                     const uint32_t* block_offs) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  const uint32_t add = block_offs[blockIdx.x];
  for (int i = threadIdx.x; i < kScanTile; i += kBlock)
    if (base + i < m) out[base + i] += add;
}

// Grand total (already reduced into *total) copied to out[m].  The input may
// have been overwritten by an in-place scan, so the total never comes from it.
__global__ void k_scan_total(const uint32_t* total, uint32_t* out, uint64_t m) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[m] = *total;
}

// ------------------------------------------------------- sorted gather
// scoord layout per job: [3*pbase + e*n + p], p = sorted position.
__global__ __launch_bounds__(kBlock) void k_gather_sorted(const DJob* __restrict__ jobs,
                                                          const uint2* __restrict__ chunks,
                                                          uint32_t n_chunks,
                                                          const float* __restrict__ X,
                                                          const uint32_t* __restrict__ idx,
                                                          float* __restrict__ scoord) {
  const uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  const int l = lane_id();
  const uint32_t d = __builtin_amdgcn_readfirstlane(J.d);
  for (int i = 0; i < kChunkItems; ++i) {
    uint32_t p = ch.y + i * kWave + l;
    if (p < J.n) {
      uint32_t k = idx[J.pbase + p];
#pragma unroll
      for (uint32_t e = 0; e < 3; ++e)
        if (e < d) scoord[3 * J.pbase + (uint64_t)e * J.n + p] = X[J.off[e] + k];
    }
  }
}

// ------------------------------------------------------------ cell starts
// start[cbase + c] = first sorted position whose key >= c, for c in
// [0, cells]; each position fills the gap of empty cells before its own key.
// Keys of points p and p - 1 (clamped to the job) for items [I0, I0 + N)
// of a chunk, all loads issued before any use.
template <int N, typename KeyT>
__device__ __forceinline__ void load_key_pairs(const DJob& J, uint32_t y, int I0, const KeyT* keys, int l,
                                               KeyT (&cur)[N], KeyT (&prev)[N]) {
  const KeyT* kk = keys + J.pbase;
  const uint32_t last = J.n - 1;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t p = y + (uint32_t)(I0 + i) * kWave + l;
    const uint32_t pc = p < last ? p : last;
    cur[i] = kk[pc];
    prev[i] = kk[pc > 0 ? pc - 1 : 0];
  }
}
// ------------------------------------------------------------ box moments
// Moment table of the 1-D jobs flagged kJobMoments.  Box b of job J, with
// origin r_b = fma(b, bs, mn) and the offsets e_j = x_j - r_b of its s0
// points (summed in FP64), stores
//   A = s0 / h^2,  B = 2 S1 / h^2,  c = s0 - S2 / h^2   (S1 = sum e, S2 = sum e^2)
// so that for a query x with d = x - r_b whose support holds the whole box,
//   sum_j (1 - ((x - x_j) / h)^2) = c + d (B - A d).
// mlist: [0] = M jobs, [1 .. M+1] = first wave of each job (prefix), then
// the M job indices; each wave covers kMomentBoxes consecutive boxes of one
// job.  The grid strides over the waves.
constexpr uint32_t kMomentBoxes = kMomentBoxesPerWave;

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v += __shfl_xor(v, m);
  return v;
}

__global__ __launch_bounds__(kBlock) void k_box_moments(const DJob* __restrict__ jobs,
                                                        const uint32_t* __restrict__ mlist,
                                                        const float* __restrict__ scoord,
                                                        const uint32_t* __restrict__ start,
                                                        float4* __restrict__ mom) {
  const uint32_t M = mlist[0];
  const uint32_t* wfirst = mlist + 1;
  const uint32_t* mjob = mlist + 2 + M;
  const uint32_t n_waves = wfirst[M];
  const int l = lane_id();
  const uint32_t stride = gridDim.x * (kBlock / kWave);
  for (uint32_t w = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; w < n_waves; w += stride) {
    // the job holding wave w: last k with wfirst[k] <= w
    uint32_t lo = 0, hi = M;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (wfirst[mid] <= w) lo = mid;
      else hi = mid;
    }
    const DJob& J = jobs[mjob[lo]];
    const float* x = scoord + 3 * J.pbase;
    const uint32_t* st = start + J.cbase;
    const uint32_t b0 = (w - wfirst[lo]) * kMomentBoxes;
    const uint32_t b1 = min(b0 + kMomentBoxes, J.nb[0]);
    const double hi2 = (double)J.hinv[0] * (double)J.hinv[0];
    for (uint32_t b = b0; b < b1; ++b) {
      const uint32_t a = st[b], e = st[b + 1];
      const float r = fmaf((float)b, J.bs[0], J.mn[0]);
      double s1 = 0.0, s2 = 0.0;
      for (uint32_t p = a + (uint32_t)l; p < e; p += kWave) {
        const double ej = (double)x[p] - (double)r;
        s1 += ej;
        s2 = fma(ej, ej, s2);
      }
      s1 = wave_sum_f64(s1);
      s2 = wave_sum_f64(s2);
      if (l == 0) {
        const double s0 = (double)(e - a);
        mom[J.mbase + b] = make_float4((float)(s0 * hi2), (float)(2.0 * s1 * hi2), (float)(s0 - s2 * hi2), 0.0f);
      }
    }
  }
}

// Moment table of the 2-D jobs flagged kJobMoments, one thread per cell
// (points staged through LDS per block, see below).
// Cell (row, cb) of job J covers dim-0 boxes [cb mcf, cb mcf + mcf) of one
// row; with origin (r0, r1) = (fma(cb mcf, bs0, mn0), fma(row, bs1, mn1)) and
// offsets (e_j, g_j) = (x_j - r0, y_j - r1) of its points, it stores the 9
// floats N[3p + q] = s0_p s1_q sum_j e_j^p g_j^q (p, q = 0..2; sums in FP64)
// with s_e = (1, 2/h_e^2, -1/h_e^2), so that for a query (x, y) with
// d = (x - r0, y - r1) whose support holds the whole cell,
//   sum_j (1 - ((x - x_j)/h0)^2)(1 - ((y - y_j)/h1)^2)
//     = sum_q b_q sum_p a_p N[3p + q],  a = (1 - d0^2/h0^2, d0, 1), b = (1 - d1^2/h1^2, d1, 1).
// mlist2: [0] = M jobs, [1 .. M+1] = first block of each job (prefix, kBlock
// cells per block), then the M job indices.  The grid strides over the blocks.
__global__ __launch_bounds__(kBlock) void k_cell_moments2(const DJob* __restrict__ jobs,
                                                          const uint32_t* __restrict__ mlist,
                                                          const float* __restrict__ scoord,
                                                          const uint32_t* __restrict__ start,
                                                          float* __restrict__ mom) {
  // The block's kBlock cells are consecutive in (row, cell) order, so their
  // points are one contiguous stretch of the sorted arrays: it is staged
  // through LDS kStage points at a time with coalesced loads, and each
  // thread sums its own cell's points from LDS.
  constexpr uint32_t kStage = 2048;
  __shared__ float2 stage[kStage];
  __shared__ uint32_t span[2];
  const uint32_t M = mlist[0];
  const uint32_t* bfirst = mlist + 1;
  const uint32_t* mjob = mlist + 2 + M;
  const uint32_t n_blocks = bfirst[M];
  for (uint32_t blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    uint32_t lo = 0, hi = M;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (bfirst[mid] <= blk) lo = mid;
      else hi = mid;
    }
    const DJob& J = jobs[mjob[lo]];
    const uint32_t cell0 = (blk - bfirst[lo]) * kBlock;
    const uint32_t cell = cell0 + threadIdx.x;
    const uint32_t ncb = J.mncb, nb0 = J.nb[0];
    const uint32_t n_cells = J.nb[1] * ncb;
    const bool valid = cell < n_cells;
    uint32_t row = 0, cb = 0, a = 0, e = 0;
    if (valid) {
      row = cell / ncb;
      cb = cell - row * ncb;
      const uint32_t f0 = cb * J.mcf;
      const uint32_t f1 = min(f0 + J.mcf, nb0);
      const uint32_t* st = start + J.cbase + (uint64_t)row * nb0;
      a = st[f0];
      e = st[f1];
    }
    const uint32_t last = min(n_cells - cell0, (uint32_t)kBlock) - 1;
    __syncthreads();  // the previous block's readers of span and stage are done
    if (threadIdx.x == 0) span[0] = a;
    if (threadIdx.x == last) span[1] = e;
    __syncthreads();
    const uint32_t A = span[0], B = span[1];
    const double r0 = (double)fmaf((float)(cb * J.mcf), J.bs[0], J.mn[0]);
    const double r1 = (double)fmaf((float)row, J.bs[1], J.mn[1]);
    const float* x = scoord + 3 * J.pbase;
    const float* y = x + J.n;
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t base = A; base < B; base += kStage) {
      const uint32_t cnt = min(B - base, kStage);
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < cnt; i += kBlock) stage[i] = make_float2(x[base + i], y[base + i]);
      __syncthreads();
      const uint32_t p0 = max(a, base), p1 = min(e, base + cnt);
      for (uint32_t p = p0; p < p1; ++p) {
        const float2 v = stage[p - base];
        const double ex = (double)v.x - r0, gy = (double)v.y - r1;
        const double g2 = gy * gy;
        const double pw[3] = {1.0, ex, ex * ex};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          m[3 * i] += pw[i];
          m[3 * i + 1] = fma(pw[i], gy, m[3 * i + 1]);
          m[3 * i + 2] = fma(pw[i], g2, m[3 * i + 2]);
        }
      }
    }
    if (valid) {
      const double k0 = (double)J.hinv[0] * (double)J.hinv[0];
      const double k1 = (double)J.hinv[1] * (double)J.hinv[1];
      const double s0[3] = {1.0, 2.0 * k0, -k0}, s1[3] = {1.0, 2.0 * k1, -k1};
      float* out = mom + (J.mbase + cell) * 9;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) out[3 * i + k] = (float)(m[3 * i + k] * s0[i] * s1[k]);
    }
  }
}

// ------------------------------------------------------------------ tiles
// A tile starts at sorted position p when p == 0, p % tq == 0, or the
// (row, dim-0 group) of p differs from that of p-1.  Tiles never cross rows,
// have at most tq queries and span at most `group` dim-0 boxes.
namespace {
template <typename KeyT>
__device__ __forceinline__ bool tile_start(const DJob& J, KeyT a, KeyT b, uint32_t p, uint32_t tq) {
  if (p == 0 || (p % tq) == 0) return true;
  if (J.flags & kJobBrute) return false;
  // key width arithmetic (32-bit division for 32-bit keys)
  const KeyT nb0 = (KeyT)J.nb[0];
  const KeyT ra = a / nb0, rb = b / nb0;
  if (ra != rb) return true;
  return (uint32_t)(a - ra * nb0) / J.group != (uint32_t)(b - rb * nb0) / J.group;
}
static_assert(kChunkItems <= 32, "tile-start flags are one 32-bit word per lane");
// Tile-start flags of a chunk: bit i = point y + 64 i + l starts a tile.
template <typename KeyT>
__device__ __forceinline__ uint32_t chunk_tile_starts(const DJob& J, uint32_t y, const KeyT* keys,
                                                      uint32_t tq, int l) {
  constexpr int kHalf = kChunkItems / 2;
  uint32_t bits = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    KeyT cur[kHalf], prev[kHalf];
    load_key_pairs<kHalf>(J, y, h * kHalf, keys, l, cur, prev);
#pragma unroll
    for (int i = 0; i < kHalf; ++i) {
      const uint32_t p = y + (uint32_t)(h * kHalf + i) * kWave + l;
      if (p < J.n && tile_start(J, prev[i], cur[i], p, tq)) bits |= 1u << (h * kHalf + i);
    }
  }
  return bits;
}
}  // namespace

// Cell-start tables (jobs with a table) and per-chunk tile counts (every
// job) in one pass over the sorted keys.  chunks are in tile order; the
// tile count of chunk c goes to tile_cnt[c].
template <typename KeyT>
__global__ __launch_bounds__(kBlock) void k_cell_starts(const DJob* __restrict__ jobs,
                                                        const uint2* __restrict__ chunks,
                                                        uint32_t n_chunks, const KeyT* __restrict__ keys,
                                                        uint32_t tq, uint32_t* __restrict__ tile_cnt,
                                                        uint32_t* __restrict__ start) {
  const uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  const bool tab = J.cbase != kNoTable;
  const int l = lane_id();
  uint32_t* st = start + (tab ? J.cbase : 0);
  uint32_t tiles = 0;  // tile starts among the lane's points
  const KeyT* kk = keys + J.pbase;
  // Point p owns the cells (key[p-1], key[p]]; cells after the last key hold
  // n.  Short runs are written by their own lane; runs longer than
  // kLaneRun (empty stretches of a sparse grid) by the whole wave, one run
  // at a time, so one long run does not serialise on a single lane.
  constexpr uint64_t kLaneRun = 16;
  constexpr int kHalf = kChunkItems / 2;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    KeyT curk[kHalf], prevk[kHalf];
    load_key_pairs<kHalf>(J, ch.y, h * kHalf, keys, l, curk, prevk);
#pragma unroll
    for (int ii = 0; ii < kHalf; ++ii) {
      const uint32_t p0 = ch.y + (uint32_t)(h * kHalf + ii) * kWave;
      if (p0 >= J.n) continue;
      const uint32_t p = p0 + (uint32_t)l;
      uint64_t first = 0, cnt = 0;
      if (p < J.n) {
        if (tile_start(J, prevk[ii], curk[ii], p, tq)) ++tiles;
        if (tab) {
          const uint64_t cur = curk[ii];
          first = p == 0 ? 0 : (uint64_t)prevk[ii] + 1;
          cnt = cur + 1 - first;
        }
      }
      const bool wide_run = cnt > kLaneRun;
      if (!wide_run)
        for (uint64_t e = 0; e < cnt; ++e) st[first + e] = p;
      uint64_t bm = __ballot(wide_run);
      while (bm) {
        const int src = __ffsll((unsigned long long)bm) - 1;
        bm &= bm - 1;
        const uint64_t f = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(first >> 32), src) << 32) |
                           __builtin_amdgcn_readlane((uint32_t)first, src);
        const uint64_t n = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(cnt >> 32), src) << 32) |
                           __builtin_amdgcn_readlane((uint32_t)cnt, src);
        const uint32_t v = p0 + (uint32_t)src;
        for (uint64_t e = (uint64_t)l; e < n; e += kWave) st[f + e] = v;
      }
      if (tab && J.n - 1 - p0 < (uint32_t)kWave) {  // this step holds the job's last point
        const uint64_t tail = (uint64_t)kk[J.n - 1] + 1;
        for (uint64_t cell = tail + (uint64_t)l; cell <= J.cells; cell += kWave) st[cell] = J.n;
      }
    }
  }
  const uint32_t total = __builtin_amdgcn_readlane(wave_incl_scan(tiles), kWave - 1);
  if (l == 0) tile_cnt[c] = total;
}


template <typename KeyT>
__global__ __launch_bounds__(kBlock) void k_tile_write(const DJob* __restrict__ jobs,
                                                       const uint2* __restrict__ chunks,
                                                       uint32_t n_chunks,
                                                       const KeyT* __restrict__ keys,
                                                       uint32_t tq,
                                                       const uint32_t* __restrict__ offs,
                                                       DTile* __restrict__ tiles) {
  const uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  const int l = lane_id();
  const uint64_t lt = lanemask_lt();
  uint32_t o = offs[c];
  const uint32_t bits = chunk_tile_starts(J, ch.y, keys, tq, l);
#pragma unroll
  for (int i = 0; i < kChunkItems; ++i) {
    const uint32_t p = ch.y + i * kWave + l;
    const bool f = (bits >> i) & 1u;
    uint64_t b = __ballot(f);
    if (f) {
      DTile t;
      t.job = ch.x;
      t.p0 = p;
      t.p1 = 0;
      t.pack = 0;
      tiles[o + __popcll(b & lt)] = t;
    }
    o += (uint32_t)__popcll(b);
  }
}

// Tile end = next tile's start within the same job, else the job's n.
__global__ void k_tile_ends(const DJob* __restrict__ jobs, DTile* __restrict__ tiles,
                            const uint32_t* __restrict__ total) {
  const uint32_t T = *total;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
    DTile a = tiles[t];
    uint32_t end = jobs[a.job].n;
    if (t + 1 < T) {
      DTile b = tiles[t + 1];
      if (b.job == a.job) end = b.p0;
    }
    tiles[t].p1 = end;
  }
}

// Packed tiles: each thread walks a run of kPackRun consecutive tiles and
// packs greedily: four tiles of one job (d >= 2) with at most 16 queries
// each become one packed tile of four sub-tiles, else two with at most 32
// each become one of two.  The first tile of a pack takes the whole span and
// the sub-tile offsets; the others are marked kTileSkip.  Packs do not cross
// run boundaries, so the threads rewrite disjoint tiles in place.
constexpr uint32_t kPackRun = 32;
__global__ void k_tile_pack(const DJob* __restrict__ jobs, DTile* __restrict__ tiles,
                            const uint32_t* __restrict__ total) {
  const uint32_t T = *total;
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r * kPackRun < T;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t t0 = (uint32_t)(r * kPackRun);
    const uint32_t t1 = T - t0 < kPackRun ? T : t0 + kPackRun;
    // tiles [a, a + cnt) of one job (d >= 2) with at most lim queries each
    auto fits = [&](uint32_t a, uint32_t cnt, uint32_t lim) {
      if (a + cnt > t1) return false;
      const DTile first = tiles[a];
      const DJob& J = jobs[first.job];
      if (J.d < 2 || (J.flags & kJobBrute)) return false;
      // a sub-tile (one row) needs up to (2 rows per h + 1)^(d-1) candidate
      // rows, one per lane of its group of 64 / cnt lanes
      uint32_t rows = 1;
      for (uint32_t e = 1; e < J.d; ++e) rows *= 2u * (uint32_t)rintf(J.h[e] / J.bs[e]) + 1u;
      if (rows > kWave / cnt) return false;
      for (uint32_t i = a; i < a + cnt; ++i) {
        const DTile ti = tiles[i];
        if (ti.job != first.job || ti.p1 - ti.p0 > lim) return false;
      }
      return true;
    };
    auto pack = [&](uint32_t a, uint32_t cnt) {
      DTile h = tiles[a];
      uint32_t code = cnt - 1;
      for (uint32_t i = 1; i < cnt; ++i) code |= (tiles[a + i].p0 - h.p0) << (8 * i);
      h.p1 = tiles[a + cnt - 1].p1;
      h.pack = code;
      for (uint32_t i = 1; i < cnt; ++i) tiles[a + i].pack = kTileSkip;
      tiles[a] = h;
    };
    uint32_t a = t0;
    while (a < t1) {
      if (fits(a, 4, 16)) {
        pack(a, 4);
        a += 4;
      } else if (fits(a, 2, 32)) {
        pack(a, 2);
        a += 2;
      } else {
        ++a;
      }
    }
  }
}

// ---------------------------------------------------------- density core
namespace {

constexpr float kSentinel = 1.0e18f;

__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }

// First position in a[0..n) whose key is >= v (keys compared as uint64).
template <typename KeyT>
__device__ __forceinline__ uint32_t lower_bound_key(const KeyT* a, uint32_t n, uint64_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Tiles whose queries span more than this many bandwidths along dim 0 (sparse
// stretches; tiles stay inside one dim-0 group, up to 64 bandwidths wide)
// evaluate u = (x_q - x_c) / h from the raw coordinates.  Narrower tiles use
// coordinates scaled about the centre of their bounding box, s = (x - o) / h,
// and u = s_q - s_c: the FP32 rounding of s is about 1.2e-7 |s|, so with
// |s| at most 17 (16 bandwidths from the centre plus the support) u stays
// within about 4e-6 of its exact value.  (With the tile's first query as the
// origin and 64-bandwidth tiles, a sparse query's density was off by 1.6e-5.)
constexpr float kLocalSpan = 32.0f;

struct WaveLds {
  float4 cand[kWave];        // staged candidates (local coordinates)
  uint2 rng[kWave];          // packed tiles: candidate ranges (first flattened index, first position)
  float4 cell[2][kWave];     // 2-D moment cells (with cand): 9 moments and the cell origin
};

// Conservative box range [lo, hi] covering [a - h, b + h] along one dim.
__device__ __forceinline__ void box_range(float a, float b, float h, float mn, float bs, uint32_t nb,
                                          uint32_t& lo, uint32_t& hi) {
  float marg_a = (fabsf(a) + h) * 1.0e-6f;
  float marg_b = (fabsf(b) + h) * 1.0e-6f;
  lo = box_of(a - h - marg_a, mn, bs, nb);
  hi = box_of(b + h + marg_b, mn, bs, nb);
}

template <int D>
__device__ __forceinline__ void stage_store(WaveLds& L, int l, const float* s) {
  float* base = reinterpret_cast<float*>(L.cand);
  if constexpr (D == 1) base[l] = s[0];
  else if constexpr (D == 2) {
    base[2 * l] = s[0];
    base[2 * l + 1] = s[1];
  } else {
    L.cand[l] = make_float4(s[0], s[1], s[2], 0.0f);
  }
}

__device__ __forceinline__ void kahan_add(float& S, float& C, float x) {
  float y = x - C;
  float t = S + y;
  C = (t - S) - y;
  S = t;
}

// Accumulate one staged chunk of 64 candidates into the Q queries of a lane.
// RAW: stage holds raw coordinates, u = (xq - xc) * hinv (plain path).
// else: stage holds (x - o) * hinv, u = sq - sc.
// NC: candidates staged at c4 (64 for a whole chunk, 16 or 32 for the
// sub-tile of a packed tile, whose lanes read their own sub-tile's stage).
template <int D, int Q, bool RAW, bool TRI = false, int NC = kWave>
__device__ __forceinline__ void accumulate_chunk(const float4* c4, const float (&sq)[3][Q],
                                                 const float (&hinv)[3], float (&S)[Q],
                                                 float (&C)[Q]) {
  float part[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) part[q] = 0.0f;
  constexpr int kPer4 = (D == 1) ? 4 : (D == 2 ? 2 : 1);
  constexpr int kReads = NC / kPer4;  // float4 LDS reads per chunk

  // the candidates of one float4 read against the lane's queries
  auto consume = [&](const float4 v, const int g) {
    float cs[4][3];
    if constexpr (D == 1) {
      cs[0][0] = v.x; cs[1][0] = v.y; cs[2][0] = v.z; cs[3][0] = v.w;
    } else if constexpr (D == 2) {
      cs[0][0] = v.x; cs[0][1] = v.y; cs[1][0] = v.z; cs[1][1] = v.w;
    } else {
      cs[0][0] = v.x; cs[0][1] = v.y; cs[0][2] = v.z;
    }
    if constexpr (D == 1) {
      // four candidates per read: the terms are summed as a small tree so
      // the accumulator sees one dependent add per four candidates (a
      // serial chain sub -> fma -> add per candidate leaves the SIMD idle
      // between dependent instructions)
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        float w4[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          float uu = sq[0][q] - cs[c][0];
          if constexpr (RAW) uu = uu * hinv[0];
          if constexpr (TRI) w4[c] = clamp01(1.0f - fabsf(uu));
          else w4[c] = clamp01(__builtin_fmaf(-uu, uu, 1.0f));
        }
        part[q] += (w4[0] + w4[1]) + (w4[2] + w4[3]);
      }
    } else {
#pragma unroll
      for (int c = 0; c < kPer4; ++c) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          float w[3];
#pragma unroll
          for (int e = 0; e < D; ++e) {
            float uu = sq[e][q] - cs[c][e];
            if constexpr (RAW) uu = uu * hinv[e];
            w[e] = clamp01(__builtin_fmaf(-uu, uu, 1.0f));
          }
          if constexpr (D == 2) part[q] = __builtin_fmaf(w[0], w[1], part[q]);
          else part[q] = __builtin_fmaf(w[0] * w[1], w[2], part[q]);
        }
      }
    }
    // compensated merge every 32 candidates (and at the end of the stage)
    if (((g + 1) * kPer4) % 32 == 0 || g + 1 == kReads) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        kahan_add(S[q], C[q], part[q]);
        part[q] = 0.0f;
      }
    }
  };

#pragma unroll
  for (int g = 0; g < kReads; ++g) consume(c4[g], g);
}

// Moment boxes [b0, b0 + nbx) of a 1-D tile added into the lanes' queries:
// 64 boxes at a time, lane k stages box k's origin and moments in LDS, then
// every lane adds c + d (B - A d), d = xq - r, per box (padding boxes are
// zero), with a compensated merge every 8 boxes.
template <int Q>
__device__ __forceinline__ void moment_boxes(const DJob& J, const float4* __restrict__ mom, uint32_t b0,
                                             uint32_t nbx, const float (&xq)[3][Q], float (&S)[Q],
                                             float (&C)[Q], WaveLds& L) {
  const int l = lane_id();
  for (uint32_t base = 0; base < nbx; base += kWave) {
    const uint32_t cnt = min(nbx - base, (uint32_t)kWave);
    float4 m = make_float4(0.f, 0.f, 0.f, 0.f);
    float r = 0.f;
    if ((uint32_t)l < cnt) {
      const uint32_t b = b0 + base + (uint32_t)l;
      m = mom[J.mbase + b];
      r = fmaf((float)b, J.bs[0], J.mn[0]);
    }
    wave_sync();
    L.cand[l] = make_float4(r, m.x, m.y, m.z);
    wave_sync();
    const uint32_t cnt8 = (cnt + 7u) & ~7u;
    for (uint32_t k = 0; k < cnt8; k += 8) {
      float part[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) part[q] = 0.0f;
#pragma unroll 2
      for (int i = 0; i < 8; ++i) {
        const float4 v = L.cand[k + i];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const float d = xq[0][q] - v.x;
          const float t = __builtin_fmaf(-v.y, d, v.z);
          part[q] += __builtin_fmaf(t, d, v.w);
        }
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) kahan_add(S[q], C[q], part[q]);
    }
  }
}

// Moment cells of a 2-D tile (rows [r0, r0 + nr) x cells [c0, c0 + nc) of
// each row) added into the lanes' queries: 64 cells at a time, lane k stages
// cell k's moments N and origin in LDS, then every lane adds
// sum_q b_q sum_p a_p N[3p + q] per cell (padding cells are zero), with a
// compensated merge every 8 cells.
template <int Q>
__device__ __forceinline__ void cell_moments(const DJob& J, const float* __restrict__ mom2, uint32_t r0,
                                             uint32_t nr, uint32_t c0, uint32_t nc, const float (&xq)[3][Q],
                                             float (&S)[Q], float (&C)[Q], WaveLds& L) {
  const int l = lane_id();
  const uint32_t cells = nr * nc;
  const float k0 = J.hinv[0] * J.hinv[0], k1 = J.hinv[1] * J.hinv[1];
  for (uint32_t base = 0; base < cells; base += kWave) {
    const uint32_t cnt = min(cells - base, (uint32_t)kWave);
    float N[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float o0 = 0.f, o1 = 0.f;
    if ((uint32_t)l < cnt) {
      const uint32_t k = base + (uint32_t)l;
      const uint32_t ri = k / nc;
      const uint32_t row = r0 + ri, cb = c0 + (k - ri * nc);
      const float* src = mom2 + (J.mbase + (uint64_t)row * J.mncb + cb) * 9;
#pragma unroll
      for (int i = 0; i < 9; ++i) N[i] = src[i];
      o0 = fmaf((float)(cb * J.mcf), J.bs[0], J.mn[0]);
      o1 = fmaf((float)row, J.bs[1], J.mn[1]);
    }
    wave_sync();
    L.cand[l] = make_float4(N[0], N[1], N[2], N[3]);
    L.cell[0][l] = make_float4(N[4], N[5], N[6], N[7]);
    L.cell[1][l] = make_float4(N[8], o0, o1, 0.f);
    wave_sync();
    const uint32_t cnt8 = (cnt + 7u) & ~7u;
    for (uint32_t k = 0; k < cnt8; k += 8) {
      float part[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) part[q] = 0.0f;
      // one cell per step: unrolled, the LDS reads of several cells would be
      // held in registers, raising the kernel's register count (and so
      // lowering its occupancy)
#pragma unroll 1
      for (int i = 0; i < 8; ++i) {
        // a = (N00, N01, N02, N10), b = (N11, N12, N20, N21), c = (N22, r0, r1, -)
        const float4 a = L.cand[k + i], b = L.cell[0][k + i], c = L.cell[1][k + i];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const float d0 = xq[0][q] - c.y, d1 = xq[1][q] - c.z;
          const float a0 = __builtin_fmaf(-(k0 * d0), d0, 1.0f);
          const float b0 = __builtin_fmaf(-(k1 * d1), d1, 1.0f);
          const float w0 = __builtin_fmaf(a0, a.x, __builtin_fmaf(d0, a.w, b.z));
          const float w1 = __builtin_fmaf(a0, a.y, __builtin_fmaf(d0, b.x, b.w));
          const float w2 = __builtin_fmaf(a0, a.z, __builtin_fmaf(d0, b.y, c.x));
          part[q] += __builtin_fmaf(b0, w0, __builtin_fmaf(d1, w1, w2));
        }
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) kahan_add(S[q], C[q], part[q]);
    }
  }
}

// Densities of one tile's queries; returns the tile's pair-by-pair candidate
// count (wave-uniform).  mstat[0..1] (jobs with a moment table): candidates
// summed from moments and the number of those boxes or cells; mstat[2]: 1
// when the tile's pair terms were taken from raw coordinates (a wide tile).
template <int D, int Q, bool RAW, typename KeyT>
__device__ uint32_t process_tile(const DJob& J, const DTile& T, const float* __restrict__ X,
                                 const float* __restrict__ scoord, const KeyT* __restrict__ keys,
                                 const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ idx,
                                 const float4* __restrict__ mom, const float* __restrict__ mom2,
                                 double* __restrict__ dens, WaveLds& L, uint32_t (&mstat)[3]) {
  const int l = lane_id();
  const uint32_t n = J.n;
  const float* src[3];
#pragma unroll
  for (int e = 0; e < D; ++e)
    src[e] = RAW ? (X + J.off[e]) : (scoord + 3 * J.pbase + (uint64_t)e * n);
  float hinv[3] = {J.hinv[0], J.hinv[1], J.hinv[2]};
  const bool tri = (J.flags & kJobTri) != 0;

  // queries, and their output slots (loaded now so the tile's end does not
  // wait for another round trip)
  float xq[3][Q];
  bool qv[Q];
  uint32_t slot[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    uint32_t p = T.p0 + q * kWave + l;
    qv[q] = p < T.p1;
    uint32_t pp = qv[q] ? p : T.p0;
#pragma unroll
    for (int e = 0; e < D; ++e) xq[e][q] = src[e][pp];
    slot[q] = RAW ? pp : idx[J.pbase + pp];
  }

  float o[3] = {0.f, 0.f, 0.f};
  bool exact = false;  // raw-coordinate pair terms (a wide tile)
  uint32_t n_ranges = 1;
  uint32_t my_lo = 0, my_len = 0;
  uint32_t m_lo = 0, m_n = 0;  // 1-D: boxes summed from the moment table
  uint32_t c_r0 = 0, c_nr = 0, c_c0 = 0, c_nc = 0, c_pts = 0;  // 2-D: moment cells and their points
  if constexpr (RAW) {
    if (l == 0) { my_lo = 0; my_len = n; }
  } else {
    // bounding box of the tile's queries; its centre is the origin
    float bmin[3], bmax[3];
#pragma unroll
    for (int e = 0; e < D; ++e) {
      float mn = xq[e][0], mx = xq[e][0];
#pragma unroll
      for (int q = 1; q < Q; ++q) {
        mn = fminf(mn, xq[e][q]);
        mx = fmaxf(mx, xq[e][q]);
      }
      bmin[e] = wave_min(mn);
      bmax[e] = wave_max(mx);
      o[e] = 0.5f * (bmin[e] + bmax[e]);
    }
    exact = (bmax[0] - bmin[0]) * J.hinv[0] > kLocalSpan;
    uint32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int e = 0; e < D; ++e) box_range(bmin[e], bmax[e], J.h[e], J.mn[e], J.bs[e], J.nb[e], lo[e], hi[e]);
    if constexpr (D == 1) {
      if (J.flags & kJobMoments) {
        // Boxes [ilo, ihi] hold only points x with bmax - h < x < bmin + h
        // (box_of is monotone and the margin covers the rounding of the box
        // arithmetic), i.e. inside the support of every query of the tile;
        // they are summed from their moments, the boxes at either end of the
        // candidate window pair by pair.
        const float h = J.h[0];
        const float marg = (fabsf(bmin[0]) + fabsf(bmax[0]) + fabsf(J.mn[0]) + h) * 1.0e-5f;
        const uint32_t ilo = box_of(bmax[0] - h + marg, J.mn[0], J.bs[0], J.nb[0]) + 1u;
        const uint32_t ihi1 = box_of(bmin[0] + h - marg, J.mn[0], J.bs[0], J.nb[0]);
        if (ilo < ihi1) {
          m_lo = ilo;
          m_n = ihi1 - ilo;
        }
      }
    }
    uint32_t n1 = (D >= 2) ? (hi[1] - lo[1] + 1) : 1u;
    uint32_t n2 = (D == 3) ? (hi[2] - lo[2] + 1) : 1u;
    n_ranges = n1 * n2;
    if (n_ranges > (uint32_t)kMaxRanges) n_ranges = kMaxRanges + 1;  // signals the slow path
    if (m_n != 0) {
      // two pair-by-pair ranges around the moment boxes (the job has a table)
      n_ranges = 2;
      if (l < 2) {
        const uint32_t* st = cstart + J.cbase;
        my_lo = st[l == 0 ? lo[0] : m_lo + m_n];
        my_len = st[l == 0 ? m_lo : hi[0] + 1] - my_lo;
      }
    } else if (n_ranges <= (uint32_t)kMaxRanges && (uint32_t)l < n_ranges) {
      uint32_t r1 = (D >= 2) ? lo[1] + (uint32_t)l % n1 : 0u;
      uint32_t r2 = (D == 3) ? lo[2] + (uint32_t)l / n1 : 0u;
      uint64_t row = (uint64_t)r2 * J.nb[1] + r1;
      uint64_t klo = row * J.nb[0] + lo[0];
      uint64_t khi = row * J.nb[0] + hi[0];
      uint32_t a, b;
      if (J.cbase != kNoTable) {
        const uint32_t* st = cstart + J.cbase;
        a = st[klo];
        b = st[khi + 1];
      } else {
        const KeyT* kk = keys + J.pbase;
        a = lower_bound_key(kk, n, klo);
        b = a + lower_bound_key(kk + a, n - a, khi + 1);
      }
      my_lo = a;
      my_len = b - a;
    }
    if constexpr (D == 2) {
      if ((J.flags & kJobMoments) && n_ranges <= (uint32_t)kMaxRanges / 2) {
        // Cells (row, cb) of rows [jlo, jhi1) and dim-0 cells [clo, chi1)
        // hold only points inside the support of every query (as for 1-D
        // boxes, per dim).  Used when they hold at least 4 points per cell
        // on average; the interior rows' pair-by-pair candidates become two
        // ranges, left and right of the cells (the right ones on lanes
        // n1 .. n1 + nir - 1).
        const uint32_t cf = J.mcf;
        const float h0 = J.h[0], h1 = J.h[1];
        const float mg0 = (fabsf(bmin[0]) + fabsf(bmax[0]) + fabsf(J.mn[0]) + h0) * 1.0e-5f;
        const float mg1 = (fabsf(bmin[1]) + fabsf(bmax[1]) + fabsf(J.mn[1]) + h1) * 1.0e-5f;
        const uint32_t ilo = box_of(bmax[0] - h0 + mg0, J.mn[0], J.bs[0], J.nb[0]) + 1u;
        const uint32_t ihi1 = box_of(bmin[0] + h0 - mg0, J.mn[0], J.bs[0], J.nb[0]);
        const uint32_t clo = (ilo + cf - 1u) / cf, chi1 = ihi1 / cf;
        const uint32_t jlo = max(box_of(bmax[1] - h1 + mg1, J.mn[1], J.bs[1], J.nb[1]) + 1u, lo[1]);
        const uint32_t jhi1 = min(box_of(bmin[1] + h1 - mg1, J.mn[1], J.bs[1], J.nb[1]), hi[1] + 1u);
        if (clo < chi1 && jlo < jhi1) {
          const uint32_t nir = jhi1 - jlo, nc = chi1 - clo;
          const uint32_t r = lo[1] + (uint32_t)l;
          const bool inr = (uint32_t)l < n_ranges && r >= jlo && r < jhi1;
          uint32_t m1 = 0, m2 = 0;
          if (inr) {
            const uint32_t* st = cstart + J.cbase + (uint64_t)r * J.nb[0];
            m1 = st[clo * cf];
            m2 = st[chi1 * cf];
          }
          const uint32_t pts = __builtin_amdgcn_readlane(wave_incl_scan(m2 - m1), kWave - 1);
          if (pts >= 4u * nir * nc) {
            const uint32_t end = my_lo + my_len;
            const int src = ((int)(jlo - lo[1]) + l - (int)n_ranges) & (kWave - 1);
            const uint32_t s_m2 = __shfl(m2, src);
            const uint32_t s_end = __shfl(end, src);
            if (inr) my_len = m1 - my_lo;
            if ((uint32_t)l >= n_ranges && (uint32_t)l < n_ranges + nir) {
              my_lo = s_m2;
              my_len = s_end - s_m2;
            }
            n_ranges += nir;
            c_r0 = jlo;
            c_nr = nir;
            c_c0 = clo;
            c_nc = nc;
            c_pts = pts;
          }
        }
      }
    }
    if (n_ranges > (uint32_t)kMaxRanges) {
      // More candidate rows than lanes: fall back to one range spanning every
      // row between the first and last candidate row (a superset).
      uint64_t row_a = (uint64_t)lo[2] * J.nb[1] + lo[1];
      uint64_t row_b = (uint64_t)hi[2] * J.nb[1] + hi[1];
      const KeyT* kk = keys + J.pbase;
      if (l == 0) {
        const bool tab = J.cbase != kNoTable;
        uint32_t a = tab ? cstart[J.cbase + row_a * J.nb[0]] : lower_bound_key(kk, n, row_a * J.nb[0]);
        uint32_t b = tab ? cstart[J.cbase + (row_b + 1) * J.nb[0]] : lower_bound_key(kk, n, (row_b + 1) * J.nb[0]);
        my_lo = a;
        my_len = b - a;
      }
      n_ranges = 1;
    }
  }

  // flattened candidate index space over the ranges: lane k < n_ranges
  // holds range k as (first sorted position, first flattened index)
  const uint32_t incl = wave_incl_scan(my_len);
  const uint32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
  const uint32_t my_ps = incl - my_len;

  float sq[3][Q];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int e = 0; e < D; ++e) sq[e][q] = (RAW || exact) ? xq[e][q] : (xq[e][q] - o[e]) * hinv[e];

  float S[Q], C[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) { S[q] = 0.0f; C[q] = 0.0f; }

  // candidate loader for the chunk of flattened indices [cb, cb + 64): lane l
  // loads index cb + l as raw coordinates (converted to local coordinates
  // only when staged, so the load's latency overlaps the previous chunk's
  // pair loop).  Range k (read from lane k) covers [ps_k, ps_k + len_k);
  // the range of an index is the last one starting at or before it (empty
  // ranges share their successor's start and are overridden by it).  Chunks
  // come in increasing order, so the scan resumes at kc, the last range
  // starting at or before the previous chunk, and stops at the first range
  // starting past this chunk: a chunk visits the ranges that start inside it.
  uint32_t kc = 0;
  // range kc's offset and the next range's start, kept between chunks: a
  // chunk that ends before the next range starts needs no scan
  uint32_t kc_off = __builtin_amdgcn_readlane(my_lo, 0) - __builtin_amdgcn_readlane(my_ps, 0);
  uint32_t next_ps = n_ranges > 1 ? __builtin_amdgcn_readlane(my_ps, 1) : 0xFFFFFFFFu;
  auto load_cand = [&](uint32_t cb, float* x) {
    const uint32_t f = cb + (uint32_t)l;
    const uint32_t cend = cb + kWave;
    const bool cend_wraps = cend < cb;
    uint32_t off = 0;  // sorted position - flattened index, mod 2^32
    if (!cend_wraps && next_ps >= cend) {
      off = kc_off;
    } else {
      for (uint32_t k = kc; k < n_ranges; ++k) {
        const uint32_t ps_k = __builtin_amdgcn_readlane(my_ps, k);
        if (!cend_wraps && ps_k >= cend) break;
        const uint32_t lo_k = __builtin_amdgcn_readlane(my_lo, k);
        if (ps_k <= cb) kc = k;
        if (f >= ps_k) off = lo_k - ps_k;
      }
      kc_off = __builtin_amdgcn_readlane(my_lo, kc) - __builtin_amdgcn_readlane(my_ps, kc);
      next_ps = kc + 1 < n_ranges ? __builtin_amdgcn_readlane(my_ps, kc + 1) : 0xFFFFFFFFu;
    }
    if (f < total) {
      const uint32_t p = f + off;
      // 32-bit byte offsets from the dims' uniform base addresses (a job
      // holds fewer than 2^30 points, so a coordinate column is below 4 GiB)
#pragma unroll
      for (int e = 0; e < D; ++e)
        x[e] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(src[e]) + (uint32_t)(p * 4u));
    } else {
#pragma unroll
      for (int e = 0; e < D; ++e) x[e] = kSentinel;
    }
  };

  // one staged chunk: convert the loaded candidates, stage them, reload the
  // buffer `ahead` chunks further on, run the pair loop
  // exact_tag: std::true_type stages raw coordinates (u from raw differences)
  auto chunk = [&](auto exact_tag, uint32_t base, float* buf, uint32_t ahead) {
    constexpr bool kRawPairs = RAW || decltype(exact_tag)::value;
    float s_c[3];
#pragma unroll
    for (int e = 0; e < D; ++e) s_c[e] = kRawPairs ? buf[e] : (buf[e] - o[e]) * hinv[e];
    wave_sync();
    stage_store<D>(L, l, s_c);
    wave_sync();
    if (base + ahead < total) load_cand(base + ahead, buf);
    if constexpr (D == 1) {
      if (tri) accumulate_chunk<1, Q, kRawPairs, true>(L.cand, sq, hinv, S, C);
      else accumulate_chunk<1, Q, kRawPairs, false>(L.cand, sq, hinv, S, C);
    } else {
      accumulate_chunk<D, Q, kRawPairs>(L.cand, sq, hinv, S, C);
    }
  };
  float nxt[3];
  load_cand(0u, nxt);
  if constexpr (D == 1 && !RAW) {
    if (m_n != 0) {
      moment_boxes<Q>(J, mom, m_lo, m_n, xq, S, C, L);
      mstat[0] = __builtin_amdgcn_readlane(my_lo, 1) -
                 (__builtin_amdgcn_readlane(my_lo, 0) + __builtin_amdgcn_readlane(my_len, 0));
      mstat[1] = m_n;
    }
  }
  if constexpr (D == 2 && !RAW) {
    if (c_nr != 0) {
      cell_moments<Q>(J, mom2, c_r0, c_nr, c_c0, c_nc, xq, S, C, L);
      mstat[0] = c_pts;
      mstat[1] = c_nr * c_nc;
    }
  }
  if (!RAW && exact) mstat[2] = 1u;
  if (!RAW && exact)
    for (uint32_t base = 0; base < total; base += kWave) chunk(std::true_type{}, base, nxt, kWave);
  else
    for (uint32_t base = 0; base < total; base += kWave) chunk(std::false_type{}, base, nxt, kWave);
  wave_sync();

#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if (qv[q]) dens[J.out_off + slot[q]] = ((double)S[q] - (double)C[q]) * J.norm;
  }
  return total;
}

// Densities of a packed tile: G sub-tiles of at most W = 64 / G queries,
// one query per lane, lanes [s W, s W + W) on sub-tile s.  Each sub-tile has
// its own bounding box, candidate ranges (held in the wave's LDS range
// table) and stage slice of W candidates; the wave steps through the
// longest sub-tile's candidates W at a time.  Returns the executed pairs.
template <int D, int G, typename KeyT>
__device__ uint64_t process_pack(const DJob& J, const DTile& T, const float* __restrict__ scoord,
                                 const KeyT* __restrict__ keys, const uint32_t* __restrict__ cstart,
                                 const uint32_t* __restrict__ idx, double* __restrict__ dens, WaveLds& L,
                                 uint64_t& lane_slots, uint64_t& raw_tiles) {
  constexpr int W = kWave / G;
  constexpr int kPer4 = D == 2 ? 2 : 1;
  const int l = lane_id();
  const int s = l / W, j = l % W;
  const uint32_t n = J.n;
  const float* src[3];
#pragma unroll
  for (int e = 0; e < D; ++e) src[e] = scoord + 3 * J.pbase + (uint64_t)e * n;
  float hinv[3] = {J.hinv[0], J.hinv[1], J.hinv[2]};

  // this lane's sub-tile and query
  const uint32_t span = T.p1 - T.p0;
  const uint32_t b0 = s == 0 ? 0u : (T.pack >> (8 * s)) & 0xFFu;
  const uint32_t b1 = s + 1 == G ? span : (T.pack >> (8 * (s + 1))) & 0xFFu;
  const uint32_t nq = b1 - b0;
  const bool qv = (uint32_t)j < nq;
  const uint32_t p = T.p0 + b0 + (qv ? (uint32_t)j : 0u);
  float xq[3];
#pragma unroll
  for (int e = 0; e < D; ++e) xq[e] = src[e][p];
  const uint32_t slot = idx[J.pbase + p];

  // sub-tile bounding box (idle lanes repeat the sub-tile's first query);
  // its centre is the sub-tile's origin.  A pack with a sub-tile wider than
  // kLocalSpan bandwidths along dim 0 uses raw-coordinate pair terms.
  float o[3] = {0.f, 0.f, 0.f};
  uint32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  float span0 = 0.f;
#pragma unroll
  for (int e = 0; e < D; ++e) {
    const float mn = sub_min<W>(xq[e]);
    const float mx = sub_max<W>(xq[e]);
    box_range(mn, mx, J.h[e], J.mn[e], J.bs[e], J.nb[e], lo[e], hi[e]);
    o[e] = 0.5f * (mn + mx);
    if (e == 0) span0 = (mx - mn) * J.hinv[0];
  }
  const bool exact = wave_max(span0) > kLocalSpan;
  const uint32_t n1 = hi[1] - lo[1] + 1;
  const uint32_t n2 = (D == 3) ? (hi[2] - lo[2] + 1) : 1u;
  uint32_t nr = n1 * n2;
  uint32_t my_lo = 0, my_len = 0;
  const bool tab = J.cbase != kNoTable;
  const KeyT* kk = keys + J.pbase;
  if (nr <= (uint32_t)W) {
    if ((uint32_t)j < nr) {
      const uint32_t r1 = lo[1] + (uint32_t)j % n1;
      const uint32_t r2 = (D == 3) ? lo[2] + (uint32_t)j / n1 : 0u;
      const uint64_t row = (uint64_t)r2 * J.nb[1] + r1;
      const uint64_t klo = row * J.nb[0] + lo[0];
      const uint64_t khi = row * J.nb[0] + hi[0];
      uint32_t a, b;
      if (tab) {
        a = cstart[J.cbase + klo];
        b = cstart[J.cbase + khi + 1];
      } else {
        a = lower_bound_key(kk, n, klo);
        b = a + lower_bound_key(kk + a, n - a, khi + 1);
      }
      my_lo = a;
      my_len = b - a;
    }
  } else {
    // more candidate rows than lanes: one range over every row between the
    // first and last candidate row (a superset)
    if (j == 0) {
      const uint64_t row_a = (uint64_t)lo[2] * J.nb[1] + lo[1];
      const uint64_t row_b = (uint64_t)hi[2] * J.nb[1] + hi[1];
      const uint32_t a = tab ? cstart[J.cbase + row_a * J.nb[0]] : lower_bound_key(kk, n, row_a * J.nb[0]);
      const uint32_t b =
          tab ? cstart[J.cbase + (row_b + 1) * J.nb[0]] : lower_bound_key(kk, n, (row_b + 1) * J.nb[0]);
      my_lo = a;
      my_len = b - a;
    }
    nr = 1;
  }
  // flattened candidate index space of the sub-tile; the range table goes
  // to LDS, entry s W + k = (first flattened index, first position) of range k
  const uint32_t incl = sub_incl_scan<W>(my_len);
  const uint32_t total = sub_sum<W>(my_len);
  wave_sync();
  L.rng[l] = make_uint2(incl - my_len, my_lo);
  wave_sync();
  const uint32_t steps = wave_max_u32((total + W - 1) / W);

  float sq[3][1];
#pragma unroll
  for (int e = 0; e < D; ++e) sq[e][0] = exact ? xq[e] : (xq[e] - o[e]) * hinv[e];
  float S[1] = {0.0f}, C[1] = {0.0f};

  // candidate f of the sub-tile: range k is the last one starting at or
  // before f; steps come in increasing order, so kc only moves forward
  uint32_t kc = 0;
  const uint2* rng = L.rng + s * W;
  auto load_cand = [&](uint32_t step, float* x) {
    const uint32_t f = step * W + (uint32_t)j;
    if (f < total) {
      while (kc + 1 < nr && rng[kc + 1].x <= f) ++kc;
      const uint2 r = rng[kc];
      const uint32_t pos = f - r.x + r.y;
#pragma unroll
      for (int e = 0; e < D; ++e) x[e] = src[e][pos];
    } else {
#pragma unroll
      for (int e = 0; e < D; ++e) x[e] = kSentinel;
    }
  };
  const float4* stage = L.cand + s * (W / kPer4);
  float nxt[3];
  load_cand(0u, nxt);
  auto run = [&](auto exact_tag) {
    constexpr bool kRawPairs = decltype(exact_tag)::value;
    for (uint32_t b = 0; b < steps; ++b) {
      float s_c[3];
#pragma unroll
      for (int e = 0; e < D; ++e) s_c[e] = kRawPairs ? nxt[e] : (nxt[e] - o[e]) * hinv[e];
      wave_sync();
      stage_store<D>(L, l, s_c);
      wave_sync();
      if (b + 1 < steps) load_cand(b + 1, nxt);
      accumulate_chunk<D, 1, kRawPairs, false, W>(stage, sq, hinv, S, C);
    }
  };
  if (exact) {
    ++raw_tiles;
    run(std::true_type{});
  } else {
    run(std::false_type{});
  }
  wave_sync();

  if (qv) dens[J.out_off + slot] = ((double)S[0] - (double)C[0]) * J.norm;
  lane_slots += (uint64_t)kWave * W * steps;
  uint64_t pr = 0;
#pragma unroll
  for (int g = 0; g < G; ++g)
    pr += (uint64_t)__builtin_amdgcn_readlane(nq, g * W) * __builtin_amdgcn_readlane(total, g * W);
  return pr;
}

// Queries per lane chosen per tile: the smallest Q <= QMAX whose 64*Q lane
// slots hold the tile (Q = 4 steps down through 3, 2 and 1), so partly
// filled tiles do not pay for idle slots.
// pairs counts every candidate pair of the tile; macc[d-1] those summed from
// moments in d-dim jobs (d = 1, 2), macc[d+1] the (query, box) / (query,
// cell) moment terms evaluated for them, macc[4] the tiles with raw-coordinate
// pair terms.
template <int D, int QMAX, bool RAW, typename KeyT>
__device__ __forceinline__ void run_tile(const DJob& J, const DTile& T, const float* __restrict__ X,
                                         const float* __restrict__ scoord, const KeyT* __restrict__ keys,
                                         const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ idx,
                                         const float4* __restrict__ mom, const float* __restrict__ mom2,
                                         double* __restrict__ dens, WaveLds& L, uint64_t& pairs,
                                         uint64_t& lane_slots, uint64_t (&macc)[5]) {
  const uint32_t nq = T.p1 - T.p0;
  if constexpr (QMAX >= 2 && !RAW) {  // plain-path tiles are full except the last
    if (nq <= (uint32_t)(kWave * (QMAX - 1))) {
      run_tile<D, QMAX - 1, RAW, KeyT>(J, T, X, scoord, keys, cstart, idx, mom, mom2, dens, L, pairs, lane_slots,
                                       macc);
      return;
    }
  }
  uint32_t mstat[3] = {0u, 0u, 0u};
  const uint32_t total =
      process_tile<D, QMAX, RAW, KeyT>(J, T, X, scoord, keys, cstart, idx, mom, mom2, dens, L, mstat);
  pairs += (uint64_t)nq * (total + mstat[0]);
  lane_slots += (uint64_t)(kWave * QMAX) * ((total + kWave - 1) / kWave * kWave);
  if constexpr (D <= 2 && !RAW) {
    macc[D - 1] += (uint64_t)nq * mstat[0];
    macc[D + 1] += (uint64_t)nq * mstat[1];
  }
  macc[4] += mstat[2];
}

}  // namespace

// Persistent density kernel: every wave pulls tiles from the claim counters
// until the tile list is exhausted.
template <int Q, bool RAW, typename KeyT>
__global__ __launch_bounds__(kBlock, ENT_DENSITY_MIN_WAVES) void k_density(const DJob* __restrict__ jobs,
                                                    const DTile* __restrict__ tiles,
                                                    const uint32_t* __restrict__ n_tiles_p,
                                                    uint32_t* __restrict__ counter,
                                                    const float* __restrict__ X,
                                                    const float* __restrict__ scoord,
                                                    const KeyT* __restrict__ keys,
                                                    const uint32_t* __restrict__ cstart,
                                                    const uint32_t* __restrict__ idx,
                                                    const float4* __restrict__ mom,
                                                    const float* __restrict__ mom2,
                                                    double* __restrict__ dens,
                                                    unsigned long long* __restrict__ pairs) {
  __shared__ WaveLds lds[kBlock / kWave];
  WaveLds& L = lds[threadIdx.x / kWave];
  const uint32_t n_tiles = *n_tiles_p;
  const int l = lane_id();
  // candidate pairs per dimension, lane slots and the moment counts, flushed
  // once per wave
  uint64_t acc[3] = {0, 0, 0}, slots = 0, macc[5] = {0, 0, 0, 0, 0};
  // Tiles are claimed from kClaimParts contiguous partitions of the tile
  // list, each with its own counter: a wave starts on the partition of its
  // workgroup's XCD (workgroups are dealt to the XCDs round robin) and moves
  // on to the others once that one is exhausted.  One shared counter
  // serialises the claims of short tiles at the memory's atomic unit.
  const uint32_t home = blockIdx.x % kClaimParts;
  for (uint32_t k = 0; k < kClaimParts; ++k) {
    const uint32_t part = (home + k) % kClaimParts;
    const uint32_t lo = (uint32_t)((uint64_t)n_tiles * part / kClaimParts);
    const uint32_t hi = (uint32_t)((uint64_t)n_tiles * (part + 1) / kClaimParts);
    uint32_t* ctr = counter + part * kClaimStride;
    for (;;) {
      uint32_t i = 0;
      if (l == 0) i = atomicAdd(ctr, 1u);
      i = __builtin_amdgcn_readfirstlane(i);
      if (i >= hi - lo) break;
      const uint32_t t = lo + i;
      const DTile T = tiles[t];
      const uint32_t pk = __builtin_amdgcn_readfirstlane(T.pack);
      if (pk == kTileSkip) continue;
      const DJob J = jobs[T.job];
      const uint32_t d = __builtin_amdgcn_readfirstlane(J.d);
      if constexpr (!RAW) {
        if (pk & 3u) {
          const bool four = (pk & 3u) == 3u;
          if (d == 2)
            acc[1] += four ? process_pack<2, 4, KeyT>(J, T, scoord, keys, cstart, idx, dens, L, slots, macc[4])
                           : process_pack<2, 2, KeyT>(J, T, scoord, keys, cstart, idx, dens, L, slots, macc[4]);
          else
            acc[2] += four ? process_pack<3, 4, KeyT>(J, T, scoord, keys, cstart, idx, dens, L, slots, macc[4])
                           : process_pack<3, 2, KeyT>(J, T, scoord, keys, cstart, idx, dens, L, slots, macc[4]);
          continue;
        }
      }
      if (d == 1)
        run_tile<1, Q, RAW, KeyT>(J, T, X, scoord, keys, cstart, idx, mom, mom2, dens, L, acc[0], slots, macc);
      else if (d == 2)
        run_tile<2, Q, RAW, KeyT>(J, T, X, scoord, keys, cstart, idx, mom, mom2, dens, L, acc[1], slots, macc);
      else
        run_tile<3, Q, RAW, KeyT>(J, T, X, scoord, keys, cstart, idx, mom, mom2, dens, L, acc[2], slots, macc);
    }
  }
  if (l == 0) {
    for (int e = 0; e < 3; ++e)
      if (acc[e]) atomicAdd(pairs + e, (unsigned long long)acc[e]);
    if (slots) atomicAdd(pairs + 3, (unsigned long long)slots);
    for (int e = 0; e < 5; ++e)
      if (macc[e]) atomicAdd(pairs + 4 + e, (unsigned long long)macc[e]);
  }
}

// ---------------------------------------------------------------- launchers
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, uint64_t n) {
  if (n == 0) return;
  uint64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(kBlock), 0, s, p, v, n);
}

void launch_init_minmax(hipStream_t s, uint32_t* mm, uint32_t n_jobs) {
  uint32_t n = n_jobs * 6;
  hipLaunchKernelGGL(k_init_minmax, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, mm, n_jobs);
}

static inline dim3 chunk_grid(uint32_t n_chunks) {
  return dim3((n_chunks + (kBlock / kWave) - 1) / (kBlock / kWave));
}

void launch_minmax(hipStream_t s, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                   const float* X, uint32_t* mm, uint32_t* bad) {
  hipLaunchKernelGGL(k_minmax, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, n_chunks, X, mm, bad);
}

void launch_cell_keys(hipStream_t s, bool wide, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                      const float* X, void* keys, uint32_t* idx) {
  if (wide)
    hipLaunchKernelGGL(k_cell_keys<uint64_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, n_chunks,
                       X, (uint64_t*)keys, idx);
  else
    hipLaunchKernelGGL(k_cell_keys<uint32_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, n_chunks,
                       X, (uint32_t*)keys, idx);
}

void launch_bin_ids(hipStream_t s, const DJob* jobs, const float* X, int32_t* box, uint32_t n) {
  uint32_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_bin_ids, dim3(blocks), dim3(kBlock), 0, s, jobs, X, box);
}

void launch_radix_hist(hipStream_t s, bool wide, const DJob* jobs, const uint2* chunks,
                       const uint32_t* chunk_first, uint32_t n_chunks, const void* keys, int shift,
                       uint32_t* hist) {
  if (wide)
    hipLaunchKernelGGL(k_radix_hist<uint64_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       chunk_first, n_chunks, (const uint64_t*)keys, shift, hist);
  else
    hipLaunchKernelGGL(k_radix_hist<uint32_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       chunk_first, n_chunks, (const uint32_t*)keys, shift, hist);
}

void launch_radix_scatter(hipStream_t s, bool wide, bool staged, const DJob* jobs, const uint2* chunks,
                          const uint32_t* chunk_first, uint32_t n_chunks, const void* kin,
                          const uint32_t* iin, void* kout, uint32_t* iout, int shift, const uint32_t* offs,
                          const float* X, float* scoord) {
  const bool gather = scoord != nullptr;
#define ENT_SCATTER(KT, G)                                                                                     \
  do {                                                                                                         \
    if (staged)                                                                                                \
      hipLaunchKernelGGL((k_radix_scatter_lds<KT, G>), dim3(n_chunks), dim3(kWave), 0, s, jobs, chunks,      \
                         chunk_first, n_chunks, (const KT*)kin, iin, (KT*)kout, iout, shift, offs, X, scoord); \
    else                                                                                                       \
      hipLaunchKernelGGL((k_radix_scatter<KT, G>), chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,   \
                         chunk_first, n_chunks, (const KT*)kin, iin, (KT*)kout, iout, shift, offs, X, scoord); \
  } while (0)
  if (wide) {
    if (gather) ENT_SCATTER(uint64_t, true);
    else ENT_SCATTER(uint64_t, false);
  } else {
    if (gather) ENT_SCATTER(uint32_t, true);
    else ENT_SCATTER(uint32_t, false);
  }
#undef ENT_SCATTER
}

void launch_major_keys(hipStream_t s, bool wide, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                       const uint64_t* bbase, const uint32_t* bidx, const float* X, void* keys,
                       uint32_t* idx, float* scoord) {
  if (wide)
    hipLaunchKernelGGL(k_major_keys<uint64_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       n_chunks, bbase, bidx, X, (uint64_t*)keys, idx, scoord);
  else
    hipLaunchKernelGGL(k_major_keys<uint32_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       n_chunks, bbase, bidx, X, (uint32_t*)keys, idx, scoord);
}

uint64_t scan_scratch_elems(uint64_t m) {
  uint64_t total = 0;
  while (m > 1) {
    uint64_t nb = (m + kScanTile - 1) / kScanTile;
    total += nb + 1;
    m = nb;
  }
  return total + 1;
}

// Exclusive scan of in[0..m) into out[0..m) (in == out allowed), total into
// out[m].  scratch holds nb+1 block sums per level.
void launch_exclusive_scan(hipStream_t s, const uint32_t* in, uint32_t* out, uint64_t m,
                           uint32_t* scratch) {
  if (m == 0) return;
  uint64_t nb = (m + kScanTile - 1) / kScanTile;
  uint32_t* sums = scratch;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(kBlock), 0, s, in, out, m, sums);
  if (nb > 1) {
    // scans the block sums in place; their total lands in sums[nb]
    launch_exclusive_scan(s, sums, sums, nb, scratch + nb + 1);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kBlock), 0, s, out, m, sums);
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, s, sums + nb, out, m);
  } else {
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, s, sums, out, m);
  }
}

void launch_gather_sorted(hipStream_t s, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                          const float* X, const uint32_t* idx, float* scoord) {
  hipLaunchKernelGGL(k_gather_sorted, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, n_chunks, X,
                     idx, scoord);
}

void launch_tile_write(hipStream_t s, bool wide, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                       const void* keys, uint32_t tq, const uint32_t* offs, DTile* tiles) {
  if (wide)
    hipLaunchKernelGGL(k_tile_write<uint64_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       n_chunks, (const uint64_t*)keys, tq, offs, tiles);
  else
    hipLaunchKernelGGL(k_tile_write<uint32_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       n_chunks, (const uint32_t*)keys, tq, offs, tiles);
}

void launch_tile_ends(hipStream_t s, const DJob* jobs, DTile* tiles, const uint32_t* total,
                      uint64_t max_tiles) {
  uint64_t blocks = (max_tiles + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_tile_ends, dim3((unsigned)blocks), dim3(kBlock), 0, s, jobs, tiles, total);
}
void launch_tile_pack(hipStream_t s, const DJob* jobs, DTile* tiles, const uint32_t* total,
                      uint64_t max_tiles) {
  uint64_t blocks = (max_tiles / kPackRun + kBlock) / kBlock;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_tile_pack, dim3((unsigned)blocks), dim3(kBlock), 0, s, jobs, tiles, total);
}

// Kernel variants: RAW (plain path, keys unused) or cell lists with 32- or
// 64-bit keys, for q = 1, 2, 4, 8 queries per lane.
template <int Q>
static const void* density_fn(bool raw, bool wide) {
  if (raw) return reinterpret_cast<const void*>(&k_density<Q, true, uint32_t>);
  if (wide) return reinterpret_cast<const void*>(&k_density<Q, false, uint64_t>);
  return reinterpret_cast<const void*>(&k_density<Q, false, uint32_t>);
}

int density_occupancy(int q, bool raw, bool wide, int* blocks_per_cu) {
  const void* fn = nullptr;
  switch (q) {
    case 1: fn = density_fn<1>(raw, wide); break;
    case 2: fn = density_fn<2>(raw, wide); break;
    case 4: fn = density_fn<4>(raw, wide); break;
    default: return -1;
  }
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, kBlock, 0);
  return e == hipSuccess ? 0 : -1;
}

template <int Q>
static void launch_density_q(hipStream_t s, bool raw, bool wide, unsigned grid, const DJob* jobs,
                             const DTile* tiles, const uint32_t* n_tiles, uint32_t* counter, const float* X,
                             const float* scoord, const void* keys, const uint32_t* cstart,
                             const uint32_t* idx, const float4* mom, const float* mom2, double* dens,
                             unsigned long long* pairs) {
  if (raw)
    hipLaunchKernelGGL((k_density<Q, true, uint32_t>), dim3(grid), dim3(kBlock), 0, s, jobs, tiles, n_tiles,
                       counter, X, scoord, (const uint32_t*)keys, cstart, idx, mom, mom2, dens, pairs);
  else if (wide)
    hipLaunchKernelGGL((k_density<Q, false, uint64_t>), dim3(grid), dim3(kBlock), 0, s, jobs, tiles, n_tiles,
                       counter, X, scoord, (const uint64_t*)keys, cstart, idx, mom, mom2, dens, pairs);
  else
    hipLaunchKernelGGL((k_density<Q, false, uint32_t>), dim3(grid), dim3(kBlock), 0, s, jobs, tiles, n_tiles,
                       counter, X, scoord, (const uint32_t*)keys, cstart, idx, mom, mom2, dens, pairs);
}

int launch_density(hipStream_t s, int q, bool raw, bool wide, unsigned grid, const DJob* jobs,
                   const DTile* tiles, const uint32_t* n_tiles, uint32_t* counter, const float* X,
                   const float* scoord, const void* keys, const uint32_t* cstart, const uint32_t* idx,
                   const float4* mom, const float* mom2, double* dens, unsigned long long* pairs) {
#define ENT_LD(QQ)                                                                                        \
  launch_density_q<QQ>(s, raw, wide, grid, jobs, tiles, n_tiles, counter, X, scoord, keys, cstart, idx, mom, \
                       mom2, dens, pairs)
  switch (q) {
    case 1: ENT_LD(1); break;
    case 2: ENT_LD(2); break;
    case 4: ENT_LD(4); break;
    default: return -1;
  }
#undef ENT_LD
  return 0;
}

void launch_box_moments(hipStream_t s, const DJob* jobs, const uint32_t* mlist, uint32_t max_waves,
                        const float* scoord, const uint32_t* start, float4* mom) {
  uint32_t blocks = (max_waves + (kBlock / kWave) - 1) / (kBlock / kWave);
  if (blocks > 2048) blocks = 2048;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_box_moments, dim3(blocks), dim3(kBlock), 0, s, jobs, mlist, scoord, start, mom);
}

void launch_cell_moments2(hipStream_t s, const DJob* jobs, const uint32_t* mlist, uint32_t max_blocks,
                          const float* scoord, const uint32_t* start, float* mom) {
  uint32_t blocks = max_blocks;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_cell_moments2, dim3(blocks), dim3(kBlock), 0, s, jobs, mlist, scoord, start, mom);
}

void launch_cell_starts(hipStream_t s, bool wide, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                        const void* keys, uint32_t tq, uint32_t* tile_cnt, uint32_t* start) {
  if (wide)
    hipLaunchKernelGGL(k_cell_starts<uint64_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       n_chunks, (const uint64_t*)keys, tq, tile_cnt, start);
  else
    hipLaunchKernelGGL(k_cell_starts<uint32_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       n_chunks, (const uint32_t*)keys, tq, tile_cnt, start);
}

}  // namespace ent
