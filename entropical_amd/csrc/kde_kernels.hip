// kde_kernels.hip -- gfx950 kernels of the KDE engine.
//
// Every chunked pass gives one wavefront one chunk of kChunk consecutive
// points of one job; a 256-thread workgroup carries four independent chunks.
// The chunk table (job id, job-local start) is built on the host.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "kde_engine.hpp"
#include "kde_kernels.hpp"

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

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
  return v;
}

// Inclusive wave scan of uint32.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = __shfl_up(v, o, kWave);
    if (l >= o) v += t;
  }
  return v;
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
__global__ __launch_bounds__(kBlock) void k_cell_keys(const DJob* __restrict__ jobs,
                                                      const uint2* __restrict__ chunks,
                                                      uint32_t n_chunks, const float* __restrict__ X,
                                                      uint64_t* __restrict__ keys,
                                                      uint32_t* __restrict__ idx) {
  const uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob J = jobs[ch.x];
  const int l = lane_id();
  for (int i = 0; i < kChunkItems; ++i) {
    uint32_t k = ch.y + i * kWave + l;
    if (k < J.n) {
      float v[3] = {0.f, 0.f, 0.f};
      for (uint32_t e = 0; e < J.d; ++e) v[e] = X[J.off[e] + k];
      keys[J.pbase + k] = cell_key(J, v);
      idx[J.pbase + k] = k;
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
__global__ __launch_bounds__(kBlock) void k_radix_hist(const DJob* __restrict__ jobs,
                                                       const uint2* __restrict__ chunks,
                                                       const uint32_t* __restrict__ chunk_first,
                                                       uint32_t n_chunks,
                                                       const uint64_t* __restrict__ keys, int shift,
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
  for (int i = 0; i < kChunkItems; ++i) {
    uint32_t k = ch.y + i * kWave + l;
    if (k < J.n) {
      uint32_t dg = (uint32_t)(keys[J.pbase + k] >> shift) & (kRadixBins - 1);
      atomicAdd(&bins[w][dg], 1u);
    }
  }
  wave_sync();
  const uint32_t c0 = chunk_first[ch.x];
  const uint32_t nch = chunk_first[ch.x + 1] - c0;
  for (int b = l; b < kRadixBins; b += kWave)
    hist[(uint64_t)kRadixBins * c0 + (uint64_t)b * nch + (c - c0)] = bins[w][b];
}

// Stable scatter: a wave walks its chunk in order, 64 elements at a time;
// equal digits inside a step are ranked by lane via 8 ballots.
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const DJob* __restrict__ jobs, const uint2* __restrict__ chunks,
    const uint32_t* __restrict__ chunk_first, uint32_t n_chunks,
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ idx_in,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ idx_out, int shift,
    const uint32_t* __restrict__ offs) {
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
    uint64_t key = 0;
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
      keys_out[dst] = key;
      idx_out[dst] = id;
    }
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
  const DJob J = jobs[ch.x];
  const int l = lane_id();
  for (int i = 0; i < kChunkItems; ++i) {
    uint32_t p = ch.y + i * kWave + l;
    if (p < J.n) {
      uint32_t k = idx[J.pbase + p];
      for (uint32_t e = 0; e < J.d; ++e)
        scoord[3 * J.pbase + (uint64_t)e * J.n + p] = X[J.off[e] + k];
    }
  }
}

// ------------------------------------------------------------------ tiles
// A tile starts at sorted position p when p == 0, p % tq == 0, or the
// (row, dim-0 group) of p differs from that of p-1.  Tiles never cross rows,
// have at most tq queries and span at most `group` dim-0 boxes.
namespace {
__device__ __forceinline__ bool tile_start(const DJob& J, const uint64_t* keys, uint32_t p,
                                           uint32_t tq) {
  if (p == 0 || (p % tq) == 0) return true;
  if (J.flags & kJobBrute) return false;
  uint64_t a = keys[J.pbase + p - 1], b = keys[J.pbase + p];
  uint64_t ra = a / J.nb[0], rb = b / J.nb[0];
  if (ra != rb) return true;
  return ((a - ra * J.nb[0]) / J.group) != ((b - rb * J.nb[0]) / J.group);
}
}  // namespace

__global__ __launch_bounds__(kBlock) void k_tile_count(const DJob* __restrict__ jobs,
                                                       const uint2* __restrict__ chunks,
                                                       uint32_t n_chunks,
                                                       const uint64_t* __restrict__ keys,
                                                       uint32_t tq, uint32_t* __restrict__ cnt) {
  const uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  const int l = lane_id();
  uint32_t total = 0;
  for (int i = 0; i < kChunkItems; ++i) {
    uint32_t p = ch.y + i * kWave + l;
    bool f = (p < J.n) && tile_start(J, keys, p, tq);
    total += (uint32_t)__popcll(__ballot(f));
  }
  if (l == 0) cnt[c] = total;
}

__global__ __launch_bounds__(kBlock) void k_tile_write(const DJob* __restrict__ jobs,
                                                       const uint2* __restrict__ chunks,
                                                       uint32_t n_chunks,
                                                       const uint64_t* __restrict__ keys,
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
  for (int i = 0; i < kChunkItems; ++i) {
    uint32_t p = ch.y + i * kWave + l;
    bool f = (p < J.n) && tile_start(J, keys, p, tq);
    uint64_t b = __ballot(f);
    if (f) {
      DTile t;
      t.job = ch.x;
      t.p0 = p;
      t.p1 = 0;
      t.pad = 0;
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

// ---------------------------------------------------------- density core
namespace {

constexpr float kSentinel = 1.0e18f;

__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }

__device__ __forceinline__ uint32_t lower_bound_u64(const uint64_t* a, uint32_t n, uint64_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

struct WaveLds {
  float4 cand[kWave];        // staged candidates (local coordinates)
  uint32_t r_lo[kMaxRanges]; // candidate range start (sorted position)
  uint32_t r_ps[kMaxRanges]; // flattened prefix start
  uint32_t r_pe[kMaxRanges]; // flattened prefix end
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
template <int D, int Q, bool RAW, bool TRI = false>
__device__ __forceinline__ void accumulate_chunk(const WaveLds& L, const float (&sq)[3][Q],
                                                 const float (&hinv)[3], float (&S)[Q],
                                                 float (&C)[Q]) {
  const float4* c4 = L.cand;
  float part[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) part[q] = 0.0f;
  constexpr int kPer4 = (D == 1) ? 4 : (D == 2 ? 2 : 1);
#pragma unroll
  for (int j4 = 0; j4 < kWave / kPer4; ++j4) {
    float4 v = c4[j4];
    float cs[4][3];
    if constexpr (D == 1) {
      cs[0][0] = v.x; cs[1][0] = v.y; cs[2][0] = v.z; cs[3][0] = v.w;
    } else if constexpr (D == 2) {
      cs[0][0] = v.x; cs[0][1] = v.y; cs[1][0] = v.z; cs[1][1] = v.w;
    } else {
      cs[0][0] = v.x; cs[0][1] = v.y; cs[0][2] = v.z;
    }
#pragma unroll
    for (int c = 0; c < kPer4; ++c) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        float w[3];
#pragma unroll
        for (int e = 0; e < D; ++e) {
          float u = sq[e][q] - cs[c][e];
          if constexpr (RAW) u = u * hinv[e];
          if constexpr (TRI) w[e] = clamp01(1.0f - fabsf(u));
          else w[e] = clamp01(__builtin_fmaf(-u, u, 1.0f));
        }
        if constexpr (D == 1) part[q] += w[0];
        else if constexpr (D == 2) part[q] = __builtin_fmaf(w[0], w[1], part[q]);
        else part[q] = __builtin_fmaf(w[0] * w[1], w[2], part[q]);
      }
    }
    // compensated merge every 32 candidates
    if (((j4 + 1) * kPer4) % 32 == 0) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        kahan_add(S[q], C[q], part[q]);
        part[q] = 0.0f;
      }
    }
  }
}

template <int D, int Q, bool RAW>
__device__ void process_tile(const DJob& J, const DTile& T, const float* __restrict__ X,
                             const float* __restrict__ scoord, const uint64_t* __restrict__ keys,
                             const uint32_t* __restrict__ idx, double* __restrict__ dens,
                             unsigned long long* __restrict__ pairs, WaveLds& L) {
  const int l = lane_id();
  const uint32_t n = J.n;
  const float* src[3];
#pragma unroll
  for (int e = 0; e < D; ++e)
    src[e] = RAW ? (X + J.off[e]) : (scoord + 3 * J.pbase + (uint64_t)e * n);
  float hinv[3] = {J.hinv[0], J.hinv[1], J.hinv[2]};
  const bool tri = (J.flags & kJobTri) != 0;

  // queries
  float xq[3][Q];
  bool qv[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    uint32_t p = T.p0 + q * kWave + l;
    qv[q] = p < T.p1;
    uint32_t pp = qv[q] ? p : T.p0;
#pragma unroll
    for (int e = 0; e < D; ++e) xq[e][q] = src[e][pp];
  }

  float o[3] = {0.f, 0.f, 0.f};
  uint32_t n_ranges = 1;
  uint32_t my_lo = 0, my_len = 0;
  if constexpr (RAW) {
    if (l == 0) { my_lo = 0; my_len = n; }
  } else {
    // origin = first query of the tile; bounding box of the tile's queries
    float bmin[3], bmax[3];
#pragma unroll
    for (int e = 0; e < D; ++e) {
      o[e] = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(xq[e][0])));
      float mn = xq[e][0], mx = xq[e][0];
#pragma unroll
      for (int q = 1; q < Q; ++q) {
        mn = fminf(mn, xq[e][q]);
        mx = fmaxf(mx, xq[e][q]);
      }
      bmin[e] = wave_min(mn);
      bmax[e] = wave_max(mx);
    }
    uint32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int e = 0; e < D; ++e) box_range(bmin[e], bmax[e], J.h[e], J.mn[e], J.bs[e], J.nb[e], lo[e], hi[e]);
    uint32_t n1 = (D >= 2) ? (hi[1] - lo[1] + 1) : 1u;
    uint32_t n2 = (D == 3) ? (hi[2] - lo[2] + 1) : 1u;
    n_ranges = n1 * n2;
    if (n_ranges > (uint32_t)kMaxRanges) n_ranges = kMaxRanges + 1;  // signals the slow path
    if (n_ranges <= (uint32_t)kMaxRanges && (uint32_t)l < n_ranges) {
      uint32_t r1 = (D >= 2) ? lo[1] + (uint32_t)l % n1 : 0u;
      uint32_t r2 = (D == 3) ? lo[2] + (uint32_t)l / n1 : 0u;
      uint64_t row = (uint64_t)r2 * J.nb[1] + r1;
      uint64_t klo = row * J.nb[0] + lo[0];
      uint64_t khi = row * J.nb[0] + hi[0];
      const uint64_t* kk = keys + J.pbase;
      uint32_t a = lower_bound_u64(kk, n, klo);
      uint32_t b = lower_bound_u64(kk, n, khi + 1);
      my_lo = a;
      my_len = b - a;
    }
    if (n_ranges > (uint32_t)kMaxRanges) {
      // More candidate rows than lanes: fall back to one range spanning every
      // row between the first and last candidate row (a superset).
      uint64_t row_a = (uint64_t)lo[2] * J.nb[1] + lo[1];
      uint64_t row_b = (uint64_t)hi[2] * J.nb[1] + hi[1];
      const uint64_t* kk = keys + J.pbase;
      if (l == 0) {
        uint32_t a = lower_bound_u64(kk, n, row_a * J.nb[0]);
        uint32_t b = lower_bound_u64(kk, n, (row_b + 1) * J.nb[0]);
        my_lo = a;
        my_len = b - a;
      }
      n_ranges = 1;
    }
  }

  // flattened candidate index space over the ranges
  uint32_t incl = wave_incl_scan(my_len);
  uint32_t total = __shfl(incl, kWave - 1, kWave);
  if (l == 0) atomicAdd(pairs + (D - 1), (unsigned long long)(T.p1 - T.p0) * total);
  if ((uint32_t)l < n_ranges) {
    L.r_lo[l] = my_lo;
    L.r_ps[l] = incl - my_len;
    L.r_pe[l] = incl;
  }
  wave_sync();

  float sq[3][Q];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int e = 0; e < D; ++e) sq[e][q] = RAW ? xq[e][q] : (xq[e][q] - o[e]) * hinv[e];

  float S[Q], C[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) { S[q] = 0.0f; C[q] = 0.0f; }

  // candidate loader: flattened index f -> staged local coordinates
  auto load_cand = [&](uint32_t f, float* s) {
    if (f < total) {
      uint32_t r = 0;
      if (n_ranges > 1) {
        // first range whose prefix end exceeds f
        uint32_t lo_r = 0, hi_r = n_ranges - 1;
        while (lo_r < hi_r) {
          uint32_t mid = (lo_r + hi_r) >> 1;
          if (L.r_pe[mid] <= f) lo_r = mid + 1;
          else hi_r = mid;
        }
        r = lo_r;
      }
      uint32_t p = L.r_lo[r] + (f - L.r_ps[r]);
#pragma unroll
      for (int e = 0; e < D; ++e) {
        float x = src[e][p];
        s[e] = RAW ? x : (x - o[e]) * hinv[e];
      }
    } else {
#pragma unroll
      for (int e = 0; e < D; ++e) s[e] = kSentinel;
    }
  };

  float nxt[3];
  load_cand((uint32_t)l, nxt);
  for (uint32_t base = 0; base < total; base += kWave) {
    wave_sync();
    stage_store<D>(L, l, nxt);
    wave_sync();
    if (base + kWave < total) load_cand(base + kWave + (uint32_t)l, nxt);
    if constexpr (D == 1) {
      if (tri) accumulate_chunk<1, Q, RAW, true>(L, sq, hinv, S, C);
      else accumulate_chunk<1, Q, RAW, false>(L, sq, hinv, S, C);
    } else {
      accumulate_chunk<D, Q, RAW>(L, sq, hinv, S, C);
    }
  }
  wave_sync();

#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if (qv[q]) {
      uint32_t p = T.p0 + q * kWave + l;
      uint32_t k = RAW ? p : idx[J.pbase + p];
      dens[J.out_off + k] = ((double)S[q] - (double)C[q]) * J.norm;
    }
  }
}

}  // namespace

// Persistent density kernel: every wave pulls tiles from a global counter
// until the tile list is exhausted.
template <int Q, bool RAW>
__global__ __launch_bounds__(kBlock) void k_density(const DJob* __restrict__ jobs,
                                                    const DTile* __restrict__ tiles,
                                                    const uint32_t* __restrict__ n_tiles_p,
                                                    uint32_t* __restrict__ counter,
                                                    const float* __restrict__ X,
                                                    const float* __restrict__ scoord,
                                                    const uint64_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ idx,
                                                    double* __restrict__ dens,
                                                    unsigned long long* __restrict__ pairs) {
  __shared__ WaveLds lds[kBlock / kWave];
  WaveLds& L = lds[threadIdx.x / kWave];
  const uint32_t n_tiles = *n_tiles_p;
  const int l = lane_id();
  for (;;) {
    uint32_t t = 0;
    if (l == 0) t = atomicAdd(counter, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= n_tiles) break;
    const DTile T = tiles[t];
    const DJob J = jobs[T.job];
    const uint32_t d = __builtin_amdgcn_readfirstlane(J.d);
    if (d == 1) process_tile<1, Q, RAW>(J, T, X, scoord, keys, idx, dens, pairs, L);
    else if (d == 2) process_tile<2, Q, RAW>(J, T, X, scoord, keys, idx, dens, pairs, L);
    else process_tile<3, Q, RAW>(J, T, X, scoord, keys, idx, dens, pairs, L);
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

void launch_cell_keys(hipStream_t s, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                      const float* X, uint64_t* keys, uint32_t* idx) {
  hipLaunchKernelGGL(k_cell_keys, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, n_chunks, X, keys, idx);
}

void launch_bin_ids(hipStream_t s, const DJob* jobs, const float* X, int32_t* box, uint32_t n) {
  uint32_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_bin_ids, dim3(blocks), dim3(kBlock), 0, s, jobs, X, box);
}

void launch_radix_hist(hipStream_t s, const DJob* jobs, const uint2* chunks, const uint32_t* chunk_first,
                       uint32_t n_chunks, const uint64_t* keys, int shift, uint32_t* hist) {
  hipLaunchKernelGGL(k_radix_hist, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, chunk_first,
                     n_chunks, keys, shift, hist);
}

void launch_radix_scatter(hipStream_t s, const DJob* jobs, const uint2* chunks,
                          const uint32_t* chunk_first, uint32_t n_chunks, const uint64_t* kin,
                          const uint32_t* iin, uint64_t* kout, uint32_t* iout, int shift,
                          const uint32_t* offs) {
  hipLaunchKernelGGL(k_radix_scatter, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, chunk_first,
                     n_chunks, kin, iin, kout, iout, shift, offs);
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

void launch_tile_count(hipStream_t s, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                       const uint64_t* keys, uint32_t tq, uint32_t* cnt) {
  hipLaunchKernelGGL(k_tile_count, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, n_chunks, keys,
                     tq, cnt);
}

void launch_tile_write(hipStream_t s, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                       const uint64_t* keys, uint32_t tq, const uint32_t* offs, DTile* tiles) {
  hipLaunchKernelGGL(k_tile_write, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, n_chunks, keys,
                     tq, offs, tiles);
}

void launch_tile_ends(hipStream_t s, const DJob* jobs, DTile* tiles, const uint32_t* total,
                      uint64_t max_tiles) {
  uint64_t blocks = (max_tiles + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_tile_ends, dim3((unsigned)blocks), dim3(kBlock), 0, s, jobs, tiles, total);
}

template <int Q, bool RAW>
static const void* density_fn() {
  return reinterpret_cast<const void*>(&k_density<Q, RAW>);
}

int density_occupancy(int q, bool raw, int* blocks_per_cu) {
  const void* fn = nullptr;
  switch (q) {
    case 1: fn = raw ? density_fn<1, true>() : density_fn<1, false>(); break;
    case 2: fn = raw ? density_fn<2, true>() : density_fn<2, false>(); break;
    case 4: fn = raw ? density_fn<4, true>() : density_fn<4, false>(); break;
    case 8: fn = raw ? density_fn<8, true>() : density_fn<8, false>(); break;
    default: return -1;
  }
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, kBlock, 0);
  return e == hipSuccess ? 0 : -1;
}

int launch_density(hipStream_t s, int q, bool raw, unsigned grid, const DJob* jobs, const DTile* tiles,
                   const uint32_t* n_tiles, uint32_t* counter, const float* X, const float* scoord,
                   const uint64_t* keys, const uint32_t* idx, double* dens,
                   unsigned long long* pairs) {
#define ENT_LAUNCH_DENSITY(QQ, RR)                                                                    \
  hipLaunchKernelGGL((k_density<QQ, RR>), dim3(grid), dim3(kBlock), 0, s, jobs, tiles, n_tiles, counter, \
                     X, scoord, keys, idx, dens, pairs)
  switch (q) {
    case 1: if (raw) ENT_LAUNCH_DENSITY(1, true); else ENT_LAUNCH_DENSITY(1, false); break;
    case 2: if (raw) ENT_LAUNCH_DENSITY(2, true); else ENT_LAUNCH_DENSITY(2, false); break;
    case 4: if (raw) ENT_LAUNCH_DENSITY(4, true); else ENT_LAUNCH_DENSITY(4, false); break;
    case 8: if (raw) ENT_LAUNCH_DENSITY(8, true); else ENT_LAUNCH_DENSITY(8, false); break;
    default: return -1;
  }
#undef ENT_LAUNCH_DENSITY
  return 0;
}

}  // namespace ent
