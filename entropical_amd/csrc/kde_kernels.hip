// kde_kernels.hip -- gfx950 kernels of the KDE engine.
//
// Every chunked pass gives one wavefront one chunk of kChunk consecutive
// points of one job; a 256-thread workgroup carries four independent chunks.
// The chunk table (job id, job-local start) is built on the host.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "kde_engine.hpp"
#include "kde_kernels.hpp"
#include "kde_device.hpp"

namespace ent {

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
    const uint32_t* __restrict__ offs, const float* __restrict__ X, float* __restrict__ scoord,
    const uint64_t* __restrict__ bbase, const uint32_t* __restrict__ bidx) {
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
      id = idx_in ? idx_in[J.pbase + k] : bidx[bbase[ch.x] + k];
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

// Exclusive prefix of each thread's v over the workgroup (wave scans, then
// the kBlock / kWave wave totals through LDS: one barrier) and the
// workgroup's total.  The caller synchronises before wsum is written again.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t& total) {
  const uint32_t incl = wave_incl_scan(v);
  const uint32_t w = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) wsum[w] = incl;
  __syncthreads();
  uint32_t off = 0;
  total = 0;
#pragma unroll
  for (uint32_t i = 0; i < (uint32_t)(kBlock / kWave); ++i) {
    const uint32_t t = wsum[i];
    off += i < w ? t : 0u;
    total += t;
  }
  return off + incl - v;
}

// LDS-staged stable scatter, one workgroup of kBlock threads per chunk.
// Phase 1 ranks the chunk's elements by digit into LDS, so the chunk sits
// there sorted by digit: wave w ranks the chunk's rounds [w R, w R + R) of
// 64 elements (R = kChunkItems / waves; ballots give each element its rank
// among the round's equal digits, a per-wave running count per digit its
// rank within the wave), and each wave's run of a digit starts after those
// of the waves before it, so equal digits keep the point order (stable).
// Phase 2 streams the chunk out in that order: consecutive threads write
// consecutive global slots within each digit's run, instead of unrelated
// addresses per store.  The chunk's per-digit counts come from the scanned
// offsets: offs[i + 1] - offs[i] (the scan is contiguous over all jobs and
// ends with the grand total at offs[m]).
// With GATHER (last pass) the full cell key is recomputed from the gathered
// coordinates, so a pass over a partial key (the two-stage sort) also ends
// with full keys.
constexpr int kScatterWaves = kBlock / kWave;
constexpr int kScatterRounds = kChunkItems / kScatterWaves;
static_assert(kChunkItems % kScatterWaves == 0, "rounds split evenly over the waves");
static_assert(kRadixBins == kBlock, "one digit per thread");
// Output phase of a gathering scatter pass for jobs of dimension D: staged
// points leave in digit order, kB per thread at a time (positions from LDS,
// then the coordinate gathers, then the stores); thread t of T.
template <int D, typename KeyT>
__device__ __forceinline__ void scatter_gather_out(const DJob& J, const KeyT* skey, const uint32_t* sid,
                                                   const uint32_t* delta, uint32_t nv, int shift,
                                                   const float* __restrict__ X, KeyT* __restrict__ keys_out,
                                                   uint32_t* __restrict__ idx_out, float* __restrict__ scoord,
                                                   uint32_t t, uint32_t T) {
  constexpr int kB = 4;
  for (uint32_t j0 = 0; j0 < nv; j0 += kB * T) {
    uint32_t id[kB], dst[kB];
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const uint32_t j = j0 + (uint32_t)i * T + t;
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
      if (j0 + (uint32_t)i * T + t >= nv) continue;
      idx_out[dst[i]] = id[i];
      const uint64_t p = dst[i] - J.pbase;
#pragma unroll
      for (int e = 0; e < D; ++e) scoord[3 * J.pbase + (uint64_t)e * J.n + p] = v[i][e];
      keys_out[dst[i]] = (KeyT)cell_key(J, v[i]);
    }
  }
}
template <typename KeyT, bool GATHER>
__global__ __launch_bounds__(kBlock) void k_radix_scatter_lds(
    const DJob* __restrict__ jobs, const uint2* __restrict__ chunks,
    const uint32_t* __restrict__ chunk_first, uint32_t n_chunks,
    const KeyT* __restrict__ keys_in, const uint32_t* __restrict__ idx_in,
    KeyT* __restrict__ keys_out, uint32_t* __restrict__ idx_out, int shift,
    const uint32_t* __restrict__ offs, const float* __restrict__ X, float* __restrict__ scoord,
    const uint64_t* __restrict__ bbase, const uint32_t* __restrict__ bidx) {
  __shared__ uint32_t wrun[kScatterWaves][kRadixBins];  // per wave and digit: count, then run start
  __shared__ uint32_t lstart[kRadixBins];               // chunk-local start of each digit
  __shared__ uint32_t delta[kRadixBins];                // global start - chunk-local start
  __shared__ uint32_t wsum[kScatterWaves];
  __shared__ KeyT skey[kChunk];
  __shared__ uint32_t sid[kChunk];
  const int l = lane_id();
  const uint32_t w = threadIdx.x / kWave;
  // the chunk is the workgroup's: the early exit is uniform over it
  const uint32_t c = GATHER ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  const uint32_t c0 = chunk_first[ch.x];
  const uint32_t nch = chunk_first[ch.x + 1] - c0;
  {
    // digit d = thread: global start, chunk count, chunk-local start
    const uint32_t d = threadIdx.x;
    const uint64_t i = (uint64_t)kRadixBins * c0 + (uint64_t)d * nch + (c - c0);
    const uint32_t g = offs[i];
    const uint32_t cnt = offs[i + 1] - g;
    uint32_t total;
    const uint32_t ls = block_excl_scan(cnt, wsum, total);
    lstart[d] = ls;
    delta[d] = g - ls;
#pragma unroll
    for (int v = 0; v < kScatterWaves; ++v) wrun[v][d] = 0u;
  }
  const uint32_t nv = min((uint32_t)kChunk, J.n - ch.y);
  // the permutation so far: the previous pass's, or (idx_in == nullptr, the
  // first pass over keys from k_major_keys) the job's base order itself
  const uint32_t* isrc = idx_in ? idx_in + J.pbase : bidx + bbase[ch.x];
  KeyT kr[kScatterRounds];
  uint32_t ir[kScatterRounds];
#pragma unroll
  for (int r = 0; r < kScatterRounds; ++r) {
    const uint32_t e = (w * kScatterRounds + (uint32_t)r) * kWave + (uint32_t)l;
    kr[r] = 0;
    ir[r] = 0;
    if (e < nv) {
      kr[r] = keys_in[J.pbase + ch.y + e];
      ir[r] = isrc[ch.y + e];
    }
  }
  __syncthreads();  // counts cleared, digit starts written
  const uint64_t lt = lanemask_lt();
  uint32_t* run = wrun[w];
  uint32_t pos[kScatterRounds];
#pragma unroll
  for (int r = 0; r < kScatterRounds; ++r) {
    const bool valid = (w * kScatterRounds + (uint32_t)r) * kWave + (uint32_t)l < nv;
    const uint32_t dg = (uint32_t)(kr[r] >> shift) & (kRadixBins - 1);
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const bool bit = (dg >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      match &= bit ? bal : ~bal;
    }
    const uint32_t rank = __popcll(match & lt);
    const uint32_t base = valid ? run[dg] : 0u;
    wave_sync();
    if (valid && rank == 0) run[dg] = base + (uint32_t)__popcll(match);
    pos[r] = base + rank;
    wave_sync();
  }
  __syncthreads();  // every wave's counts are final
  {
    // runs of digit d: the chunk-local start, then the waves in order
    const uint32_t d = threadIdx.x;
    uint32_t acc = lstart[d];
#pragma unroll
    for (int v = 0; v < kScatterWaves; ++v) {
      const uint32_t t = wrun[v][d];
      wrun[v][d] = acc;
      acc += t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kScatterRounds; ++r) {
    if ((w * kScatterRounds + (uint32_t)r) * kWave + (uint32_t)l < nv) {
      const uint32_t dg = (uint32_t)(kr[r] >> shift) & (kRadixBins - 1);
      const uint32_t p = run[dg] + pos[r];
      skey[p] = kr[r];
      sid[p] = ir[r];
    }
  }
  __syncthreads();  // the chunk sits in LDS in digit order
  if constexpr (GATHER) {
    const uint32_t d = __builtin_amdgcn_readfirstlane(J.d);
    if (d == 1) scatter_gather_out<1>(J, skey, sid, delta, nv, shift, X, keys_out, idx_out, scoord, threadIdx.x, kBlock);
    else if (d == 2) scatter_gather_out<2>(J, skey, sid, delta, nv, shift, X, keys_out, idx_out, scoord, threadIdx.x, kBlock);
    else scatter_gather_out<3>(J, skey, sid, delta, nv, shift, X, keys_out, idx_out, scoord, threadIdx.x, kBlock);
  } else {
    for (uint32_t j = threadIdx.x; j < nv; j += kBlock) {
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
                                                 uint32_t* __restrict__ idx, float* __restrict__ scoord, int l,
                                                 uint32_t* bins) {
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
      if (FINAL || idx) idx[J.pbase + k] = id[i];
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
        if (bins) atomicAdd(&bins[(uint32_t)r & (kRadixBins - 1)], 1u);
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
                                                       float* __restrict__ scoord,
                                                       const uint32_t* __restrict__ chunk_first,
                                                       uint32_t* __restrict__ hist) {
  // with hist: the digit counts of the first radix pass over the keys written
  // here (digit = key & 0xFF), in k_radix_hist's layout
  __shared__ uint32_t bins[kBlock / kWave][kRadixBins];
  const int w = threadIdx.x / kWave;
  const int l = lane_id();
  if (hist) {
    for (int b = l; b < kRadixBins; b += kWave) bins[w][b] = 0u;
    wave_sync();
  }
  const uint32_t c = xcd_order(blockIdx.x, gridDim.x) * (kBlock / kWave) + (uint32_t)w;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob J = jobs[ch.x];
  const uint32_t* bi = bidx + bbase[ch.x];
  const uint32_t d = __builtin_amdgcn_readfirstlane(J.d);
  uint32_t* wb = hist ? bins[w] : nullptr;
  if (scoord) {
    if (d == 1) major_keys_chunk<1, true>(J, ch.y, bi, X, keys, idx, scoord, l, nullptr);
    else if (d == 2) major_keys_chunk<2, true>(J, ch.y, bi, X, keys, idx, scoord, l, nullptr);
    else major_keys_chunk<3, true>(J, ch.y, bi, X, keys, idx, scoord, l, nullptr);
  } else {
    if (d == 1) major_keys_chunk<1, false>(J, ch.y, bi, X, keys, idx, scoord, l, wb);
    else if (d == 2) major_keys_chunk<2, false>(J, ch.y, bi, X, keys, idx, scoord, l, wb);
    else major_keys_chunk<3, false>(J, ch.y, bi, X, keys, idx, scoord, l, wb);
  }
  if (hist && !scoord) {
    wave_sync();
    const uint32_t c0 = chunk_first[ch.x];
    const uint32_t nch = chunk_first[ch.x + 1] - c0;
    for (int b = l; b < kRadixBins; b += kWave)
      hist[(uint64_t)kRadixBins * c0 + (uint64_t)b * nch + (c - c0)] = bins[w][b];
  }
}

// --------------------------------------------------------- exclusive scan
// Three-phase device-wide exclusive scan of uint32 (in place allowed).
// kScanItems elements per thread, kBlock threads per block.
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

__global__ __launch_bounds__(kBlock) void k_scan_blocks(const uint32_t* in, uint32_t* out,
                                                        uint64_t m, uint32_t* block_sums) {
  __shared__ uint32_t wsum[kBlock / kWave];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t tsum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < m) ? in[base + i] : 0u;
    tsum += v[i];
  }
  uint32_t total;
  uint32_t run = block_excl_scan(tsum, wsum, total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < m) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Adds each tile's offset (the scanned tile sums, whose total follows them
// at block_offs[nb]); block 0 also writes the total to out[m].
__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* out, uint64_t m,
                                                     const uint32_t* block_offs) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  const uint32_t add = block_offs[blockIdx.x];
  for (int i = threadIdx.x; i < kScanTile; i += kBlock)
    if (base + i < m) out[base + i] += add;
  if (blockIdx.x == 0 && threadIdx.x == 0) out[m] = block_offs[gridDim.x];
}

// Exclusive scan of up to kScanSmallTiles tiles by one workgroup, tile after
// tile with a running carry (in place allowed: each thread reads its items
// before it writes them); the total goes to out[m].  One launch where the
// tiled scan takes three or more.
constexpr int kScanSmallTiles = 16;
__global__ __launch_bounds__(kBlock) void k_scan_small(const uint32_t* in, uint32_t* out, uint64_t m) {
  __shared__ uint32_t wsum[kBlock / kWave];
  uint32_t carry = 0;
  for (uint64_t t0 = 0; t0 < m; t0 += kScanTile) {
    const uint64_t base = t0 + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t tsum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      v[i] = (base + i < m) ? in[base + i] : 0u;
      tsum += v[i];
    }
    uint32_t total;
    uint32_t run = carry + block_excl_scan(tsum, wsum, total);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      if (base + i < m) out[base + i] = run;
      run += v[i];
    }
    carry += total;
    __syncthreads();  // wsum is rewritten by the next tile
  }
  if (threadIdx.x == 0) out[m] = carry;
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
    // the block's cells leave through LDS (the point stage, free now) so
    // that its 9 x kBlock floats are stored contiguously, one float per lane
    // and instruction, instead of nine stores 36 bytes apart per lane
    __syncthreads();
    float* ostage = reinterpret_cast<float*>(stage);
    if (valid) {
      const double k0 = (double)J.hinv[0] * (double)J.hinv[0];
      const double k1 = (double)J.hinv[1] * (double)J.hinv[1];
      const double s0[3] = {1.0, 2.0 * k0, -k0}, s1[3] = {1.0, 2.0 * k1, -k1};
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) ostage[9 * threadIdx.x + 3 * i + k] = (float)(m[3 * i + k] * s0[i] * s1[k]);
    }
    __syncthreads();
    float* out = mom + (J.mbase + cell0) * 9;
    const uint32_t nf = 9 * (last + 1);
    for (uint32_t i = threadIdx.x; i < nf; i += kBlock) out[i] = ostage[i];
  }
}

// ------------------------------------------------------------------ tiles
// A tile starts at sorted position p when p == 0, p % tq == 0, or the
// (row, dim-0 group) of p differs from that of p-1.  Tiles never cross rows,
// have at most tq queries and span at most `group` dim-0 boxes.
namespace {
template <typename KeyT>
__device__ __forceinline__ bool tile_start(const DJob& J, KeyT a, KeyT b, uint32_t p) {
  if (p == 0 || (p % J.tq) == 0) return true;
  if (J.flags & kJobBrute) return false;
  // key width arithmetic (32-bit division for 32-bit keys)
  const KeyT nb0 = (KeyT)J.nb[0];
  const KeyT ra = a / nb0, rb = b / nb0;
  if (ra != rb) return true;
  return (uint32_t)(a - ra * nb0) / J.group != (uint32_t)(b - rb * nb0) / J.group;
}
static_assert(kChunkItems <= 32, "tile-start flags are one 32-bit word per lane");
}  // namespace

// Cell-start tables (jobs with a table) and per-chunk tile counts (every
// job) in one pass over the sorted keys.  chunks are in tile order; the
// tile count of chunk c goes to tile_cnt[c].
template <typename KeyT>
__global__ __launch_bounds__(kBlock) void k_cell_starts(const DJob* __restrict__ jobs,
                                                        const uint2* __restrict__ chunks,
                                                        uint32_t n_chunks, const KeyT* __restrict__ keys,
                                                        uint32_t* __restrict__ tile_cnt,
                                                        uint32_t* __restrict__ start,
                                                        uint32_t* __restrict__ tbits) {
  const uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const DJob& J = jobs[ch.x];
  const bool tab = J.cbase != kNoTable;
  const int l = lane_id();
  uint32_t* st = start + (tab ? J.cbase : 0);
  uint32_t tiles = 0;  // tile starts among the lane's points
  uint32_t bits = 0;   // bit i: point ch.y + 64 i + lane starts a tile
  const KeyT* kk = keys + J.pbase;
  // Row-start tables index rows of boxes (key / nb0), cell-start tables cells
  // (the key itself).
  const bool rowtab = (J.flags & kJobRowTable) != 0;
  const KeyT nb0 = (KeyT)J.nb[0];
  auto unit = [&](KeyT k) -> uint64_t { return rowtab ? (uint64_t)(k / nb0) : (uint64_t)k; };
  const uint64_t units = rowtab ? J.cells / J.nb[0] : J.cells;
  // Point p owns the units (unit(key[p-1]), unit(key[p])]; units after the
  // last key hold n.  Short runs are written by their own lane; runs longer than
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
        if (tile_start(J, prevk[ii], curk[ii], p)) {
          ++tiles;
          bits |= 1u << (h * kHalf + ii);
        }
        if (tab) {
          const uint64_t cur = unit(curk[ii]);
          first = p == 0 ? 0 : unit(prevk[ii]) + 1;
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
        const uint64_t tail = unit(kk[J.n - 1]) + 1;
        for (uint64_t u = tail + (uint64_t)l; u <= units; u += kWave) st[u] = J.n;
      }
    }
  }
  const uint32_t total = __builtin_amdgcn_readlane(wave_incl_scan(tiles), kWave - 1);
  if (l == 0) tile_cnt[c] = total;
  tbits[(uint64_t)c * kWave + l] = bits;
}


__global__ __launch_bounds__(kBlock) void k_tile_write(const DJob* __restrict__ jobs,
                                                       const uint2* __restrict__ chunks,
                                                       uint32_t n_chunks,
                                                       const uint32_t* __restrict__ tbits,
                                                       const uint32_t* __restrict__ offs,
                                                       DTile* __restrict__ tiles) {
  const uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= n_chunks) return;
  const uint2 ch = chunks[c];
  const int l = lane_id();
  const uint64_t lt = lanemask_lt();
  uint32_t o = offs[c];
  const uint32_t bits = tbits[(uint64_t)c * kWave + l];
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
                          const float* X, float* scoord, const uint64_t* bbase, const uint32_t* bidx) {
  const bool gather = scoord != nullptr;
#define ENT_SCATTER(KT, G)                                                                                     \
  do {                                                                                                         \
    if (staged)                                                                                                \
      hipLaunchKernelGGL((k_radix_scatter_lds<KT, G>), dim3(n_chunks), dim3(kBlock), 0, s, jobs, chunks,     \
                         chunk_first, n_chunks, (const KT*)kin, iin, (KT*)kout, iout, shift, offs, X, scoord, \
                         bbase, bidx);                                                                 \
    else                                                                                                       \
      hipLaunchKernelGGL((k_radix_scatter<KT, G>), chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,   \
                         chunk_first, n_chunks, (const KT*)kin, iin, (KT*)kout, iout, shift, offs, X, scoord, \
                         bbase, bidx);                                                                 \
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
                       uint32_t* idx, float* scoord, const uint32_t* chunk_first, uint32_t* hist) {
  if (wide)
    hipLaunchKernelGGL(k_major_keys<uint64_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       n_chunks, bbase, bidx, X, (uint64_t*)keys, idx, scoord, chunk_first, hist);
  else
    hipLaunchKernelGGL(k_major_keys<uint32_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       n_chunks, bbase, bidx, X, (uint32_t*)keys, idx, scoord, chunk_first, hist);
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
  if (nb <= (uint64_t)kScanSmallTiles) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kBlock), 0, s, in, out, m);
    return;
  }
  uint32_t* sums = scratch;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(kBlock), 0, s, in, out, m, sums);
  // scans the block sums in place; their total lands in sums[nb], which
  // k_scan_add copies to out[m]
  launch_exclusive_scan(s, sums, sums, nb, scratch + nb + 1);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kBlock), 0, s, out, m, sums);
}

void launch_gather_sorted(hipStream_t s, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                          const float* X, const uint32_t* idx, float* scoord) {
  hipLaunchKernelGGL(k_gather_sorted, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, n_chunks, X,
                     idx, scoord);
}

void launch_tile_write(hipStream_t s, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                       const uint32_t* tbits, const uint32_t* offs, DTile* tiles) {
  hipLaunchKernelGGL(k_tile_write, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks, n_chunks, tbits, offs,
                     tiles);
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

// Density kernel variants for q = 1, 2, 4 queries per lane, plain path or
// cell lists with 32- or 64-bit keys, and for q = 1 the variant with 3-D
// sub-tiles (compiled in kde_density_*.hip).
template <int Q, int SUBD>
static int occupancy_q(bool raw, bool wide, int* bpc) {
  if (raw) return density_occupancy_v<Q, true, uint32_t, 2>(bpc);
  return wide ? density_occupancy_v<Q, false, uint64_t, SUBD>(bpc) : density_occupancy_v<Q, false, uint32_t, SUBD>(bpc);
}

int density_occupancy(int q, bool raw, bool wide, bool sub3, int* blocks_per_cu) {
  switch (q) {
    case 1: return sub3 ? occupancy_q<1, 3>(raw, wide, blocks_per_cu) : occupancy_q<1, 2>(raw, wide, blocks_per_cu);
    case 2: return occupancy_q<2, 2>(raw, wide, blocks_per_cu);
    case 4: return occupancy_q<4, 2>(raw, wide, blocks_per_cu);
    default: return -1;
  }
}

int launch_density(hipStream_t s, int q, bool raw, bool wide, bool sub3, unsigned grid, const DJob* jobs,
                   const DTile* tiles, const uint32_t* n_tiles, uint32_t* counter, const float* X,
                   const float* scoord, const void* keys, const uint32_t* cstart, const uint32_t* idx,
                   const float4* mom, const float* mom2, double* dens, unsigned long long* pairs) {
#define ENT_LD(QQ, SD)                                                                                            \
  do {                                                                                                            \
    if (raw)                                                                                                      \
      launch_density_v<QQ, true, uint32_t, 2>(s, grid, jobs, tiles, n_tiles, counter, X, scoord, keys, cstart,   \
                                              idx, mom, mom2, dens, pairs);                                      \
    else if (wide)                                                                                                \
      launch_density_v<QQ, false, uint64_t, SD>(s, grid, jobs, tiles, n_tiles, counter, X, scoord, keys, cstart, \
                                                idx, mom, mom2, dens, pairs);                                    \
    else                                                                                                          \
      launch_density_v<QQ, false, uint32_t, SD>(s, grid, jobs, tiles, n_tiles, counter, X, scoord, keys, cstart, \
                                                idx, mom, mom2, dens, pairs);                                    \
  } while (0)
  switch (q) {
    case 1:
      if (sub3) ENT_LD(1, 3);
      else ENT_LD(1, 2);
      break;
    case 2: ENT_LD(2, 2); break;
    case 4: ENT_LD(4, 2); break;
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
                        const void* keys, uint32_t* tile_cnt, uint32_t* start, uint32_t* tbits) {
  if (wide)
    hipLaunchKernelGGL(k_cell_starts<uint64_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       n_chunks, (const uint64_t*)keys, tile_cnt, start, tbits);
  else
    hipLaunchKernelGGL(k_cell_starts<uint32_t>, chunk_grid(n_chunks), dim3(kBlock), 0, s, jobs, chunks,
                       n_chunks, (const uint32_t*)keys, tile_cnt, start, tbits);
}

}  // namespace ent
