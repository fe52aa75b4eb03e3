// kde_density.hpp -- the persistent density kernel k_density<Q, RAW, KeyT>
// and its tile paths.  Included by kde_density_q{1,2,4}.hip, which
// instantiate one queries-per-lane setting each (so the three compile in
// parallel).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>
#include "kde_device.hpp"
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

// Sorted positions [a, b) of the points with keys klo..khi, which lie in one
// row of boxes (row = key / nb0): read from the job's cell-start table, or
// found by binary search within the row that its row-start table locates,
// or (a job with neither table) within the job.
template <typename KeyT>
__device__ __forceinline__ void key_range(const DJob& J, const uint32_t* __restrict__ cstart,
                                          const KeyT* __restrict__ kk, uint64_t row, uint64_t klo, uint64_t khi,
                                          uint32_t& a, uint32_t& b) {
  if (J.cbase == kNoTable) {
    a = lower_bound_key(kk, J.n, klo);
    b = a + lower_bound_key(kk + a, J.n - a, khi + 1);
    return;
  }
  const uint32_t* t = cstart + J.cbase;
  if (J.flags & kJobRowTable) {
    const uint32_t r0 = t[row], r1 = t[row + 1];
    a = r0 + lower_bound_key(kk + r0, r1 - r0, klo);
    b = a + lower_bound_key(kk + a, r1 - a, khi + 1);
  } else {
    a = t[klo];
    b = t[khi + 1];
  }
}
// Sorted positions [a, b) of the rows row_a..row_b.
template <typename KeyT>
__device__ __forceinline__ void rows_range(const DJob& J, const uint32_t* __restrict__ cstart,
                                           const KeyT* __restrict__ kk, uint64_t row_a, uint64_t row_b,
                                           uint32_t& a, uint32_t& b) {
  if (J.cbase == kNoTable) {
    const uint64_t nb0 = J.nb[0];
    a = lower_bound_key(kk, J.n, row_a * nb0);
    b = a + lower_bound_key(kk + a, J.n - a, (row_b + 1) * nb0);
    return;
  }
  const uint32_t* t = cstart + J.cbase;
  if (J.flags & kJobRowTable) {
    a = t[row_a];
    b = t[row_b + 1];
  } else {
    a = t[row_a * J.nb[0]];
    b = t[(row_b + 1) * J.nb[0]];
  }
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

// LDS of one wave.  Whole tiles stage 64 candidates in cand and 2-D moment
// cells in cand + cell.  Sub-tiles (process_subt) of 2-D tiles stage their
// moment cells in gcell (sub-tile s from kSubCellStride * s) and 64
// candidates each in gcand (sub-tile s from kSubCandStride * s; 3-D tiles 32
// candidates as float4): the strides are one float4 longer than the data, so
// the addresses that the sub-tiles' lane groups read with one instruction
// fall on different LDS banks.
constexpr int kSubCellStride = kWave / 4 + 1;  // float4 per sub-tile of 16 (a 32-lane one: 2x)
constexpr int kSubCandStride = kWave / 2 + 1;  // float4 per sub-tile (64 2-D or 32 3-D candidates)
struct WaveLds {
  union {
    struct {
      float4 cand[kWave];     // staged candidates (local coordinates)
      float4 cell[2][kWave];  // 2-D moment cells (with cand): 9 moments and the cell origin
    };
    float4 gcell[3][4 * kSubCellStride];
    float4 gcand[4 * kSubCandStride];
  };
  uint2 rng[2 * kWave];  // candidate ranges (first flattened index, first position): packed
                         // tiles W per sub-tile, sub-tiles (process_subt) 2W per sub-tile
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

// Pair statistics of a wave, one counter per lane: lane i holds counter i
// (0-2 candidate pairs of 1-, 2-, 3-D jobs, 3 lane slots, 4-5 moment pairs of
// 1-D / 2-D jobs, 6-7 their moment terms, 8 tiles with raw-coordinate pair
// terms), so the persistent loop holds no 64-bit scalar accumulators.
enum : int { kStPairs = 0, kStSlots = 3, kStMomPairs = 4, kStMomTerms = 6, kStRaw = 8, kStCount = 9 };
struct WaveStats {
  uint64_t v = 0;
  __device__ __forceinline__ void add(int i, uint64_t x) { v += lane_id() == i ? x : 0ull; }
};

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

// 3-D candidates staged as structure of arrays (x of NC candidates, then y,
// then z; NC a multiple of 32) against one query per lane: one float4 read
// per coordinate serves four candidates (three reads per four candidates
// instead of four), compensated merge every 32.
template <bool RAW, int NC>
__device__ __forceinline__ void accumulate_soa3(const float4* st, const float (&sq)[3][1], const float (&hinv)[3],
                                                float& S, float& C) {
  static_assert(NC % 32 == 0, "merge every 32 candidates");
  float part = 0.0f;
#pragma unroll
  for (int g = 0; g < NC / 4; ++g) {
    const float4 v[3] = {st[g], st[NC / 4 + g], st[NC / 2 + g]};
    float cs[3][4];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      cs[e][0] = v[e].x;
      cs[e][1] = v[e].y;
      cs[e][2] = v[e].z;
      cs[e][3] = v[e].w;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float w[3];
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        float uu = sq[e][0] - cs[e][c];
        if constexpr (RAW) uu = uu * hinv[e];
        w[e] = clamp01(__builtin_fmaf(-uu, uu, 1.0f));
      }
      part = __builtin_fmaf(w[0] * w[1], w[2], part);
    }
    if ((g + 1) % 8 == 0) {
      kahan_add(S, C, part);
      part = 0.0f;
    }
  }
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
      uint32_t a, b;
      key_range(J, cstart, keys + J.pbase, row, row * J.nb[0] + lo[0], row * J.nb[0] + hi[0], a, b);
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
      if (l == 0) {
        uint32_t a, b;
        rows_range(J, cstart, keys + J.pbase, row_a, row_b, a, b);
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
                                 WaveStats& ws) {
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
  const KeyT* kk = keys + J.pbase;
  if (nr <= (uint32_t)W) {
    if ((uint32_t)j < nr) {
      const uint32_t r1 = lo[1] + (uint32_t)j % n1;
      const uint32_t r2 = (D == 3) ? lo[2] + (uint32_t)j / n1 : 0u;
      const uint64_t row = (uint64_t)r2 * J.nb[1] + r1;
      uint32_t a, b;
      key_range(J, cstart, kk, row, row * J.nb[0] + lo[0], row * J.nb[0] + hi[0], a, b);
      my_lo = a;
      my_len = b - a;
    }
  } else {
    // more candidate rows than lanes: one range over every row between the
    // first and last candidate row (a superset)
    if (j == 0) {
      const uint64_t row_a = (uint64_t)lo[2] * J.nb[1] + lo[1];
      const uint64_t row_b = (uint64_t)hi[2] * J.nb[1] + hi[1];
      uint32_t a, b;
      rows_range(J, cstart, kk, row_a, row_b, a, b);
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
    ws.add(kStRaw, 1);
    run(std::true_type{});
  } else {
    run(std::false_type{});
  }
  wave_sync();

  if (qv) dens[J.out_off + slot] = ((double)S[0] - (double)C[0]) * J.norm;
  ws.add(kStSlots, (uint64_t)kWave * W * steps);
  uint64_t pr = 0;
#pragma unroll
  for (int g = 0; g < G; ++g)
    pr += (uint64_t)__builtin_amdgcn_readlane(nq, g * W) * __builtin_amdgcn_readlane(total, g * W);
  return pr;
}

// Densities of a D-dim tile (D = 2, 3) run as G sub-tiles side by side,
// W = 64 / G lanes and one query per lane each: a packed tile (sub-tile
// offsets in T.pack) or a tile of a kJobSub16 / kJobSub32 job split into runs
// of W consecutive queries (a short tile leaves its last sub-tiles empty).  Fewer queries
// share a candidate stream than in a whole tile, so the candidates that lie
// in the support of only some of them -- the margin between the union and
// the intersection of their supports, which grows with the extent of the
// query set -- are fewer.  Each sub-tile has its own bounding box, moment
// cells (2-D jobs with a moment table: the cells of its interior rows and
// interior dim-0 span, as for a whole tile), candidate ranges (two entries
// per lane in the wave's LDS range table, 2Ws + 2j and 2Ws + 2j + 1 for lane
// j of sub-tile s: in 2-D candidate row j, split in two around the moment
// cells; in 3-D the candidate rows j and j + W, a row being one (dim-1,
// dim-2) pair of boxes) and candidate stream.  The wave steps through the
// longest stream 64 candidates (3-D: 32) per sub-tile at a time, each lane
// loading G (G / 2) candidates of its sub-tile per step.
template <int D, int G, typename KeyT>
__device__ void process_subt(const DJob& J, const DTile& T, bool packed, const float* __restrict__ scoord,
                             const KeyT* __restrict__ keys, const uint32_t* __restrict__ cstart,
                             const uint32_t* __restrict__ idx, const float* __restrict__ mom2,
                             double* __restrict__ dens, WaveLds& L, WaveStats& ws) {
  constexpr int W = kWave / G;
  // candidates per sub-tile and step, and per lane (3-D: half as many, so
  // fewer coordinates are in flight per lane)
  constexpr int kNC = D == 3 ? kWave / 2 : kWave;
  constexpr int kLd = kNC / W;
  constexpr int kCellStride = W + 1;
  static_assert(G * kCellStride <= 4 * kSubCellStride, "moment-cell stage");
  const int l = lane_id();
  const int s = l / W, j = l % W;
  const uint32_t n = J.n;
  static_assert(D == 2 || D == 3, "sub-tiles of 2-D and 3-D tiles");
  const float* src0 = scoord + 3 * J.pbase;
  const float* src1 = src0 + n;
  const float* src2 = src1 + n;
  const float hinv[3] = {J.hinv[0], J.hinv[1], D == 3 ? J.hinv[2] : 1.0f};

  // this lane's sub-tile and query; idle lanes repeat their sub-tile's first
  // query, the lanes of an empty sub-tile the tile's first (they get no
  // candidates and no moment cells)
  const uint32_t span = T.p1 - T.p0;
  uint32_t b0, b1;
  if (packed) {
    b0 = s == 0 ? 0u : (T.pack >> (8 * s)) & 0xFFu;
    b1 = s + 1 == G ? span : (T.pack >> (8 * (s + 1))) & 0xFFu;
  } else {
    b0 = min((uint32_t)(s * W), span);
    b1 = min((uint32_t)((s + 1) * W), span);
  }
  const uint32_t nq = b1 - b0;
  const bool qv = (uint32_t)j < nq;
  const uint32_t p = T.p0 + (nq == 0 ? 0u : b0 + (qv ? (uint32_t)j : 0u));
  const float xq[3] = {src0[p], src1[p], D == 3 ? src2[p] : 0.0f};
  const uint32_t slot = idx[J.pbase + p];

  // bounding box (its centre is the origin of the local coordinates) and the
  // boxes of the candidate window
  float bmin[3], bmax[3], o[3];
  uint32_t lo[3], hi[3];
#pragma unroll
  for (int e = 0; e < D; ++e) {
    bmin[e] = sub_min<W>(xq[e]);
    bmax[e] = sub_max<W>(xq[e]);
    o[e] = 0.5f * (bmin[e] + bmax[e]);
    box_range(bmin[e], bmax[e], J.h[e], J.mn[e], J.bs[e], J.nb[e], lo[e], hi[e]);
  }
  const bool exact = wave_max((bmax[0] - bmin[0]) * J.hinv[0]) > kLocalSpan;

  // this lane's candidate row r and its range [ra, rb) of sorted positions
  // (3-D: also row r + W's, [ra2, rb2)); with more candidate rows than the
  // lanes take, one range over every row between the first and the last (a
  // superset) on the sub-tile's first lane
  const uint32_t n1 = hi[1] - lo[1] + 1;
  const uint32_t nr = D == 3 ? n1 * (hi[2] - lo[2] + 1) : n1;
  const bool rows_fit = nq != 0 && nr <= (uint32_t)(D == 3 ? 2 * W : W);
  const uint32_t* st = cstart + J.cbase;  // the cell-start table of a job with moments
  const uint32_t r = lo[1] + (uint32_t)j;
  const KeyT* kk = keys + J.pbase;
  uint32_t ra = 0, rb = 0, ra2 = 0, rb2 = 0;
  if constexpr (D == 2) {
    if (rows_fit && (uint32_t)j < n1) {
      key_range(J, cstart, kk, r, (uint64_t)r * J.nb[0] + lo[0], (uint64_t)r * J.nb[0] + hi[0], ra, rb);
    } else if (nq != 0 && !rows_fit && j == 0) {
      rows_range(J, cstart, kk, lo[1], hi[1], ra, rb);
    }
  } else {
    auto row_of = [&](uint32_t i) { return (uint64_t)(lo[2] + i / n1) * J.nb[1] + lo[1] + i % n1; };
    if (rows_fit) {
      if ((uint32_t)j < nr) {
        const uint64_t row = row_of((uint32_t)j);
        key_range(J, cstart, kk, row, row * J.nb[0] + lo[0], row * J.nb[0] + hi[0], ra, rb);
      }
      if ((uint32_t)j + W < nr) {
        const uint64_t row = row_of((uint32_t)j + W);
        key_range(J, cstart, kk, row, row * J.nb[0] + lo[0], row * J.nb[0] + hi[0], ra2, rb2);
      }
    } else if (nq != 0 && j == 0) {
      rows_range(J, cstart, kk, (uint64_t)lo[2] * J.nb[1] + lo[1], (uint64_t)hi[2] * J.nb[1] + hi[1], ra, rb);
    }
  }

  // moment cells (rows [jlo, jlo + nir) x cells [clo, clo + nc)): inside the
  // support of every query of the sub-tile (the margins as for whole tiles),
  // used when they hold at least 4 points per cell on average
  uint32_t m1 = 0, m2 = 0, jlo = 0, nir = 0, clo = 0, nc = 0;
  bool inr = false;
  if (D == 2 && (J.flags & kJobMoments)) {
    const uint32_t cf = J.mcf;
    const float h0 = J.h[0], h1 = J.h[1];
    const float mg0 = (fabsf(bmin[0]) + fabsf(bmax[0]) + fabsf(J.mn[0]) + h0) * 1.0e-5f;
    const float mg1 = (fabsf(bmin[1]) + fabsf(bmax[1]) + fabsf(J.mn[1]) + h1) * 1.0e-5f;
    const uint32_t ilo = box_of(bmax[0] - h0 + mg0, J.mn[0], J.bs[0], J.nb[0]) + 1u;
    const uint32_t ihi1 = box_of(bmin[0] + h0 - mg0, J.mn[0], J.bs[0], J.nb[0]);
    clo = (ilo + cf - 1u) / cf;
    const uint32_t chi1 = ihi1 / cf;
    jlo = max(box_of(bmax[1] - h1 + mg1, J.mn[1], J.bs[1], J.nb[1]) + 1u, lo[1]);
    const uint32_t jhi1 = min(box_of(bmin[1] + h1 - mg1, J.mn[1], J.bs[1], J.nb[1]), hi[1] + 1u);
    if (rows_fit && clo < chi1 && jlo < jhi1) {
      nir = jhi1 - jlo;
      nc = chi1 - clo;
      inr = (uint32_t)j < n1 && r >= jlo && r < jhi1;
      if (inr) {
        const uint32_t* rs = st + (uint64_t)r * J.nb[0];
        m1 = rs[clo * cf];
        m2 = rs[chi1 * cf];
      }
    }
  }
  const uint32_t pts = sub_sum<W>(inr ? m2 - m1 : 0u);
  const bool use_m = nir != 0 && pts >= 4u * nir * nc;

  // range table: [ra, rb), or [ra, m1) and [m2, rb) for an interior row;
  // 3-D: [ra, rb) and [ra2, rb2)
  uint32_t a_len = rb - ra, b_pos = D == 3 ? ra2 : rb, b_len = rb2 - ra2;
  if (use_m && inr) {
    a_len = m1 - ra;
    b_pos = m2;
    b_len = rb - m2;
  }
  const uint32_t tot = a_len + b_len;
  const uint32_t incl = sub_incl_scan<W>(tot);
  const uint32_t total = sub_sum<W>(tot);
  wave_sync();
  L.rng[2 * W * s + 2 * j] = make_uint2(incl - tot, ra);
  L.rng[2 * W * s + 2 * j + 1] = make_uint2(incl - tot + a_len, b_pos);
  wave_sync();

  float S[1] = {0.0f}, C[1] = {0.0f};
  // Moment cells, merged by rows: lane j of a sub-tile with moment cells sums
  // the nc cells of its interior row jlo + j (j < nir) into one cell about the
  // origin of the row's middle interior cell (short offsets keep the
  // polynomial's terms small against their sum).  A cell's dim-0 moments move to
  // an origin delta further on as N'0 = N0, N'1 = N1 - 2 k0 delta N0,
  // N'2 = N2 + delta N1 - k0 delta^2 N0 (for each dim-1 index); the merged cell
  // holds only points inside every query's support, so each query adds one
  // polynomial term per interior row instead of one per cell.
  const bool mrow = use_m && (uint32_t)j < nir;
  if (D == 2 && wave_max_u32(mrow ? 1u : 0u) != 0) {
    const float k0 = J.hinv[0] * J.hinv[0], k1 = J.hinv[1] * J.hinv[1];
    float M[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float mo0 = 0.f, mo1 = 0.f;
    if (mrow) {
      const uint32_t row = jlo + (uint32_t)j;
      mo0 = fmaf((float)((clo + nc / 2) * J.mcf), J.bs[0], J.mn[0]);
      mo1 = fmaf((float)row, J.bs[1], J.mn[1]);
      const float* srcm = mom2 + (J.mbase + (uint64_t)row * J.mncb + clo) * 9;
#pragma unroll 2
      for (uint32_t c = 0; c < nc; ++c) {
        float N[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) N[i] = srcm[9 * c + i];
        const float dl = mo0 - fmaf((float)((clo + c) * J.mcf), J.bs[0], J.mn[0]);
        const float t1 = -2.0f * k0 * dl, t2 = -k0 * dl * dl;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          M[q] += N[q];
          M[3 + q] += __builtin_fmaf(t1, N[q], N[3 + q]);
          M[6 + q] += __builtin_fmaf(t2, N[q], __builtin_fmaf(dl, N[3 + q], N[6 + q]));
        }
      }
    }
    const int cs0 = s * (W + 1);
    wave_sync();
    L.gcell[0][cs0 + j] = make_float4(M[0], M[1], M[2], M[3]);
    L.gcell[1][cs0 + j] = make_float4(M[4], M[5], M[6], M[7]);
    L.gcell[2][cs0 + j] = make_float4(M[8], mo0, mo1, 0.f);
    wave_sync();
    // every query of a sub-tile adds its merged rows (padding rows are zero)
    const uint32_t rows = wave_max_u32(use_m ? nir : 0u);
    for (uint32_t i8 = 0; i8 < rows; i8 += 8) {
      float part = 0.0f;
#pragma unroll 1
      for (uint32_t i = i8; i < i8 + 8 && i < rows; ++i) {
        // a = (N00, N01, N02, N10), b = (N11, N12, N20, N21), c = (N22, r0, r1, -)
        const int ci = cs0 + (int)i;
        const float4 a = L.gcell[0][ci], b = L.gcell[1][ci], c = L.gcell[2][ci];
        const float d0 = xq[0] - c.y, d1 = xq[1] - c.z;
        const float a0 = __builtin_fmaf(-(k0 * d0), d0, 1.0f);
        const float bb0 = __builtin_fmaf(-(k1 * d1), d1, 1.0f);
        const float w0 = __builtin_fmaf(a0, a.x, __builtin_fmaf(d0, a.w, b.z));
        const float w1 = __builtin_fmaf(a0, a.y, __builtin_fmaf(d0, b.x, b.w));
        const float w2 = __builtin_fmaf(a0, a.z, __builtin_fmaf(d0, b.y, c.x));
        part += __builtin_fmaf(bb0, w0, __builtin_fmaf(d1, w1, w2));
      }
      kahan_add(S[0], C[0], part);
    }
  }

  // pair by pair: candidate f of the sub-tile lies in the last range that
  // starts at or before f; each lane's f only grows, so its cursor kc only
  // moves forward.  The lane keeps range kc's offset (sorted position less
  // flattened index) and the next range's first index in registers and
  // reads the table only when it moves on.
  float sq[3][1] = {{0.0f}, {0.0f}, {0.0f}};
#pragma unroll
  for (int e = 0; e < D; ++e) sq[e][0] = exact ? xq[e] : (xq[e] - o[e]) * hinv[e];
  const uint32_t steps = wave_max_u32((total + kNC - 1) / kNC);
  const uint2* rt = L.rng + 2 * W * s;
  uint32_t kc = 0;
  uint32_t off = rt[0].y - rt[0].x;
  uint32_t nx = rt[1].x;
  // 32-bit byte offsets from the columns' uniform bases (a job holds fewer
  // than 2^30 points)
  const char* c0 = reinterpret_cast<const char*>(src0);
  const char* c1 = reinterpret_cast<const char*>(src1);
  const char* c2 = reinterpret_cast<const char*>(src2);
  auto load = [&](uint32_t step, float (&x)[kLd][D]) {
#pragma unroll
    for (int k = 0; k < kLd; ++k) {
      // 2-D: lane j loads candidates k W + j of the step; 3-D: kLd
      // consecutive ones, kLd j + k (one vector store per coordinate)
      const uint32_t f = step * kNC + (uint32_t)(D == 3 ? kLd * j + k : k * W + j);
      if (f < total) {
        while (f >= nx) {
          ++kc;
          off = rt[kc].y - rt[kc].x;
          nx = kc + 1 < 2u * W ? rt[kc + 1].x : 0xFFFFFFFFu;
        }
        const uint32_t byte = (f + off) * 4u;
        x[k][0] = *reinterpret_cast<const float*>(c0 + byte);
        x[k][1] = *reinterpret_cast<const float*>(c1 + byte);
        if constexpr (D == 3) x[k][2] = *reinterpret_cast<const float*>(c2 + byte);
      } else {
#pragma unroll
        for (int e = 0; e < D; ++e) x[k][e] = kSentinel;
      }
    }
  };
  float nxt[kLd][D];
  load(0u, nxt);
  // the sub-tile's stage: 2-D candidates as float2, 3-D ones as float4
  float4* stage4 = L.gcand + s * kSubCandStride;
  float2* stage2 = reinterpret_cast<float2*>(stage4);
  auto run = [&](auto exact_tag) {
    constexpr bool kRawPairs = decltype(exact_tag)::value;
    for (uint32_t b = 0; b < steps; ++b) {
      float c[kLd][D];
#pragma unroll
      for (int k = 0; k < kLd; ++k)
#pragma unroll
        for (int e = 0; e < D; ++e) c[k][e] = kRawPairs ? nxt[k][e] : (nxt[k][e] - o[e]) * hinv[e];
      wave_sync();
#pragma unroll
      for (int k = 0; k < kLd; ++k) {
        if constexpr (D == 2) stage2[k * W + j] = make_float2(c[k][0], c[k][1]);
      }
      if constexpr (D == 3) {
        // structure of arrays: x, y, z of the step's kNC candidates
        float* st3 = reinterpret_cast<float*>(stage4);
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          if constexpr (kLd == 2)
            reinterpret_cast<float2*>(st3 + e * kNC)[j] = make_float2(c[0][e], c[1][e]);
          else
            st3[e * kNC + j] = c[0][e];
        }
      }
      wave_sync();
      if (b + 1 < steps) load(b + 1, nxt);
      if constexpr (D == 3) accumulate_soa3<kRawPairs, kNC>(stage4, sq, hinv, S[0], C[0]);
      else accumulate_chunk<D, 1, kRawPairs, false, kNC>(stage4, sq, hinv, S, C);
    }
  };
  if (exact) run(std::true_type{});
  else run(std::false_type{});
  wave_sync();

  if (qv) dens[J.out_off + slot] = ((double)S[0] - (double)C[0]) * J.norm;
  // statistics: the sub-tiles' candidate pairs, moment pairs and terms
  // (per-lane products, summed over each sub-tile's first lane)
  const uint32_t mp = use_m ? pts : 0u;
  const bool first = j == 0;
  uint64_t cp = first ? (uint64_t)nq * (total + mp) : 0ull;
  uint64_t mpp = first ? (uint64_t)nq * mp : 0ull;
  uint64_t mtt = first ? (uint64_t)nq * (use_m ? nir : 0u) : 0ull;
  uint64_t sc = 0, sm = 0, st2 = 0;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    sc += (uint64_t)__builtin_amdgcn_readlane((uint32_t)cp, g * W) |
          ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(cp >> 32), g * W) << 32);
    sm += (uint64_t)__builtin_amdgcn_readlane((uint32_t)mpp, g * W) |
          ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(mpp >> 32), g * W) << 32);
    st2 += (uint64_t)__builtin_amdgcn_readlane((uint32_t)mtt, g * W) |
           ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(mtt >> 32), g * W) << 32);
  }
  ws.add(kStPairs + D - 1, sc);
  ws.add(kStSlots, (uint64_t)kWave * kNC * steps);
  if constexpr (D == 2) {
    ws.add(kStMomPairs + 1, sm);
    ws.add(kStMomTerms + 1, st2);
  }
  if (exact) ws.add(kStRaw, 1);
}

// Queries per lane chosen per tile: the smallest Q <= QMAX whose 64*Q lane
// slots hold the tile (Q = 4 steps down through 3, 2 and 1), so partly
// filled tiles do not pay for idle slots.
// Statistics: every candidate pair of the tile, those summed from moments in
// d-dim jobs (d = 1, 2) and the (query, box) / (query, cell) moment terms
// evaluated for them, the tiles with raw-coordinate pair terms.
template <int D, int QMAX, bool RAW, typename KeyT>
__device__ __forceinline__ void run_tile(const DJob& J, const DTile& T, const float* __restrict__ X,
                                         const float* __restrict__ scoord, const KeyT* __restrict__ keys,
                                         const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ idx,
                                         const float4* __restrict__ mom, const float* __restrict__ mom2,
                                         double* __restrict__ dens, WaveLds& L, WaveStats& ws) {
  const uint32_t nq = T.p1 - T.p0;
  if constexpr (QMAX >= 2 && !RAW) {  // plain-path tiles are full except the last
    if (nq <= (uint32_t)(kWave * (QMAX - 1))) {
      run_tile<D, QMAX - 1, RAW, KeyT>(J, T, X, scoord, keys, cstart, idx, mom, mom2, dens, L, ws);
      return;
    }
  }
  uint32_t mstat[3] = {0u, 0u, 0u};
  const uint32_t total =
      process_tile<D, QMAX, RAW, KeyT>(J, T, X, scoord, keys, cstart, idx, mom, mom2, dens, L, mstat);
  ws.add(kStPairs + D - 1, (uint64_t)nq * (total + mstat[0]));
  ws.add(kStSlots, (uint64_t)(kWave * QMAX) * ((total + kWave - 1) / kWave * kWave));
  if constexpr (D <= 2 && !RAW) {
    ws.add(kStMomPairs + D - 1, (uint64_t)nq * mstat[0]);
    ws.add(kStMomTerms + D - 1, (uint64_t)nq * mstat[1]);
  }
  ws.add(kStRaw, mstat[2]);
}

}  // namespace

// Persistent density kernel: every wave pulls tiles from the claim counters
// until the tile list is exhausted.  SUBD: the job dimension whose tiles the
// one-query-per-lane variant runs as sub-tiles (2 or 3; tiles of the other
// dimension run whole or packed), so that each variant holds the registers
// of one sub-tile path only.
template <int Q, bool RAW, typename KeyT, int SUBD>
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
  WaveStats ws;
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
      const DJob& J = jobs[__builtin_amdgcn_readfirstlane(T.job)];  // uniform: scalar loads
      const uint32_t d = __builtin_amdgcn_readfirstlane(J.d);
      if constexpr (!RAW) {
        if (pk & 3u) {
          const bool four = (pk & 3u) == 3u;
          if (d == 2) {
            // the variant with one query per lane and 2-D sub-tiles runs
            // packed 2-D tiles as sub-tiles (with moments); the others keep
            // the lean pack path, so the sub-tile code does not add to their
            // register pressure
            if constexpr (Q == 1 && SUBD == 2) {
              if (four) process_subt<2, 4, KeyT>(J, T, true, scoord, keys, cstart, idx, mom2, dens, L, ws);
              else process_subt<2, 2, KeyT>(J, T, true, scoord, keys, cstart, idx, mom2, dens, L, ws);
            } else {
              ws.add(kStPairs + 1, four ? process_pack<2, 4, KeyT>(J, T, scoord, keys, cstart, idx, dens, L, ws)
                                        : process_pack<2, 2, KeyT>(J, T, scoord, keys, cstart, idx, dens, L, ws));
            }
          } else {
            // packed 3-D tiles of jobs with sub-tiles, in the variant with
            // 3-D sub-tiles: 32 candidates per sub-tile and step (the pack
            // path: W)
            bool done = false;
            if constexpr (Q == 1 && SUBD == 3) {
              if (__builtin_amdgcn_readfirstlane(J.flags) & (kJobSub16 | kJobSub32)) {
                if (four) process_subt<3, 4, KeyT>(J, T, true, scoord, keys, cstart, idx, mom2, dens, L, ws);
                else process_subt<3, 2, KeyT>(J, T, true, scoord, keys, cstart, idx, mom2, dens, L, ws);
                done = true;
              }
            }
            if (!done)
              ws.add(kStPairs + 2, four ? process_pack<3, 4, KeyT>(J, T, scoord, keys, cstart, idx, dens, L, ws)
                                        : process_pack<3, 2, KeyT>(J, T, scoord, keys, cstart, idx, dens, L, ws));
          }
          continue;
        }
        if constexpr (Q == 1) {  // sub-tiles run with one query per lane
          const uint32_t jf = __builtin_amdgcn_readfirstlane(J.flags);
          if (d == (uint32_t)SUBD && (jf & (kJobSub16 | kJobSub32))) {
            if (jf & kJobSub16) process_subt<SUBD, 4, KeyT>(J, T, false, scoord, keys, cstart, idx, mom2, dens, L, ws);
            else process_subt<SUBD, 2, KeyT>(J, T, false, scoord, keys, cstart, idx, mom2, dens, L, ws);
            continue;
          }
        }
      }
      if (d == 1)
        run_tile<1, Q, RAW, KeyT>(J, T, X, scoord, keys, cstart, idx, mom, mom2, dens, L, ws);
      else if (d == 2)
        run_tile<2, Q, RAW, KeyT>(J, T, X, scoord, keys, cstart, idx, mom, mom2, dens, L, ws);
      else
        run_tile<3, Q, RAW, KeyT>(J, T, X, scoord, keys, cstart, idx, mom, mom2, dens, L, ws);
    }
  }
  // counter i of the statistics (on lane i) into pairs[i]
  if (l < kStCount && ws.v != 0) atomicAdd(pairs + l, (unsigned long long)ws.v);
}

// One kernel variant: Q queries per lane at most, RAW (plain path, keys
// unused) or cell lists with KeyT keys, sub-tiles of SUBD-dim jobs.  Each
// variant is instantiated in one of the kde_density_*.hip files (so the slow
// ones compile in parallel).
template <int Q, bool RAW, typename KeyT, int SUBD>
int density_occupancy_v(int* blocks_per_cu) {
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocks_per_cu, reinterpret_cast<const void*>(&k_density<Q, RAW, KeyT, SUBD>), kBlock, 0);
  return e == hipSuccess ? 0 : -1;
}

template <int Q, bool RAW, typename KeyT, int SUBD>
void launch_density_v(hipStream_t s, unsigned grid, const DJob* jobs, const DTile* tiles, const uint32_t* n_tiles,
                      uint32_t* counter, const float* X, const float* scoord, const void* keys,
                      const uint32_t* cstart, const uint32_t* idx, const float4* mom, const float* mom2,
                      double* dens, unsigned long long* pairs) {
  hipLaunchKernelGGL((k_density<Q, RAW, KeyT, SUBD>), dim3(grid), dim3(kBlock), 0, s, jobs, tiles, n_tiles, counter, X,
                     scoord, (const KeyT*)keys, cstart, idx, mom, mom2, dens, pairs);
}

// explicit instantiation of one variant
#define ENT_DENSITY_VARIANT(Q, RAW, KeyT, SUBD)                                                              \
  template int density_occupancy_v<Q, RAW, KeyT, SUBD>(int*);                                               \
  template void launch_density_v<Q, RAW, KeyT, SUBD>(hipStream_t, unsigned, const DJob*, const DTile*,            \
                                               const uint32_t*, uint32_t*, const float*, const float*,      \
                                               const void*, const uint32_t*, const uint32_t*, const float4*, \
                                               const float*, double*, unsigned long long*);

}  // namespace ent
