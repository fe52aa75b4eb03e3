// kde_engine.hpp -- shared host/device declarations of the MI355X KDE engine.
//
// Density definition (densities_omp.cpp:19-26,87-93; density_1d.cuh:39-46;
// densities_cuda.cu:95-103):
//   p(x_i) = 1/n_rows * sum_j prod_e K_{h_e}(x_{i,e} - x_{j,e}),
//   K_h(u) = 0.75/h * (1 - (u/h)^2) for |u| <= h, else 0,
// over every point of the job including the self pair, divided by the
// ORIGINAL row count even for lagged jobs.
//
// Engine stages for a batch of jobs (one stream, capturable in a hipGraph):
//   minmax -> (host: boxes per dim) -> cell keys -> stable LSD radix sort by
//   cell key -> gather sorted coordinates -> query tiles -> one persistent
//   density launch over every (job, tile) -> reductions.
#pragma once
#include <cstddef>
#include <cstdint>

namespace ent {

constexpr int kWave = 64;          // CDNA wavefront
constexpr int kBlock = 256;        // 4 waves per workgroup
constexpr int kChunkItems = 32;    // items per lane in chunked passes
constexpr int kChunk = kWave * kChunkItems;  // points per chunk (one wave)
constexpr int kRadixBits = 8;
constexpr int kRadixBins = 1 << kRadixBits;
constexpr int kMaxRanges = 64;     // candidate row ranges per tile (one per lane)

// Job flags
constexpr uint32_t kJobBrute = 1u;  // all-pairs over the job's points (plain path)
constexpr uint32_t kJobTri = 2u;    // 1-D triangular kernel max(0, 1-|u|): the
                                    // Epanechnikov self-convolution of epa_convolution
constexpr uint32_t kJobBase2 = 8u;  // two-stage sort of a 3-D job: its base is sorted by
                                    // dims 0 and 1, its own passes are over dim 2
constexpr uint32_t kJobMoments = 16u;  // 1-D or 2-D job with a moment table: the boxes (1-D)
                                       // or cells (2-D) that lie inside every query's
                                       // support are summed from their moments instead
                                       // of pair by pair
constexpr uint32_t kJobSub16 = 32u;  // 2-D or 3-D job whose tiles run as four sub-tiles of 16
                                     // queries, each with its own candidate stream
constexpr uint32_t kJobSub32 = 64u;  // the same with two sub-tiles of 32 queries
constexpr uint32_t kJobRowTable = 128u;  // cbase locates a row-start table (the first sorted
                                         // position of every row of boxes, row = key / nb[0])
                                         // instead of a cell-start table

// Device job descriptor.  Points of job j, dim e, index k in [0, n):
//   X[off[e] + k].
struct DJob {
  uint64_t off[3];
  float h[3];
  float hinv[3];
  float mn[3];       // bin origin (min of the job's points, dim e)
  float bs[3];       // box size, dim e
  uint32_t nb[3];    // boxes per dim (>= 1)
  uint32_t d;
  uint32_t n;        // points
  uint32_t n_rows;   // normalisation count
  uint32_t flags;
  uint32_t group;    // dim-0 boxes per tile group (tile extent bound)
  uint64_t pbase;    // first slot in the concatenated point arrays
  uint64_t out_off;  // first density slot in the output array
  double norm;       // prod_e(0.75/h_e) / n_rows
  uint64_t cells;    // prod_e nb[e]
  uint64_t cbase;    // first entry of this job's cell-start table (or row-start table,
                     // kJobRowTable), or kNoTable
  uint64_t mbase;    // first entry of this job's box-moment table (kJobMoments)
  uint32_t mcf;      // 2-D moment cells: dim-0 boxes per cell
  uint32_t mncb;     // 2-D moment cells per row
  uint32_t tq;       // maximum queries per tile
  uint32_t pad_;
};

// Jobs whose cell count exceeds this multiple of their points (sparse grids,
// heavy tails) get a row-start table instead of a cell-start table and locate
// candidate ranges by binary search within the row.
constexpr uint64_t kNoTable = ~0ull;
constexpr uint64_t kTableCellsPerPoint = 8;

// Tile: queries [p0, p1) of job `job` (sorted positions, job-local).
// `pack` (cell-list tiles only): bits 0-1 hold G - 1 for a packed tile of G
// = 2 or 4 consecutive sub-tiles of one job processed side by side by one
// wave (64 / G lanes each); bits 8 s .. 8 s + 7 hold the offset from p0 of
// sub-tile s >= 1.  kTileSkip marks a tile absorbed into the packed tile
// before it.
struct DTile {
  uint32_t job;
  uint32_t p0;
  uint32_t p1;
  uint32_t pack;
};
constexpr uint32_t kTileSkip = 2u;  // bits 0-1 = 2: no G of its own

struct Tuning {
  int q = 2;            // maximum queries per lane (tile = 64*q queries)
  int m0 = 8;           // dim-0 boxes per bandwidth
  int mr = 1;           // row-dim boxes per bandwidth
  int group_h = 4;      // tile extent bound along dim 0, in bandwidths
  int sub = 0;          // 2-D jobs: 0 = whole tiles of up to 64*q queries; 16 or 32 = tiles
                        // of up to 64 queries run as sub-tiles of that many queries, each
                        // with its own candidate ranges, moment cells and stream
};

}  // namespace ent
