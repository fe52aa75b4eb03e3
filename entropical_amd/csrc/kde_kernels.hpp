// kde_kernels.hpp -- host-side launchers of the kernels in kde_kernels.hip.
// `wide` selects 64-bit cell keys (else 32-bit); key buffers are void*.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "kde_engine.hpp"

namespace ent {

void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, uint64_t n);
void launch_init_minmax(hipStream_t s, uint32_t* mm, uint32_t n_jobs);
void launch_minmax(hipStream_t s, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                   const float* X, uint32_t* mm, uint32_t* bad);
void launch_cell_keys(hipStream_t s, bool wide, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                      const float* X, void* keys, uint32_t* idx);
void launch_bin_ids(hipStream_t s, const DJob* jobs, const float* X, int32_t* box, uint32_t n);
void launch_radix_hist(hipStream_t s, bool wide, const DJob* jobs, const uint2* chunks,
                       const uint32_t* chunk_first, uint32_t n_chunks, const void* keys, int shift,
                       uint32_t* hist);
// scoord != nullptr: also write sorted coordinates and full cell keys (last
// pass).  staged: LDS-staged variant (one workgroup of four waves per chunk,
// coalesced stores).
// iin == nullptr: the input order is each job's base order, bidx + bbase[job]
// (the first pass after launch_major_keys, which then writes no permutation).
void launch_radix_scatter(hipStream_t s, bool wide, bool staged, const DJob* jobs, const uint2* chunks,
                          const uint32_t* chunk_first, uint32_t n_chunks, const void* kin,
                          const uint32_t* iin, void* kout, uint32_t* iout, int shift, const uint32_t* offs,
                          const float* X, float* scoord, const uint64_t* bbase = nullptr,
                          const uint32_t* bidx = nullptr);
// Two-stage sort, stage 2 input: job order from its base's sorted points
// (bidx + bbase[job]); key = box key of dims 1.. (or, with scoord, the full
// key plus the sorted coordinates; idx == nullptr without scoord: no
// permutation, the next pass reads the base order).  hist != nullptr (without scoord): also
// the per-chunk digit counts of the first radix pass over these keys (shift
// 0), as launch_radix_hist writes them.
void launch_major_keys(hipStream_t s, bool wide, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                       const uint64_t* bbase, const uint32_t* bidx, const float* X, void* keys,
                       uint32_t* idx, float* scoord, const uint32_t* chunk_first, uint32_t* hist);
uint64_t scan_scratch_elems(uint64_t m);
void launch_exclusive_scan(hipStream_t s, const uint32_t* in, uint32_t* out, uint64_t m,
                           uint32_t* scratch);
void launch_gather_sorted(hipStream_t s, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                          const float* X, const uint32_t* idx, float* scoord);
// Tile boundaries: every job.tq points, at row changes and between dim-0
// groups (flagged and counted per chunk by launch_cell_starts: tbits holds
// one 32-bit word per lane and chunk, offs the scanned counts).
void launch_tile_write(hipStream_t s, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                       const uint32_t* tbits, const uint32_t* offs, DTile* tiles);
void launch_tile_ends(hipStream_t s, const DJob* jobs, DTile* tiles, const uint32_t* total,
                      uint64_t max_tiles);
// Packs runs of small tiles of one job for side-by-side processing (in place).
void launch_tile_pack(hipStream_t s, const DJob* jobs, DTile* tiles, const uint32_t* total,
                      uint64_t max_tiles);
// Claim counters of the persistent density kernel: kClaimParts counters
// kClaimStride words apart (separate memory channels), one per partition of
// the tile list; words 2..19 hold the pair statistics (nine uint64).  The buffer has
// kCounterWords words, all cleared before each launch.
#ifndef ENT_CLAIM_PARTS
#define ENT_CLAIM_PARTS 8
#endif
constexpr uint32_t kClaimPartsMax = 8;
constexpr uint32_t kClaimParts = ENT_CLAIM_PARTS;
constexpr uint32_t kClaimStride = 1024;
constexpr uint32_t kCounterWords = kClaimPartsMax * kClaimStride;
static_assert(kClaimParts >= 1 && kClaimParts <= kClaimPartsMax, "claim partitions");
// sub3: the one-query-per-lane variant that runs 3-D jobs' tiles as
// sub-tiles (2-D ones whole), for batches with 3-D jobs and sub-tiles
int density_occupancy(int q, bool raw, bool wide, bool sub3, int* blocks_per_cu);
// One density kernel variant (instantiated in the kde_density_*.hip files).
template <int Q, bool RAW, typename KeyT, int SUBD>
int density_occupancy_v(int* blocks_per_cu);
template <int Q, bool RAW, typename KeyT, int SUBD>
void launch_density_v(hipStream_t s, unsigned grid, const DJob* jobs, const DTile* tiles, const uint32_t* n_tiles,
                      uint32_t* counter, const float* X, const float* scoord, const void* keys,
                      const uint32_t* cstart, const uint32_t* idx, const float4* mom, const float* mom2,
                      double* dens, unsigned long long* pairs);
int launch_density(hipStream_t s, int q, bool raw, bool wide, bool sub3, unsigned grid, const DJob* jobs,
                   const DTile* tiles, const uint32_t* n_tiles, uint32_t* counter, const float* X,
                   const float* scoord, const void* keys, const uint32_t* cstart, const uint32_t* idx,
                   const float4* mom, const float* mom2, double* dens, unsigned long long* pairs);
// Box-moment table of the kJobMoments jobs listed in mlist ([0] = M, [1..M+1]
// = first wave per job, then the M job indices); max_waves bounds the list's
// waves (sets the grid, which strides over them).
void launch_box_moments(hipStream_t s, const DJob* jobs, const uint32_t* mlist, uint32_t max_waves,
                        const float* scoord, const uint32_t* start, float4* mom);
// Boxes per wave of k_box_moments (a wave sums its boxes one after another;
// few per wave keep a single job's table from waiting on a few long waves).
constexpr uint32_t kMomentBoxesPerWave = 4;
// Moment table of the 2-D kJobMoments jobs listed in mlist ([0] = M, [1..M+1]
// = first block per job, kBlock cells per block, then the M job indices);
// max_blocks bounds the list's blocks.
void launch_cell_moments2(hipStream_t s, const DJob* jobs, const uint32_t* mlist, uint32_t max_blocks,
                          const float* scoord, const uint32_t* start, float* mom);
// Cell-start tables of the jobs that have one, and the tile count of every
// chunk (chunks in tile order) into tile_cnt, in one pass over the keys; the
// tile-start flags into tbits (kWave words per chunk: bit i of lane l's word
// is point 64 i + l of the chunk).
void launch_cell_starts(hipStream_t s, bool wide, const DJob* jobs, const uint2* chunks, uint32_t n_chunks,
                        const void* keys, uint32_t* tile_cnt, uint32_t* start, uint32_t* tbits);

}  // namespace ent
