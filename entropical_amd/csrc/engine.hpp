// engine.hpp -- host side of the MI355X KDE engine (one context per device).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <map>
#include <string>
#include <vector>
#include "kde_engine.hpp"

namespace ent {

enum Status : int {
  kOk = 0,
  kErrArg = 1,
  kErrNoDevice = 2,
  kErrHip = 3,
  kErrOom = 4,
  kErrNonFinite = 5,
  kErrUnsupported = 6,
};

void set_error(const std::string& msg);
const char* last_error();

// One density estimate.  Point k, dim e: X[off[e] + k], k in [0, n).
struct JobSpec {
  uint32_t d = 1;
  uint64_t off[3] = {0, 0, 0};
  float h[3] = {1.f, 1.f, 1.f};
  uint32_t n = 0;       // points
  uint32_t n_rows = 0;  // normalisation count (original rows)
  uint64_t out_off = 0; // first slot of this job's densities in the output
  uint32_t flags = 0;   // kJobTri: triangular kernel (epa self-convolution)
  double norm = 0.0;    // 0: prod_e(0.75/h_e)/n_rows; else this factor
};

enum Stop : int { kStopNone = 0, kStopAfterPlan = 1, kStopAfterSort = 2 };

enum Method : int { kMethodCell = 0, kMethodBrute = 1 };

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

// Shape of a captured launch sequence: a graph is replayed only for an
// identical shape, identical pointers and unchanged workspace allocations.
struct PlanKey {
  const void* X = nullptr;
  const void* dens = nullptr;
  uint64_t P = 0, gen = 0, PB = 0;
  uint32_t J = 0, NC = 0, NCB = 0;
  int npass = 0, raw = 0, q = 0, m0 = 0, mr = 0, gh = 0, sub = 0, grid = 0, variant = 0;
  bool operator==(const PlanKey& o) const {
    return X == o.X && dens == o.dens && P == o.P && gen == o.gen && PB == o.PB && J == o.J &&
           NC == o.NC && NCB == o.NCB && npass == o.npass && raw == o.raw && q == o.q && m0 == o.m0 &&
           mr == o.mr && gh == o.gh && sub == o.sub && grid == o.grid && variant == o.variant;
  }
};

struct RunStats {
  uint32_t graph_replays = 0;  // captured graphs replayed in the last run
  double density_ms = 0.0;   // persistent density launch (HIP events)
  double total_ms = 0.0;     // whole pipeline on the stream (HIP events)
  double plan_ms = 0.0;      // minmax + host planning round trip
  uint64_t pairs = 0;        // executed (query, candidate) evaluations
  uint64_t pairs_by_d[3] = {0, 0, 0};  // the same, per job dimension
  uint64_t lane_pairs = 0;   // lane-slot evaluations issued (incl. idle lanes / padding)
  uint64_t moment_pairs[2] = {0, 0};  // candidate pairs (within pairs_by_d) of 1-D / 2-D jobs
                                      // summed from box / cell moments
  uint64_t moment_terms[2] = {0, 0};  // (query, box) terms of 1-D jobs, (query, cell) of 2-D jobs
  uint64_t tiles = 0;
  uint64_t raw_tiles = 0;  // tiles (and packed tiles) whose pair terms used raw coordinates
  uint32_t radix_passes = 0;  // base passes + job passes
  uint32_t sort_bases = 0;    // shared dim-0 sorts (0: one sort per job by full key)
  uint64_t table_cells = 0;  // cell-start table entries of the last plan
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  int num_cus() const { return num_cus_; }

  // Densities of every job over device-resident X into device dens.
  int densities(const float* X, const std::vector<JobSpec>& jobs, double* dens, int method,
                int stop = kStopNone);

  Tuning tuning;
  RunStats last;

  // Autotuner: full grid over (q, m0, mr, group_h) on a subsample of
  // the jobs, timing whole runs (planning round trip, sort, tiles, density
  // launch) with HIP events.  Sets `tuning` to the best and caches it under a
  // signature of the workload (in memory, and in the file named by
  // ENT_TUNE_CACHE when set).
  int autotune(const float* X, const std::vector<JobSpec>& jobs, double* dens, Tuning* best,
               double* best_ms);

  // grow-only device workspace
  int ensure(DevBuf& b, size_t bytes);
  int ensure_host(size_t bytes);
  void* host_stage() { return host_stage_; }

  // Fixed-arena mode (for callers that cannot allocate): every workspace
  // buffer of a run is carved afresh from the device range [base, base + cap)
  // (256-byte aligned, in a fixed order), the host staging buffer of
  // host_cap bytes is allocated here, once, and launches are issued without
  // graph capture.  A run whose workspace does not fit fails with
  // kErrUnsupported; nothing is allocated or freed after this call.
  int use_arena(void* base, size_t cap, size_t host_cap);
  bool arena() const { return arena_base_ != nullptr; }
  size_t arena_capacity() const { return arena_cap_; }

 private:
  int plan_and_run_cell(const float* X, const std::vector<JobSpec>& jobs, double* dens, int stop);
  int run_brute(const float* X, const std::vector<JobSpec>& jobs, double* dens);
  int launch_density_common(bool raw, const float* X, double* dens, uint64_t max_tiles);
  PlanKey plan_key(const float* X, const double* dens, uint32_t J, uint32_t NC, uint64_t P, int npass,
                   bool raw) const;
  struct GraphCache {
    std::vector<std::pair<PlanKey, hipGraphExec_t>> entries;  // most recent last
    static constexpr size_t kMax = 8;
  };
  template <class F>
  int run_captured(GraphCache& cache, const PlanKey& key, F&& issue);

  void reset_arena();
  char* arena_base_ = nullptr;
  size_t arena_cap_ = 0, arena_off_ = 0;

  GraphCache pre_graphs_, dens_graphs_;
  uint64_t generation_ = 0;  // bumped on every workspace reallocation
  std::map<std::string, Tuning> tune_cache_;
  bool tune_cache_loaded_ = false;

 public:
  bool use_graph = true;

 private:

  int device_ = 0;
  int num_cus_ = 0;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_[6] = {};
  void* host_stage_ = nullptr;
  size_t host_stage_cap_ = 0;

 public:
  // workspace buffers
  DevBuf d_jobs, d_chunks, d_chunk_first, d_mm, d_keys[2], d_idx[2], d_hist,
      d_scan_scratch, d_scoord, d_tile_cnt, d_tiles, d_counter, d_cellstart;
  // two-stage sort: base (shared dim-0 sort) tables and buffers
  DevBuf d_bjobs, d_bchunks, d_bchunk_first, d_bbase, d_bkeys[2], d_bidx[2];
  // min/max windows: the distinct (column offset, length) of all job dims
  DevBuf d_wjobs, d_wchunks;
  // chunk list in tile order (jobs with the most candidates per query first)
  DevBuf d_tchunks;
  DevBuf d_moments, d_mlist;    // box-moment table of the 1-D jobs and its wave list
  DevBuf d_moments2, d_mlist2;  // cell-moment table of the 2-D jobs and its block list
  // host copies
  std::vector<DJob> h_jobs;
  std::vector<uint2> h_chunks;
  std::vector<uint32_t> h_chunk_first;
  std::vector<DJob> h_bjobs;
  std::vector<uint2> h_bchunks;
  std::vector<uint32_t> h_bchunk_first;
  std::vector<uint64_t> h_bbase;  // per job: first slot of its base
  std::vector<DJob> h_wjobs;      // one 1-D job per min/max window
  std::vector<uint2> h_wchunks;
  std::vector<uint32_t> h_jw;     // per job and dim (3 per job): its window
  std::vector<uint2> h_tchunks;      // chunk list in tile order (last upload)
  std::vector<uint32_t> h_tord;      // job order of the last upload
  std::vector<uint32_t> h_tsizes;    // job sizes of the last upload
  std::vector<uint32_t> h_mlist;     // box-moment wave list (k_box_moments)
  std::vector<uint32_t> h_mlist2;    // cell-moment block list (k_cell_moments2)
  uint64_t total_points = 0;
  int sorted_buf = 0;  // which of d_keys/d_idx holds the sorted result
  bool keys_wide = false;  // cell keys stored as uint64 (else uint32)
  bool sub3 = false;       // 3-D jobs with sub-tiles: the q = 1 variant with 3-D sub-tiles
};

}  // namespace ent
