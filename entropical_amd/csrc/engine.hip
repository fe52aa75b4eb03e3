// engine.hip -- host side of the KDE engine: workspace, planning, launches.
#include "engine.hpp"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include "kde_kernels.hpp"

namespace ent {

namespace {
thread_local std::string g_last_error;

inline float o2f(uint32_t o) {
  uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

#define ENT_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call);     \
      return e_ == hipErrorOutOfMemory ? kErrOom : kErrHip;                             \
    }                                                                                   \
  } while (0)

#define ENT_TRY(expr)              \
  do {                             \
    int r_ = (expr);               \
    if (r_ != kOk) return r_;      \
  } while (0)

Engine::Engine(int device) : device_(device) {
  if (hipSetDevice(device) != hipSuccess) return;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) num_cus_ = prop.multiProcessorCount;
  // failures leave stream_/events null; the first use reports the error
  (void)hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
  for (auto& e : ev_) (void)hipEventCreate(&e);
}

Engine::~Engine() {
  // teardown: errors here have no one to report to
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  DevBuf* bufs[] = {&d_jobs,  &d_chunks, &d_chunk_first, &d_mm,     &d_keys[0],
                    &d_keys[1], &d_idx[0], &d_idx[1],     &d_hist,   &d_scan_scratch, &d_scoord,
                    &d_tile_cnt, &d_tiles, &d_counter, &d_cellstart, &d_bjobs, &d_bchunks,
                    &d_bchunk_first, &d_bbase, &d_bkeys[0], &d_bkeys[1], &d_bidx[0], &d_bidx[1],
                    &d_wjobs, &d_wchunks, &d_tchunks, &d_moments, &d_mlist, &d_moments2, &d_mlist2};
  for (GraphCache* c : {&pre_graphs_, &dens_graphs_})
    for (auto& e : c->entries) (void)hipGraphExecDestroy(e.second);
  if (!arena())  // arena buffers belong to the caller's range
    for (DevBuf* b : bufs)
      if (b->p) (void)hipFree(b->p);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (host_stage_) (void)hipHostFree(host_stage_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::use_arena(void* base, size_t cap, size_t host_cap) {
  // the last 64 KiB stay unassigned, so a kernel's read-ahead past the end
  // of the last buffer stays inside the range
  constexpr size_t kGuard = 64 * 1024;
  if (!base || cap <= 16 * kGuard) {
    set_error("fixed arena: null or too small range");
    return kErrArg;
  }
  ENT_TRY(ensure_host(host_cap));
  arena_base_ = (char*)base;
  arena_cap_ = cap - kGuard;
  use_graph = false;
  reset_arena();
  return kOk;
}

// Forget every buffer of the previous run: this run carves them again from
// the start of the arena.
void Engine::reset_arena() {
  DevBuf* bufs[] = {&d_jobs,  &d_chunks, &d_chunk_first, &d_mm,     &d_keys[0],
                    &d_keys[1], &d_idx[0], &d_idx[1],     &d_hist,   &d_scan_scratch, &d_scoord,
                    &d_tile_cnt, &d_tiles, &d_counter, &d_cellstart, &d_bjobs, &d_bchunks,
                    &d_bchunk_first, &d_bbase, &d_bkeys[0], &d_bkeys[1], &d_bidx[0], &d_bidx[1],
                    &d_wjobs, &d_wchunks, &d_tchunks, &d_moments, &d_mlist, &d_moments2, &d_mlist2};
  for (DevBuf* b : bufs) *b = DevBuf{};
  arena_off_ = 0;
  ++generation_;
}

int Engine::ensure(DevBuf& b, size_t bytes) {
  if (bytes <= b.cap) return kOk;
  if (arena()) {
    const size_t off = (arena_off_ + 255) & ~(size_t)255;
    const size_t want = std::max<size_t>(bytes, 256);
    if (off > arena_cap_ || want > arena_cap_ - off) {
      set_error("workspace exceeds the fixed scratch range of " + std::to_string(arena_cap_ >> 20) + " MiB");
      return kErrUnsupported;
    }
    b.p = arena_base_ + off;
    b.cap = want;
    arena_off_ = off + want;
    return kOk;
  }
  ENT_HIP(hipStreamSynchronize(stream_));
  if (b.p) ENT_HIP(hipFree(b.p));
  b.p = nullptr;
  b.cap = 0;
  // captured graphs hold the old address: none is replayed from here on,
  // whether or not the new allocation succeeds
  ++generation_;
  size_t want = std::max<size_t>(bytes + bytes / 8, 4096);
  ENT_HIP(hipMalloc(&b.p, want));
  b.cap = want;
  return kOk;
}

int Engine::ensure_host(size_t bytes) {
  if (bytes <= host_stage_cap_) return kOk;
  if (arena()) {
    set_error("host staging exceeds the fixed buffer of " + std::to_string(host_stage_cap_) + " bytes");
    return kErrUnsupported;
  }
  ENT_HIP(hipStreamSynchronize(stream_));
  if (host_stage_) ENT_HIP(hipHostFree(host_stage_));
  host_stage_ = nullptr;
  host_stage_cap_ = 0;
  size_t want = std::max<size_t>(bytes, 1 << 20);
  ENT_HIP(hipHostMalloc(&host_stage_, want, hipHostMallocDefault));
  host_stage_cap_ = want;
  return kOk;
}

int Engine::densities(const float* X, const std::vector<JobSpec>& jobs, double* dens, int method,
                      int stop) {
  if (!stream_) {
    set_error("engine has no stream (device initialisation failed)");
    return kErrNoDevice;
  }
  ENT_HIP(hipSetDevice(device_));
  last = RunStats{};
  if (arena()) reset_arena();
  if (jobs.empty()) return kOk;
  for (const JobSpec& j : jobs) {
    if (j.d < 1 || j.d > 3) {
      set_error("job dimension must be 1, 2 or 3");
      return kErrArg;
    }
    if (j.n == 0) {
      set_error("job with zero points");
      return kErrArg;
    }
    if (j.n >= (1u << 30)) {  // the density kernel addresses a job's columns with 32-bit byte offsets
      set_error("job exceeds 2^30 points");
      return kErrUnsupported;
    }
    for (uint32_t e = 0; e < j.d; ++e)
      if (!(j.h[e] > 0.0f) || !std::isfinite(j.h[e])) {
        set_error("bandwidths must be finite and > 0");
        return kErrArg;
      }
  }
  // host job table
  h_jobs.assign(jobs.size(), DJob{});
  h_chunks.clear();
  h_chunk_first.assign(jobs.size() + 1, 0);
  uint64_t pbase = 0;
  for (size_t j = 0; j < jobs.size(); ++j) {
    const JobSpec& s = jobs[j];
    DJob& J = h_jobs[j];
    std::memset(&J, 0, sizeof(J));
    J.d = s.d;
    J.n = s.n;
    J.n_rows = s.n_rows;
    J.flags = (method == kMethodBrute ? kJobBrute : 0u) | s.flags;
    if ((s.flags & kJobTri) && s.d != 1) {
      set_error("triangular-kernel jobs must be 1-D");
      return kErrArg;
    }
    J.pbase = pbase;
    J.out_off = s.out_off;
    double norm = 1.0 / (double)s.n_rows;
    for (uint32_t e = 0; e < 3; ++e) {
      J.off[e] = e < s.d ? s.off[e] : 0;
      J.h[e] = e < s.d ? s.h[e] : 1.0f;
      J.hinv[e] = 1.0f / J.h[e];
      J.nb[e] = 1;
      J.bs[e] = 1.0f;
      if (e < s.d) norm *= 0.75 / (double)s.h[e];
    }
    J.norm = s.norm != 0.0 ? s.norm : norm;
    h_chunk_first[j] = (uint32_t)h_chunks.size();
    for (uint32_t st = 0; st < s.n; st += kChunk) h_chunks.push_back(make_uint2((uint32_t)j, st));
    pbase += s.n;
  }
  h_chunk_first[jobs.size()] = (uint32_t)h_chunks.size();
  total_points = pbase;
  if (total_points >= (1ull << 32)) {
    set_error("batch exceeds 2^32 points");
    return kErrUnsupported;
  }
  if (method == kMethodBrute) return run_brute(X, jobs, dens);
  return plan_and_run_cell(X, jobs, dens, stop);
}

int Engine::plan_and_run_cell(const float* X, const std::vector<JobSpec>& jobs, double* dens,
                              int stop) {
  const uint32_t J = (uint32_t)jobs.size();
  const uint32_t NC = (uint32_t)h_chunks.size();
  const uint64_t P = total_points;
  // Min/max per distinct 1-D window (column offset, length) rather than per
  // job and dim: the jobs of a batch share columns (the MI matrix's 4160
  // job dims read 64 windows).
  {
    std::map<std::pair<uint64_t, uint32_t>, uint32_t> wid;
    h_wjobs.clear();
    h_wchunks.clear();
    h_jw.assign(3 * (size_t)J, 0u);
    for (uint32_t j = 0; j < J; ++j)
      for (uint32_t e = 0; e < h_jobs[j].d; ++e) {
        const auto key = std::make_pair(h_jobs[j].off[e], h_jobs[j].n);
        auto ins = wid.emplace(key, (uint32_t)h_wjobs.size());
        if (ins.second) {
          DJob W;
          std::memset(&W, 0, sizeof(W));
          W.d = 1;
          W.n = key.second;
          W.off[0] = key.first;
          for (uint32_t st = 0; st < W.n; st += kChunk)
            h_wchunks.push_back(make_uint2((uint32_t)h_wjobs.size(), st));
          h_wjobs.push_back(W);
        }
        h_jw[3 * (size_t)j + e] = ins.first->second;
      }
  }
  const uint32_t NW = (uint32_t)h_wjobs.size();
  const uint32_t NWC = (uint32_t)h_wchunks.size();
  ENT_TRY(ensure(d_jobs, sizeof(DJob) * J));
  ENT_TRY(ensure(d_chunks, sizeof(uint2) * NC));
  ENT_TRY(ensure(d_chunk_first, sizeof(uint32_t) * (J + 1)));
  ENT_TRY(ensure(d_wjobs, sizeof(DJob) * NW));
  ENT_TRY(ensure(d_wchunks, sizeof(uint2) * NWC));
  // min/max slots of the windows, then the non-finite flag (one read-back)
  ENT_TRY(ensure(d_mm, sizeof(uint32_t) * (6 * NW + 16)));
  // host staging: jobs + chunks + chunk_first + windows + minmax readback
  size_t stage_bytes = sizeof(DJob) * J + sizeof(uint2) * NC + sizeof(uint32_t) * (J + 1) +
                       8 + sizeof(DJob) * NW + sizeof(uint2) * NWC + sizeof(uint32_t) * 6 * NW + 64;
  ENT_TRY(ensure_host(stage_bytes));
  char* hs = (char*)host_stage_;
  DJob* hs_jobs = (DJob*)hs;
  uint2* hs_chunks = (uint2*)(hs_jobs + J);
  uint32_t* hs_cf = (uint32_t*)(hs_chunks + NC);
  DJob* hs_wjobs = (DJob*)(((uintptr_t)(hs_cf + (J + 1)) + 7) & ~(uintptr_t)7);
  uint2* hs_wchunks = (uint2*)(hs_wjobs + NW);
  uint32_t* hs_mm = (uint32_t*)(hs_wchunks + NWC);
  uint32_t* hs_flags = hs_mm + 6 * NW;
  std::memcpy(hs_jobs, h_jobs.data(), sizeof(DJob) * J);
  std::memcpy(hs_chunks, h_chunks.data(), sizeof(uint2) * NC);
  std::memcpy(hs_cf, h_chunk_first.data(), sizeof(uint32_t) * (J + 1));
  std::memcpy(hs_wjobs, h_wjobs.data(), sizeof(DJob) * NW);
  std::memcpy(hs_wchunks, h_wchunks.data(), sizeof(uint2) * NWC);

  ENT_HIP(hipEventRecord(ev_[0], stream_));
  ENT_HIP(hipMemcpyAsync(d_jobs.p, hs_jobs, sizeof(DJob) * J, hipMemcpyHostToDevice, stream_));
  ENT_HIP(hipMemcpyAsync(d_chunks.p, hs_chunks, sizeof(uint2) * NC, hipMemcpyHostToDevice, stream_));
  ENT_HIP(hipMemcpyAsync(d_chunk_first.p, hs_cf, sizeof(uint32_t) * (J + 1), hipMemcpyHostToDevice, stream_));
  const DJob* dj = (const DJob*)d_jobs.p;
  const uint2* dch = (const uint2*)d_chunks.p;
  const uint32_t* dcf = (const uint32_t*)d_chunk_first.p;
  ENT_HIP(hipMemcpyAsync(d_wjobs.p, hs_wjobs, sizeof(DJob) * NW, hipMemcpyHostToDevice, stream_));
  ENT_HIP(hipMemcpyAsync(d_wchunks.p, hs_wchunks, sizeof(uint2) * NWC, hipMemcpyHostToDevice, stream_));
  launch_init_minmax(stream_, (uint32_t*)d_mm.p, NW);
  uint32_t* d_flag = (uint32_t*)d_mm.p + 6 * NW;
  launch_fill_u32(stream_, d_flag, 0u, 16);
  launch_minmax(stream_, (const DJob*)d_wjobs.p, (const uint2*)d_wchunks.p, NWC, X, (uint32_t*)d_mm.p, d_flag);
  ENT_HIP(hipGetLastError());
  ENT_HIP(hipMemcpyAsync(hs_mm, d_mm.p, sizeof(uint32_t) * (6 * NW + 1), hipMemcpyDeviceToHost, stream_));
  ENT_HIP(hipStreamSynchronize(stream_));
  if (hs_flags[0] != 0) {
    set_error("non-finite coordinate in density input");
    return kErrNonFinite;
  }
  // host planning: boxes per dim (boxedsearch.cpp:21-27 with the total fixes)
  uint64_t max_key = 0;
  uint64_t table_total = 0;
  // Sub-tile shapes in use (others keep whole and packed tiles): 3-D jobs
  // with rows h high (sub-tiles of 16 or 32) or h/2 (sub-tiles of 16); 2-D
  // jobs with rows h/2 or finer (sub-tiles of 16) or h/4 or finer (of 32),
  // and at any row height in a batch whose 3-D jobs run as sub-tiles (its
  // 2-D tiles then run whole, with the sub-tile moment table).  The sub-tile
  // code handles every shape; the others are held back because a randomized
  // sweep over them faulted the GPU for a reason not yet found.
  const bool sub3_shape = (tuning.sub == 16 && tuning.mr <= 2) || (tuning.sub == 32 && tuning.mr == 1);
  bool batch3 = false;
  for (const DJob& D : h_jobs) batch3 = batch3 || D.d == 3;
  const bool sub2_shape = (tuning.sub == 16 && tuning.mr >= 2) || (tuning.sub == 32 && tuning.mr >= 4) ||
                          (batch3 && sub3_shape);
  for (uint32_t j = 0; j < J; ++j) {
    DJob& D = h_jobs[j];
    uint64_t cells = 1;
    for (uint32_t e = 0; e < D.d; ++e) {
      const uint32_t w = h_jw[3 * (size_t)j + e];
      float mn = o2f(hs_mm[w * 6]);
      float mx = o2f(hs_mm[w * 6 + 1]);
      float m = (float)(e == 0 ? tuning.m0 : tuning.mr);
      volatile float bs = D.h[e] / m;
      volatile float range = mx - mn;
      volatile float q = range / bs;
      float nbf = std::ceil((float)q);
      if (!(nbf < 16777216.0f)) {
        set_error("bandwidth too small for the data range (more than 2^24 boxes per dim)");
        return kErrUnsupported;
      }
      uint32_t nb = nbf < 1.0f ? 1u : (uint32_t)nbf;
      D.mn[e] = mn;
      D.bs[e] = bs;
      D.nb[e] = nb;
      cells *= nb;
      if (cells >= (1ull << 62)) {
        set_error("too many cells for the bandwidths given");
        return kErrUnsupported;
      }
    }
    D.group = (uint32_t)std::max(1, tuning.group_h * tuning.m0);
    // 2-D jobs, and 3-D jobs with rows at least h/2 high (a sub-tile's
    // candidate rows then fit its lanes' two range entries each), with
    // sub-tiles: tiles of up to 64 queries (four of 16 or two of 32 side by
    // side in one wave)
    D.tq = (uint32_t)(kWave * tuning.q);
    if ((D.d == 2 && sub2_shape) || (D.d == 3 && sub3_shape)) {
      D.tq = kWave;
      D.flags |= tuning.sub == 16 ? kJobSub16 : kJobSub32;
    }
    D.cells = cells;
    D.cbase = table_total;
    const uint64_t table_cap = kTableCellsPerPoint * (uint64_t)D.n + 4096;
    if (cells + 1 <= table_cap) {
      table_total += cells + 1;
    } else if (cells / D.nb[0] + 1 <= table_cap) {
      // one entry per row of boxes: a binary search over one row's points
      // instead of the job's
      D.flags |= kJobRowTable;
      table_total += cells / D.nb[0] + 1;
    } else {
      // more rows than the table bound (sparse heavy-tailed 3-D data with
      // narrow bandwidths): no table, each range is a binary search over the
      // job's keys
      D.flags |= kJobRowTable;
      D.cbase = kNoTable;
    }
    max_key = std::max<uint64_t>(max_key, cells - 1);
  }
  // Box moments for 1-D Epanechnikov jobs with a cell-start table and no more
  // boxes than points (on unless ENT_BOX_MOMENTS=0): each job's boxes get a
  // moment table, listed for k_box_moments in waves of kMomentBoxesPerWave
  // boxes.
  bool moments = true;
  if (const char* m = std::getenv("ENT_BOX_MOMENTS"))
    if (!std::strcmp(m, "0")) moments = false;
  // 2-D jobs get a table of cells one row high and about h/4 wide in dim 0
  // (a power of two near m0/4 dim-0 boxes), when the job has a cell-start
  // table and at most a quarter as many cells as points.
  uint64_t mom_total = 0, mom2_total = 0;
  h_mlist.assign(1, 0u);
  h_mlist2.assign(1, 0u);
  {
    std::vector<uint32_t> mj, mj2;
    uint32_t waves = 0, blocks2 = 0;
    for (uint32_t j = 0; j < J; ++j) {
      DJob& D = h_jobs[j];
      D.flags &= ~kJobMoments;
      D.mbase = 0;
      D.mcf = 1;
      D.mncb = D.nb[0];
      if (!moments || (D.flags & (kJobRowTable | kJobTri))) continue;
      if (D.d == 1 && D.nb[0] <= D.n) {
        D.flags |= kJobMoments;
        D.mbase = mom_total;
        mom_total += D.nb[0];
        mj.push_back(j);
        h_mlist.push_back(waves);
        waves += (D.nb[0] + kMomentBoxesPerWave - 1) / kMomentBoxesPerWave;
      } else if (D.d == 2) {
        // dim-0 boxes per cell: the power of two nearest below m0 / 4 (cells
        // about h/4 wide); jobs with sub-tiles merge their cells by rows, so
        // narrower cells (m0 / 8, about h/8) cost them little and leave fewer
        // pairs to the pair loop
        const bool sub = (D.flags & (kJobSub16 | kJobSub32)) != 0;
        uint32_t cf = 1;
        while (cf * 2 <= (uint32_t)std::max(1, tuning.m0 / (sub ? 8 : 4))) cf *= 2;
        const uint32_t ncb = (D.nb[0] + cf - 1) / cf;
        const uint64_t cells = (uint64_t)ncb * D.nb[1];
        // cells hold 4 points on average at least (at most 9 B of table per
        // point); jobs with sub-tiles, whose cells are smaller, 1 point
        const uint64_t per = sub ? 1 : 4;
        if (per * cells > D.n) continue;
        D.flags |= kJobMoments;
        D.mbase = mom2_total;
        D.mcf = cf;
        D.mncb = ncb;
        mom2_total += cells;
        mj2.push_back(j);
        h_mlist2.push_back(blocks2);
        blocks2 += (uint32_t)((cells + kBlock - 1) / kBlock);
      }
    }
    h_mlist[0] = (uint32_t)mj.size();
    h_mlist.push_back(waves);
    h_mlist.insert(h_mlist.end(), mj.begin(), mj.end());
    h_mlist2[0] = (uint32_t)mj2.size();
    h_mlist2.push_back(blocks2);
    h_mlist2.insert(h_mlist2.end(), mj2.begin(), mj2.end());
  }
  bool has_mom = mom_total != 0;
  bool has_mom2 = mom2_total != 0;
  // grid bound of k_cell_moments2 (kBlock cells per block; cells <= points;
  // the kernel strides over the listed blocks)
  const uint32_t mom2_blocks = (uint32_t)std::min<uint64_t>(P / kBlock + J, 1u << 30);
  // grid bound of k_box_moments from the plan's shape: the listed boxes are at
  // most the points
  const uint32_t mom_waves = (uint32_t)std::min<uint64_t>(P / kMomentBoxesPerWave + J, 1u << 30);
  auto passes_for = [](uint64_t max_value) {
    int bits = 0;
    while (bits < 64 && (max_value >> bits) != 0) ++bits;
    return (bits + kRadixBits - 1) / kRadixBits;
  };
  int bits = 0;
  while (bits < 64 && (max_key >> bits) != 0) ++bits;
  const int npass_full = passes_for(max_key);
  const bool wide = bits > 32;
  keys_wide = wide;
  sub3 = false;
  for (const DJob& D : h_jobs)
    if (D.d == 3 && (D.flags & (kJobSub16 | kJobSub32))) sub3 = true;

  // Two-stage sort: jobs with the same dim-0 window (offset, length) and
  // dim-0 bandwidth have the same dim-0 boxes; one stable sort of that window
  // by dim-0 box (a base) serves all of them, and each job then needs only
  // the passes over its remaining dims' key (k_major_keys).  3-D jobs that
  // share their dims 0 and 1 with another job (the TE set: one pair of lagged
  // windows per series and lag) get a base sorted by those two dims and one
  // pass over dim 2.  Chosen when the bases hold at most half the batch's
  // points; ENT_SORT=full|shared forces a mode (all give the same permutation).
  h_bjobs.clear();
  h_bchunks.clear();
  h_bchunk_first.clear();
  h_bbase.assign(J, 0);
  uint64_t PB = 0, max_b0 = 0, max_major = 0;
  {
    // (off0, off1, n, h0 bits, h1 bits); off1 = ~0 for a dim-0 base
    using BaseKey = std::tuple<uint64_t, uint64_t, uint32_t, uint32_t, uint32_t>;
    auto key_of = [&](const DJob& D, bool two) {
      uint32_t h0, h1 = 0;
      std::memcpy(&h0, &D.h[0], 4);
      if (two) std::memcpy(&h1, &D.h[1], 4);
      return BaseKey(D.off[0], two ? D.off[1] : ~0ull, D.n, h0, h1);
    };
    std::map<BaseKey, uint32_t> count2;
    for (uint32_t j = 0; j < J; ++j)
      if (h_jobs[j].d == 3) ++count2[key_of(h_jobs[j], true)];
    std::map<BaseKey, uint32_t> seen;
    for (uint32_t j = 0; j < J; ++j) {
      DJob& D = h_jobs[j];
      const bool two = D.d == 3 && count2[key_of(D, true)] >= 2;
      if (two) D.flags |= kJobBase2;
      auto ins = seen.emplace(key_of(D, two), (uint32_t)h_bjobs.size());
      if (ins.second) {
        DJob B = D;
        B.d = two ? 2 : 1;
        for (int e = (int)B.d; e < 3; ++e) {
          B.off[e] = 0;
          B.h[e] = B.hinv[e] = B.bs[e] = 1.0f;
          B.mn[e] = 0.0f;
          B.nb[e] = 1;
        }
        B.flags = 0;
        B.cells = two ? (uint64_t)D.nb[0] * D.nb[1] : D.nb[0];
        B.cbase = kNoTable;
        B.pbase = PB;
        const uint32_t bi = (uint32_t)h_bjobs.size();
        h_bchunk_first.push_back((uint32_t)h_bchunks.size());
        for (uint32_t st = 0; st < D.n; st += kChunk) h_bchunks.push_back(make_uint2(bi, st));
        PB += D.n;
        h_bjobs.push_back(B);
      }
      h_bbase[j] = h_bjobs[ins.first->second].pbase;
      max_b0 = std::max<uint64_t>(max_b0, h_bjobs[ins.first->second].cells - 1);
      max_major = std::max<uint64_t>(max_major, two ? D.nb[2] - 1 : D.cells / D.nb[0] - 1);
    }
    h_bchunk_first.push_back((uint32_t)h_bchunks.size());
  }
  bool shared = 2 * PB <= P;
  if (const char* m = std::getenv("ENT_SORT")) {
    if (!std::strcmp(m, "full")) shared = false;
    if (!std::strcmp(m, "shared")) shared = true;
  }
  bool staged = true;
  if (const char* m = std::getenv("ENT_SCATTER"))
    if (!std::strcmp(m, "direct")) staged = false;
  // packing of small tiles (on unless ENT_TILE_PACK=0; only 2-D and 3-D
  // jobs pack, so a batch of 1-D jobs skips the pass)
  bool pack = std::any_of(h_jobs.begin(), h_jobs.end(), [](const DJob& D) { return D.d >= 2; });
  if (const char* m = std::getenv("ENT_TILE_PACK"))
    if (!std::strcmp(m, "0")) pack = false;
  const uint32_t NB = (uint32_t)h_bjobs.size();
  const uint32_t NCB = (uint32_t)h_bchunks.size();
  const int npass0 = shared ? passes_for(max_b0) : 0;
  const int npass = shared ? passes_for(max_major) : npass_full;
  last.radix_passes = (uint32_t)(npass0 + npass);
  last.sort_bases = shared ? NB : 0;

  // Tile order: the tile list follows this chunk order, jobs with the most
  // candidates per query (n times the fraction of the range within +-h per
  // dim) first, so the longest tiles are claimed early and the kernel's
  // tail holds short ones.  Uploaded only when it changes.
  bool upload_tchunks = false;
  {
    std::vector<double> est(J);
    for (uint32_t j = 0; j < J; ++j) {
      const DJob& D = h_jobs[j];
      double c = (double)D.n;
      for (uint32_t e = 0; e < D.d; ++e) {
        const double range = (double)D.nb[e] * (double)D.bs[e];
        c *= range > 0.0 ? std::min(1.0, 2.0 * (double)D.h[e] / range) : 1.0;
      }
      est[j] = c;
    }
    std::vector<uint32_t> ord(J);
    for (uint32_t j = 0; j < J; ++j) ord[j] = j;
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return est[a] > est[b]; });
    const size_t tchunks_cap = d_tchunks.cap;
    ENT_TRY(ensure(d_tchunks, sizeof(uint2) * NC));
    // a reallocated buffer holds nothing; an unchanged order (same jobs,
    // chunks and order as the last upload) needs no upload
    std::vector<uint32_t> sizes(J);
    for (uint32_t j = 0; j < J; ++j) sizes[j] = h_jobs[j].n;  // the chunk list follows from these
    upload_tchunks = d_tchunks.cap != tchunks_cap || ord != h_tord || sizes != h_tsizes;
    if (upload_tchunks) {
      h_tchunks.clear();
      h_tchunks.reserve(NC);
      for (uint32_t j : ord)
        for (uint32_t c = h_chunk_first[j]; c < h_chunk_first[j + 1]; ++c) h_tchunks.push_back(h_chunks[c]);
      h_tord = ord;
      h_tsizes = sizes;
    }
  }

  // upload the final job table (and the base tables)
  const size_t base_bytes = shared ? sizeof(DJob) * NB + sizeof(uint2) * NCB + sizeof(uint32_t) * (NB + 1) +
                                         sizeof(uint64_t) * J
                                   : 0;
  const size_t mlist_bytes = sizeof(uint32_t) * h_mlist.size();
  const size_t mlist2_bytes = sizeof(uint32_t) * h_mlist2.size();
  ENT_TRY(ensure_host(sizeof(DJob) * J + base_bytes + (upload_tchunks ? sizeof(uint2) * NC : 0) + mlist_bytes +
                      mlist2_bytes + 80));
  {
    char* p = (char*)host_stage_;
    std::memcpy(p, h_jobs.data(), sizeof(DJob) * J);
    ENT_HIP(hipMemcpyAsync(d_jobs.p, p, sizeof(DJob) * J, hipMemcpyHostToDevice, stream_));
    p += sizeof(DJob) * J;
    if (has_mom) {
      ENT_TRY(ensure(d_mlist, mlist_bytes));
      std::memcpy(p, h_mlist.data(), mlist_bytes);
      ENT_HIP(hipMemcpyAsync(d_mlist.p, p, mlist_bytes, hipMemcpyHostToDevice, stream_));
      p += (mlist_bytes + 7) & ~(size_t)7;
    }
    if (has_mom2) {
      ENT_TRY(ensure(d_mlist2, mlist2_bytes));
      std::memcpy(p, h_mlist2.data(), mlist2_bytes);
      ENT_HIP(hipMemcpyAsync(d_mlist2.p, p, mlist2_bytes, hipMemcpyHostToDevice, stream_));
      p += (mlist2_bytes + 7) & ~(size_t)7;
    }
    if (upload_tchunks) {
      std::memcpy(p, h_tchunks.data(), sizeof(uint2) * NC);
      ENT_HIP(hipMemcpyAsync(d_tchunks.p, p, sizeof(uint2) * NC, hipMemcpyHostToDevice, stream_));
      p += sizeof(uint2) * NC;
    }
    if (shared) {
      ENT_TRY(ensure(d_bjobs, sizeof(DJob) * NB));
      ENT_TRY(ensure(d_bchunks, sizeof(uint2) * NCB));
      ENT_TRY(ensure(d_bchunk_first, sizeof(uint32_t) * (NB + 1)));
      ENT_TRY(ensure(d_bbase, sizeof(uint64_t) * J));
      auto up = [&](DevBuf& b, const void* src, size_t n) {
        std::memcpy(p, src, n);
        hipError_t e = hipMemcpyAsync(b.p, p, n, hipMemcpyHostToDevice, stream_);
        p += n;
        return e;
      };
      ENT_HIP(up(d_bjobs, h_bjobs.data(), sizeof(DJob) * NB));
      ENT_HIP(up(d_bchunks, h_bchunks.data(), sizeof(uint2) * NCB));
      ENT_HIP(up(d_bchunk_first, h_bchunk_first.data(), sizeof(uint32_t) * (NB + 1)));
      ENT_HIP(up(d_bbase, h_bbase.data(), sizeof(uint64_t) * J));
    }
  }
  if (stop == kStopAfterPlan) return kOk;

  // buffers sized by the plan
  const uint64_t HM = (uint64_t)kRadixBins * NC;
  const uint64_t HMB = (uint64_t)kRadixBins * NCB;
  const uint64_t scan_m = std::max<uint64_t>(std::max<uint64_t>(HM, NC), shared ? HMB : 0);
  const size_t key_bytes = wide ? sizeof(uint64_t) : sizeof(uint32_t);
  ENT_TRY(ensure(d_keys[0], key_bytes * P));
  ENT_TRY(ensure(d_keys[1], key_bytes * P));
  ENT_TRY(ensure(d_idx[0], sizeof(uint32_t) * P));
  ENT_TRY(ensure(d_idx[1], sizeof(uint32_t) * P));
  ENT_TRY(ensure(d_hist, sizeof(uint32_t) * (scan_m + 1)));
  ENT_TRY(ensure(d_scan_scratch, sizeof(uint32_t) * scan_scratch_elems(scan_m)));
  ENT_TRY(ensure(d_scoord, sizeof(float) * 3 * P));
  ENT_TRY(ensure(d_tile_cnt, sizeof(uint32_t) * (NC + 1)));
  ENT_TRY(ensure(d_tiles, sizeof(DTile) * P));
  ENT_TRY(ensure(d_counter, sizeof(uint32_t) * kCounterWords));
  ENT_TRY(ensure(d_cellstart, sizeof(uint32_t) * std::max<uint64_t>(table_total, 1)));
  if (has_mom || has_mom2) {
    // ENT_MOMENT_TABLE_MAX (bytes) caps the moment tables' memory
    uint64_t cap = ~0ull;
    if (const char* m = std::getenv("ENT_MOMENT_TABLE_MAX")) cap = std::strtoull(m, nullptr, 10);
    const uint64_t bytes1 = sizeof(float4) * mom_total, bytes2 = sizeof(float) * 9 * mom2_total;
    const bool ok = bytes1 + bytes2 <= cap && (!has_mom || ensure(d_moments, bytes1) == kOk) &&
                    (!has_mom2 || ensure(d_moments2, bytes2) == kOk);
    if (!ok) {
      // moment tables over the cap or beyond the device memory left: every
      // pair is evaluated one by one (a failed allocation is not a sticky error)
      (void)hipGetLastError();
      ENT_HIP(hipStreamSynchronize(stream_));
      for (DJob& D : h_jobs) D.flags &= ~kJobMoments;
      std::memcpy(host_stage_, h_jobs.data(), sizeof(DJob) * J);
      ENT_HIP(hipMemcpyAsync(d_jobs.p, host_stage_, sizeof(DJob) * J, hipMemcpyHostToDevice, stream_));
      has_mom = has_mom2 = false;
    }
  }
  if (shared) {
    for (int b = 0; b < 2; ++b) {
      ENT_TRY(ensure(d_bkeys[b], key_bytes * PB));
      ENT_TRY(ensure(d_bidx[b], sizeof(uint32_t) * PB));
    }
  }
  last.table_cells = table_total;

  void* keys[2] = {d_keys[0].p, d_keys[1].p};
  uint32_t* idx[2] = {(uint32_t*)d_idx[0].p, (uint32_t*)d_idx[1].p};
  uint32_t* hist = (uint32_t*)d_hist.p;
  uint32_t* scratch = (uint32_t*)d_scan_scratch.p;
  float* scoord = (float*)d_scoord.p;

  sorted_buf = npass & 1;
  uint32_t* tcnt = (uint32_t*)d_tile_cnt.p;
  // stable LSD passes over `np` digits of a chunked batch; the last one of
  // the main batch gathers the coordinates and writes full keys
  const uint32_t* bidx_final = nullptr;  // the bases' sorted order (shared sorts)
  // first_counted: pass 0's digit counts and keys come from k_major_keys,
  // and its input order is each job's base order (bidx + bbase[job])
  auto radix = [&](const DJob* jb, const uint2* cb, const uint32_t* cf, uint32_t nc, void** kk, uint32_t** ii,
                   int np, float* sc_last, bool first_counted) {
    const uint64_t hm = (uint64_t)kRadixBins * nc;
    int cur = 0;
    for (int pass = 0; pass < np; ++pass) {
      const int shift = pass * kRadixBits;
      // the first pass's digit counts may come with the keys (k_major_keys)
      if (pass > 0 || !first_counted) launch_radix_hist(stream_, wide, jb, cb, cf, nc, kk[cur], shift, hist);
      launch_exclusive_scan(stream_, hist, hist, hm, scratch);
      const bool from_base = pass == 0 && first_counted;
      launch_radix_scatter(stream_, wide, staged, jb, cb, cf, nc, kk[cur], from_base ? nullptr : ii[cur],
                           kk[cur ^ 1], ii[cur ^ 1], shift, hist, X, pass == np - 1 ? sc_last : nullptr,
                           from_base ? (const uint64_t*)d_bbase.p : nullptr, from_base ? bidx_final : nullptr);
      cur ^= 1;
    }
    return cur;
  };
  // keys -> radix passes [-> gather -> tiles]; a fixed launch sequence for a
  // given plan shape, replayed from a captured graph when the shape repeats
  auto issue_pre = [&](bool with_tiles) {
    int cur = 0;
    if (shared) {
      const DJob* dbj = (const DJob*)d_bjobs.p;
      void* bk[2] = {d_bkeys[0].p, d_bkeys[1].p};
      uint32_t* bi[2] = {(uint32_t*)d_bidx[0].p, (uint32_t*)d_bidx[1].p};
      launch_cell_keys(stream_, wide, dbj, (const uint2*)d_bchunks.p, NCB, X, bk[0], bi[0]);
      const int bc = radix(dbj, (const uint2*)d_bchunks.p, (const uint32_t*)d_bchunk_first.p, NCB, bk, bi,
                           npass0, nullptr, false);
      bidx_final = bi[bc];
      launch_major_keys(stream_, wide, dj, dch, NC, (const uint64_t*)d_bbase.p, bi[bc], X, keys[0],
                        npass == 0 ? idx[0] : nullptr, npass == 0 ? scoord : nullptr, dcf,
                        npass > 0 ? hist : nullptr);
      cur = radix(dj, dch, dcf, NC, keys, idx, npass, scoord, npass > 0);
    } else {
      launch_cell_keys(stream_, wide, dj, dch, NC, X, keys[0], idx[0]);
      cur = radix(dj, dch, dcf, NC, keys, idx, npass, with_tiles ? scoord : nullptr, false);
      if (with_tiles && npass == 0) launch_gather_sorted(stream_, dj, dch, NC, X, idx[cur], scoord);
    }
    if (!with_tiles) return;
    const uint2* tch = (const uint2*)d_tchunks.p;
    // the tile-start flags go to the radix histogram buffer (kRadixBins
    // words per chunk, free once the sort is done; the flags take kWave)
    launch_cell_starts(stream_, wide, dj, tch, NC, keys[cur], tcnt, (uint32_t*)d_cellstart.p, hist);
    if (has_mom)
      launch_box_moments(stream_, dj, (const uint32_t*)d_mlist.p, mom_waves, scoord, (const uint32_t*)d_cellstart.p,
                         (float4*)d_moments.p);
    if (has_mom2)
      launch_cell_moments2(stream_, dj, (const uint32_t*)d_mlist2.p, mom2_blocks, scoord,
                           (const uint32_t*)d_cellstart.p, (float*)d_moments2.p);
    launch_exclusive_scan(stream_, tcnt, tcnt, NC, scratch);
    launch_tile_write(stream_, dj, tch, NC, hist, tcnt, (DTile*)d_tiles.p);
    launch_tile_ends(stream_, dj, (DTile*)d_tiles.p, tcnt + NC, P);
    if (pack) launch_tile_pack(stream_, dj, (DTile*)d_tiles.p, tcnt + NC, P);
  };
  if (stop == kStopAfterSort) {
    issue_pre(false);
    ENT_HIP(hipGetLastError());
    return kOk;
  }
  PlanKey key = plan_key(X, dens, J, NC, P, npass, false);
  key.NCB = shared ? NCB : 0;
  key.PB = shared ? PB : 0;
  // the key width selects the kernel variants of every captured launch (and
  // the buffers' element size), so it is part of the shape: a batch whose
  // keys change width at unchanged pass counts must not replay the other
  // width's graph
  key.variant = (shared ? 1 : 0) | (staged ? 2 : 0) | (pack ? 4 : 0) | (npass0 << 3) | (has_mom ? 256 : 0) |
                (has_mom2 ? 512 : 0) | (wide ? 1024 : 0);
  ENT_TRY(run_captured(pre_graphs_, key, [&] { issue_pre(true); }));
  return launch_density_common(false, X, dens, P);
}

PlanKey Engine::plan_key(const float* X, const double* dens, uint32_t J, uint32_t NC, uint64_t P, int npass,
                         bool raw) const {
  PlanKey k;
  k.X = X;
  k.dens = dens;
  k.J = J;
  k.NC = NC;
  k.P = P;
  k.npass = npass;
  k.raw = raw ? 1 : 0;
  k.q = tuning.q;
  k.m0 = tuning.m0;
  k.mr = tuning.mr;
  k.gh = tuning.group_h;
  k.sub = tuning.sub;
  k.gen = generation_;
  return k;
}

template <class F>
int Engine::run_captured(GraphCache& cache, const PlanKey& key, F&& issue) {
  if (!use_graph) {
    issue();
    ENT_HIP(hipGetLastError());
    return kOk;
  }
  for (size_t i = 0; i < cache.entries.size(); ++i) {
    if (cache.entries[i].first == key) {
      auto hit = cache.entries[i];
      cache.entries.erase(cache.entries.begin() + i);
      cache.entries.push_back(hit);
      ENT_HIP(hipGraphLaunch(hit.second, stream_));
      ++last.graph_replays;
      return kOk;
    }
  }
  if (cache.entries.size() >= GraphCache::kMax) {
    // the stream may still run the oldest graph; wait before destroying it
    ENT_HIP(hipStreamSynchronize(stream_));
    (void)hipGraphExecDestroy(cache.entries.front().second);
    cache.entries.erase(cache.entries.begin());
  }
  hipGraphExec_t exec = nullptr;
  hipGraph_t g = nullptr;
  ENT_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  issue();
  hipError_t le = hipGetLastError();
  hipError_t ce = hipStreamEndCapture(stream_, &g);
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    set_error(std::string("graph capture failed: ") + hipGetErrorString(le != hipSuccess ? le : ce));
    return kErrHip;
  }
  hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) {
    set_error(std::string("graph instantiation failed: ") + hipGetErrorString(ie));
    return kErrHip;
  }
  cache.entries.emplace_back(key, exec);
  ENT_HIP(hipGraphLaunch(exec, stream_));
  return kOk;
}

int Engine::run_brute(const float* X, const std::vector<JobSpec>& jobs, double* dens) {
  const uint32_t J = (uint32_t)jobs.size();
  const uint32_t tq = (uint32_t)(kWave * tuning.q);
  std::vector<DTile> tiles;
  for (uint32_t j = 0; j < J; ++j)
    for (uint32_t p = 0; p < jobs[j].n; p += tq) {
      DTile t;
      t.job = j;
      t.p0 = p;
      t.p1 = std::min(jobs[j].n, p + tq);
      t.pack = 0;
      tiles.push_back(t);
    }
  const uint64_t T = tiles.size();
  ENT_TRY(ensure(d_jobs, sizeof(DJob) * J));
  ENT_TRY(ensure(d_tiles, sizeof(DTile) * T));
  ENT_TRY(ensure(d_tile_cnt, sizeof(uint32_t) * 2));
  ENT_TRY(ensure(d_counter, sizeof(uint32_t) * kCounterWords));
  size_t stage_bytes = sizeof(DJob) * J + sizeof(DTile) * T + 16;
  ENT_TRY(ensure_host(stage_bytes));
  char* hs = (char*)host_stage_;
  std::memcpy(hs, h_jobs.data(), sizeof(DJob) * J);
  std::memcpy(hs + sizeof(DJob) * J, tiles.data(), sizeof(DTile) * T);
  uint32_t* hs_T = (uint32_t*)(hs + sizeof(DJob) * J + sizeof(DTile) * T);
  hs_T[0] = (uint32_t)T;
  ENT_HIP(hipEventRecord(ev_[0], stream_));
  ENT_HIP(hipMemcpyAsync(d_jobs.p, hs, sizeof(DJob) * J, hipMemcpyHostToDevice, stream_));
  ENT_HIP(hipMemcpyAsync(d_tiles.p, hs + sizeof(DJob) * J, sizeof(DTile) * T, hipMemcpyHostToDevice, stream_));
  ENT_HIP(hipMemcpyAsync(d_tile_cnt.p, hs_T, sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
  // the density kernel reads the tile count from n_tiles
  return launch_density_common(true, X, dens, T);
}

int Engine::launch_density_common(bool raw, const float* X, double* dens, uint64_t max_tiles) {
  (void)max_tiles;
  int bpc = 0;
  const bool wide = !raw && keys_wide;
  const bool s3 = !raw && tuning.q == 1 && sub3;
  if (density_occupancy(tuning.q, raw, wide, s3, &bpc) != 0 || bpc <= 0) {
    set_error("unsupported queries-per-lane setting");
    return kErrUnsupported;
  }
  unsigned grid = (unsigned)std::max(1, num_cus_ * bpc);
  const uint32_t NC = (uint32_t)h_chunks.size();
  const uint32_t* n_tiles = raw ? (const uint32_t*)d_tile_cnt.p : (const uint32_t*)d_tile_cnt.p + NC;
  uint32_t* counter = (uint32_t*)d_counter.p;
  int rc = 0;
  // the claim counters and pair counters are cleared by a fill kernel (not a
  // memset node) inside the same graph as the persistent launch
  auto issue = [&] {
    launch_fill_u32(stream_, counter, 0u, kCounterWords);
    rc = launch_density(stream_, tuning.q, raw, wide, s3, grid, (const DJob*)d_jobs.p, (const DTile*)d_tiles.p,
                        n_tiles, counter, X, raw ? nullptr : (const float*)d_scoord.p,
                        raw ? nullptr : d_keys[sorted_buf].p, raw ? nullptr : (const uint32_t*)d_cellstart.p,
                        raw ? nullptr : (const uint32_t*)d_idx[sorted_buf].p,
                        raw ? nullptr : (const float4*)d_moments.p, raw ? nullptr : (const float*)d_moments2.p,
                        dens, (unsigned long long*)(counter + 2));
  };
  ENT_HIP(hipEventRecord(ev_[1], stream_));
  if (raw) {
    issue();
    ENT_HIP(hipGetLastError());
  } else {
    PlanKey key = plan_key(X, dens, (uint32_t)h_jobs.size(), NC, total_points, sorted_buf, false);
    key.grid = (int)grid;
    key.raw = wide ? 2 : 0;  // key width selects the kernel variant
    key.variant = s3 ? 1 : 0;
    ENT_TRY(run_captured(dens_graphs_, key, issue));
  }
  if (rc != 0) {
    set_error("density launch: unsupported configuration");
    return kErrUnsupported;
  }
  ENT_HIP(hipEventRecord(ev_[2], stream_));
  // stats readback
  ENT_TRY(ensure_host(96));
  uint32_t* hc = (uint32_t*)host_stage_;
  ENT_HIP(hipMemcpyAsync(hc, counter, 80, hipMemcpyDeviceToHost, stream_));
  ENT_HIP(hipMemcpyAsync(hc + 20, n_tiles, 4, hipMemcpyDeviceToHost, stream_));
  ENT_HIP(hipStreamSynchronize(stream_));
  uint64_t pd[9];
  std::memcpy(pd, hc + 2, 72);
  last.lane_pairs = pd[3];
  last.pairs = pd[0] + pd[1] + pd[2];
  for (int d = 0; d < 3; ++d) last.pairs_by_d[d] = pd[d];
  for (int d = 0; d < 2; ++d) {
    last.moment_pairs[d] = pd[4 + d];
    last.moment_terms[d] = pd[6 + d];
  }
  last.raw_tiles = pd[8];
  last.tiles = hc[20];
  float ms = 0.f;
  ENT_HIP(hipEventElapsedTime(&ms, ev_[1], ev_[2]));
  last.density_ms = ms;
  ENT_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[2]));
  last.total_ms = ms;
  return kOk;
}

}  // namespace ent
