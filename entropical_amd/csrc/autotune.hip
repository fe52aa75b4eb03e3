// autotune.hip -- the engine's tile-shape autotuner.
//
// Search space: maximum queries per lane q (tile = 64*q queries), dim-0 boxes
// per bandwidth m0, row boxes per bandwidth mr, tile extent along dim 0 in
// bandwidths group_h.  Coordinate descent from the current setting, two
// sweeps, each candidate timed on a subsample of the batch (second of two
// runs, so graph capture and workspace growth are excluded).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include "engine.hpp"

namespace ent {

namespace {

// Workload signature: per dimension, the number of jobs and their total
// points rounded to two significant digits.
std::string signature(const std::vector<JobSpec>& jobs) {
  uint64_t cnt[3] = {0, 0, 0}, pts[3] = {0, 0, 0};
  for (const JobSpec& j : jobs) {
    cnt[j.d - 1]++;
    pts[j.d - 1] += j.n;
  }
  std::ostringstream os;
  for (int d = 0; d < 3; ++d) {
    double p = (double)pts[d];
    double r = p > 0 ? std::round(p / std::pow(10.0, std::floor(std::log10(p)) - 1)) *
                           std::pow(10.0, std::floor(std::log10(p)) - 1)
                     : 0.0;
    os << "d" << d + 1 << ":" << cnt[d] << ":" << (uint64_t)r << ";";
  }
  return os.str();
}

const char* cache_path() { return std::getenv("ENT_TUNE_CACHE"); }

}  // namespace

int Engine::autotune(const float* X, const std::vector<JobSpec>& jobs, double* dens, Tuning* best_out,
                     double* best_ms_out) {
  if (jobs.empty()) {
    set_error("autotune: empty job list");
    return kErrArg;
  }
  // subsample: every k-th job, about an eighth of the batch but at least 8
  // jobs (all jobs when there are few)
  std::vector<JobSpec> sample;
  const size_t want = std::max<size_t>(8, jobs.size() / 8);
  const size_t stride = std::max<size_t>(1, jobs.size() / want);
  for (size_t i = 0; i < jobs.size(); i += stride) sample.push_back(jobs[i]);
  const std::string sig = signature(sample);

  if (!tune_cache_loaded_) {
    tune_cache_loaded_ = true;
    if (const char* p = cache_path()) {
      std::ifstream f(p);
      std::string s;
      Tuning t;
      while (f >> s >> t.q >> t.m0 >> t.mr >> t.group_h) tune_cache_[s] = t;
    }
  }
  auto hit = tune_cache_.find(sig);
  if (hit != tune_cache_.end()) {
    tuning = hit->second;
    if (best_out) *best_out = tuning;
    if (best_ms_out) *best_ms_out = -1.0;
    return kOk;
  }

  auto eval = [&](const Tuning& t, double* ms) -> int {
    tuning = t;
    double best = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      int rc = densities(X, sample, dens, kMethodCell);
      if (rc != kOk) return rc;
      if (rep == 1) best = std::min(best, last.total_ms);
    }
    *ms = best;
    return kOk;
  };
  const std::vector<int> qs = {1, 2, 4}, m0s = {4, 8, 16}, mrs = {1, 2}, ghs = {2, 4, 8};
  Tuning cur = tuning;
  double cur_ms = 0.0;
  int rc = eval(cur, &cur_ms);
  if (rc != kOk) return rc;
  for (int sweep = 0; sweep < 2; ++sweep) {
    bool improved = false;
    for (int axis = 0; axis < 4; ++axis) {
      const std::vector<int>& vals = axis == 0 ? qs : axis == 1 ? m0s : axis == 2 ? mrs : ghs;
      for (int v : vals) {
        Tuning t = cur;
        int* field = axis == 0 ? &t.q : axis == 1 ? &t.m0 : axis == 2 ? &t.mr : &t.group_h;
        if (*field == v) continue;
        *field = v;
        double ms = 0.0;
        rc = eval(t, &ms);
        if (rc != kOk) return rc;
        if (ms < cur_ms * 0.99) {
          cur = t;
          cur_ms = ms;
          improved = true;
        }
      }
    }
    if (!improved) break;
  }
  tuning = cur;
  tune_cache_[sig] = cur;
  if (const char* p = cache_path()) {
    std::ofstream f(p, std::ios::app);
    f << sig << " " << cur.q << " " << cur.m0 << " " << cur.mr << " " << cur.group_h << "\n";
  }
  if (best_out) *best_out = cur;
  if (best_ms_out) *best_ms_out = cur_ms;
  return kOk;
}

}  // namespace ent
