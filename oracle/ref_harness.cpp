// ref_harness.cpp -- TEST INFRASTRUCTURE.  Thin extern "C" wrappers around
// reference functions that compile from their own sources with g++ alone:
//   prob_dens_coord_prep  (densities_common.hxx:7-50, header-only template)
//   Transs::BoxedSearch::Boxes  (boxedsearch.cpp:15-62)
// Built by oracle/Makefile into oracle/_ref/ from the reference tree; used by
// tests only, to pin the oracle's restatements.  Never shipped, never timed.
#include <cstddef>
#include <cstdint>
#include <vector>
#include "densities_common.hpp"
#include "boxedsearch.hpp"

namespace {
// Boxes keeps the per-state box id protected; a derived probe reads it.
struct BoxesProbe : public Transs::BoxedSearch::Boxes {
  BoxesProbe(const float* c, std::size_t n, std::size_t col, float bs) : Boxes(c, n, col, bs) {}
  const std::vector<std::size_t>& box_of_state() const { return _box_of_state; }
};
}  // namespace

extern "C" {

// Calls prob_dens_coord_prep<float>; out receives (n_rows - tau_max) * d floats.
long ref_coord_prep(const float* coords, std::size_t n_rows, const unsigned* i_cols,
                    const float* h, const unsigned* tau, unsigned d, int row_major, float* out) {
  std::vector<unsigned int> ic(i_cols, i_cols + d), tv(tau, tau + d);
  std::vector<float> hv(h, h + d);
  std::vector<float> r = prob_dens_coord_prep<float>(coords, n_rows, ic, hv, tv, row_major != 0);
  for (std::size_t i = 0; i < r.size(); ++i) out[i] = r[i];
  return (long)(r.size() / d);
}

// Bins column `col` of a column-major array with the reference Boxes class.
// Callers must avoid the inputs on which that class is undefined (data whose
// max equals min + k*box_size exactly, all-equal data).
long ref_boxes(const float* coords, std::size_t n_rows, std::size_t col, float box_size,
               int64_t* box_of_state) {
  BoxesProbe b(coords, n_rows, col, box_size);
  const std::vector<std::size_t>& bs = b.box_of_state();
  for (std::size_t i = 0; i < n_rows; ++i) box_of_state[i] = (int64_t)bs[i];
  return (long)b.n_boxes();
}

// Neighbour list of state i (sorted union of boxes b-1..b+1).
long ref_neighbors(const float* coords, std::size_t n_rows, std::size_t col, float box_size,
                   std::size_t state, int64_t* out, std::size_t cap) {
  BoxesProbe b(coords, n_rows, col, box_size);
  std::vector<std::size_t> nb = b.neighbors_of_state_ordered(state);
  for (std::size_t i = 0; i < nb.size() && i < cap; ++i) out[i] = (int64_t)nb[i];
  return (long)nb.size();
}

}  // extern "C"
