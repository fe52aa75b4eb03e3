// densities_shim.cpp -- the reference's C++ density call surface
// (include/entropical/densities.hpp) over the C ABI.
#include "entropical/densities.hpp"

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <mutex>
#include <string>
#include <vector>
#include "entropical_kde.h"

namespace {

// Process-wide context on ENT_DEVICE (default 0), created on first use.
ent_ctx* shim_ctx() {
  static ent_ctx* ctx = nullptr;
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0;
    if (const char* s = std::getenv("ENT_DEVICE")) dev = std::atoi(s);
    if (ent_ctx_create(dev, &ctx) != ENT_OK) {
      std::cerr << ent_last_error() << std::endl;
      std::exit(EXIT_FAILURE);
    }
  });
  return ctx;
}

int shim_method() {
  const char* s = std::getenv("ENT_METHOD");
  return (s && std::strcmp(s, "brute") == 0) ? ENT_METHOD_BRUTE : ENT_METHOD_CELL;
}

[[noreturn]] void die() {
  std::cerr << ent_last_error() << std::endl;
  std::exit(EXIT_FAILURE);
}

// The reference checks these before touching the data
// (densities_common.hxx:15-33) and exits with these texts.
void check_sizes(std::size_t d, std::size_t nh, std::size_t nt) {
  if (d < 1 || 3 < d) {
    std::cerr << "error: can only compute combined probabilities in 1, 2 or 3 dimensions." << std::endl;
    std::exit(EXIT_FAILURE);
  }
  if (nh != d) {
    std::cerr << "error: number of bandwidth parameters does not match number of selected columns."
              << std::endl;
    std::exit(EXIT_FAILURE);
  }
  if (nt != d) {
    std::cerr << "error: number of lagtimes (tau) does not match number of selected dimensions." << std::endl;
    std::exit(EXIT_FAILURE);
  }
}

}  // namespace

std::vector<double> combined_densities(const float* coords, std::size_t n_rows,
                                       std::vector<unsigned int> i_cols, std::vector<float> h) {
  if (i_cols.size() < 1 || i_cols.size() > 3) {
    std::cerr << "error: unsupported number of dimensions. this should never happen!" << std::endl;
    std::exit(EXIT_FAILURE);
  }
  return combined_densities(coords, n_rows, i_cols, h, std::vector<unsigned int>(i_cols.size(), 0u));
}

std::vector<double> combined_densities(const float* coords, std::size_t n_rows,
                                       std::vector<unsigned int> i_cols, std::vector<float> h,
                                       std::vector<unsigned int> tau) {
  check_sizes(i_cols.size(), h.size(), tau.size());
  unsigned int tau_max = 0, col_max = 0;
  for (unsigned int t : tau) tau_max = t > tau_max ? t : tau_max;
  for (unsigned int c : i_cols) col_max = c > col_max ? c : col_max;
  std::vector<double> out(n_rows > tau_max ? n_rows - tau_max : 0);
  // the reference signature carries no column count: the highest selected
  // column bounds the array, as the reference's indexing assumes
  int rc = ent_combined_densities(shim_ctx(), coords, n_rows, (std::size_t)col_max + 1, i_cols.data(), h.data(),
                                  tau.data(), (uint32_t)i_cols.size(), shim_method(), out.data());
  if (rc != ENT_OK) die();
  return out;
}

float epa_convolution(const std::vector<float>& sorted_coords, float h) {
  float out = 0.0f;
  if (ent_epa_convolution(shim_ctx(), sorted_coords.data(), sorted_coords.size(), h, &out) != ENT_OK) die();
  return out;
}

template <typename NUM>
std::vector<NUM> prob_dens_coord_prep(const NUM* coords, std::size_t n_rows, std::vector<unsigned int> i_cols,
                                      std::vector<NUM> h, std::vector<unsigned int> tau, bool row_major_result) {
  check_sizes(i_cols.size(), h.size(), tau.size());
  unsigned int tau_max = 0;
  for (unsigned int t : tau) tau_max = t > tau_max ? t : tau_max;
  std::vector<NUM> out(n_rows > tau_max ? (n_rows - tau_max) * i_cols.size() : 0);
  std::size_t n_sel = 0;
  if (ent_coord_prep(coords, n_rows, i_cols.data(), h.data(), tau.data(), (uint32_t)i_cols.size(),
                     row_major_result ? 1 : 0, out.data(), &n_sel) != ENT_OK)
    die();
  return out;
}

template std::vector<float> prob_dens_coord_prep<float>(const float*, std::size_t, std::vector<unsigned int>,
                                                        std::vector<float>, std::vector<unsigned int>, bool);
