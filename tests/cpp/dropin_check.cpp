// dropin_check.cpp -- exercises the C++ drop-in (include/entropical/densities.hpp)
// the way the reference drivers call it.  Reads a column-major float32 file,
// writes densities / MI / epa_convolution results as raw doubles for the
// pytest that compares them with the oracle.
//   dropin_check <coords.f32> <n_rows> <n_cols> <h> <outdir>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>
#include "entropical/densities.hpp"

static void dump(const std::string& path, const std::vector<double>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(double));
}

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s coords.f32 n_rows n_cols h outdir\n", argv[0]);
    return 2;
  }
  const std::size_t n_rows = std::strtoul(argv[2], nullptr, 10);
  const unsigned n_cols = (unsigned)std::strtoul(argv[3], nullptr, 10);
  const float h = std::strtof(argv[4], nullptr);
  const std::string out = argv[5];
  std::vector<float> coords(n_rows * n_cols);
  std::ifstream in(argv[1], std::ios::binary);
  in.read(reinterpret_cast<char*>(coords.data()), coords.size() * sizeof(float));
  if (!in) return 3;

  dump(out + "/p0.f64", combined_densities(coords.data(), n_rows, {0}, {h}));
  dump(out + "/p01.f64", combined_densities(coords.data(), n_rows, {0, 1}, {h, h}));
  dump(out + "/p012_lag.f64", combined_densities(coords.data(), n_rows, {0, 1, 2}, {h, h, h}, {0, 2, 1}));

  // the loop of Mi::main (mi.cpp:36-57) over the drop-in, FP32 accumulation
  std::vector<float> mutinf(n_cols * n_cols);
  for (unsigned int i = 0; i < n_cols; ++i) {
    std::vector<double> p_i = combined_densities(coords.data(), n_rows, {i}, {h});
    for (unsigned int j = i + 1; j < n_cols; ++j) {
      std::vector<double> p_j = combined_densities(coords.data(), n_rows, {j}, {h});
      std::vector<double> p_ij = combined_densities(coords.data(), n_rows, {i, j}, {h, h});
      mutinf[i * n_cols + j] = 0.0f;
      for (unsigned int k = 0; k < n_rows; ++k) {
        float r = (float)(p_ij[k] / p_i[k] / p_j[k]);
        mutinf[i * n_cols + j] += p_ij[k] * (r == 0 ? 0.0f : std::log(r));
      }
      mutinf[j * n_cols + i] = mutinf[i * n_cols + j];
    }
  }
  dump(out + "/mi.f64", std::vector<double>(mutinf.begin(), mutinf.end()));

  std::vector<float> sorted(coords.begin(), coords.begin() + n_rows);
  std::sort(sorted.begin(), sorted.end());
  dump(out + "/epa.f64", {(double)epa_convolution(sorted, h)});
  std::vector<float> prep = prob_dens_coord_prep<float>(coords.data(), n_rows, {1, 0}, {h, h}, {3, 0}, true);
  dump(out + "/prep.f64", std::vector<double>(prep.begin(), prep.end()));
  std::printf("ok\n");
  return 0;
}
