// densities.hpp -- C++ drop-in for entropical's density backend header.
//
// Same declarations and semantics as the reference's densities_cuda.hpp:10-31
// and densities_common.hpp:11-18, implemented over the C ABI in
// entropical_kde.h.  A caller of the reference (mi.cpp, transs.cpp, ...)
// includes this header instead of densities_cuda.hpp and links
// libentropical_kde.so.
//
// Backend selection at run time (the reference chooses at build time with
// USE_CUDA, CMakeLists.txt:56-71):
//   ENT_DEVICE=<n>          HIP device (default 0)
//   ENT_METHOD=cell|brute   candidate search (default cell)
// Errors print the message to std::cerr and exit(EXIT_FAILURE), as the
// reference does (densities_common.hxx:16-33, tools.cuh:33-42).
#pragma once

#include <cstddef>
#include <vector>

// densities_cuda.hpp:10-16 (tau defaults to zero)
std::vector<double>
combined_densities(const float* coords
                 , std::size_t n_rows
                 , std::vector<unsigned int> i_cols
                 , std::vector<float> h);

// densities_cuda.hpp:18-24
std::vector<double>
combined_densities(const float* coords
                 , std::size_t n_rows
                 , std::vector<unsigned int> i_cols
                 , std::vector<float> h
                 , std::vector<unsigned int> tau);

// densities_cuda.hpp:29-31 (OpenMP semantics, densities_omp.cpp:235-277)
float
epa_convolution(const std::vector<float>& sorted_coords
              , float h);

// densities_common.hpp:11-18
template <typename NUM>
std::vector<NUM>
prob_dens_coord_prep(const NUM* coords
                   , std::size_t n_rows
                   , std::vector<unsigned int> i_cols
                   , std::vector<NUM> h
                   , std::vector<unsigned int> tau
                   , bool row_major_result);

extern template std::vector<float>
prob_dens_coord_prep<float>(const float*, std::size_t, std::vector<unsigned int>, std::vector<float>,
                            std::vector<unsigned int>, bool);
