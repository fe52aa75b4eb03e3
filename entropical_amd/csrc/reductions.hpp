// reductions.hpp -- per-result FP64 reductions over device density arrays.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ent {

enum RedOp : uint32_t {
  kRedMI = 0,       // sum p_ij log(p_ij / p_i / p_j)           (mi.cpp:53-56)
  kRedLogMean = 1,  // sum ln p                                 (H = -mean ln p)
  kRedTE = 2,       // sum p3 pref log(p3 p1 inv(p2) inv(ptau))  (transs.cpp:77-100)
  kRedAutoEnt = 3,  // sum p2 h h log(p1 inv(p2 h)), pref = h   (autoent.cpp:52-56)
  kRedLoo = 4,      // sum ln max(p_{-i}, floor), p_{-i} = (n p - K0) / (n - 1):
                    // leave-one-out log-likelihood (p includes the self pair
                    // K0 = prod 0.75/h); pref = K0, scale = floor density
  kRedSum = 5,      // sum p (epa self-convolution rows, densities_omp.cpp:271-276)
  kRedKL = 6,       // sum q_a log(q_a / q_b), q_a = p_a * pref, q_b = p_b * pref2
                    // (sum1-normalised densities, kldiv.cpp:49-51)
  kRedNegs = 7,     // sum log(q) q, q = p * pref (negs.cpp:43-45)
};

// off[0..3] index density arrays in one device buffer D (doubles).
struct RedItem {
  uint64_t off[4];
  uint64_t n;
  uint32_t op;
  uint32_t pad;
  double pref;   // TE prefactor / autoent bandwidth / normalisation of off[0]
  double scale;  // 1 for ln, 1/ln 2 for bits
  double pref2;  // normalisation of off[1] (KL)
};

void launch_reduce(hipStream_t s, const RedItem* items, uint32_t n_items, const double* D,
                   double* partial, uint32_t blocks_per_item, double* out);

// Per column of a column-major float32 array X[col*n_rows + row]: mean and
// population variance in FP64 (two passes: sum, then centred squares), with
// a fixed partition so results are deterministic.  partial holds
// n_cols*blocks_per_col doubles; mean and var n_cols each.
void launch_col_moments(hipStream_t s, const float* X, uint64_t n_rows, uint32_t n_cols, double* partial,
                        uint32_t blocks_per_col, double* mean, double* var);

}  // namespace ent
