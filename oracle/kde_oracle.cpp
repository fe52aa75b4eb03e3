// kde_oracle.cpp -- CPU oracle for the Epanechnikov KDE hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (entropical_amd/, the
// C-ABI library, the C++ drop-in) links, loads or calls this file.  It is
// loaded by tests/ (as the checker), by __graft_entry__.smoke() (as the
// checker) and by bench.py's cpu_baseline leg (the timed host OpenMP path).
//
// Contents, each restating a reference function (file:line relative to the
// lettis/entropical source tree):
//   orc_coord_prep            densities_common.hxx:7-50   (lag gather)
//   orc_density_fp64          densities_cuda.cu:86-104 + density_1d.cuh:39-46
//                             semantics, evaluated exactly in FP64 (brute force,
//                             Neumaier sums)
//   orc_density_fp64_window   same values, O(N*W) via a sorted dim-0 window
//   orc_density_cuda_replica  FP32 replica of the plain CUDA path, 1-D/2-D
//                             (density_1d.cuh:32-90, density_2d.cuh:32-100,
//                             densities_cuda.cu:86-104)
//   orc_density_omp_packed    restatement of the OpenMP backend
//                             (densities_omp.cpp:19-159, tools.hxx:113-150);
//                             this is the host OpenMP baseline
//   orc_bin / orc_cell_sort   Transs::BoxedSearch::Boxes binning
//                             (boxedsearch.cpp:15-34) with the min/max/edge
//                             fixes, plus a stable counting sort by cell
//   orc_epa_conv_*            densities_omp.cpp:235-277
//   orc_thumb / orc_amise     hestimate.cpp:19-104
//   orc_mi_ref / orc_te_ref / orc_autoent_ref / orc_kldiv_ref / orc_negs_ref
//                             mi.cpp:53-56, transs.cpp:75-113,
//                             autoent.cpp:52-57, kldiv.cpp:37-53, negs.cpp:37-47
//   orc_count_in_support      pairs with a nonzero kernel term (measurement)
//
// Pinning (what checks each function against something other than itself)
// is listed in tests/test_oracle.py.  Functions whose bit-level behaviour
// depends on Boost (kahan_sum, variance(lazy)) are restated from the textbook
// algorithm; see the notes at orc_thumb and orc_epa_conv_omp.

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <vector>
#include <array>
#include <omp.h>

namespace {

struct Neumaier {
  double s = 0.0, c = 0.0;
  inline void add(double x) {
    double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  inline double value() const { return s + c; }
};

// Textbook Kahan summation in FP32 (Boost sum_kahan restated).
inline float kahan_f32(const float* xs, std::size_t n) {
  float sum = 0.0f, c = 0.0f;
  for (std::size_t i = 0; i < n; ++i) {
    float y = xs[i] - c;
    float t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

inline double kahan_f64(const double* xs, std::size_t n) {
  double sum = 0.0, c = 0.0;
  for (std::size_t i = 0; i < n; ++i) {
    double y = xs[i] - c;
    double t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

// Validation shared by the density entry points (densities_common.hxx:15-33).
// Returns n_sel (> 0) or a negative error code.
long validate(std::size_t n_rows, const unsigned* tau, unsigned d) {
  if (d < 1 || d > 3) return -1;
  unsigned tau_max = 0;
  for (unsigned e = 0; e < d; ++e) tau_max = std::max(tau_max, tau[e]);
  if (tau_max >= n_rows) return -2;
  return (long)(n_rows - tau_max);
}

// point k, dim e of a density job: coords[i_cols[e]*n_rows + k + tau[e]]
inline float pt(const float* coords, std::size_t n_rows, const unsigned* i_cols,
                const unsigned* tau, unsigned e, std::size_t k) {
  return coords[(std::size_t)i_cols[e] * n_rows + k + tau[e]];
}

inline double term_fp64(const double* xi, const double* xj, const double* h, unsigned d) {
  double t = 1.0;
  for (unsigned e = 0; e < d; ++e) {
    double u = (xi[e] - xj[e]) / h[e];
    double u2 = u * u;
    if (!(u2 <= 1.0)) return 0.0;
    t *= 0.75 * (1.0 - u2) / h[e];
  }
  return t;
}

}  // namespace

extern "C" {

int orc_version() { return 1; }

void orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

int orc_max_threads() { return omp_get_max_threads(); }

// densities_common.hxx:7-50.  out has (n_rows - tau_max) * d floats.
long orc_coord_prep(const float* coords, std::size_t n_rows, const unsigned* i_cols,
                    const unsigned* tau, unsigned d, int row_major, float* out) {
  long n_sel = validate(n_rows, tau, d);
  if (n_sel < 0) return n_sel;
  for (long i = 0; i < n_sel; ++i)
    for (unsigned j = 0; j < d; ++j) {
      float v = pt(coords, n_rows, i_cols, tau, j, (std::size_t)i);
      if (row_major) out[(std::size_t)i * d + j] = v;
      else out[(std::size_t)j * n_sel + i] = v;
    }
  return n_sel;
}

// Exact (FP64, Neumaier) density at every point of the job, divided by the
// ORIGINAL n_rows (densities_cuda.cu:101-103).  Brute force O(n^2).
long orc_density_fp64(const float* coords, std::size_t n_rows, const unsigned* i_cols,
                      const float* h, const unsigned* tau, unsigned d, double* out) {
  long n = validate(n_rows, tau, d);
  if (n < 0) return n;
  double hd[3];
  for (unsigned e = 0; e < d; ++e) hd[e] = (double)h[e];
  std::vector<double> pts((std::size_t)n * d);
  for (long k = 0; k < n; ++k)
    for (unsigned e = 0; e < d; ++e)
      pts[(std::size_t)k * d + e] = (double)pt(coords, n_rows, i_cols, tau, e, (std::size_t)k);
#pragma omp parallel for schedule(dynamic, 16)
  for (long i = 0; i < n; ++i) {
    Neumaier acc;
    const double* xi = &pts[(std::size_t)i * d];
    for (long j = 0; j < n; ++j) acc.add(term_fp64(xi, &pts[(std::size_t)j * d], hd, d));
    out[i] = acc.value() / (double)n_rows;
  }
  return n;
}

// Same values as orc_density_fp64 (identical per-term predicate and value),
// but only samples inside a dim-0 window are visited: terms outside it are
// exactly zero.  O(n * window).
long orc_density_fp64_window(const float* coords, std::size_t n_rows, const unsigned* i_cols,
                             const float* h, const unsigned* tau, unsigned d, double* out) {
  long n = validate(n_rows, tau, d);
  if (n < 0) return n;
  double hd[3];
  for (unsigned e = 0; e < d; ++e) hd[e] = (double)h[e];
  std::vector<std::size_t> order((std::size_t)n);
  for (long k = 0; k < n; ++k) order[k] = (std::size_t)k;
  std::vector<double> x0((std::size_t)n);
  for (long k = 0; k < n; ++k) x0[k] = (double)pt(coords, n_rows, i_cols, tau, 0, (std::size_t)k);
  std::sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
    return x0[a] < x0[b] || (x0[a] == x0[b] && a < b);
  });
  std::vector<double> spts((std::size_t)n * d);
  std::vector<double> sx0((std::size_t)n);
  for (long p = 0; p < n; ++p) {
    for (unsigned e = 0; e < d; ++e)
      spts[(std::size_t)p * d + e] = (double)pt(coords, n_rows, i_cols, tau, e, order[p]);
    sx0[p] = spts[(std::size_t)p * d];
  }
  const double widen = hd[0] * (1.0 + 1e-9);
#pragma omp parallel for schedule(dynamic, 64)
  for (long p = 0; p < n; ++p) {
    const double* xi = &spts[(std::size_t)p * d];
    std::size_t lo = std::lower_bound(sx0.begin(), sx0.end(), xi[0] - widen) - sx0.begin();
    std::size_t hi = std::upper_bound(sx0.begin(), sx0.end(), xi[0] + widen) - sx0.begin();
    Neumaier acc;
    for (std::size_t j = lo; j < hi; ++j) acc.add(term_fp64(xi, &spts[j * d], hd, d));
    out[order[p]] = acc.value() / (double)n_rows;
  }
  return n;
}

// FP32 replica of the reference's plain CUDA path for d = 1, 2
// (density_1d.cuh:39-46,61-89,143-153; density_2d.cuh:39-52,68-99;
// densities_cuda.cu:87-103).  Compile with -ffp-contract=off.
long orc_density_cuda_replica(const float* coords, std::size_t n_rows, const unsigned* i_cols,
                              const float* h, const unsigned* tau, unsigned d, double* out) {
  long n = validate(n_rows, tau, d);
  if (n < 0) return n;
  if (d > 2) return -3;  // the reference 3-D kernel is defective; not replicated
  float hinv[2];
  for (unsigned e = 0; e < d; ++e) hinv[e] = 1.0f / h[e];
  std::vector<float> s((std::size_t)n * d);
  for (long k = 0; k < n; ++k)
    for (unsigned e = 0; e < d; ++e)
      s[(std::size_t)k * d + e] = hinv[e] * pt(coords, n_rows, i_cols, tau, e, (std::size_t)k);
  const long B = 128;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    float res = 0.0f;
    for (long t0 = 0; t0 < n; t0 += B) {
      long t1 = std::min(n, t0 + B);
      float acc = 0.0f;
      for (long j = t0; j < t1; ++j) {
        float term;
        if (d == 1) {
          float p = s[i] - s[j];
          p = p * p;
          term = (p <= 1.0f) ? (float)((double)hinv[0] * std::fma((double)p, -0.75, 0.75)) : 0.0f;
        } else {
          float p1 = s[2 * i] - s[2 * j];
          float p2 = s[2 * i + 1] - s[2 * j + 1];
          p1 = p1 * p1;
          p2 = p2 * p2;
          if (p1 <= 1.0f && p2 <= 1.0f) {
            p1 = (float)((double)hinv[0] * std::fma((double)p1, -0.75, 0.75));
            p2 = (float)((double)hinv[1] * std::fma((double)p2, -0.75, 0.75));
            term = p1 * p2;
          } else {
            term = 0.0f;
          }
        }
        acc = acc + term;
      }
      res = res + acc;
    }
    out[i] = (double)res / (double)n_rows;
  }
  return n;
}

// Restatement of the OpenMP backend (densities_omp.cpp:19-159 with
// Tools::dim1_sorted_coords, tools.hxx:113-150), fed PACKED columns
// (column-major [d][n], i.e. i_cols = {0..d-1}, tau = 0; the reference code
// is only correct for that input).  Output: float P[i] = sum / n, in input
// order.  This is the host OpenMP baseline: same sort, same O(n^2) linear
// window scans (boxes, densities_omp.cpp:33-65), same FP32 per-row order and
// the same OpenMP schedules.
}  // extern "C"
namespace {
struct Row1 { float v[1]; };
struct Row2 { float v[2]; };
struct Row3 { float v[3]; };

float epa_ref(float ref_val, float val, float h) {  // densities_omp.cpp:19-26
  float u_squared = (ref_val - val) / h;
  u_squared *= u_squared;
  return (float)((u_squared <= 1.0f ? 0.75 * (double)(1.0f - u_squared) : 0.0) / (double)h);
}

template <class ROW, unsigned D>
std::vector<float> sorted_dim1(const float* packed, std::size_t n) {
  std::vector<float> out(n * D);
  if (D == 1) {
    for (std::size_t i = 0; i < n; ++i) out[i] = packed[i];
    std::sort(out.begin(), out.end());
    return out;
  }
  std::vector<ROW> tmp(n);
  for (std::size_t i = 0; i < n; ++i)
    for (unsigned j = 0; j < D; ++j) tmp[i].v[j] = packed[j * n + i];
  std::sort(tmp.begin(), tmp.end(), [](const ROW& a, const ROW& b) { return a.v[0] < b.v[0]; });
  for (std::size_t i = 0; i < n; ++i)
    for (unsigned j = 0; j < D; ++j) out[j * n + i] = tmp[i].v[j];
  return out;
}
}  // namespace
extern "C" {

long orc_density_omp_packed(const float* packed, std::size_t n, unsigned d, const float* h,
                            float* out) {
  if (d < 1 || d > 3 || n == 0) return -1;
  std::vector<float> sc;
  if (d == 1) sc = sorted_dim1<Row1, 1>(packed, n);
  else if (d == 2) sc = sorted_dim1<Row2, 2>(packed, n);
  else sc = sorted_dim1<Row3, 3>(packed, n);
  std::vector<std::array<std::size_t, 2>> bxs(n);
  const float h0 = h[0];
  long i;
#pragma omp parallel for schedule(dynamic, 8)
  for (i = 0; i < (long)n; ++i) {
    float ref_val = packed[i];
    for (std::size_t k = 0; k < n; ++k)
      if (ref_val - sc[k] < h0) { bxs[i][0] = k; break; }
    for (std::size_t k = n; k > 0; --k)
      if (sc[k - 1] - ref_val < h0) { bxs[i][1] = k - 1; break; }
  }
  if (d == 1) {
#pragma omp parallel for schedule(dynamic, 1)
    for (i = 0; i < (long)n; ++i) {
      float ref_val = packed[i];
      float P = 0.0f;
      for (std::size_t k = bxs[i][0]; k <= bxs[i][1]; ++k) P += epa_ref(ref_val, sc[k], h0);
      out[i] = P / (float)n;
    }
  } else if (d == 2) {
#pragma omp parallel for schedule(static)
    for (i = 0; i < (long)n; ++i) {
      float r1 = packed[i], r2 = packed[n + i];
      float P = 0.0f;
      for (std::size_t k = bxs[i][0]; k <= bxs[i][1]; ++k)
        P += epa_ref(r1, sc[k], h[0]) * epa_ref(r2, sc[n + k], h[1]);
      out[i] = P / (float)n;
    }
  } else {
#pragma omp parallel for schedule(static)
    for (i = 0; i < (long)n; ++i) {
      float r1 = packed[i], r2 = packed[n + i], r3 = packed[2 * n + i];
      float P = 0.0f;
      for (std::size_t k = bxs[i][0]; k <= bxs[i][1]; ++k)
        P += epa_ref(r1, sc[k], h[0]) * epa_ref(r2, sc[n + k], h[1]) * epa_ref(r3, sc[2 * n + k], h[2]);
      out[i] = P / (float)n;
    }
  }
  return (long)n;
}

// ---------------------------------------------------------------- binning
// Transs::BoxedSearch::Boxes (boxedsearch.cpp:21-34) with the fixes that make
// it total: max initialised to lowest() (not min()), n_boxes >= 1, and the
// box of x == max clamped to n_boxes-1.  Arithmetic: box = (size_t)floorf(
// fl(fl(x - mn) / bs)), n_boxes = (size_t)ceilf(fl(fl(mx - mn) / bs)).
// Returns 0, or -1 for n == 0 / non-finite input / bs <= 0.
int orc_bin(const float* x, std::size_t n, float bs, int32_t* box, uint32_t* n_boxes,
            float* mn_out, float* mx_out) {
  if (n == 0 || !(bs > 0.0f) || !std::isfinite(bs)) return -1;
  float mn = std::numeric_limits<float>::max();
  float mx = std::numeric_limits<float>::lowest();
  for (std::size_t i = 0; i < n; ++i) {
    if (!std::isfinite(x[i])) return -1;
    mn = std::min(mn, x[i]);
    mx = std::max(mx, x[i]);
  }
  volatile float range = mx - mn;
  volatile float q = range / bs;
  float nb = std::ceil((float)q);
  uint32_t nbox = nb < 1.0f ? 1u : (uint32_t)nb;
  for (std::size_t i = 0; i < n; ++i) {
    volatile float dx = x[i] - mn;
    volatile float qq = dx / bs;
    uint32_t b = (uint32_t)std::floor((float)qq);
    if (b > nbox - 1) b = nbox - 1;
    box[i] = (int32_t)b;
  }
  *n_boxes = nbox;
  if (mn_out) *mn_out = mn;
  if (mx_out) *mx_out = mx;
  return 0;
}

// Stable counting sort of indices 0..n-1 by key (0 <= key < n_keys).
// perm[p] = original index at sorted position p; start has n_keys+1 entries.
int orc_counting_sort(const uint32_t* key, std::size_t n, uint32_t n_keys, uint32_t* perm,
                      uint32_t* start) {
  std::vector<uint32_t> cnt((std::size_t)n_keys + 1, 0);
  for (std::size_t i = 0; i < n; ++i) {
    if (key[i] >= n_keys) return -1;
    cnt[key[i] + 1]++;
  }
  for (uint32_t k = 0; k < n_keys; ++k) cnt[k + 1] += cnt[k];
  if (start) for (uint32_t k = 0; k <= n_keys; ++k) start[k] = cnt[k];
  for (std::size_t i = 0; i < n; ++i) perm[cnt[key[i]]++] = (uint32_t)i;
  return 0;
}

// --------------------------------------------------------- epa_convolution
// densities_omp.cpp:235-277: (1/N^2) sum_i sum_j [d<=2] (4.5 - 2.25 d),
// d = |x_i - x_j| / h.  FP32 replica: d = fl(|fl(x_i - x_j)| / h),
// c = fl(fma((double)d, -2.25, 4.5)); rows and total summed with textbook
// Kahan in FP32 (Boost sum_kahan: parity unpinned beyond the textbook
// algorithm); N*N formed in size_t then converted to float.
float orc_epa_conv_omp(const float* sorted, std::size_t n, float h) {
  std::vector<float> outer(n);
  std::vector<float> inner(n);
  for (std::size_t i = 0; i < n; ++i) {
    std::fill(inner.begin(), inner.end(), 0.0f);
    for (std::size_t j = 0; j < n; ++j) {
      float d = std::fabs(sorted[i] - sorted[j]) / h;
      if (d <= 2.0f) inner[j] = (float)std::fma((double)d, -2.25, 4.5);
    }
    outer[i] = kahan_f32(inner.data(), n);
  }
  return kahan_f32(outer.data(), n) / (float)(n * n);
}

// Exact FP64 value of the same functional (Neumaier, O(N*W) on sorted input).
double orc_epa_conv_fp64(const float* sorted, std::size_t n, float h) {
  double hd = (double)h;
  std::vector<double> rows(n);
#pragma omp parallel for schedule(dynamic, 64)
  for (long i = 0; i < (long)n; ++i) {
    Neumaier acc;
    double xi = (double)sorted[i];
    for (std::size_t j = 0; j < n; ++j) {
      double d = std::fabs(xi - (double)sorted[j]) / hd;
      if (d <= 2.0) acc.add(4.5 - 2.25 * d);
    }
    rows[i] = acc.value();
  }
  Neumaier tot;
  for (std::size_t i = 0; i < n; ++i) tot.add(rows[i]);
  return tot.value() / ((double)n * (double)n);
}

// ------------------------------------------------------------- hestimate
// Thumb (hestimate.cpp:19-30): h = N^(-1/7) * sqrt(var_pop).  Boost's lazy
// variance over an FP32 accumulator is restated as the textbook population
// variance in FP64 (parity unpinned at the FP32-accumulation level).
float orc_thumb(const float* x, std::size_t n) {
  Neumaier s;
  for (std::size_t i = 0; i < n; ++i) s.add((double)x[i]);
  double mean = s.value() / (double)n;
  Neumaier v;
  for (std::size_t i = 0; i < n; ++i) { double dx = (double)x[i] - mean; v.add(dx * dx); }
  double var = v.value() / (double)n;
  return (float)(std::pow((double)n, -1.0 / 7.0) * std::sqrt(var));
}

}  // extern "C"
namespace {
// hestimate.cpp:42-53, with the convolution value supplied by the caller.
float amise_from_conv(float conv, float h, std::size_t n) {
  float a = (float)(0.04 * 1.0 / std::pow((double)h, 5.0) * (double)conv);
  a = (float)std::pow((double)a, 0.2);
  a = (float)((double)a * 1.25 * std::pow(0.6, 0.8) / std::pow((double)(float)n, 0.8));
  return a;
}

float interval_h(const std::vector<float>& sorted, float prec, float lo, float hi, unsigned iter,
                 int exact) {
  float delta = hi - lo;
  float pivot = (float)((double)lo + 0.5 * (double)delta);
  if (iter == 0 || delta <= prec) return pivot;
  float pl = (float)((double)pivot - 0.25 * (double)delta);
  float ph = (float)((double)pivot + 0.25 * (double)delta);
  std::size_t n = sorted.size();
  float cl = exact ? (float)orc_epa_conv_fp64(sorted.data(), n, pl) : orc_epa_conv_omp(sorted.data(), n, pl);
  float ch = exact ? (float)orc_epa_conv_fp64(sorted.data(), n, ph) : orc_epa_conv_omp(sorted.data(), n, ph);
  float al = amise_from_conv(cl, pl, n);
  float ah = amise_from_conv(ch, ph, n);
  if (al < ah) return interval_h(sorted, prec, lo, pivot, iter - 1, exact);
  return interval_h(sorted, prec, pivot, hi, iter - 1, exact);
}
}  // namespace
extern "C" {

float orc_amise_value(float conv, float h, std::size_t n) { return amise_from_conv(conv, h, n); }

// AmiseMin::Kernel::operator() (hestimate.cpp:89-104).  exact != 0 uses the
// FP64 convolution, else the FP32 replica.
float orc_amise(const float* x, std::size_t n, int exact) {
  std::vector<float> sorted(x, x + n);
  std::sort(sorted.begin(), sorted.end());
  float h0 = orc_thumb(x, n);
  return interval_h(sorted, (float)(0.01 * (double)h0), (float)(0.1 * (double)h0), 2.0f * h0, 20, exact);
}

// ------------------------------------------------------- driver formulas
// select_log (tools.cpp:18-37): 0 -> 0, else ln or log2 in float.
static inline float ref_log(float x, int bits) {
  if (x == 0.0f) return 0.0f;
  return bits ? std::log2(x) : std::log(x);
}
static inline double inv_d(double x) { return x == 0.0 ? x : 1.0 / x; }  // tools.hxx:45-53

// mi.cpp:53-56: m = 0f; m += p_ij * log(float(p_ij / p_i / p_j)) (FP32 acc).
float orc_mi_ref(const double* p_ij, const double* p_i, const double* p_j, std::size_t n, int bits) {
  float m = 0.0f;
  for (std::size_t k = 0; k < n; ++k)
    m = (float)((double)m + p_ij[k] * (double)ref_log((float)(p_ij[k] / p_i[k] / p_j[k]), bits));
  return m;
}

// Same functional, FP64 throughout (0 * log 0 -> 0).
double orc_mi_fp64(const double* p_ij, const double* p_i, const double* p_j, std::size_t n, int bits) {
  Neumaier m;
  double sc = bits ? 1.0 / std::log(2.0) : 1.0;
  for (std::size_t k = 0; k < n; ++k) {
    double r = p_ij[k] / p_i[k] / p_j[k];
    if (r != 0.0) m.add(p_ij[k] * std::log(r) * sc);
  }
  return m.value();
}

// transs.cpp:75-113, one direction: sum over k < n of
// p3*pref*log(float(p3 * p1 * inv(p2) * inv(ptau))) where p3 > 0, the term is
// finite and p3 <= 1.  Returns the FP64 accumulator; *n_strange counts the
// rejected terms (the reference prints them).
double orc_te_ref(const double* p3, const double* p1, const double* p2, const double* ptau,
                  std::size_t n, double pref, int bits, long* n_strange) {
  double te_sum = 0.0;
  long strange = 0;
  for (std::size_t k = 0; k < n; ++k) {
    if (p3[k] > 0) {
      double te = p3[k] * pref * (double)ref_log((float)(p3[k] * p1[k] * inv_d(p2[k]) * inv_d(ptau[k])), bits);
      if (std::isfinite(te) && p3[k] <= 1.0) te_sum += te;
      else ++strange;
    }
  }
  if (n_strange) *n_strange = strange;
  return te_sum;
}

double orc_te_fp64(const double* p3, const double* p1, const double* p2, const double* ptau,
                   std::size_t n, double pref, int bits) {
  Neumaier s;
  double sc = bits ? 1.0 / std::log(2.0) : 1.0;
  for (std::size_t k = 0; k < n; ++k) {
    if (p3[k] > 0) {
      double r = p3[k] * p1[k] * inv_d(p2[k]) * inv_d(ptau[k]);
      if (r == 0.0) continue;
      double te = p3[k] * pref * std::log(r) * sc;
      if (std::isfinite(te) && p3[k] <= 1.0) s.add(te);
    }
  }
  return s.value();
}

// autoent.cpp:52-57: sum_k [p2>0] p2*h*h*log(float(p1 * inv(p2*h))) in FP64.
double orc_autoent_ref(const double* p1, const double* p2, std::size_t n, float h, int bits) {
  double a = 0.0;
  for (std::size_t k = 0; k < n; ++k)
    if (p2[k] > 0) a += p2[k] * h * h * (double)ref_log((float)(p1[k] * inv_d(p2[k] * h)), bits);
  return a;
}

// sum1_normalized (tools.hxx:33-43) with an FP64 Kahan sum.
void orc_sum1_normalized(const double* p, std::size_t n, double* out) {
  double s = kahan_f64(p, n);
  for (std::size_t k = 0; k < n; ++k) out[k] = (s != 0) ? p[k] / s : p[k];
}

// kldiv.cpp:49-51, both directions, from sum1-normalised densities.
void orc_kldiv_ref(const double* pi, const double* pj, std::size_t n, int bits, double* kl_ij,
                   double* kl_ji) {
  double a = 0.0, b = 0.0;
  for (std::size_t k = 0; k < n; ++k) {
    a += pi[k] * (double)ref_log((float)(pi[k] / pj[k]), bits);
    b += pj[k] * (double)ref_log((float)(pj[k] / pi[k]), bits);
  }
  *kl_ij = a;
  *kl_ji = b;
}

// negs.cpp:37-47 from sum1-normalised densities: FP32 accumulate, negated.
float orc_negs_ref(const double* p, std::size_t n, int bits) {
  float neg = 0.0f;
  for (std::size_t k = 0; k < n; ++k) neg = (float)((double)neg + (double)ref_log((float)p[k], bits) * p[k]);
  return -neg;
}

// Resubstitution entropy H = -(1/n) sum_k ln p_k (FP64).
double orc_entropy(const double* p, std::size_t n) {
  Neumaier s;
  for (std::size_t k = 0; k < n; ++k) s.add(std::log(p[k]));
  return -s.value() / (double)n;
}

// Pairs in the kernel's support: the number of ordered (query, point) pairs
// (i, j), the self pair included, with |x_ie - x_je| < h_e in every dim --
// exactly the pairs with a nonzero Epanechnikov term -- for d = 1 or 2
// (packed column-major input of n points; differences in FP64).  O(n log n):
// the points sorted by dim 0; a window of the points within h_0 along dim 0
// slides over them; for d = 2 a Fenwick tree over the dim-1 ranks of the
// window's points counts those within h_1.  (Measurement only: the share of
// the evaluated pairs that lie inside the support.)
uint64_t orc_count_in_support(const float* packed, std::size_t n, unsigned d, const float* h) {
  if (n == 0 || d < 1 || d > 2) return 0;
  std::vector<uint32_t> ord(n);
  for (std::size_t i = 0; i < n; ++i) ord[i] = (uint32_t)i;
  const float* x = packed;
  std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return x[a] < x[b]; });
  std::vector<double> xs(n);
  for (std::size_t i = 0; i < n; ++i) xs[i] = (double)x[ord[i]];
  const double h0 = (double)h[0];
  uint64_t total = 0;
  if (d == 1) {
    std::size_t lo = 0, hi = 0;
    for (std::size_t i = 0; i < n; ++i) {
      while (xs[i] - xs[lo] >= h0) ++lo;
      while (hi < n && xs[hi] - xs[i] < h0) ++hi;
      total += hi - lo;
    }
    return total;
  }
  const float* y = packed + n;
  const double h1 = (double)h[1];
  std::vector<double> yv(n);
  for (std::size_t i = 0; i < n; ++i) yv[i] = (double)y[i];
  std::vector<double> uy(yv);
  std::sort(uy.begin(), uy.end());
  uy.erase(std::unique(uy.begin(), uy.end()), uy.end());
  const std::size_t R = uy.size();
  std::vector<uint32_t> rank(n);
  for (std::size_t i = 0; i < n; ++i)
    rank[i] = (uint32_t)(std::lower_bound(uy.begin(), uy.end(), yv[i]) - uy.begin());
  std::vector<uint32_t> fen(R + 1, 0);
  auto add = [&](uint32_t r, int v) {
    for (std::size_t k = r + 1; k <= R; k += k & (~k + 1)) fen[k] += (uint32_t)v;
  };
  auto prefix = [&](std::size_t r) {  // points with rank < r
    uint64_t c = 0;
    for (std::size_t k = r; k > 0; k -= k & (~k + 1)) c += fen[k];
    return c;
  };
  std::size_t lo = 0, hi = 0;
  for (std::size_t i = 0; i < n; ++i) {
    while (xs[i] - xs[lo] >= h0) add(rank[ord[lo++]], -1);
    while (hi < n && xs[hi] - xs[i] < h0) add(rank[ord[hi++]], +1);
    const double yi = yv[ord[i]];
    // ranks r with uy[r] - yi > -h1 and uy[r] - yi < h1
    const std::size_t a = (std::size_t)(std::upper_bound(uy.begin(), uy.end(), yi,
                                                         [&](double v, double u) { return u - v > -h1; }) -
                                        uy.begin());
    const std::size_t b = (std::size_t)(std::lower_bound(uy.begin(), uy.end(), yi,
                                                         [&](double u, double v) { return u - v < h1; }) -
                                        uy.begin());
    total += prefix(b) - prefix(a);
  }
  return total;
}

// The same summed over jobs of d = 1 or 2 columns of one column-major matrix
// (job k: columns cols[k*d .. k*d + d), bandwidths h[k*d ..]), OpenMP over
// the jobs.  The jobs of a batch share columns (the MI matrix: 2016 pairs of
// 64), so each column's order by value and each (column, bandwidth)'s value
// ranks and rank bounds are computed once; a 2-D job is then one sweep.
namespace {
struct SortedCol {  // dim-0 role: values in ascending order and the order
  std::vector<uint32_t> ord;
  std::vector<double> xs;
};
struct RankedCol {  // dim-1 role at bandwidth h: per point its value rank and
  std::vector<uint32_t> rank, lo, hi;  // the ranks [lo, hi) within h of it
  std::size_t R = 0;
};
SortedCol sort_col(const float* x, std::size_t n) {
  SortedCol c;
  c.ord.resize(n);
  for (std::size_t i = 0; i < n; ++i) c.ord[i] = (uint32_t)i;
  std::stable_sort(c.ord.begin(), c.ord.end(), [&](uint32_t a, uint32_t b) { return x[a] < x[b]; });
  c.xs.resize(n);
  for (std::size_t i = 0; i < n; ++i) c.xs[i] = (double)x[c.ord[i]];
  return c;
}
RankedCol rank_col(const float* y, std::size_t n, double h1) {
  RankedCol c;
  std::vector<double> uy(n);
  for (std::size_t i = 0; i < n; ++i) uy[i] = (double)y[i];
  std::sort(uy.begin(), uy.end());
  uy.erase(std::unique(uy.begin(), uy.end()), uy.end());
  c.R = uy.size();
  c.rank.resize(n);
  c.lo.resize(n);
  c.hi.resize(n);
  for (std::size_t i = 0; i < n; ++i) {
    const double yi = (double)y[i];
    c.rank[i] = (uint32_t)(std::lower_bound(uy.begin(), uy.end(), yi) - uy.begin());
    c.lo[i] = (uint32_t)(std::upper_bound(uy.begin(), uy.end(), yi, [&](double v, double u) { return u - v > -h1; }) -
                         uy.begin());
    c.hi[i] = (uint32_t)(std::lower_bound(uy.begin(), uy.end(), yi, [&](double u, double v) { return u - v < h1; }) -
                         uy.begin());
  }
  return c;
}
uint64_t sweep_1d(const SortedCol& c, double h0) {
  const std::size_t n = c.xs.size();
  uint64_t total = 0;
  std::size_t lo = 0, hi = 0;
  for (std::size_t i = 0; i < n; ++i) {
    while (c.xs[i] - c.xs[lo] >= h0) ++lo;
    while (hi < n && c.xs[hi] - c.xs[i] < h0) ++hi;
    total += hi - lo;
  }
  return total;
}
uint64_t sweep_2d(const SortedCol& c, double h0, const RankedCol& r) {
  const std::size_t n = c.xs.size(), R = r.R;
  std::vector<uint32_t> fen(R + 1, 0);
  auto add = [&](uint32_t k0, uint32_t v) {
    for (std::size_t k = k0 + 1; k <= R; k += k & (~k + 1)) fen[k] += v;
  };
  auto prefix = [&](std::size_t k0) {
    uint64_t t = 0;
    for (std::size_t k = k0; k > 0; k -= k & (~k + 1)) t += fen[k];
    return t;
  };
  uint64_t total = 0;
  std::size_t lo = 0, hi = 0;
  for (std::size_t i = 0; i < n; ++i) {
    while (c.xs[i] - c.xs[lo] >= h0) add(r.rank[c.ord[lo++]], ~0u);  // -1 mod 2^32
    while (hi < n && c.xs[hi] - c.xs[i] < h0) add(r.rank[c.ord[hi++]], 1u);
    const uint32_t q = c.ord[i];
    total += prefix(r.hi[q]) - prefix(r.lo[q]);
  }
  return total;
}
}  // namespace

uint64_t orc_count_in_support_jobs(const float* X, std::size_t n_rows, unsigned n_jobs, unsigned d,
                                   const unsigned* cols, const float* h) {
  if (d < 1 || d > 2 || n_rows == 0) return 0;
  std::map<unsigned, std::size_t> s_id;
  std::map<std::pair<unsigned, uint32_t>, std::size_t> r_id;
  for (unsigned k = 0; k < n_jobs; ++k) {
    s_id.emplace(cols[k * d], s_id.size());
    if (d == 2) {
      uint32_t hb;
      std::memcpy(&hb, &h[k * d + 1], 4);
      r_id.emplace(std::make_pair(cols[k * d + 1], hb), r_id.size());
    }
  }
  std::vector<SortedCol> sc(s_id.size());
  std::vector<RankedCol> rc(r_id.size());
  std::vector<std::pair<unsigned, std::size_t>> sl(s_id.begin(), s_id.end());
  std::vector<std::pair<std::pair<unsigned, uint32_t>, std::size_t>> rl(r_id.begin(), r_id.end());
#pragma omp parallel for schedule(dynamic, 1)
  for (long i = 0; i < (long)sl.size(); ++i) sc[sl[i].second] = sort_col(X + (std::size_t)sl[i].first * n_rows, n_rows);
#pragma omp parallel for schedule(dynamic, 1)
  for (long i = 0; i < (long)rl.size(); ++i) {
    float hv;
    std::memcpy(&hv, &rl[i].first.second, 4);
    rc[rl[i].second] = rank_col(X + (std::size_t)rl[i].first.first * n_rows, n_rows, (double)hv);
  }
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
  for (long k = 0; k < (long)n_jobs; ++k) {
    const SortedCol& c = sc[s_id.at(cols[k * d])];
    const double h0 = (double)h[k * d];
    if (d == 1) {
      total += sweep_1d(c, h0);
    } else {
      uint32_t hb;
      std::memcpy(&hb, &h[k * d + 1], 4);
      total += sweep_2d(c, h0, rc[r_id.at(std::make_pair(cols[k * d + 1], hb))]);
    }
  }
  return total;
}

// Leave-one-out log-likelihood for a set of candidate isotropic scale
// factors: per dim the bandwidth is hs[c] * scale[e].  p_{-i} = (S_i - K(0))
// / (n-1); isolated points (p_{-i} <= 0) contribute log(floor).  Brute force.
void orc_loo_loglik(const float* pts_rowmajor, std::size_t n, unsigned d, const float* scale,
                    const float* hs, unsigned n_h, double floor_p, double* out) {
  std::vector<double> lp(n);
  for (unsigned c = 0; c < n_h; ++c) {
    double h[3];
    for (unsigned e = 0; e < d; ++e) h[e] = (double)hs[c] * (double)scale[e];
#pragma omp parallel for schedule(dynamic, 32)
    for (long i = 0; i < (long)n; ++i) {
      double xi[3], xj[3];
      for (unsigned e = 0; e < d; ++e) xi[e] = pts_rowmajor[(std::size_t)i * d + e];
      Neumaier s;
      for (std::size_t j = 0; j < n; ++j) {
        if ((long)j == i) continue;
        for (unsigned e = 0; e < d; ++e) xj[e] = pts_rowmajor[j * d + e];
        s.add(term_fp64(xi, xj, h, d));
      }
      double p = s.value() / (double)(n - 1);
      lp[i] = std::log(p > floor_p ? p : floor_p);
    }
    Neumaier tot;
    for (std::size_t i = 0; i < n; ++i) tot.add(lp[i]);
    out[c] = tot.value();
  }
}

}  // extern "C"
