/* entropical_kde.h -- C ABI of the MI355X (gfx950) Epanechnikov KDE hot path.
 *
 * Drop-in boundary for lettis/entropical's density backend.  Each entry
 * point names the reference interface it replaces (file:line in the
 * lettis/entropical source tree).  All functions return an ENT_* status;
 * on failure ent_last_error() holds a message (per calling thread).
 *
 * Data layout: coordinates are column-major float32, coords[col*n_rows+row],
 * as Tools::IO::read_coords(..., 'C') produces them (tools.hxx:226-228).
 * Densities are returned as double, divided by the ORIGINAL n_rows, with
 * n_rows - max(tau) entries (densities_cuda.cu:95-103).
 *
 * Threading: one context per thread/device; a context is not re-entrant.
 */
#ifndef ENTROPICAL_KDE_H
#define ENTROPICAL_KDE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ENT_OK 0
#define ENT_ERR_ARG 1         /* bad dims, |h| != d, |tau| != d, tau >= n_rows, column out of range, h <= 0 */
#define ENT_ERR_NO_DEVICE 2   /* no HIP device (the reference exits: density_1d.cuh:190-193) */
#define ENT_ERR_HIP 3         /* HIP runtime error (replaces check_error, tools.cuh:33-42) */
#define ENT_ERR_OOM 4
#define ENT_ERR_NONFINITE 5   /* NaN/Inf coordinate */
#define ENT_ERR_UNSUPPORTED 6 /* configuration outside the engine's limits: a job of 2^30
                                 points or more, a batch of 2^32 points or more, more than
                                 2^24 boxes along one dim */

#define ENT_METHOD_CELL 0     /* cell-list candidates (default) */
#define ENT_METHOD_BRUTE 1    /* all pairs, as the reference's CUDA backend does */

#define ENT_HEST_THUMB 0      /* Hestimate::Thumb (hestimate.cpp:19-30) */
#define ENT_HEST_AMISE 1      /* Hestimate::AmiseMin (hestimate.cpp:41-104) */

typedef struct ent_ctx ent_ctx;

/* One density estimate in a batch.  Point k (0 <= k < n), dim e:
 * X[off[e] + k].  Densities go to dens[out_off + k], divided by n_rows. */
typedef struct ent_job {
  uint32_t n_dim;
  uint32_t n;
  uint32_t n_rows;
  uint32_t reserved;
  uint64_t off[3];
  float h[3];
  float reserved2;
  uint64_t out_off;
} ent_job;

/* Library version (no reference counterpart). */
int ent_version(void);

/* Message of the last failing call on this thread.  Replaces the reference's
 * std::cerr + exit(EXIT_FAILURE) error path (densities_common.hxx:16-33,
 * tools.cuh:33-42). */
const char* ent_last_error(void);

/* Number of visible HIP devices (replaces cudaGetDeviceCount,
 * density_1d.cuh:188-193). */
int ent_device_count(int* n);

/* Context = device + stream + grow-only workspace.  Replaces the per-call
 * cudaSetDevice/cudaMalloc/cudaFree of density_Nd_per_gpu
 * (density_1d.cuh:105-111,163-164). */
int ent_ctx_create(int device, ent_ctx** out);
int ent_ctx_destroy(ent_ctx* ctx);

/* Engine tuning: maximum queries per lane (1,2,4; smaller tiles use fewer), dim-0 boxes per bandwidth,
 * row-dim boxes per bandwidth, tile extent bound in bandwidths, and the
 * sub-tile size (0: whole tiles; 16 or 32, with q = 1: tiles of up to 64
 * queries run as sub-tiles of that many queries, each with its own candidate
 * stream: 2-D jobs at mr >= 2 with 16, mr >= 4 with 32; 3-D jobs at mr <= 2
 * with 16, mr = 1 with 32, the batch's 2-D jobs then as whole tiles; other
 * shapes keep whole tiles).  Replaces the reference's compile-time BSIZE
 * (densities_cuda.cu:14). */
int ent_ctx_set_tuning(ent_ctx* ctx, int q, int m0, int mr, int group_h, int sub);

/* Statistics of the last density run: persistent-kernel time and whole
 * pipeline time (HIP events, ms), candidate pair evaluations (the pairs of
 * every tile's candidate window), tiles. */
int ent_ctx_last_stats(ent_ctx* ctx, double* density_ms, double* total_ms, uint64_t* pairs,
                       uint64_t* tiles);

/* Sort shape of the last cell-list run: out[0] = radix passes launched (over
 * the shared dim-0 sorts plus over the jobs), out[1] = shared dim-0 sorts
 * (0 when every job was sorted by its full cell key).  Jobs with the same
 * dim-0 window and bandwidth share one dim-0 sort when that halves the
 * points to sort; ENT_SORT=full|shared in the environment forces a mode.
 * No reference counterpart. */
int ent_ctx_last_sort(ent_ctx* ctx, uint32_t* out2);

/* Capture the post-planning launch sequence (cell keys, radix passes, gather,
 * tiles) and the persistent density launch in hipGraphs that are replayed
 * while the plan shape repeats (default on); 0 launches eagerly.  No
 * reference counterpart. */
int ent_ctx_set_graph(ent_ctx* ctx, int enable);

/* Candidate pair evaluations of the last density run, split by job dimension
 * (out[d-1] for d = 1, 2, 3), and out[3] = lane-slot evaluations issued
 * pair by pair (64*Q lanes per tile times candidates rounded up to chunks;
 * the pairs evaluated pair by pair, sum(out[0..2]) less the moment pairs of
 * ent_ctx_last_moments, over out[3] is the lane utilisation). */
int ent_ctx_last_pairs(ent_ctx* ctx, uint64_t* out4);

/* Moment sums of the last density run: out[d-1] = candidate pairs of d-dim
 * jobs (d = 1, 2; counted in ent_ctx_last_pairs) summed from box or cell
 * moments instead of pair by pair, out[2] = (query, box) terms evaluated for
 * them in 1-D jobs, out[3] = (query, cell) terms in 2-D jobs.  A 1-D box, or
 * a 2-D cell (one row high,
 * about h/4 wide), that lies inside the support of every query of a tile
 * contributes a polynomial in the query's offset from its origin, whose
 * coefficients are the box's point count and moments up to second order
 * per dim: it equals the sum of the pairs' kernel values.
 * ENT_BOX_MOMENTS=0 in the environment turns this off, and
 * ENT_MOMENT_TABLE_MAX=<bytes> caps the tables' device memory (a run whose
 * tables would exceed it, or the memory left, evaluates every pair one by
 * one).  No reference counterpart. */
int ent_ctx_last_moments(ent_ctx* ctx, uint64_t* out4);

/* Tiles of the last cell-list run: out[0] = tiles claimed by the density
 * kernel (packed tiles count once), out[1] = those whose pair terms were
 * taken from raw coordinates, u = (x_q - x_c) / h, because their queries
 * span more than 32 bandwidths along dim 0 (sparse stretches); the others
 * use coordinates scaled about the tile's centre.  No reference counterpart. */
int ent_ctx_last_tiles(ent_ctx* ctx, uint64_t* out2);

/* Captured graphs (launch sequences of a plan shape) replayed in the last
 * density run instead of launched anew; 0 when every launch was captured
 * or issued eagerly.  No reference counterpart. */
int ent_ctx_last_graph_replays(ent_ctx* ctx, uint32_t* out);

/* Device memory helpers on the context's device and stream. */
int ent_dev_alloc(ent_ctx* ctx, size_t bytes, void** ptr);
int ent_dev_free(ent_ctx* ctx, void* ptr);
int ent_h2d(ent_ctx* ctx, void* dst, const void* src, size_t bytes);
int ent_d2h(ent_ctx* ctx, void* dst, const void* src, size_t bytes);
int ent_ctx_sync(ent_ctx* ctx);

/* prob_dens_coord_prep<float> (densities_common.hpp:11-18,
 * densities_common.hxx:7-50): lagged gather of the selected columns into a
 * (n_rows - tau_max) x n_dim array, row- or column-major.  Host only. */
int ent_coord_prep(const float* coords, size_t n_rows, const uint32_t* i_cols, const float* h,
                   const uint32_t* tau, uint32_t n_dim, int row_major, float* out, size_t* n_sel);

/* combined_densities(coords, n_rows, i_cols, h, tau) -> vector<double>
 * (densities_cuda.hpp:10-24, densities_cuda.cu:16-105).  Host pointers;
 * n_cols bounds i_cols; tau may be NULL (all zero, densities_cuda.cu:17-44).
 * out receives n_rows - max(tau) doubles. */
int ent_combined_densities(ent_ctx* ctx, const float* coords, size_t n_rows, size_t n_cols,
                           const uint32_t* i_cols, const float* h, const uint32_t* tau,
                           uint32_t n_dim, int method, double* out);

/* Same, with coords and out in device memory of the context's device. */
int ent_combined_densities_dev(ent_ctx* ctx, const float* coords_dev, size_t n_rows, size_t n_cols,
                               const uint32_t* i_cols, const float* h, const uint32_t* tau,
                               uint32_t n_dim, int method, double* out_dev);

/* Limits of ent_combined_densities_dev_noalloc: the bytes of its fixed
 * device scratch array, and the most points (n_rows - max(tau)) a call may
 * have. */
#define ENT_NOALLOC_SCRATCH_BYTES (128u * 1024u * 1024u)
#define ENT_NOALLOC_MAX_POINTS 1000000u

/* ent_combined_densities_dev that allocates nothing (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as
 * serving loops and harnesses that forbid hipMalloc/hipFree.  Same arguments,
 * results and errors as ent_combined_densities_dev (bit-identical densities
 * for the same tuning), with coords_dev and out_dev owned by the caller; the
 * workspace is a fixed device array of ENT_NOALLOC_SCRATCH_BYTES per GPU,
 * located when the context is created and held until the process ends.
 * More than ENT_NOALLOC_MAX_POINTS points return ENT_ERR_UNSUPPORTED (as
 * does a workspace that would not fit).  No device or pinned host memory is
 * allocated or freed and no graph is captured during the call.  All contexts
 * on one device share the array: calls that overlap on one GPU race. */
int ent_combined_densities_dev_noalloc(ent_ctx* ctx, const float* coords_dev, size_t n_rows, size_t n_cols,
                                       const uint32_t* i_cols, const float* h, const uint32_t* tau,
                                       uint32_t n_dim, int method, double* out_dev);

/* Many density estimates in one engine run (no reference counterpart: the
 * reference makes one combined_densities call per density, mi.cpp:37-52).
 * X_dev and dens_dev are device pointers; jobs is a host array. */
int ent_batch_densities_dev(ent_ctx* ctx, const float* X_dev, const ent_job* jobs, size_t n_jobs,
                            int method, double* dens_dev);

/* epa_convolution(sorted_coords, h) (densities_cuda.hpp:29-31) with the
 * OpenMP semantics (densities_omp.cpp:235-277):
 * (1/N^2) sum_i sum_j [d <= 2] (4.5 - 2.25 d), d = |x_i - x_j| / h.
 * Host input (any order), result in *out. */
int ent_epa_convolution(ent_ctx* ctx, const float* x, size_t n, float h, float* out);

/* Transs::BoxedSearch::Boxes binning of one column (boxedsearch.cpp:15-34):
 * box = floor((x - min) / box_size), n_boxes = ceil((max - min) / box_size),
 * with max initialised to the lowest float, n_boxes >= 1 and x == max
 * clamped to n_boxes - 1.  Computed on the GPU; host in/out. */
int ent_bin(ent_ctx* ctx, const float* x, size_t n, float box_size, int32_t* box,
            uint32_t* n_boxes, float* mn, float* mx);

/* Cell list of one density job (no reference counterpart; the multi-dim
 * generalisation of Boxes): stable sort of the job's points by cell key
 * ((b2*nb1 + b1)*nb0 + b0) with box sizes h_e/m (m = tuning m0 for dim 0,
 * mr for the others).  Host in/out: keys_sorted and perm have
 * n_rows - max(tau) entries; nb, mn, bs have n_dim entries. */
int ent_cell_sort(ent_ctx* ctx, const float* coords, size_t n_rows, size_t n_cols,
                  const uint32_t* i_cols, const float* h, const uint32_t* tau, uint32_t n_dim,
                  uint64_t* keys_sorted, uint32_t* perm, uint32_t* nb, float* mn, float* bs);

/* Mi::main (mi.cpp:14-69) as one batch: n_cols 1-D marginals (once each;
 * the reference recomputes p_j per pair, mi.cpp:43-46) and n_cols*(n_cols-1)/2
 * 2-D joints in one engine run, then
 * MI[i][j] = sum_k p_ij log(p_ij / p_i / p_j) reduced on the device in FP64
 * (the reference accumulates in FP32 with logf, mi.cpp:53-56).  Mirrored,
 * diagonal 0.  X_dev: device column-major coordinates; h: host, n_cols
 * bandwidths; bits != 0 selects log2 (select_log, tools.cpp:18-37); mi: host
 * n_cols*n_cols doubles. */
int ent_mi_matrix_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h,
                      int bits, int method, double* mi);

/* Transs::main (transs.cpp:16-125) for several lags in one batch.
 * te[t*n_cols*n_cols + i*n_cols + j] has the layout of the reference's
 * transs matrix for lag taus[t] (TE(j->i) at [i][j], transs.cpp:84,102),
 * including its prefactors (h_i h_i h_j and h_i h_j h_i) and its rejection of
 * non-finite terms and p3 > 1.  FP64 reduction on the device. */
int ent_te_matrix_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h,
                      const uint32_t* taus, size_t n_taus, int bits, int method, double* te);

/* Resubstitution entropy H = -(1/n) sum_k ln p(x_k) of every job in one
 * batch (no reference counterpart; the closest reference output is
 * Negs::main, negs.cpp:37-47).  H: host, n_jobs doubles. */
int ent_batch_entropy_dev(ent_ctx* ctx, const float* X_dev, const ent_job* jobs, size_t n_jobs,
                          int method, double* H);

/* Leave-one-out log-likelihood sum_k ln max(p_{-k}(x_k), floor) of every job
 * (bandwidth labels of the learned model; no reference counterpart).
 * p_{-k} excludes the self pair; floor = floor_rel * prod(0.75/h) / (n-1),
 * i.e. floor_rel of one coincident neighbour.  jobs[j].n_rows is ignored. */
int ent_batch_loo_dev(ent_ctx* ctx, const float* X_dev, const ent_job* jobs, size_t n_jobs, double floor_rel,
                      int method, double* loglik);

/* AutoEnt::main (autoent.cpp:16-63) for every column of device-resident X in
 * one batch: out[i] = sum_k [p2 > 0] p2 h h log(p1 inv(p2 h)), p1 = 1-D
 * density of x_{t+tau}, p2 = 2-D density of (x_t, x_{t+tau}), both over the
 * n_rows - tau lagged rows (FP64 reduction; the reference accumulates float
 * logs in double).  bits: log2 instead of ln (select_log, tools.cpp:18-37). */
int ent_autoent_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h,
                    uint32_t tau, int bits, int method, double* out);

/* KLDiv::main (kldiv.cpp:14-63): out[i*n_cols + j] = sum_k p_i log(p_i / p_j)
 * over sum1-normalised 1-D densities (tools.hxx:33-43), diagonal 0. */
int ent_kldiv_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h, int bits,
                  int method, double* out);

/* Negs::main (negs.cpp:14-53): out[i] = -sum_k p log p over the
 * sum1-normalised 1-D density of column i (FP64; the reference accumulates
 * in float). */
int ent_negs_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h, int bits,
                 int method, double* out);

/* GROMACS XTC trajectory (CoordsFile::XtcHandler, coords_file/coords_file.cpp:72-127,
 * over the vendored xdrfile library) read on the host into the column-major
 * coordinate layout: frame f is row f, coordinate e of atom a is column
 * 3a + e (the row layout of XtcHandler::next).  Call with out == NULL for
 * *n_frames and *natoms, then with capacity >= n_frames * 3 * natoms floats.
 * Own decoder of the format (no xdrfile); all frames must have the same
 * atom count. */
int ent_read_xtc(const char* path, float* out, size_t capacity, size_t* n_frames, uint32_t* natoms);

/* Bandwidth per column of device-resident column-major X, as Hestimate::main
 * computes it for every selected column (hestimate.hxx:6-24):
 * ENT_HEST_THUMB: N^(-1/7) * population sd (Thumb::Kernel,
 * hestimate.cpp:19-30; variance in FP64);
 * ENT_HEST_AMISE: AmiseMin::Kernel (hestimate.cpp:41-104): interval search
 * from [0.1 h0, 2 h0] to 0.01 h0 (at most 20 halvings) for the minimum of the
 * AMISE built on epa_convolution.  All columns search in lockstep: each step
 * evaluates both pivots of every unfinished column in one batched launch.
 * Zero-variance columns are an error for AMISE. */
int ent_hestimate_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, int kind, float* h_out);

/* Tile-shape autotuner (no reference counterpart; the reference fixes its
 * tile at compile time, densities_cuda.cu:14).  Times every setting of the
 * grid q in {1,2,4}, m0 in {8,16,32}, mr in {1,2}, group_h in {4,16,64} with
 * whole tiles, for batches with 2-D jobs also q = 1, m0 in {32,64},
 * mr in {2,3,4}, group_h = 64 with sub-tiles of 16, and for batches with
 * 3-D jobs q = 1, m0 in {16,32}, mr in {1,2}, group_h = 64 with sub-tiles of
 * 16 (3-D jobs; their 2-D jobs then run whole tiles), on a
 * subsample of the job set (the four fastest re-timed on the whole set),
 * keeps the fastest on the context (and returns it in tuning5 = (q, m0, mr,
 * group_h, sub) when non-NULL),
 * and caches it per workload signature (file named by ENT_TUNE_CACHE, if set).
 * The MI/TE variants tune on the job sets of ent_mi_matrix_dev /
 * ent_te_matrix_dev; the generic one on an arbitrary batch (best_ms: time of
 * the winning setting on the subsample, -1 on a cache hit). */
int ent_mi_autotune_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h,
                        int* tuning5);
int ent_te_autotune_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const float* h,
                        const uint32_t* taus, size_t n_taus, int* tuning5);
int ent_autotune_dev(ent_ctx* ctx, const float* X_dev, const ent_job* jobs, size_t n_jobs, int* tuning5,
                     double* best_ms);

/* Current engine tuning (q, m0, mr, group_h, sub) into tuning5. */
int ent_ctx_get_tuning(ent_ctx* ctx, int* tuning5);

/* Wall time (ms) of the last driver call on this context. */
int ent_ctx_last_driver_ms(ent_ctx* ctx, double* ms);

/* ---- learned bandwidth model (no reference counterpart; the reference
 * chooses h with Hestimate::Thumb / AmiseMin, hestimate.cpp:19-104, which
 * stay available as the fallback and the yardstick).  A multilayer perceptron
 * n_in -> hidden^n_hidden_layers -> 1 with tanh units, BF16 weights and FP32
 * accumulation, trained with minibatch Adam by this library's kernels. */
/* Per-column sample statistics of the learned bandwidth's features (no
 * reference counterpart; the rule of thumb it refines is
 * Hestimate::Thumb::Kernel, hestimate.cpp:19-30), computed on the device
 * for every column of device-resident column-major X: out[c * WIDTH + ...] =
 * mean, population variance (FP64), the column's order statistics at the
 * ENT_COLSTATS_RANKS given 0-based ranks (exact, as float values), and with
 * z = (x - mean) / sd: sum z^3, sum z^4, sum |z|, the count of |z| > 3 and
 * the 32-bin histogram of z over [-4, 4] (edges -4 + k/4, last bin closed).
 * out is host memory of n_cols * ENT_COLSTATS_WIDTH doubles. */
#define ENT_COLSTATS_RANKS 18
#define ENT_COLSTATS_WIDTH 56
int ent_column_stats_dev(ent_ctx* ctx, const float* X_dev, size_t n_rows, size_t n_cols, const uint32_t* ranks,
                         double* out);

typedef struct ent_mlp ent_mlp;
int ent_mlp_create(ent_ctx* ctx, int n_in, int hidden, int n_hidden_layers, uint64_t seed, ent_mlp** out);
int ent_mlp_destroy(ent_mlp* mlp);
int ent_mlp_num_params(ent_mlp* mlp, size_t* n);
/* FP32 master parameters: per layer W[n_in][n_out] then b[n_out]. */
int ent_mlp_get_params(ent_ctx* ctx, ent_mlp* mlp, float* out);
int ent_mlp_set_params(ent_ctx* ctx, ent_mlp* mlp, const float* in);
/* y[i] = mlp(X[i, :]) for n rows of n_in floats (host arrays). */
int ent_mlp_predict(ent_ctx* ctx, ent_mlp* mlp, const float* X, size_t n, float* y);
/* Gradient of the batch mean squared error at the current parameters. */
int ent_mlp_gradient(ent_ctx* ctx, ent_mlp* mlp, const float* X, const float* t, size_t n, float* grads,
                     float* loss);
/* Minibatch Adam: order holds epochs*n row indices (one permutation per
 * epoch); losses receives the mean training loss of each epoch. */
int ent_mlp_train(ent_ctx* ctx, ent_mlp* mlp, const float* X, const float* t, size_t n,
                  const uint32_t* order, int epochs, int batch, float lr, float* losses);

#ifdef __cplusplus
}
#endif

#endif /* ENTROPICAL_KDE_H */
