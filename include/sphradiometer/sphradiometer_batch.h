/*
 * sphradiometer_batch.h -- batched, device-resident form of the
 * frequency-domain network integrator plus map synthesis.  No counterpart
 * in the reference: it runs, for many segments in one set of launches,
 * what a caller of the reference does per segment,
 *
 *   for each detector: correlator_tseries_to_fseries   (correlator.c:224)
 *   correlator_network_integrate_power_fd              (correlator.c:993)
 *   sh_series_rotate_z(sky, sky, gmst)                 (caller's step)
 *   sh_series_to_mesh(sky)                             (sh_series_trans.c:796)
 *
 * with the plan built by correlator_network_plan_fd_new / _set_l.  Inputs
 * and outputs are device buffers; nothing is copied to or from the host.
 *
 * Layouts (row-major):
 *   tseries  [n_det][n_segments][N]        float64 (or float32)
 *   window   [N] float64, applied before the FFT, may be NULL
 *   gmst     [n_segments] float64 radians, may be NULL (no rotate_z)
 *   alm      [n_segments][(L+1)^2]         complex (re, im) in the
 *                                          reference coefficient order
 *   map      [n_segments][2(L+1)][2(L+1)]  complex, rows theta ascending
 *                                          (Gauss-Legendre), columns phi
 * Complex values are float64 pairs for SPHR_FP64 and float32 pairs for
 * SPHR_MIXED.
 */
#ifndef SPHRADIOMETER_BATCH_H
#define SPHRADIOMETER_BATCH_H

#include <stddef.h>
#include <sphradiometer/correlator.h>

#ifdef __cplusplus
extern "C" {
#endif

enum sphr_precision {
	/* every stage in float64: the drop-in arithmetic */
	SPHR_FP64 = 0,
	/* FFT, projection, rotation and synthesis in float32 with float32
	 * accumulation; plan tables rounded from the float64 plan */
	SPHR_MIXED = 1
};

struct sphr_batch;

/* Device tables from a network FD plan whose baselines are all polar (a
 * real-data plan, before any mult_by_projection), segment length N a
 * power of two, every baseline's l_max <= sky_l_max.  max_segments sizes
 * the scratch; NULL on failure (sphr_last_error() says why). */
struct sphr_batch *sphr_batch_new(const struct correlator_network_plan_fd *plan, unsigned sky_l_max, int precision, int max_segments);
void sphr_batch_free(struct sphr_batch *);

/* One pass over n_segments <= max_segments segments, asynchronous on
 * stream (hipStream_t; NULL = the library stream).  alm may be NULL only
 * when map is NULL too.  Returns 0 when the launches were queued. */
int sphr_batch_run(struct sphr_batch *, const void *tseries, int tseries_is_f32, int n_segments, const double *window, const double *gmst, void *alm, void *map, void *stream);

/* sphr_batch_run without allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory while they run (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as sphr_batch_run; the spectra,
 * K-split partials and ring sums go to a static 64 MiB device array that
 * the library holds on each GPU it runs on until the process ends, so
 * n_segments is bounded by that size instead of max_segments (at three
 * sites, N = 1024, l_max 32: 1024 segments in float64).  Larger calls
 * return -1 with the byte count in sphr_last_error().  No allocation, no
 * profiling events.  Every call on one GPU shares the array: calls that
 * overlap on one GPU (other streams, other threads) race; calls queued on
 * one stream do not.  The batch object itself is made beforehand by
 * sphr_batch_new. */
int sphr_batch_run_static(struct sphr_batch *, const void *tseries, int tseries_is_f32, int n_segments, const double *window, const double *gmst, void *alm, void *map, void *stream);

/* Per-stage timing: while enabled, every run records HIP events on its
 * stream around its stages (FFT, projection, a_lm + Legendre ring sums,
 * phi DFT); _read synchronises and returns the summed milliseconds per
 * stage and the number of runs since enabling. */
int sphr_batch_profile(struct sphr_batch *, int enable);
int sphr_batch_profile_read(struct sphr_batch *, double ms_per_stage[4], int *runs);

/* Peak of each segment's sky: the maximum (find_max != 0) or minimum of
 * Re f for the n_segments series in alm, as sh_series_real_minimum
 * (sh_series_extremize.c:273-360) finds it, one wavefront per segment.
 * alm and map are the device outputs of sphr_batch_run (same precision);
 * with map the seed is the map's extreme pixel, with map NULL it is the
 * extreme pixel of the l = 6 low-pass mesh, as in the reference.  The
 * refinement evaluates the series in float64 whatever its precision.
 * theta, phi, value: device float64 [n_segments], theta in [0, pi], phi
 * in [0, 2 pi).  Asynchronous on stream; 0 when queued. */
int sphr_batch_peaks(struct sphr_batch *, const void *alm, const void *map, int n_segments, int find_max, double *theta, double *phi, double *value, void *stream);

/*
 * Batched localization skies (no counterpart in the reference): for many
 * events, what RapidLocalization_.sphcoeff does per event
 * (RapidLocalization.py:246-298): preprocess_SNRTimeSeries, the c2c FFT
 * per detector, generate_alm_sky for the beta = +1 and beta = -1 plans
 * (inspiral_sky_map.c:869-939) and sh_series_scale_power_l(power,
 * overall).  plansp / plansn have been through
 * correlator_network_plan_mult_by_projection (beta = +1 / -1) and set_l;
 * logprior has order l_max (the sky order); N of the plans must be
 * precalculated_TimeSeries_length(length, delta_t).
 *
 * Layouts (row-major, device):
 *   series [n_det][n_events][length]   complex (float32 pairs, or float64
 *                                       pairs with series_f64)
 *   gmst   [n_events] float64, may be NULL (no rotate_z)
 *   alm    [2][n_events][(L+1)^2]      complex, beta = +1 first
 *   map    [2][n_events][2(L+1)][2(L+1)] complex, may be NULL
 *   theta, phi, value [2][n_events]    float64: the maximum of each sky
 */
struct sphr_loc;
struct sphr_loc *sphr_loc_new(const struct correlator_network_plan_fd *plansp, const struct correlator_network_plan_fd *plansn, const struct sh_series *logprior, unsigned l_max, double power, double overall, int length, double delta_t, int precision, int max_events);
void sphr_loc_free(struct sphr_loc *);
int sphr_loc_run(struct sphr_loc *, const void *series, int series_f64, int n_events, const double *gmst, void *alm, void *map, void *stream);
/* peaks of the 2 n_events skies of the last run's alm (and map, which
 * seeds the search when given) */
int sphr_loc_peaks(struct sphr_loc *, const void *alm, const void *map, int n_events, double *theta, double *phi, double *value, void *stream);
/* [0] detectors, [1] baselines, [2] N, [3] length, [4] l_max,
 * [5] precision, [6] max_events */
void sphr_loc_info(const struct sphr_loc *, int info[8]);

/*
 * HEALPix maps of many series on the GPU (the pixel-domain output of
 * sh_series_healpix.cc:247-286 and healpix.py:141-171, and the posterior
 * of RapidLocalization.py:498-530).  A plan holds the Legendre tables of
 * order l_max for nside (any nside up to 1024; sphr_healpix_nside gives the
 * reference's lmax2nside rule).  Layouts (device, row-major):
 *   alm   [n][(l_max+1)^2] complex (float64 pairs for SPHR_FP64, float32
 *         pairs for SPHR_MIXED), the reference coefficient order
 *   map   [n][12 nside^2] real: Re f at the RING pixel centres
 *   maps  [2][n_events][12 nside^2], the beta = +1 and -1 skies (as
 *         sphr_loc_run's alm through sphr_healpix_maps with n = 2 n_events)
 *   prob  [n_events][12 nside^2]: (e_+ / sum e_+ + e_- / sum e_-) / 2,
 *         e = exp(map - max map)
 * Calls are asynchronous on stream (NULL = the library stream).
 */
struct sphr_healpix;
int sphr_healpix_nside(unsigned l_max);
struct sphr_healpix *sphr_healpix_new(unsigned l_max, int nside, int precision);
void sphr_healpix_free(struct sphr_healpix *);
long sphr_healpix_npix(const struct sphr_healpix *);
int sphr_healpix_maps(const struct sphr_healpix *, const void *alm, int n, void *map, void *stream);
int sphr_healpix_posterior(const struct sphr_healpix *, const void *maps, int n_events, void *prob, void *stream);

/*
 * Learned peak proposal (no counterpart in the reference): a multilayer
 * perceptron from the (L+1)^2 real-form SH coefficients of a sky
 * (sphr_features) to K candidate directions, trained and evaluated with
 * this library's kernels; sphr_refine_peaks then refines candidates on
 * the exact series with the extremizer.  All pointers except host
 * parameter arrays are device pointers; calls are asynchronous on stream
 * (NULL = the library stream).
 */
struct sphr_mlp;
/* depth hidden layers of width (GELU), n_candidates outputs x 3; bf16
 * selects bfloat16 operands (float32 accumulation), else float32;
 * parameters initialised from seed */
struct sphr_mlp *sphr_mlp_new(int n_in, int width, int depth, int n_candidates, int max_batch, int bf16, unsigned long long seed);
void sphr_mlp_free(struct sphr_mlp *);
long sphr_mlp_n_params(const struct sphr_mlp *);
/* flat float32 parameters W_1 [in][out], b_1, ..., W_out, b_out (host) */
int sphr_mlp_set_params(struct sphr_mlp *, const float *host);
int sphr_mlp_get_params(const struct sphr_mlp *, float *host);
int sphr_mlp_get_grads(const struct sphr_mlp *, float *host);
/* y [B][3 K] (may be NULL: kept inside for sphr_mlp_seeds) */
int sphr_mlp_forward(struct sphr_mlp *, const float *x, int B, float *y, void *stream);
/* forward + loss (per row, may be NULL) + gradients of the mean loss */
int sphr_mlp_loss_grad(struct sphr_mlp *, const float *x, const float *target, int B, float lambda, float *loss, void *stream);
int sphr_mlp_adam(struct sphr_mlp *, float lr, float beta1, float beta2, float eps, int step, void *stream);
int sphr_mlp_train_step(struct sphr_mlp *, const float *x, const float *target, int B, float lambda, float lr, int step, float *loss, void *stream);
/* (theta, phi) of the B x K candidates of the last forward: seeds [B][K][2] */
int sphr_mlp_seeds(struct sphr_mlp *, int B, double *seeds, void *stream);
/* real-form coefficients at unit RMS: out [S][(L+1)^2] float32 */
int sphr_features(const void *alm, int alm_f32, int S, int L, float *out, void *stream);
/* extremum refinement of series alm[s / alm_div] from seeds [n][2];
 * results float64 [n] */
int sphr_refine_peaks(const void *alm, int alm_f32, int n_series, int alm_div, int l_max, const double *seeds, int find_max, double *theta, double *phi, double *value, void *stream);
/* synthetic burst data for training and evaluation: out [n_det][S][n]
 * float32 = noise * N(0,1) (Philox4x32-10 keyed by seed) + per-segment
 * sine-Gaussian amp sin(2 pi f0 u) exp(-(u / tau)^2), u = (k - n/2)
 * delta_t + delay_d; params [S][3 + n_det] = amp, f0, tau, delay_0.. */
int sphr_inject_bursts(const float *params, int n_det, int n_segments, int n, double delta_t, double noise, unsigned long long seed, float *out, void *stream);
/* test hook: C = op(A) op(B) of the network's GEMM (device float32) */
int sphr_mlp_gemm_check(int bf16, int ta, int tb, int M, int N, int K, const float *A, const float *B, float *C);
/* test hook of the split-FP16 product: C [M][N] = A [M][K] B[N][K]^T
 * (float32 device arrays) through the fp16 pair split and three MFMA
 * products; a_f32: A split while staged (the form the localization meshes
 * use) instead of split beforehand; synchronous */
int sphr_gemm_h3_check(int a_f32, int M, int N, int K, const float *A, const float *B, float *C);
/* test and timing hook of the real GEMM: C [M][N] = A [M][K] B [K][N]
 * (row-major device arrays, float64 when f64 else float32) on the
 * library's kernel (impl 0) or rocBLAS (impl 1), queued on stream */
int sphr_gemm_run(int impl, int f64, int M, int N, int K, const void *A, long lda, const void *B, long ldb, void *C, long ldc, void *stream);

/*
 * Batches of independent jobs (no counterpart in the reference): job j
 * has its own network of n_det[j] sites (positions in light-seconds,
 * [sum n_det][3]), segment length N[j] and sky order l_max[j], with one
 * real segment each; the result per job is what
 * correlator_network_plan_fd_new + _set_l(l_max) + tseries_to_fseries +
 * correlator_network_integrate_power_fd + to_mesh give for it.  Plans are
 * generated on the GPU by sphr_jobs_new (the lag-domain form of the
 * frequency-domain plan, (2 n_p - 1) x (l + 1) values per baseline); each
 * baseline's tap count n_p must not exceed its job's N.
 *
 * Layouts (device, packed in job order; offsets from sphr_jobs_offsets):
 *   in   float32, job j at in[j]: [n_det][N]
 *   alm  complex (float64 or float32 pairs), job j at alm[j]: (L+1)^2
 *   map  complex, job j at map[j]: [2(L+1)][2(L+1)], may be NULL
 * sphr_jobs_run replays a captured HIP graph while the buffers and the
 * stream stay the same.
 */
struct sphr_jobs;
struct sphr_jobs *sphr_jobs_new(int n_jobs, const int *n_det, const double *positions, const int *N, const int *l_max, double delta_t, int precision);
void sphr_jobs_free(struct sphr_jobs *);
int sphr_jobs_run(struct sphr_jobs *, const float *in, void *alm, void *map, void *stream);
/* autotuner: times the lag-correlation tiles on these buffers, keeps the
 * fastest; returns the chosen candidate index, its time per run */
int sphr_jobs_tune(struct sphr_jobs *, const float *in, void *alm, double *best_ms);
/* sets the lag tile to one of the tuner's candidates (512/1024/2048
 * samples x 128/256 threads); -1 for any other */
int sphr_jobs_set_lag_tile(struct sphr_jobs *, int samples, int threads);
/* [0] input floats, [1] a_lm elements, [2] map elements, [3] baselines,
 * [4] l_max classes, [5..6] lag tile, [7] plan bytes */
void sphr_jobs_info(const struct sphr_jobs *, long info[8]);
void sphr_jobs_offsets(const struct sphr_jobs *, long *in, long *alm, long *map);

/* Autotuner: runs the batch on these buffers with each candidate shape of
 * the projection (tiles per wave, segments per workgroup, K split) and,
 * while the lag-domain route is in use, with each of that route's kernel
 * forms too, and keeps the fastest (so the route may change to the
 * frequency-domain one; a batch set to the frequency-domain route keeps
 * it), then tries the fused sky kernel's other forms (two workgroups per
 * CU; one segment per workgroup at three per CU; four segments per
 * workgroup, mixed precision);
 * returns the projection shape's index (time per run in best_ms), or -1
 * with the previous settings kept */
int sphr_batch_tune(struct sphr_batch *, const void *tseries, int tseries_is_f32, int n_segments, const double *window, const double *gmst, void *alm, void *map, double *best_ms);
/* Select one of the autotuner's shapes by its index (0..27, as
 * sphr_batch_tune returns it: 7 x K split + shape); 0, or -1 for an index
 * out of range */
int sphr_batch_set_shape(struct sphr_batch *, int index);

/* Shape information: [0] detectors, [1] baselines, [2] N, [3] sky l_max,
 * [4] projection columns, [5] K splits, [6] precision, [7] K per split,
 * [8] 1 when the sky stage runs as the one fused kernel (orders it is
 * compiled for), 0 for the two-kernel form, [9] projection tiles per wave,
 * [10] projection segments per workgroup, [11] fused sky workgroups per
 * CU requested (1-3, picked by the autotuner), [12] fused sky segments per
 * workgroup requested (0 for the precision's default, 1, 2 or 4) */
void sphr_batch_info(const struct sphr_batch *, int info[13]);

/* Select the sky stage's form: 1 (default) the fused kernel where it is
 * compiled for the batch's order, 0 the two-kernel form (a_lm and ring
 * sums, then the phi DFT as a GEMM).  Both give the same results to
 * rounding. */
void sphr_batch_set_fused(struct sphr_batch *, int on);
/* Fused sky kernel's persistent workgroups per CU: 1 (default), 2 (the
 * rotation rows then stay in global memory when both would not fit the
 * CU's LDS) or 3 (one segment per workgroup); fewer where the LDS does
 * not allow as many; same results; -1 for other values */
int sphr_batch_set_sky_occupancy(struct sphr_batch *, int workgroups_per_cu);
/* Fused sky kernel's segments per workgroup: 0 (default: two in mixed
 * precision, one in FP64, one at three workgroups per CU), 1, 2, or 4
 * (mixed precision: the four share one LDS copy of the tables and the
 * rotation rows; one workgroup per CU); a form whose LDS does not fit runs
 * the unfused sky stage; same results; -1 for other values */
int sphr_batch_set_sky_segments(struct sphr_batch *, int segments_per_workgroup);

/* The projection's route.  For real series the frequency-domain product
 * (1/N) sum_f dp[f] X[f] equals a sum over lags of the segments' circular
 * cross-correlation with the inverse DFT of the delay product, which
 * vanishes beyond 2 transient samples.  Where the batch's shape allows it
 * (two or three detectors at 1024 samples, two at 4096 in mixed
 * precision) and its plan passes the checks of that identity (the table
 * vanishes past 2 transient, is real, and has the baseline's mirror
 * symmetry, each to 1e-12 of its largest value), sphr_batch_new selects
 * this lag-domain route: one kernel forms the FFTs, the cross spectra, the
 * cross-correlations and the projection of every segment pair, and the
 * spectra and K splits are not used.  sphr_batch_set_lag(0) selects the
 * frequency-domain route; 1 the lag-domain route, or -1 where it does not
 * apply.  sphr_batch_get_lag: 1 while the lag-domain route is in use.
 * Results agree to rounding. */
int sphr_batch_set_lag(struct sphr_batch *, int on);
int sphr_batch_get_lag(const struct sphr_batch *);
/* Bytes of the projection table the route in use reads per run (the
 * folded delay products, or the lag-domain table) */
size_t sphr_batch_table_bytes(const struct sphr_batch *);
/* The lag-domain kernel's form (picked by the autotuner): 1 (default), a
 * radix-4 Stockham FFT over 256 threads per transform with the projection
 * table held in registers (one workgroup per CU); 2, the same with the
 * table read from L2 (two per CU); 3 (mixed precision at 1024 samples),
 * each transform in one wavefront, two segment pairs per workgroup; 4
 * (mixed precision at 1024 samples), two segment pairs at once, 128
 * threads per transform, the table in registers.  Same results to rounding; -1 for a form not compiled for the
 * batch.  sphr_batch_get_xcorr_form: the form in use, 0 on the
 * frequency-domain route. */
int sphr_batch_set_xcorr_form(struct sphr_batch *, int form);
int sphr_batch_get_xcorr_form(const struct sphr_batch *);

#ifdef __cplusplus
}
#endif

#endif
