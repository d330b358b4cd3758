/*
 * Interface between the host C library (object management and plan-time
 * numerics) and the HIP side (kernels and device state).  Plain C types
 * only; complex arrays are interleaved (re, im) doubles.
 */
#ifndef SPHR_INTERNAL_H
#define SPHR_INTERNAL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- host numerics (numerics.c) ---- */

/* n-point Gauss-Legendre rule on [-1, 1], nodes ascending */
void sphr_gl_rule(int n, double *x, double *w);
/* the rule of the (2(l_max+1))^2 pixel mesh: cos(theta) descending (theta
 * ascending), cached per l_max; returned arrays are owned by the cache */
void sphr_mesh_rule(unsigned l_max, const double **cos_theta, const double **weights);
void sphr_rules_prefetch(const int *mesh_orders, int n_mesh, const int *gl_counts, int n_gl);
/* the n-point Gauss-Legendre rule of sphr_gl_rule, computed once per n
 * and kept (thread-safe); the arrays belong to the cache */
void sphr_gl_rule_cached(int n, const double **x, const double **w);
/* orthonormal associated Legendre functions (Condon-Shortley phase),
 * out[l - m] for l = m .. l_max */
void sphr_plm(unsigned l_max, int m, double x, double *out);
/* tables of the GPU time-domain correlator (td_tables.c) */
struct sh_series_array;
int sphr_td_tables(const struct sh_series_array *pa, const struct sh_series_array *pb, int nq, int P, double *Ta, double *Tb, double *WP);
/* HEALPix map FITS file (healpix_fits.c) */
int sphr_fits_write_healpix_map(const char *filename, const double *map, int nside);
/* all 3j (j1 j2 j; m1 m2 -m1-m2) for j = jmin .. j1+j2 into out[j - jmin];
 * returns jmin, or -1 when every symbol vanishes */
int sphr_3j_range(int j1, int j2, int m1, int m2, double *out);
/* in-place unnormalised complex DFT, sign = -1 forward, +1 backward */
void sphr_host_dft(double *data, int n, int sign);

/* LAL's Tukey window of n samples, taper fraction beta (localize.c) */
void sphr_tukey_window(int n, double beta, double *w);

/* packed Wigner-D blocks of a rotation plan (l = 1 first, each
 * (2l+1)^2 in storage order); NULL when l_max = 0 */
struct sh_series_rotation_plan;
const double *sphr_rotation_blocks(const struct sh_series_rotation_plan *);

/* a rotation plan of order l_max with its D^l blocks allocated (contents
 * undefined), as sh_series_rotation_plan_free releases them */
struct sh_series_rotation_plan *sphr_rotation_plan_alloc(unsigned l_max);

/* ---- GPU side (runtime.hip, dropin.hip) ---- */

/* 0 on success; on failure -1 with a message for sphr_last_error() */
int sphr_gpu_ensure(void);
void sphr_set_error(const char *fmt, ...);

/* result[c] = sum_i coeff[i * stride + c] * vec[i] over n rows */
int sphr_gpu_dotc(const double *coeff, int n, int stride, const double *vec, int vec_is_real, double *result);

/* batched localization tables (host-built, localize.c) */
struct sphr_loc;
struct sphr_loc_desc {
	int precision, D, nb, N, len, L, max_events;
	int ia[16], ib[16];
	const double *W;	/* complex [2][nb][N][(L+1)^2]: sky per unit X */
	const double *prior;	/* complex [(L+1)^2], scale_power_l applied */
	const double *dft;	/* complex [len][N]: preprocessing + c2c DFT */
};
struct sphr_loc *sphr_loc_create(const struct sphr_loc_desc *);

/* complex float64 C[M][N] = A[M][K] B[K][N] (row-major host arrays) on
 * the MFMA GEMM; plan-time use */
int sphr_gpu_zgemm(int M, int N, int K, const double *A, const double *B, double *C);

/* synthesis of one series onto the 2(l+1) x 2(l+1) mesh (row-major
 * [theta][phi], complex) */
int sphr_gpu_to_mesh(unsigned l_max, int polar, const double *coeff, double *mesh);

/* rotation.  For a polar input only D^l_{0 m'} is used.  blocks: packed
 * (2l+1)^2 blocks for l = 1..l_max in the reference storage order. */
int sphr_gpu_rotate(unsigned l_max, int polar, const double *coeff, const double *blocks, double *result);
int sphr_gpu_rotate_z(unsigned l_max, const double *coeff, double phi, double *result);

/* network frequency-domain integrate.
 * n_inst detector spectra fseries[i] (length n each, full complex);
 * per baseline k: index_a/b, delay product dp[k] (n x stride_k complex,
 * l_max_k, polar_k), rotation blocks blocks[k] (packed as above; polar
 * baselines may pass only the m = 0 rows, see sphr_gpu_network_fd);
 * outputs: power_1d[k] (stride_k complex), fsp[k] (n complex, the
 * fseries_product scratch of the reference plan, may be NULL) and
 * sky (non-polar, sky_l) = (1/nb) sum_k rotate(power_1d[k]). */
struct sphr_fd_baseline {
	const void *key;           /* the plan: its device copy is kept across calls */
	int index_a, index_b;
	unsigned l_max;
	int polar;
	int stride;
	const double *dp;          /* n * stride complex */
	const double *blocks;      /* packed D blocks l = 1..l_max */
	double *power_1d;          /* out: stride complex */
	double *fsp;               /* out: n complex, optional */
};
int sphr_gpu_network_fd(int n_inst, const double *const *fseries, int n, int nb, const struct sphr_fd_baseline *bl, unsigned sky_l, double scale, double *sky);
/* drops the device copy of a plan's delay product and rotation table
 * (called by every function that changes or frees them) */
void sphr_gpu_plan_forget(const void *key);

/* a network FD plan flattened for the batched device pipeline */
#define SPHR_DESC_MAX 16
struct sphr_net_desc {
	int n_det, nb, N;
	int ia[SPHR_DESC_MAX], ib[SPHR_DESC_MAX], Lb[SPHR_DESC_MAX], polar[SPHR_DESC_MAX];
	int transient[SPHR_DESC_MAX];           /* the plan's transient: taps 2 t + 1 */
	const double *dp[SPHR_DESC_MAX];       /* N x stride complex */
	const double *blocks[SPHR_DESC_MAX];   /* packed D blocks */
};
struct sphr_batch;
struct sphr_batch *sphr_batch_create(const struct sphr_net_desc *d, unsigned sky_l, int precision, int max_segments);

/* extremum of Re f (find_max = 0: minimum) of a non-polar series of
 * order l_max, seeded like sh_series_real_minimum */
int sphr_gpu_extremize(unsigned l_max, const double *coeff, int find_max, double *theta, double *phi, double *value);
/* device copy of the seed mesh's ring nodes */
const double *sphr_gpu_seed_nodes(void);

/* FFT helpers of correlator.h */
struct sphr_fft;
struct sphr_fft *sphr_fft_new(int n, int real_input, double *in, double *out);
int sphr_fft_exec(struct sphr_fft *);
void sphr_fft_free(struct sphr_fft *);

#ifdef __cplusplus
}
#endif

#endif
