/*
 * oracle.h -- FP64 CPU restatement of sphradiometer's time-resolved
 * full-sky cross-power path.
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the checker the GPU library is
 * compared against and the CPU baseline that bench.py times.  Nothing in
 * the shipped library links, loads or calls it.
 *
 * Every function follows one reference function (cited file:line,
 * relative to the sphradiometer source tree) in algorithm and loop order.
 * Third-party numerics the reference takes from GSL and FFTW are restated
 * from their published definitions:
 *   - Gauss-Legendre nodes/weights (gsl_integration_glfixed): Newton
 *     iteration on P_n with the Tricomi start, ordered like glfixed_point.
 *   - gsl_sf_legendre_sphPlm(_array): orthonormal associated Legendre
 *     functions with the Condon-Shortley phase, standard three-term
 *     recurrence in l from P_mm.
 *   - gsl_sf_coupling_3j: Racah's closed form, summed in binary128.
 *   - FFTW c2c / r2c / c2r: unnormalised DFTs, mixed-radix recursive
 *     Cooley-Tukey with exactly reduced twiddle angles; Bluestein for
 *     lengths with a prime factor above 64.
 *   - GSL BFGS2 (extremize): not restated; the peak refiner is judged by
 *     the peak it finds, not by its iterates.
 *
 * Pins (tests/test_oracle_*.py): closed-form Y_lm table, numpy leggauss,
 * sympy exact 3j, Y_lm projection and mesh round trip, rotated impulse,
 * pointwise product at random points, microcode == pixel product, the
 * analytic white-noise cross-power of the correlator accuracy test, and
 * brute-force DFTs.
 */
#ifndef SPHR_ORACLE_H
#define SPHR_ORACLE_H

#include <complex.h>
#include <stddef.h>

/* Same field meaning and coefficient order as struct sh_series
 * (sh_series.h:86-90): m = 0 block, then m = 1..L, then m = -L..-1. */
struct or_series {
	unsigned l_max;
	int polar;
	double complex *coeff;
};

/* struct sh_series_array (sh_series.h:135-142) */
struct or_array {
	unsigned l_max;
	int polar;
	int n;
	int stride;
	double complex *coeff;
};

/* struct sh_series_rotation_plan (sh_series.h:110-113): D[l] points at the
 * centre of a (2l+1)^2 block, element (m, m') at D[l][(2l+1) m' + m]. */
struct or_rotation_plan {
	unsigned l_max;
	double complex **D;
};

struct or_product_op {
	double factor;
	int dest_offset, a_offset, b_offset;
};

struct or_product_plan {
	unsigned a_l_max, b_l_max, dest_l_max;
	int a_polar, b_polar, dest_polar;
	int plan_length;
	struct or_product_op *microcode;
};

/* struct correlator_plan_fd (correlator.h:111-119), minus the baseline
 * pointer; geometry is kept inline. */
struct or_plan_fd {
	double d[3], theta, phi, delta_t;
	int transient;
	struct or_array *delay_product;
	double complex *fseries_product;
	struct or_series *power_1d;
	struct or_rotation_plan *rotation_plan;
};

struct or_network {
	int n_instruments;
	int n_baselines;
	int *index_a, *index_b;
	struct or_plan_fd **plans;
};

/* offsets: sh_series.h:172-227 */
size_t or_moffset(unsigned l_max, int m);
size_t or_lmoffset(unsigned l_max, unsigned l, int m);
size_t or_length(unsigned l_max, int polar);

/* series basics: sh_series.c */
struct or_series *or_series_new(unsigned l_max, int polar);
struct or_series *or_series_new_zero(unsigned l_max, int polar);
struct or_series *or_series_copy(const struct or_series *);
void or_series_free(struct or_series *);
struct or_series *or_series_zero(struct or_series *);
struct or_series *or_series_assign(struct or_series *dst, const struct or_series *src);
struct or_series *or_series_resize(struct or_series *, unsigned l_max);
struct or_series *or_series_set_polar(struct or_series *, int polar);
double complex or_get(const struct or_series *, unsigned l, int m);
void or_set(struct or_series *, unsigned l, int m, double complex v);
struct or_series *or_series_add(struct or_series *a, double complex z, const struct or_series *b);
struct or_series *or_series_scale(struct or_series *a, double complex z);
struct or_series *or_series_conj(struct or_series *a);
struct or_series *or_series_real(struct or_series *a);
struct or_series *or_series_scale_power_l(struct or_series *a, double power, double overall);
double or_rms_error(const struct or_series *measured, const struct or_series *exact);

/* arrays: sh_series_array.c */
struct or_array *or_array_new(int n, unsigned l_max, int polar);
void or_array_free(struct or_array *);
struct or_array *or_array_set_len(struct or_array *, int n);
struct or_array *or_array_set_l(struct or_array *, unsigned l_max);
struct or_array *or_array_scale(struct or_array *, double complex z);
struct or_series *or_array_dotc(struct or_series *result, const struct or_array *, const double complex *v);
struct or_array *or_array_forward_fft(struct or_array *);
void or_array_series(const struct or_array *, int i, struct or_series *view);

/* numerics restated from GSL / FFTW */
void or_gl_nodes(int n, double *x, double *w);           /* ascending x */
void or_sphPlm_array(unsigned l_max, int m, double x, double *out); /* out[l-m] */
double or_sphPlm(unsigned l, int m, double x);
double or_wigner_3j(int ja, int jb, int jc, int ma, int mb, int mc);
void or_dft(double complex *data, int n, int sign);    /* in place, sign=-1 forward */

/* transforms: sh_series_trans.c */
void or_mesh_nodes(unsigned l_max, int *ntheta, int *nphi, double *cos_theta, double *weights);
double complex or_Y(unsigned l, int m, double theta, double phi);
double complex or_eval(const struct or_series *, double theta, double phi);
/* time-domain correlator (or_td.c) */
struct or_plan_td;
struct or_network_td {
	int n_instruments, n_baselines;
	int *index_a, *index_b;
	struct or_plan_td **plans;
};
struct or_plan_td *or_plan_td_new(const double ra[3], const double rb[3], double delta_t);
void or_plan_td_free(struct or_plan_td *);
struct or_series *or_baseline_integrate_power_td(const double *ta, const double *tb, const double *window, int n, struct or_plan_td *);
struct or_network_td *or_network_td_new(int n_instruments, const double *pos, double delta_t);
void or_network_td_free(struct or_network_td *);
struct or_series *or_network_integrate_power_td(struct or_series *sky, double **ts, int n, double **windows, struct or_network_td *);
int or_plan_td_transient(const struct or_plan_td *);
unsigned or_plan_td_l_max(const struct or_plan_td *);
const double complex *or_plan_td_power_1d(const struct or_plan_td *);

/* HEALPix output (or_healpix.c) */
int or_healpix_nside_c(unsigned l_max);
int or_healpix_nside_py(unsigned l_max);
void or_healpix_pix2ang(int nside, long pix, double *theta, double *phi);
void or_healpix_map(const struct or_series *, int nside, double *map);
void or_healpix_posterior(const double *map_p, const double *map_n, long npix, double *prob);
struct or_series *or_from_mesh(struct or_series *, const double complex *mesh);
struct or_series *or_from_realmesh(struct or_series *, const double *mesh);
double complex *or_to_mesh(const struct or_series *, int *ntheta, int *nphi);
double complex *or_to_mesh_threads(const struct or_series *, int *ntheta, int *nphi, int threads);
extern int or_mesh_threads;
struct or_series *or_from_realfunc(struct or_series *, double (*f)(double, double, void *), void *data);
struct or_series *or_from_func(struct or_series *, double complex (*f)(double, double, void *), void *data);
struct or_series *or_impulse(unsigned l_max, double theta, double phi);

/* rotation: sh_series_rotation.c */
void or_euler_rotation_matrix(double gamma, double beta, double alpha, double R[9]);
void or_rot_matrix(double theta, double phi, double R[9]);
void or_invrot_matrix(double theta, double phi, double R[9]);
struct or_rotation_plan *or_rotation_plan_new(unsigned l_max, const double R[9]);
void or_rotation_plan_free(struct or_rotation_plan *);
struct or_rotation_plan *or_rotation_plan_set_l(struct or_rotation_plan *, unsigned l_max);
double complex or_wigner_D(const struct or_rotation_plan *, unsigned l, int m, int mp);
struct or_series *or_rotate(struct or_series *result, const struct or_series *s, const struct or_rotation_plan *);
struct or_series *or_rotate_z(struct or_series *result, const struct or_series *s, double phi);

/* extremizer and derivative: sh_series_extremize.c, sh_series_deriv.c */
double or_find_near_minimum(const struct or_series *, unsigned l_lowpass, double *theta, double *phi);
double or_real_minimum(const struct or_series *, double *theta, double *phi);
double or_refine_minimum(const struct or_series *, double th, double ph, double size, double *theta, double *phi);
struct or_series *or_sintheta_d_by_dtheta(const struct or_series *);

/* localization: inspiral_sky_map.c, RapidLocalization.py */
void or_det_response(const double resp[9], double ra, double dec, double psi, double gmst, double *fp, double *fx);
double complex or_cbc_projection(double theta, double phi, int i, int j, const double *resp, int n, double beta, double psi, const double *psds);
int or_network_mult_by_projection(struct or_network *, const double *resp, double beta, double psi, const double *psd);
struct or_series *or_log_uniformsky_prior(unsigned l_max);
double or_tukey(int n, int i, double beta);
long or_precalc_length(long length, double delta_t);
void or_preprocess_snr(const double complex *series, int n_series, int len, double delta_t, double complex *out);
struct or_series *or_generate_alm_sky(struct or_series *sky, struct or_network *, double complex **fseries, double gmst, const struct or_series *logprior);

/* products: sh_series_product.c */
struct or_product_plan *or_product_plan_new(const struct or_series *dest, const struct or_series *a, const struct or_series *b);
struct or_product_plan *or_product_plan_new_forced(const struct or_series *dest, const struct or_series *a, const struct or_series *b, int microcode);
void or_product_plan_free(struct or_product_plan *);
struct or_series *or_product(struct or_series *dest, const struct or_series *a, const struct or_series *b, const struct or_product_plan *);

/* geometry, projection and correlator: instrument.c, projection.c, correlator.c */
double or_vector_r_dot_s(const double v[3], double theta, double phi);
void or_vector_direction(const double v[3], double *theta, double *phi);
unsigned or_power_l_max_naive(double D, double delta_t);
unsigned or_projection_l_max(double r, double delta_t);
int or_projection_n_elements(double r, double delta_t);
struct or_array *or_projection_matrix_delay(unsigned n, unsigned l_max, const double r[3], double delta_t);
struct or_plan_fd *or_plan_fd_new(const double ra[3], const double rb[3], int n, double delta_t, int target_l);
void or_plan_fd_free(struct or_plan_fd *);
struct or_plan_fd *or_plan_fd_set_l(struct or_plan_fd *, unsigned l_max);
struct or_series *or_baseline_integrate_power_fd(const double complex *fa, const double complex *fb, struct or_plan_fd *);
struct or_network *or_network_new(int n_instruments, const double *positions, int n, double delta_t, int target_l);
void or_network_free(struct or_network *);
struct or_network *or_network_set_l(struct or_network *, unsigned l_max);
struct or_series *or_network_integrate_power_fd(struct or_series *sky, double complex **fseries, struct or_network *);
void or_tseries_to_fseries(const double *tseries, double complex *fseries, int n);
void or_ctseries_to_fseries(const double complex *tseries, double complex *fseries, int n);

#endif
