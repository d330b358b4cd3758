/*
 * sh_series.h -- spherical harmonic series objects of the MI355X
 * cross-power library.  Drop-in for sphradiometer's
 * src/include/sphradiometer/sh_series.h: the same struct layouts, inline
 * offset helpers and prototypes, so code compiled against the reference
 * header links against this library unchanged.
 *
 * Coefficient order (sh_series.h:53-83 of the reference): for each m the
 * coefficients with l = |m| .. l_max are contiguous; the m blocks run
 * m = 0, 1, ..., l_max, then m = -l_max, ..., -1.  A "polar" series holds
 * only the m = 0 block.
 *
 * Functions on the hot path run on the GPU: sh_series_array_dotc,
 * sh_series_rotate, sh_series_rotate_z and sh_series_to_mesh.  They are
 * synchronous (results are in host memory on return).  The rest are
 * host-side object management and plan-time numerics.
 */
#ifndef __RADIOMETER_SH_SERIES_H__
#define __RADIOMETER_SH_SERIES_H__

#include <complex.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#define complex _Complex
#endif

/* sh_series.h:86-90 */
struct sh_series {
	unsigned int l_max;
	int polar;
	double complex *coeff;
};

/* sh_series.h:99-109 */
struct sh_series_product_plan {
	unsigned int a_l_max, b_l_max, dest_l_max;
	int a_polar, b_polar, dest_polar;
	int plan_length;
	struct _sh_series_product_plan_op {
		double factor;
		int dest_offset;
		int a_offset;
		int b_offset;
	} *microcode;
};

/* sh_series.h:110-113: D[l] (l = 1 .. l_max) points at the centre element
 * of a (2l+1) x (2l+1) block; element (m, m') is D[l][(2l+1) m' + m]. */
struct sh_series_rotation_plan {
	unsigned int l_max;
	double complex **D;
};

/* sh_series.h:135-142 */
struct sh_series_array {
	unsigned int l_max;
	int polar;
	int n;
	struct sh_series *series;
	double complex *coeff;
	int stride;
};

struct sh_series_eval_interp;

/* Offset helpers; identical arithmetic to sh_series.h:172-227. */
#ifndef SWIG
static size_t sh_series_moffset(unsigned int, int) __attribute__ ((unused));
#endif
static size_t sh_series_moffset(unsigned int l_max, int m)
{
	/* the product is formed before the halving so that the integer
	 * division is exact */
	const int x = m * (m - 1);
	l_max += 1;
	if(m >= 0)
		return l_max * m - x / 2;
	return l_max * (l_max + m) + x / 2;
}

#ifndef SWIG
static size_t sh_series_params_lmoffset(unsigned int, unsigned int, int) __attribute__ ((unused));
#endif
static size_t sh_series_params_lmoffset(unsigned int l_max, unsigned int l, int m)
{
	return sh_series_moffset(l_max, m) + l - abs(m);
}

#ifndef SWIG
static size_t sh_series_lmoffset(const struct sh_series *, unsigned int, int) __attribute__ ((unused));
#endif
static size_t sh_series_lmoffset(const struct sh_series *series, unsigned int l, int m)
{
	return sh_series_params_lmoffset(series->l_max, l, m);
}

#ifndef SWIG
static size_t sh_series_length(unsigned int, int) __attribute__ ((unused));
#endif
static size_t sh_series_length(unsigned int l_max, int polar)
{
	l_max += 1;
	if(polar)
		return l_max;
	return l_max * l_max;
}

double sh_series_wigner_3j(int, int, int, int, int, int);

double complex sh_series_Y(unsigned int, int, double, double);
double complex sh_series_Yconj(unsigned int, int, double, double);
double complex *sh_series_Y_array(double complex *, unsigned int, int, double, double);
double complex *sh_series_Yconj_array(double complex *, unsigned int, int, double, double);

struct sh_series *sh_series_new(unsigned int, int);
struct sh_series *sh_series_new_zero(unsigned int, int);
struct sh_series *sh_series_copy(const struct sh_series *);
struct sh_series *sh_series_assign(struct sh_series *, const struct sh_series *);
struct sh_series *sh_series_resize(struct sh_series *, unsigned int);
struct sh_series *sh_series_zero(struct sh_series *);
struct sh_series *sh_series_set_polar(struct sh_series *, int);
void sh_series_free(struct sh_series *);

double complex sh_series_get(const struct sh_series *, unsigned int, int);
double complex sh_series_set(struct sh_series *, unsigned int, int, double complex);

double complex sh_series_eval(const struct sh_series *, double, double);

complex double *sh_series_mesh_new(unsigned int, int *, int *, double **, double **);
double *sh_series_real_mesh_new(unsigned int, int *, int *, double **, double **);
struct sh_series *sh_series_from_mesh(struct sh_series *, double complex *);
struct sh_series *sh_series_from_realmesh(struct sh_series *, double *);
complex double *sh_series_to_mesh(const struct sh_series *, int *, int *, double **);

struct sh_series *sh_series_from_func(struct sh_series *, double complex (*)(double, double, void *), void *);
struct sh_series *sh_series_from_realfunc(struct sh_series *, double (*)(double, double, void *), void *);

struct sh_series *sh_series_impulse(unsigned int, double, double);

struct sh_series *sh_series_add(struct sh_series *, double complex, const struct sh_series *);
struct sh_series *sh_series_scale(struct sh_series *, double complex);
struct sh_series *sh_series_neg(struct sh_series *);
struct sh_series *sh_series_conj(struct sh_series *);
struct sh_series *sh_series_real(struct sh_series *);

/*
 * HEALPix FITS output (sh_series_healpix.cc:131-286, declared at
 * sh_series.h:380-382 of the reference): the a_lm table of the real part
 * of the series, and the RING-ordered map of Re f at nside =
 * ceilpow2(ceil(l_max / 2 + 0.75)), computed on the GPU.  The files are
 * written directly in the healpix binary-table layout (no healpix_cxx).
 * Return 0 / the series on success, -1 / NULL on failure.
 */
int sh_series_write_healpix_map(const struct sh_series *, const char *);
int sh_series_write_healpix_alm(const struct sh_series *, const char *);
struct sh_series *sh_series_read_healpix_alm(const char *);
struct sh_series *sh_series_imag(struct sh_series *);
struct sh_series *sh_series_clip(struct sh_series *, double);
struct sh_series *sh_series_scale_power_l(struct sh_series *, double, double);
double complex sh_series_dot(const struct sh_series *, const struct sh_series *);

struct sh_series_product_plan *sh_series_product_plan_new(const struct sh_series *, const struct sh_series *, const struct sh_series *);
void sh_series_product_plan_free(struct sh_series_product_plan *);
struct sh_series *sh_series_product(struct sh_series *, const struct sh_series *, const struct sh_series *, const struct sh_series_product_plan *);
int sh_series_is_nan(const struct sh_series *);
int sh_series_is_inf(const struct sh_series *);

double *euler_rotation_matrix(double, double, double);
double *euler_inv_rotation_matrix(double, double, double);
double *sh_series_rot_matrix(double, double);
double *sh_series_invrot_matrix(double, double);
struct sh_series_rotation_plan *sh_series_rotation_plan_new(const struct sh_series *, const double *);
void sh_series_rotation_plan_free(struct sh_series_rotation_plan *);
struct sh_series_rotation_plan *sh_series_rotation_plan_copy(const struct sh_series_rotation_plan *);
struct sh_series_rotation_plan *sh_series_rotation_plan_set_l(struct sh_series_rotation_plan *, unsigned);
double complex sh_series_rotation_plan_wigner_D(const struct sh_series_rotation_plan *, unsigned, int, int);
struct sh_series *sh_series_rotate(struct sh_series *, const struct sh_series *, const struct sh_series_rotation_plan *);
struct sh_series *sh_series_rotate_z(struct sh_series *, const struct sh_series *, double);

/* derivatives used by the extremizer (sh_series_deriv.c:241-543) */
struct sh_series *sh_series_d_by_dphi(struct sh_series *);
struct sh_series *sh_series_sintheta_d_by_dtheta(const struct sh_series *);
const struct sh_series *sh_series_sintheta_grad(const struct sh_series *, struct sh_series **, struct sh_series **);

/* minimum of Re f: value returned, co-ordinates stored (GPU search);
 * sh_series_extremize.c:273-360 */
double sh_series_real_minimum(const struct sh_series *, double *, double *);

struct sh_series_array *sh_series_array_new(int, unsigned int, int);
void sh_series_array_free(struct sh_series_array *);
struct sh_series_array *sh_series_array_set_len(struct sh_series_array *, int);
struct sh_series_array *sh_series_array_set_polar(struct sh_series_array *, int);
struct sh_series_array *sh_series_array_set_l(struct sh_series_array *, unsigned int);
struct sh_series_array *sh_series_array_copy(const struct sh_series_array *);
struct sh_series_array *sh_series_array_assign(struct sh_series_array *, const struct sh_series_array *);
struct sh_series_array *sh_series_array_scale(struct sh_series_array *, double complex);

struct sh_series *sh_series_array_dot(struct sh_series *, const struct sh_series_array *, const double *);
struct sh_series *sh_series_array_dotc(struct sh_series *, const struct sh_series_array *, const double complex *);

struct sh_series_array *sh_series_array_window(struct sh_series_array *, const double *);
struct sh_series_array *sh_series_array_windowc(struct sh_series_array *, const double complex *);

struct sh_series_array *sh_series_array_forward_fft(struct sh_series_array *);
struct sh_series_array *sh_series_array_reverse_fft(struct sh_series_array *);

#ifdef __cplusplus
#undef complex
}
#endif

#endif
