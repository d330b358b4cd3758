/*
 * The gradient operators the extremizer uses (sphradiometer
 * src/lib/sh_series_deriv.c:241-543) and sh_series_real_minimum
 * (sh_series_extremize.c:273-360), whose search runs on the GPU.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sh_series.h>
#include "../sphr_internal.h"

/* sh_series_deriv.c:241-283: d/dphi multiplies a_lm by i m, in place */
struct sh_series *sh_series_d_by_dphi(struct sh_series *s)
{
	const int mm = s->polar ? 0 : (int) s->l_max;
	for(int m = -mm; m <= mm; m++) {
		double complex *c = s->coeff + sh_series_moffset(s->l_max, m);
		for(unsigned i = 0; i + abs(m) <= s->l_max; i++)
			c[i] *= I * m;
	}
	return s;
}

/*
 * sin(theta) d/dtheta as an operator on coefficients
 * (sh_series_deriv.c:288-~520), result at order l_max + 1.  From the
 * recurrences of the normalised Legendre functions,
 *   sin(theta) dY_lm/dtheta = l c_{l+1,m} Y_{l+1,m} - (l+1) c_{l,m} Y_{l-1,m},
 *   c_{l,m} = sqrt((l^2 - m^2) / ((2l - 1)(2l + 1))),
 * so the coefficient of Y_{l,m} in the result is
 *   (l-1) c_{l,m} a_{l-1,m} - (l+2) c_{l+1,m} a_{l+1,m}.
 * (The reference evaluates the same operator through Wigner 3j symbols.)
 */
static double cfac(int l, int m)
{
	if(l <= 0 || abs(m) >= l)
		return 0.0;
	return sqrt(((double) l * l - (double) m * m) / ((2.0 * l - 1.0) * (2.0 * l + 1.0)));
}

struct sh_series *sh_series_sintheta_d_by_dtheta(const struct sh_series *s)
{
	const unsigned L = s->l_max, Lo = L + 1;
	struct sh_series *r = sh_series_new_zero(Lo, s->polar);
	if(!r)
		return NULL;
	const int mm = s->polar ? 0 : (int) L;
	for(int m = -mm; m <= mm; m++)
		for(int l = abs(m); l <= (int) L; l++) {
			const double complex a = s->coeff[sh_series_lmoffset(s, l, m)];
			/* a_lm contributes to (l+1, m) and (l-1, m) */
			r->coeff[sh_series_params_lmoffset(Lo, l + 1, m)] += l * cfac(l + 1, m) * a;
			if(l - 1 >= abs(m))
				r->coeff[sh_series_params_lmoffset(Lo, l - 1, m)] -= (l + 1) * cfac(l, m) * a;
		}
	return r;
}

/* sh_series_deriv.c:531-543: u = sin(theta) d/dtheta, v = d/dphi */
const struct sh_series *sh_series_sintheta_grad(const struct sh_series *s, struct sh_series **u, struct sh_series **v)
{
	*u = sh_series_sintheta_d_by_dtheta(s);
	*v = sh_series_copy(s);
	if(!*u || !*v) {
		sh_series_free(*u);
		sh_series_free(*v);
		*u = *v = NULL;
		return NULL;
	}
	sh_series_d_by_dphi(*v);
	return s;
}

/*
 * sh_series_extremize.c:273-360: value and co-ordinates of the minimum of
 * Re f.  Seed from the l = 6 low-pass mesh as the reference does; the
 * refinement (GPU, Newton in the tangent plane) replaces GSL's BFGS2.
 * The returned co-ordinates are in [0, pi] x [0, 2 pi).  NaN on failure.
 */
double sh_series_real_minimum(const struct sh_series *series, double *theta, double *phi)
{
	struct sh_series *full = NULL;
	const struct sh_series *s = series;
	if(series->polar) {
		full = sh_series_copy(series);
		if(!full || !sh_series_set_polar(full, 0)) {
			sh_series_free(full);
			return NAN;
		}
		s = full;
	}
	double t = NAN, p = NAN, v = NAN;
	const int rc = sphr_gpu_extremize(s->l_max, (const double *) s->coeff, 0, &t, &p, &v);
	sh_series_free(full);
	if(rc)
		return NAN;
	*theta = t;
	*phi = p;
	return v;
}
