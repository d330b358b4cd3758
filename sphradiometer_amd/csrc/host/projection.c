/*
 * SH expansion of the sinc delay kernel (sphradiometer src/lib/
 * projection.c).  Plan-time, host side.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/instrument.h>
#include <sphradiometer/projection.h>
#include "../sphr_internal.h"

/* projection.c:56-66: 60-sample guard band; always odd */
int projection_matrix_n_elements(double r, double delta_t)
{
	const int guard = 60;
	return 2 * (int) (guard + r / delta_t + 1) + 1;
}

/* projection.c:75-87 */
unsigned int projection_matrix_l_max(double r, double delta_t)
{
	const int l = M_PI * r / delta_t + 1;
	return (unsigned) l;
}

/* projection.c:99-107: sinc(j - k - r.s / delta_t) */
double projection_delay_element(double theta, double phi, void *vd)
{
	const struct projection_delay_element_data *d = vd;
	const double px = M_PI * (d->j - d->k - vector_r_dot_s(d->r, theta, phi) / d->delta_t);
	return px == 0.0 ? 1.0 : sin(px) / px;
}

/*
 * projection.c:110-131.  Tap i (k = i - (n-1)/2) is the expansion of
 * sinc(-k - r.s/dt) at order l_max.  For a phase centre on the z axis the
 * kernel does not depend on phi and the series is polar: the analysis of
 * sh_series_from_realfunc then reduces to one Gauss-Legendre sum per l,
 *   a_l = sum_j (2 pi w_j) f(x_j) Pbar_l0(x_j),
 * which is evaluated here as a small matrix product over all taps.
 */
struct sh_series_array *projection_matrix_delay(unsigned int n, unsigned int l_max, const gsl_vector *r, double delta_t)
{
	const int polar = gsl_vector_get(r, 0) == 0.0 && gsl_vector_get(r, 1) == 0.0;
	struct sh_series_array *m = sh_series_array_new((int) n, l_max, polar);
	if(!m)
		return NULL;
	struct projection_delay_element_data dd = { .r = r, .delta_t = delta_t, .j = 0, .k = 0 };
	if(!polar) {
		for(unsigned i = 0; i < n; i++) {
			dd.k = -((int) n - 1) / 2 + (int) i;
			if(!sh_series_from_realfunc(&m->series[i], projection_delay_element, &dd)) {
				sh_series_array_free(m);
				return NULL;
			}
		}
		return m;
	}
	const int ng = 2 * ((int) l_max + 1);
	const double *x, *w;
	sphr_mesh_rule(l_max, &x, &w);
	double *P = malloc((size_t) ng * (l_max + 1) * sizeof(*P));
	double *F = malloc((size_t) n * ng * sizeof(*F));
	if(!P || !F) {
		free(P);
		free(F);
		sh_series_array_free(m);
		return NULL;
	}
	for(int j = 0; j < ng; j++)
		sphr_plm(l_max, 0, x[j], P + (size_t) j * (l_max + 1));
	for(unsigned i = 0; i < n; i++) {
		dd.k = -((int) n - 1) / 2 + (int) i;
		for(int j = 0; j < ng; j++)
			F[(size_t) i * ng + j] = 2.0 * M_PI * w[j] * projection_delay_element(acos(x[j]), 0.0, &dd);
	}
	#pragma omp parallel for schedule(static)
	for(int i = 0; i < (int) n; i++) {
		double complex *c = m->coeff + (size_t) i * m->stride;
		for(unsigned l = 0; l <= l_max; l++) {
			double acc = 0.0;
			for(int j = 0; j < ng; j++)
				acc += F[(size_t) i * ng + j] * P[(size_t) j * (l_max + 1) + l];
			c[l] = acc;
		}
	}
	free(P);
	free(F);
	return m;
}
