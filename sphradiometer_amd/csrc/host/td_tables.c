/*
 * Tables of the GPU time-domain correlator (td.hip): the delay-tap
 * projections synthesised at the nodes of a Gauss-Legendre rule, and the
 * weighted Legendre functions that project a product back to polar
 * coefficients.  Host code (OpenMP over the taps), run once per plan.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <sphradiometer/sh_series.h>
#include "../sphr_internal.h"

/* Pbar_l0(x), l = 0..lmax */
static void pbar_l0(double x, int lmax, double *row)
{
	double p0 = 0.28209479177387814347, p1 = 0.0, am1 = 1.0;
	row[0] = p0;
	for(int l = 1; l <= lmax; l++) {
		const double al = sqrt((4.0 * l * l - 1.0) / ((double) l * l));
		const double pn = l == 1 ? al * x * p0 : al * (x * p0 - p1 / am1);
		p1 = p0;
		p0 = pn;
		am1 = al;
		row[l] = pn;
	}
}

/*
 * nq-node rule (x_q, w_q); Ta[k][q] = sum_l Re pa[k][l] Pbar_l0(x_q),
 * Tb likewise; WP[q][d] = 2 pi w_q Pbar_d0(x_q), d <= P.  The delay
 * projections are real functions of theta, so their coefficients are
 * real.  Returns 0, or -1 when out of memory.
 */
int sphr_td_tables(const struct sh_series_array *pa, const struct sh_series_array *pb, int nq, int P, double *Ta, double *Tb, double *WP)
{
	const int np = pa->n, la = (int) pa->l_max, lb = (int) pb->l_max;
	int lq = la > lb ? la : lb;
	lq = lq > P ? lq : P;
	const double *x, *w;
	sphr_gl_rule_cached(nq, &x, &w);
	double *row = malloc((size_t) nq * (lq + 1) * sizeof(double));
	if(!row)
		return -1;
	#pragma omp parallel for schedule(static)
	for(int q = 0; q < nq; q++)
		pbar_l0(x[q], lq, row + (size_t) q * (lq + 1));
	#pragma omp parallel for schedule(static)
	for(int k = 0; k < np; k++)
		for(int q = 0; q < nq; q++) {
			const double *r = row + (size_t) q * (lq + 1);
			double sa = 0.0, sb = 0.0;
			for(int l = 0; l <= la; l++)
				sa += creal(pa->coeff[(size_t) k * pa->stride + l]) * r[l];
			for(int l = 0; l <= lb; l++)
				sb += creal(pb->coeff[(size_t) k * pb->stride + l]) * r[l];
			Ta[(size_t) k * nq + q] = sa;
			Tb[(size_t) k * nq + q] = sb;
		}
	for(int q = 0; q < nq; q++)
		for(int d = 0; d <= P; d++)
			WP[(size_t) q * (P + 1) + d] = 2. * M_PI * w[q] * row[(size_t) q * (lq + 1) + d];
	free(row);
	return 0;
}
