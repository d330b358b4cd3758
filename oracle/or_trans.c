/*
 * Oracle (test infrastructure only): spherical harmonic evaluation and
 * transforms, restating sh_series_trans.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* pixels_from_l_max (sh_series_trans.c:149-196): 2(l_max+1) GL nodes in
 * cos(theta), reversed so theta ascends; nphi = 2(l_max+1). */
void or_mesh_nodes(unsigned l_max, int *ntheta, int *nphi, double *cos_theta, double *weights)
{
	const int n = 2 * ((int) l_max + 1);
	double *x = malloc(n * sizeof(*x)), *w = malloc(n * sizeof(*w));
	or_gl_nodes(n, x, w);
	for(int i = 0; i < n; i++) {
		if(cos_theta)
			cos_theta[i] = x[n - 1 - i];
		if(weights)
			weights[i] = w[n - 1 - i];
	}
	*ntheta = n;
	*nphi = n;
	free(x);
	free(w);
}

/* sh_series_trans.c:65-69 */
double complex or_Y(unsigned l, int m, double theta, double phi)
{
	double s = (m < 0 && (m & 1)) ? -1. : 1.;
	return s * (cos(m * phi) + I * sin(m * phi)) * or_sphPlm(l, abs(m), cos(theta));
}

/* sh_series_trans.c:257-291: per m block, the -m terms use
 * Y_l(-m) = (-1)^m conj(Y_lm). */
double complex or_eval(const struct or_series *s, double theta, double phi)
{
	const int m_max = s->polar ? 0 : (int) s->l_max;
	double *P = malloc((s->l_max + 1) * sizeof(*P));
	double complex val = 0.;
	for(int m = 0; m <= m_max; m++) {
		or_sphPlm_array(s->l_max, m, cos(theta), P);
		const double complex e = cos(m * phi) + I * sin(m * phi);
		double complex x = 0., xm = 0.;
		for(unsigned l = m; l <= s->l_max; l++) {
			const double complex y = e * P[l - m];
			x += or_get(s, l, m) * y;
			if(m)
				xm += or_get(s, l, -m) * conj(y);
		}
		val += x + ((m & 1) ? -xm : xm);
	}
	free(P);
	return val;
}

/* sh_series_trans.c:575-660 */
struct or_series *or_from_mesh(struct or_series *s, const double complex *mesh)
{
	int nt, np;
	const unsigned L = s->l_max;
	const int m_max = s->polar ? 0 : (int) L;
	double *x = malloc(2 * (L + 1) * sizeof(*x)), *w = malloc(2 * (L + 1) * sizeof(*w));
	or_mesh_nodes(L, &nt, &np, x, w);
	double complex *F = malloc((size_t) nt * np * sizeof(*F));
	memcpy(F, mesh, (size_t) nt * np * sizeof(*F));
	for(int i = 0; i < nt; i++) {
		or_dft(F + (size_t) i * np, np, -1);
		const double wt = w[i] * (2. * M_PI / np);
		for(int j = 0; j < np; j++)
			F[(size_t) i * np + j] *= wt;
	}
	double *P = malloc((size_t) nt * (L + 1) * sizeof(*P));
	for(int m = 0; m <= m_max; m++) {
		for(int i = 0; i < nt; i++)
			or_sphPlm_array(L, m, x[i], P + (size_t) i * (L + 1));
		for(unsigned l = m; l <= L; l++) {
			double complex c = 0.;
			for(int i = 0; i < nt; i++)
				c += F[(size_t) i * np + m] * P[(size_t) i * (L + 1) + l - m];
			or_set(s, l, m, c);
		}
		if(m) {
			const double sg = (m & 1) ? -1. : 1.;
			for(unsigned l = m; l <= L; l++) {
				double complex c = 0.;
				for(int i = 0; i < nt; i++)
					c += F[(size_t) i * np + np - m] * P[(size_t) i * (L + 1) + l - m];
				or_set(s, l, -m, sg * c);
			}
		}
	}
	free(P);
	free(F);
	free(x);
	free(w);
	return s;
}

/* sh_series_trans.c:663-734: r2c along phi, a_l(-m) = (-1)^m conj(a_lm);
 * for m = 0 the coefficient is stored conjugated (the second store of the
 * reference's paired set). */
struct or_series *or_from_realmesh(struct or_series *s, const double *mesh)
{
	int nt, np;
	const unsigned L = s->l_max;
	const int m_max = s->polar ? 0 : (int) L;
	double *x = malloc(2 * (L + 1) * sizeof(*x)), *w = malloc(2 * (L + 1) * sizeof(*w));
	or_mesh_nodes(L, &nt, &np, x, w);
	const int nh = np / 2 + 1;
	double complex *F = malloc((size_t) nt * nh * sizeof(*F));
	double complex *row = malloc(np * sizeof(*row));
	for(int i = 0; i < nt; i++) {
		for(int j = 0; j < np; j++)
			row[j] = mesh[(size_t) i * np + j];
		or_dft(row, np, -1);
		const double wt = w[i] * (2. * M_PI / np);
		for(int j = 0; j < nh; j++)
			F[(size_t) i * nh + j] = row[j] * wt;
	}
	double *P = malloc((size_t) nt * (L + 1) * sizeof(*P));
	for(int m = 0; m <= m_max; m++) {
		for(int i = 0; i < nt; i++)
			or_sphPlm_array(L, m, x[i], P + (size_t) i * (L + 1));
		for(unsigned l = m; l <= L; l++) {
			double complex c = 0.;
			for(int i = 0; i < nt; i++)
				c += F[(size_t) i * nh + m] * P[(size_t) i * (L + 1) + l - m];
			or_set(s, l, m, c);
			or_set(s, l, -m, ((m & 1) ? -1. : 1.) * conj(c));
		}
	}
	free(P);
	free(row);
	free(F);
	free(x);
	free(w);
	return s;
}

/* sh_series_trans.c:796-871: Legendre sums per ring and m, -m stored at
 * column nphi - m with the (-1)^m factor, then an unnormalised backward
 * DFT along phi. */
/* threads of the ring loop in or_to_mesh: the reference runs it on 4
 * (sh_series_trans.c:829) */
int or_mesh_threads = 4;

double complex *or_to_mesh_threads(const struct or_series *s, int *ntheta, int *nphi, int threads)
{
	int nt, np;
	const unsigned L = s->l_max;
	const int m_max = s->polar ? 0 : (int) L;
	double *x = malloc(2 * (L + 1) * sizeof(*x));
	or_mesh_nodes(L, &nt, &np, x, NULL);
	double complex *F = calloc((size_t) nt * np, sizeof(*F));
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
	for(int i = 0; i < nt; i++) {
		double P[L + 1];
		or_sphPlm_array(L, 0, x[i], P);
		double complex acc = 0.;
		for(unsigned l = 0; l <= L; l++)
			acc += or_get(s, l, 0) * P[l];
		F[(size_t) i * np] = acc;
		for(int m = 1; m <= m_max; m++) {
			or_sphPlm_array(L, m, x[i], P);
			double complex xp = 0., xm = 0.;
			for(unsigned l = m; l <= L; l++) {
				xp += or_get(s, l, m) * P[l - m];
				xm += or_get(s, l, -m) * P[l - m];
			}
			F[(size_t) i * np + m] = xp;
			F[(size_t) (i + 1) * np - m] = (m & 1) ? -xm : xm;
		}
	}
	for(int i = 0; i < nt; i++)
		or_dft(F + (size_t) i * np, np, +1);
	free(x);
	if(ntheta)
		*ntheta = nt;
	if(nphi)
		*nphi = np;
	return F;
}

double complex *or_to_mesh(const struct or_series *s, int *ntheta, int *nphi)
{
	return or_to_mesh_threads(s, ntheta, nphi, or_mesh_threads);
}

/* sh_series_trans.c:538-565 + 777-787 */
struct or_series *or_from_realfunc(struct or_series *s, double (*f)(double, double, void *), void *data)
{
	int nt, np;
	const unsigned L = s->l_max;
	double *x = malloc(2 * (L + 1) * sizeof(*x));
	or_mesh_nodes(L, &nt, &np, x, NULL);
	double *mesh = malloc((size_t) nt * np * sizeof(*mesh));
	const double dphi = 2. * M_PI / np;
	for(int i = 0; i < nt; i++) {
		const double theta = acos(x[i]);
		for(int j = 0; j < np; j++)
			mesh[(size_t) i * np + j] = f(theta, dphi * j, data);
	}
	or_from_realmesh(s, mesh);
	free(mesh);
	free(x);
	return s;
}

/* sh_series_trans.c:508-535 + 750-762 */
struct or_series *or_from_func(struct or_series *s, double complex (*f)(double, double, void *), void *data)
{
	int nt, np;
	const unsigned L = s->l_max;
	double *x = malloc(2 * (L + 1) * sizeof(*x));
	or_mesh_nodes(L, &nt, &np, x, NULL);
	double complex *mesh = malloc((size_t) nt * np * sizeof(*mesh));
	const double dphi = 2. * M_PI / np;
	for(int i = 0; i < nt; i++) {
		const double theta = acos(x[i]);
		for(int j = 0; j < np; j++)
			mesh[(size_t) i * np + j] = f(theta, dphi * j, data);
	}
	or_from_mesh(s, mesh);
	free(mesh);
	free(x);
	return s;
}

/* sh_series_trans.c:885-899: coefficients are Y*_lm(theta, phi) */
struct or_series *or_impulse(unsigned l_max, double theta, double phi)
{
	struct or_series *s = or_series_new(l_max, 0);
	for(int m = -(int) l_max; m <= (int) l_max; m++)
		for(unsigned l = abs(m); l <= l_max; l++)
			or_set(s, l, m, conj(or_Y(l, m, theta, phi)));
	return s;
}
