/*
 * Spherical harmonics, the Gauss-Legendre x equispaced pixel mesh and the
 * transforms between series and mesh (sphradiometer
 * src/lib/sh_series_trans.c).  Synthesis (sh_series_to_mesh) runs on the
 * GPU; analysis and evaluation are host-side plan-time helpers.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sh_series.h>
#include "../sphr_internal.h"

static inline double complex eiphi(double x)
{
	return cos(x) + I * sin(x);
}

/* sh_series_trans.c:65-69: Y_lm with (-1)^m for odd negative m */
double complex sh_series_Y(unsigned int l, int m, double theta, double phi)
{
	if((unsigned) abs(m) > l)
		return 0.0;
	double *p = malloc((l + 1) * sizeof(*p));
	sphr_plm(l, abs(m), cos(theta), p);
	const double v = p[l - abs(m)];
	free(p);
	return ((m < 0 && (m & 1)) ? -1.0 : 1.0) * eiphi(m * phi) * v;
}

double complex sh_series_Yconj(unsigned int l, int m, double theta, double phi)
{
	return conj(sh_series_Y(l, m, theta, phi));
}

/* sh_series_trans.c:85-116: l = |m| .. l_max */
double complex *sh_series_Y_array(double complex *out, unsigned int l_max, int m, double theta, double phi)
{
	const int am = abs(m);
	double *p = malloc((l_max + 1) * sizeof(*p));
	sphr_plm(l_max, am, cos(theta), p);
	const double complex f = ((m < 0 && (m & 1)) ? -1.0 : 1.0) * eiphi(m * phi);
	for(unsigned i = 0; i + am <= l_max; i++)
		out[i] = f * p[i];
	free(p);
	return out;
}

double complex *sh_series_Yconj_array(double complex *out, unsigned int l_max, int m, double theta, double phi)
{
	sh_series_Y_array(out, l_max, m, theta, phi);
	for(unsigned i = 0; i + abs(m) <= l_max; i++)
		out[i] = conj(out[i]);
	return out;
}

/* sh_series_trans.c:149-240: 2(l_max+1) rings at the GL nodes in
 * descending cos(theta), 2(l_max+1) equispaced phi columns */
static int mesh_params(unsigned l_max, int *nt, int *np, double **cos_theta, double **weights)
{
	const int n = 2 * ((int) l_max + 1);
	const double *x, *w;
	sphr_mesh_rule(l_max, &x, &w);
	*nt = n;
	*np = n;
	*cos_theta = malloc(n * sizeof(double));
	if(!*cos_theta)
		return -1;
	memcpy(*cos_theta, x, n * sizeof(double));
	if(weights) {
		*weights = malloc(n * sizeof(double));
		if(!*weights) {
			free(*cos_theta);
			*cos_theta = NULL;
			return -1;
		}
		memcpy(*weights, w, n * sizeof(double));
	}
	return 0;
}

complex double *sh_series_mesh_new(unsigned int l_max, int *ntheta, int *nphi, double **cos_theta, double **weights)
{
	if(mesh_params(l_max, ntheta, nphi, cos_theta, weights))
		return NULL;
	complex double *m = malloc((size_t) *ntheta * *nphi * sizeof(*m));
	if(!m) {
		free(*cos_theta);
		*cos_theta = NULL;
		if(weights) {
			free(*weights);
			*weights = NULL;
		}
	}
	return m;
}

double *sh_series_real_mesh_new(unsigned int l_max, int *ntheta, int *nphi, double **cos_theta, double **weights)
{
	if(mesh_params(l_max, ntheta, nphi, cos_theta, weights))
		return NULL;
	double *m = malloc((size_t) *ntheta * *nphi * sizeof(*m));
	if(!m) {
		free(*cos_theta);
		*cos_theta = NULL;
		if(weights) {
			free(*weights);
			*weights = NULL;
		}
	}
	return m;
}

/* sh_series_trans.c:257-291: sum over m blocks; the -m terms use
 * Y_l(-m) = (-1)^m conj(Y_lm) */
double complex sh_series_eval(const struct sh_series *s, double theta, double phi)
{
	const int mm = s->polar ? 0 : (int) s->l_max;
	double *p = malloc((s->l_max + 1) * sizeof(*p));
	const double x = cos(theta);
	double complex v = 0.0;
	for(int m = 0; m <= mm; m++) {
		sphr_plm(s->l_max, m, x, p);
		const double complex e = eiphi(m * phi);
		const double complex *cp = s->coeff + sh_series_moffset(s->l_max, m);
		double complex sp = 0.0, sn = 0.0;
		for(unsigned i = 0; i + m <= s->l_max; i++)
			sp += cp[i] * p[i];
		if(m) {
			const double complex *cn = s->coeff + sh_series_moffset(s->l_max, -m);
			for(unsigned i = 0; i + m <= s->l_max; i++)
				sn += cn[i] * p[i];
			v += e * sp + ((m & 1) ? -1.0 : 1.0) * conj(e) * sn;
		} else
			v += sp;
	}
	free(p);
	return v;
}

/* sh_series_trans.c:575-660 (complex mesh) and 663-734 (real mesh):
 * phi DFT of every ring, times w_j 2 pi / nphi, then the Legendre
 * quadrature per m */
static struct sh_series *analyse(struct sh_series *s, const double *mesh, int is_real)
{
	const unsigned L = s->l_max;
	const int n = 2 * ((int) L + 1);
	const int mm = s->polar ? 0 : (int) L;
	const double *x, *w;
	sphr_mesh_rule(L, &x, &w);
	double *F = malloc(2 * (size_t) n * n * sizeof(*F));
	double *P = malloc((size_t) n * (L + 1) * sizeof(*P));
	if(!F || !P) {
		free(F);
		free(P);
		return NULL;
	}
	for(int j = 0; j < n; j++) {
		double *row = F + 2 * (size_t) j * n;
		for(int k = 0; k < n; k++) {
			row[2 * k] = is_real ? mesh[(size_t) j * n + k] : mesh[2 * ((size_t) j * n + k)];
			row[2 * k + 1] = is_real ? 0.0 : mesh[2 * ((size_t) j * n + k) + 1];
		}
		sphr_host_dft(row, n, -1);
		const double wt = w[j] * (2.0 * M_PI / n);
		for(int k = 0; k < 2 * n; k++)
			row[k] *= wt;
	}
	for(int m = 0; m <= mm; m++) {
		for(int j = 0; j < n; j++)
			sphr_plm(L, m, x[j], P + (size_t) j * (L + 1));
		const double sg = (m & 1) ? -1.0 : 1.0;
		for(unsigned l = m; l <= L; l++) {
			double complex cp = 0.0, cn = 0.0;
			for(int j = 0; j < n; j++) {
				const double pv = P[(size_t) j * (L + 1) + l - m];
				const double *row = F + 2 * (size_t) j * n;
				cp += (row[2 * m] + I * row[2 * m + 1]) * pv;
				if(m && !is_real)
					cn += (row[2 * (n - m)] + I * row[2 * (n - m) + 1]) * pv;
			}
			if(is_real) {
				/* real mesh: a_l(-m) = (-1)^m conj(a_lm); the m = 0
				 * coefficient is stored conjugated, as the reference's
				 * paired store leaves it */
				s->coeff[sh_series_lmoffset(s, l, m)] = cp;
				s->coeff[sh_series_lmoffset(s, l, -m)] = sg * conj(cp);
			} else {
				s->coeff[sh_series_lmoffset(s, l, m)] = cp;
				if(m)
					s->coeff[sh_series_lmoffset(s, l, -m)] = sg * cn;
			}
		}
	}
	free(P);
	free(F);
	return s;
}

struct sh_series *sh_series_from_mesh(struct sh_series *s, double complex *mesh)
{
	return analyse(s, (const double *) mesh, 0);
}

struct sh_series *sh_series_from_realmesh(struct sh_series *s, double *mesh)
{
	return analyse(s, mesh, 1);
}

/* sh_series_trans.c:508-565 + 750-787: sample at theta_j = acos(x_j),
 * phi_k = 2 pi k / nphi, then analyse */
struct sh_series *sh_series_from_func(struct sh_series *s, double complex (*f)(double, double, void *), void *data)
{
	const int n = 2 * ((int) s->l_max + 1);
	const double *x, *w;
	sphr_mesh_rule(s->l_max, &x, &w);
	double complex *mesh = malloc((size_t) n * n * sizeof(*mesh));
	if(!mesh)
		return NULL;
	for(int j = 0; j < n; j++) {
		const double th = acos(x[j]);
		for(int k = 0; k < n; k++)
			mesh[(size_t) j * n + k] = f(th, 2.0 * M_PI / n * k, data);
	}
	struct sh_series *r = analyse(s, (const double *) mesh, 0);
	free(mesh);
	return r;
}

struct sh_series *sh_series_from_realfunc(struct sh_series *s, double (*f)(double, double, void *), void *data)
{
	const int n = 2 * ((int) s->l_max + 1);
	const double *x, *w;
	sphr_mesh_rule(s->l_max, &x, &w);
	double *mesh = malloc((size_t) n * n * sizeof(*mesh));
	if(!mesh)
		return NULL;
	for(int j = 0; j < n; j++) {
		const double th = acos(x[j]);
		for(int k = 0; k < n; k++)
			mesh[(size_t) j * n + k] = f(th, 2.0 * M_PI / n * k, data);
	}
	struct sh_series *r = analyse(s, mesh, 1);
	free(mesh);
	return r;
}

/* sh_series_trans.c:796-871 (GPU synthesis) */
complex double *sh_series_to_mesh(const struct sh_series *s, int *ntheta, int *nphi, double **cos_theta)
{
	const int n = 2 * ((int) s->l_max + 1);
	complex double *mesh = malloc((size_t) n * n * sizeof(*mesh));
	double *ct = NULL;
	if(cos_theta) {
		const double *x;
		sphr_mesh_rule(s->l_max, &x, NULL);
		ct = malloc(n * sizeof(*ct));
		if(ct)
			memcpy(ct, x, n * sizeof(*ct));
	}
	if(!mesh || (cos_theta && !ct) || sphr_gpu_to_mesh(s->l_max, s->polar, (const double *) s->coeff, (double *) mesh)) {
		free(mesh);
		free(ct);
		if(ntheta)
			*ntheta = 0;
		if(nphi)
			*nphi = 0;
		if(cos_theta)
			*cos_theta = NULL;
		return NULL;
	}
	if(ntheta)
		*ntheta = n;
	if(nphi)
		*nphi = n;
	if(cos_theta)
		*cos_theta = ct;
	return mesh;
}

/* sh_series_trans.c:885-899: coefficients conj(Y_lm(theta, phi)) */
struct sh_series *sh_series_impulse(unsigned int l_max, double theta, double phi)
{
	struct sh_series *s = sh_series_new(l_max, 0);
	if(!s)
		return NULL;
	for(int m = -(int) l_max; m <= (int) l_max; m++)
		sh_series_Yconj_array(s->coeff + sh_series_moffset(l_max, m), l_max, m, theta, phi);
	return s;
}
