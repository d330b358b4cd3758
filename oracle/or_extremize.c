/*
 * Oracle (test infrastructure only): the extremizer and the gradient
 * operator it uses, restating sh_series_extremize.c and
 * sh_series_deriv.c:288-543.
 *
 * The seed search (find_near_minimum, sh_series_extremize.c:168-241) is
 * restated as is.  GSL's BFGS2 refinement is not: the oracle refines with
 * a Nelder-Mead simplex in the tangent plane of the seed, shrunk to a
 * 1e-12 rad diameter, a method independent of the library's Newton steps.
 * Both find the stationary point the seed leads to; parity on the BFGS
 * iterates themselves is not claimed (parity unpinned for the iterates).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* sh_series_extremize.c:63-87 */
static void condition_theta_phi(double *theta, double *phi)
{
	*theta = fmod(*theta, 2 * M_PI);
	if(*theta < 0.)
		*theta += 2 * M_PI;
	if(*theta > M_PI) {
		*theta = 2. * M_PI - *theta;
		*phi += M_PI;
	}
	*phi = fmod(*phi, 2 * M_PI);
	if(*phi < 0.)
		*phi += 2 * M_PI;
}

/* sh_series_extremize.c:168-241: minimum pixel of the l_lowpass copy on
 * its own mesh, first minimum in row-major order */
double or_find_near_minimum(const struct or_series *s, unsigned l_lowpass, double *theta, double *phi)
{
	struct or_series *lo = or_series_copy(s);
	if(lo->l_max > l_lowpass)
		or_series_resize(lo, l_lowpass);
	int nt, np;
	double complex *mesh = or_to_mesh(lo, &nt, &np);
	double *x = malloc(nt * sizeof(*x));
	or_mesh_nodes(lo->l_max, &nt, &np, x, NULL);
	double min = INFINITY;
	for(int j = 0; j < nt; j++)
		for(int i = 0; i < np; i++) {
			const double v = creal(mesh[(size_t) j * np + i]);
			if(v < min) {
				min = v;
				*theta = acos(x[j]);
				*phi = 2. * M_PI * i / (double) np;
			}
		}
	free(x);
	free(mesh);
	or_series_free(lo);
	return min;
}

struct nm_ctx {
	const struct or_series *s;
	double th0, ph0;
};

/* value at tangent offset (d1, d2) from the seed */
static double f_at(const struct nm_ctx *c, const double d[2])
{
	const double st = sin(c->th0), ct = cos(c->th0), sp = sin(c->ph0), cp = cos(c->ph0);
	const double x = st * cp + d[0] * ct * cp - d[1] * sp;
	const double y = st * sp + d[0] * ct * sp + d[1] * cp;
	const double z = ct - d[0] * st;
	double th = atan2(sqrt(x * x + y * y), z), ph = atan2(y, x);
	condition_theta_phi(&th, &ph);
	return creal(or_eval(c->s, th, ph));
}

/* Nelder-Mead in the tangent plane from (th, ph), initial simplex size
 * `size` rad, restarted around the best vertex with shrinking simplices */
double or_refine_minimum(const struct or_series *s, double th, double ph, double size, double *theta, double *phi)
{
	struct nm_ctx c = { s, th, ph };
	double p[3][2] = { { 0, 0 }, { size, 0 }, { 0, size } }, v[3];
	for(int restart = 0; restart < 6; restart++) {
		for(int k = 0; k < 3; k++)
			v[k] = f_at(&c, p[k]);
		for(int it = 0; it < 4000; it++) {
			/* order */
			for(int a = 0; a < 3; a++)
				for(int b = a + 1; b < 3; b++)
					if(v[b] < v[a]) {
						double t = v[a];
						v[a] = v[b];
						v[b] = t;
						for(int q = 0; q < 2; q++) {
							t = p[a][q];
							p[a][q] = p[b][q];
							p[b][q] = t;
						}
					}
			const double size = fabs(p[1][0] - p[0][0]) + fabs(p[1][1] - p[0][1]) + fabs(p[2][0] - p[0][0]) + fabs(p[2][1] - p[0][1]);
			if(size < 1e-12)
				break;
			double cen[2] = { (p[0][0] + p[1][0]) / 2, (p[0][1] + p[1][1]) / 2 };
			double r[2] = { 2 * cen[0] - p[2][0], 2 * cen[1] - p[2][1] };
			const double fr = f_at(&c, r);
			if(fr < v[0]) {
				double e[2] = { 3 * cen[0] - 2 * p[2][0], 3 * cen[1] - 2 * p[2][1] };
				const double fe = f_at(&c, e);
				if(fe < fr) {
					memcpy(p[2], e, sizeof(e));
					v[2] = fe;
				} else {
					memcpy(p[2], r, sizeof(r));
					v[2] = fr;
				}
			} else if(fr < v[1]) {
				memcpy(p[2], r, sizeof(r));
				v[2] = fr;
			} else {
				double k2[2] = { (cen[0] + p[2][0]) / 2, (cen[1] + p[2][1]) / 2 };
				const double fk = f_at(&c, k2);
				if(fk < v[2]) {
					memcpy(p[2], k2, sizeof(k2));
					v[2] = fk;
				} else {
					for(int q = 1; q < 3; q++) {
						p[q][0] = (p[q][0] + p[0][0]) / 2;
						p[q][1] = (p[q][1] + p[0][1]) / 2;
						v[q] = f_at(&c, p[q]);
					}
				}
			}
		}
		/* re-centre the tangent frame on the best vertex and restart */
		const double st = sin(c.th0), ct = cos(c.th0), sp = sin(c.ph0), cp = cos(c.ph0);
		const double x = st * cp + p[0][0] * ct * cp - p[0][1] * sp;
		const double y = st * sp + p[0][0] * ct * sp + p[0][1] * cp;
		const double z = ct - p[0][0] * st;
		c.th0 = atan2(sqrt(x * x + y * y), z);
		c.ph0 = atan2(y, x);
		condition_theta_phi(&c.th0, &c.ph0);
		const double sz = 1e-3 / (1 << (2 * restart));
		double np[3][2] = { { 0, 0 }, { sz, 0 }, { 0, sz } };
		memcpy(p, np, sizeof(np));
	}
	*theta = c.th0;
	*phi = c.ph0;
	return creal(or_eval(s, c.th0, c.ph0));
}

/* sh_series_extremize.c:273-360 with the refinement described above */
double or_real_minimum(const struct or_series *s, double *theta, double *phi)
{
	double th, ph;
	or_find_near_minimum(s, 6, &th, &ph);
	return or_refine_minimum(s, th, ph, 0.05, theta, phi);
}

/* sh_series_deriv.c:288-~520: sin(theta) d/dtheta through the reference's
 * Wigner 3j form; result at l_max + 1 */
struct or_series *or_sintheta_d_by_dtheta(const struct or_series *s)
{
	struct or_series *r = or_series_new_zero(s->l_max + 1, s->polar);
	for(int l = 1; l <= (int) s->l_max; l++) {
		double x = 2. * sqrt(l + 0.5);
		const double g1 = x * sqrt(l - 0.5) * or_wigner_3j(1, l, l - 1, 0, 0, 0);
		const double g2 = x * sqrt(l + 1.5) * or_wigner_3j(1, l, l + 1, 0, 0, 0);
		const int mm = s->polar ? 0 : l;
		double w_lm1_m = or_wigner_3j(1, l, l - 1, 0, -mm, mm);
		double w_lp1_m = or_wigner_3j(1, l, l + 1, 0, -mm, mm);
		for(int m = -mm; m <= mm; m++) {
			const double complex a = (m & 1) ? -or_get(s, l, m) : or_get(s, l, m);
			const int A = l - m, B = l + m, C = A * (B + 1);
			const double w_lm1_mp1 = or_wigner_3j(1, l, l - 1, 0, m + 1, -(m + 1));
			const double w_lp1_mp1 = or_wigner_3j(1, l, l + 1, 0, m + 1, -(m + 1));
			x = m - C;
			const double f1 = x * w_lm1_m - sqrt((double) C * (A - 1) * B) * w_lm1_mp1;
			const double f2 = x * w_lp1_m - sqrt((double) C * (A + 1) * (B + 2)) * w_lp1_mp1;
			w_lm1_m = w_lm1_mp1;
			w_lp1_m = w_lp1_mp1;
			if(l - 1 >= abs(m))
				r->coeff[or_lmoffset(r->l_max, l - 1, m)] += a * f1 * g1;
			r->coeff[or_lmoffset(r->l_max, l + 1, m)] += a * f2 * g2;
		}
	}
	return r;
}

void orx_real_minimum(unsigned l, int polar, const double complex *coeff, double *out)
{
	struct or_series s = { l, polar, (double complex *) coeff };
	double th = NAN, ph = NAN;
	out[2] = or_real_minimum(&s, &th, &ph);
	out[0] = th;
	out[1] = ph;
}

void orx_find_near_minimum(unsigned l, int polar, const double complex *coeff, unsigned l_lowpass, double *out)
{
	struct or_series s = { l, polar, (double complex *) coeff };
	double th = NAN, ph = NAN;
	out[2] = or_find_near_minimum(&s, l_lowpass, &th, &ph);
	out[0] = th;
	out[1] = ph;
}

void orx_sintheta_d_by_dtheta(unsigned l, int polar, const double complex *coeff, double complex *out)
{
	struct or_series s = { l, polar, (double complex *) coeff };
	struct or_series *r = or_sintheta_d_by_dtheta(&s);
	memcpy(out, r->coeff, or_length(r->l_max, r->polar) * sizeof(*out));
	or_series_free(r);
}

void orx_refine_minimum(unsigned l, int polar, const double complex *coeff, double theta, double phi, double size, double *out)
{
	struct or_series s = { l, polar, (double complex *) coeff };
	double th = NAN, ph = NAN;
	out[2] = or_refine_minimum(&s, theta, phi, size, &th, &ph);
	out[0] = th;
	out[1] = ph;
}
