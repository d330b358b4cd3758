/*
 * Oracle (test infrastructure only): Euler matrices and the Choi, Ivanic,
 * Gordon & Ruedenberg Wigner-D recursion, restating sh_series_rotation.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* sh_series_rotation.c:104-153: R = Rz(alpha) Ry(beta) Rz(gamma), C order */
void or_euler_rotation_matrix(double gamma, double beta, double alpha, double R[9])
{
	const double ca = cos(alpha), sa = sin(alpha), cb = cos(beta), sb = sin(beta), cg = cos(gamma), sg = sin(gamma);
	R[0] = ca * cb * cg - sa * sg;
	R[1] = -ca * cb * sg - sa * cg;
	R[2] = ca * sb;
	R[3] = sa * cb * cg + ca * sg;
	R[4] = -sa * cb * sg + ca * cg;
	R[5] = sa * sb;
	R[6] = -sb * cg;
	R[7] = sb * sg;
	R[8] = cb;
}

/* sh_series_rotation.c:170-173 */
void or_rot_matrix(double theta, double phi, double R[9])
{
	or_euler_rotation_matrix(0., theta, phi, R);
}

/* sh_series_rotation.c:156-159, 184-187 */
void or_invrot_matrix(double theta, double phi, double R[9])
{
	or_euler_rotation_matrix(-phi, -theta, -0., R);
}

/* element (m, m') of block l; sh_series_rotation.c:279 */
#define DE(Dl, l, m, mp) (Dl)[(2 * (l) + 1) * (mp) + (m)]

static double complex *block_new(int l)
{
	double complex *D = malloc((size_t) (2 * l + 1) * (2 * l + 1) * sizeof(*D));
	return D + (2 * l + 1) * l + l;
}

static void block_free(double complex *D, int l)
{
	if(D)
		free(D - (2 * l + 1) * l - l);
}

/* sh_series_rotation.c:287-310 (Choi et al. 5.3-5.5) */
static double complex *or_D1(const double *R)
{
	enum { x = 0, y = 1, z = 2 };
	double complex *D = block_new(1);
	const double r2 = sqrt(2.);
	DE(D, 1, -1, -1) = (R[3 * y + y] + R[3 * x + x] - I * (R[3 * x + y] - R[3 * y + x])) / 2;
	DE(D, 1, -1, 0) = (R[3 * z + x] - I * R[3 * z + y]) / r2;
	DE(D, 1, -1, +1) = (R[3 * y + y] - R[3 * x + x] + I * (R[3 * x + y] + R[3 * y + x])) / 2;
	DE(D, 1, 0, -1) = (R[3 * x + z] + I * R[3 * y + z]) / r2;
	DE(D, 1, 0, 0) = R[3 * z + z];
	DE(D, 1, 0, +1) = (-R[3 * x + z] + I * R[3 * y + z]) / r2;
	DE(D, 1, +1, -1) = (R[3 * y + y] - R[3 * x + x] - I * (R[3 * x + y] + R[3 * y + x])) / 2;
	DE(D, 1, +1, 0) = (-R[3 * z + x] - I * R[3 * z + y]) / r2;
	DE(D, 1, +1, +1) = (R[3 * y + y] + R[3 * x + x] - I * (R[3 * y + x] - R[3 * x + y])) / 2;
	return D;
}

/* Choi et al. (6.2), (6.3), (6.10), (6.11); sh_series_rotation.c:318-343 */
static double ca(int l, int m, int mp) { return sqrt((l + m) * (l - m) / (double) ((l + mp) * (l - mp))); }
static double cb(int l, int m, int mp) { return sqrt((l + m) * (l + m - 1) / (double) (2 * (l + mp) * (l - mp))); }
static double cc(int l, int m, int mp) { return sqrt(2 * (l + m) * (l - m) / (double) ((l + mp) * (l + mp - 1))); }
static double cd(int l, int m, int mp) { return sqrt((l + m) * (l + m - 1) / (double) ((l + mp) * (l + mp - 1))); }

/* sh_series_rotation.c:346-369 (Choi et al. 6.1, 6.9, 6.14) */
static double complex *or_DL(int l, const double complex *P, const double complex *D1)
{
	double complex *D = block_new(l);
	for(int m = -l; m <= l; m++) {
		for(int mp = -l + 1; mp <= l - 1; mp++)
			DE(D, l, m, mp) = (abs(m) == l ? 0 : ca(l, m, mp) * DE(D1, 1, 0, 0) * DE(P, l - 1, m, mp))
				+ (m <= -l + 1 ? 0 : cb(l, m, mp) * DE(D1, 1, 1, 0) * DE(P, l - 1, m - 1, mp))
				+ (-m <= -l + 1 ? 0 : cb(l, -m, mp) * DE(D1, 1, -1, 0) * DE(P, l - 1, m + 1, mp));
		int mp = -l;
		DE(D, l, m, mp) = (abs(m) == l ? 0 : cc(l, m, -mp) * DE(D1, 1, 0, -1) * DE(P, l - 1, m, mp + 1))
			+ (m <= -l + 1 ? 0 : cd(l, m, -mp) * DE(D1, 1, 1, -1) * DE(P, l - 1, m - 1, mp + 1))
			+ (-m <= -l + 1 ? 0 : cd(l, -m, -mp) * DE(D1, 1, -1, -1) * DE(P, l - 1, m + 1, mp + 1));
		mp = l;
		DE(D, l, m, mp) = (abs(m) == l ? 0 : cc(l, m, mp) * DE(D1, 1, 0, 1) * DE(P, l - 1, m, mp - 1))
			+ (m <= -l + 1 ? 0 : cd(l, m, mp) * DE(D1, 1, 1, 1) * DE(P, l - 1, m - 1, mp - 1))
			+ (-m <= -l + 1 ? 0 : cd(l, -m, mp) * DE(D1, 1, -1, 1) * DE(P, l - 1, m + 1, mp - 1));
	}
	return D;
}

/* sh_series_rotation.c:386-422 */
struct or_rotation_plan *or_rotation_plan_new(unsigned l_max, const double R[9])
{
	struct or_rotation_plan *p = malloc(sizeof(*p));
	p->l_max = l_max;
	p->D = calloc(l_max + 1, sizeof(*p->D));
	if(l_max >= 1)
		p->D[1] = or_D1(R);
	for(unsigned l = 2; l <= l_max; l++)
		p->D[l] = or_DL(l, p->D[l - 1], p->D[1]);
	return p;
}

void or_rotation_plan_free(struct or_rotation_plan *p)
{
	if(p) {
		for(unsigned l = 1; l <= p->l_max; l++)
			block_free(p->D[l], l);
		free(p->D);
	}
	free(p);
}

/* sh_series_rotation.c:473-513: drop the blocks above l_max */
struct or_rotation_plan *or_rotation_plan_set_l(struct or_rotation_plan *p, unsigned l_max)
{
	if(l_max > p->l_max)
		return NULL;
	for(unsigned l = l_max + 1; l <= p->l_max; l++) {
		block_free(p->D[l], l);
		p->D[l] = NULL;
	}
	p->l_max = l_max;
	return p;
}

/* sh_series_rotation.c:523-529 */
double complex or_wigner_D(const struct or_rotation_plan *p, unsigned l, int m, int mp)
{
	if(l > p->l_max || (unsigned) abs(m) > l || (unsigned) abs(mp) > l)
		return NAN + I * NAN;
	return l == 0 ? 1.0 : DE(p->D[l], (int) l, m, mp);
}

/* sh_series_rotation.c:546-589: F'_{l m'} = sum_m D^l_{m m'} F_{l m};
 * polar input keeps only the m = 0 row. */
struct or_series *or_rotate(struct or_series *r, const struct or_series *s, const struct or_rotation_plan *p)
{
	if(r->l_max != s->l_max || r->polar || r == s)
		return NULL;
	or_set(r, 0, 0, or_get(s, 0, 0));
	for(unsigned l = 1; l <= s->l_max; l++) {
		const double complex *D = p->D[l];
		const int L = l;
		for(int mp = -L; mp <= L; mp++) {
			if(s->polar) {
				or_set(r, l, mp, DE(D, L, 0, mp) * or_get(s, l, 0));
			} else {
				double complex c = 0.;
				for(int m = -L; m <= L; m++)
					c += DE(D, L, m, mp) * or_get(s, l, m);
				or_set(r, l, mp, c);
			}
		}
	}
	return r;
}

/* sh_series_rotation.c:600-625 */
struct or_series *or_rotate_z(struct or_series *r, const struct or_series *s, double phi)
{
	if(r->l_max != s->l_max)
		return NULL;
	if(s->polar) {
		if(r != s)
			or_series_assign(r, s);
		return r;
	}
	if(r->polar)
		return NULL;
	for(int m = -(int) s->l_max; m <= (int) s->l_max; m++) {
		const double complex f = cexp(-I * m * phi);
		for(unsigned l = abs(m); l <= s->l_max; l++)
			or_set(r, l, m, or_get(s, l, m) * f);
	}
	return r;
}
