/*
 * Rotations of SH series (sphradiometer src/lib/sh_series_rotation.c).
 *
 * Plan construction (host, OpenMP over m rows) builds the Wigner-D blocks
 * with the recursion of Choi, Ivanic, Gordon & Ruedenberg, J. Chem. Phys.
 * 111, 8825 (1999), eqs. (5.3)-(5.5), (6.1)-(6.14), in the reference's
 * convention: the input rotation matrix is the transpose of the paper's,
 * and a coefficient transforms as F'_{l m'} = sum_m D^l_{m m'} F_{l m},
 * with D^l_{m m'} stored at D[l][(2l+1) m' + m].  The blocks of one plan
 * live in one contiguous allocation (l = 1 first), which is also the
 * packed form the GPU rotation kernel reads.
 *
 * sh_series_rotate and sh_series_rotate_z run on the GPU.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sh_series.h>
#include "../sphr_internal.h"

/* sh_series_rotation.c:104-153: Rz(alpha) Ry(beta) Rz(gamma), C order */
double *euler_rotation_matrix(double gamma, double beta, double alpha)
{
	double *R = malloc(9 * sizeof(*R));
	if(!R)
		return NULL;
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
	return R;
}

double *euler_inv_rotation_matrix(double gamma, double beta, double alpha)
{
	return euler_rotation_matrix(-alpha, -beta, -gamma);
}

/* rotates z-hat to (theta, phi): sh_series_rotation.c:170-173 */
double *sh_series_rot_matrix(double theta, double phi)
{
	return euler_rotation_matrix(0., theta, phi);
}

double *sh_series_invrot_matrix(double theta, double phi)
{
	return euler_inv_rotation_matrix(0., theta, phi);
}

static size_t block_offset(unsigned l)
{
	/* sum_{k=1}^{l-1} (2k+1)^2 = (4 l^3 - l - 3) / 3 */
	return ((size_t) 4 * l * l * l - l - 3) / 3;
}

static size_t blocks_total(unsigned l_max)
{
	return block_offset(l_max + 1);
}

/* element (m, m') of the centred block */
#define DM(Dl, l, m, mp) (Dl)[(2 * (l) + 1) * (mp) + (m)]

static void fill_D1(double complex *D, const double *R)
{
	enum { X = 0, Y = 1, Z = 2 };
	const double s = sqrt(0.5);
	DM(D, 1, -1, -1) = 0.5 * (R[3 * Y + Y] + R[3 * X + X]) - 0.5 * I * (R[3 * X + Y] - R[3 * Y + X]);
	DM(D, 1, -1, 0) = s * (R[3 * Z + X] - I * R[3 * Z + Y]);
	DM(D, 1, -1, 1) = 0.5 * (R[3 * Y + Y] - R[3 * X + X]) + 0.5 * I * (R[3 * X + Y] + R[3 * Y + X]);
	DM(D, 1, 0, -1) = s * (R[3 * X + Z] + I * R[3 * Y + Z]);
	DM(D, 1, 0, 0) = R[3 * Z + Z];
	DM(D, 1, 0, 1) = s * (-R[3 * X + Z] + I * R[3 * Y + Z]);
	DM(D, 1, 1, -1) = 0.5 * (R[3 * Y + Y] - R[3 * X + X]) - 0.5 * I * (R[3 * X + Y] + R[3 * Y + X]);
	DM(D, 1, 1, 0) = s * (-R[3 * Z + X] - I * R[3 * Z + Y]);
	DM(D, 1, 1, 1) = 0.5 * (R[3 * Y + Y] + R[3 * X + X]) - 0.5 * I * (R[3 * Y + X] - R[3 * X + Y]);
}

/*
 * Block l from block l-1 and block 1.  Interior columns |m'| < l use
 * (6.1) with the coefficients (6.2)/(6.3); the edge columns m' = -l and
 * m' = +l use (6.9)/(6.14) with (6.10)/(6.11).  Each term is present only
 * when its row of block l-1 exists.
 */
static void fill_DL(double complex *D, int l, const double complex *P, const double complex *D1)
{
	const double complex d00 = DM(D1, 1, 0, 0), dp0 = DM(D1, 1, 1, 0), dm0 = DM(D1, 1, -1, 0);
	const double complex d0m = DM(D1, 1, 0, -1), dpm = DM(D1, 1, 1, -1), dmm = DM(D1, 1, -1, -1);
	const double complex d0p = DM(D1, 1, 0, 1), dpp = DM(D1, 1, 1, 1), dmp = DM(D1, 1, -1, 1);
	#pragma omp parallel for schedule(static) if(l > 48)
	for(int m = -l; m <= l; m++) {
		const int has0 = abs(m) < l, hasm = m > -l + 1, hasp = -m > -l + 1;
		for(int mp = -l + 1; mp <= l - 1; mp++) {
			const double den = (double) (l + mp) * (l - mp);
			double complex v = 0.0;
			if(has0)
				v += sqrt((double) (l + m) * (l - m) / den) * d00 * DM(P, l - 1, m, mp);
			if(hasm)
				v += sqrt((double) (l + m) * (l + m - 1) / (2.0 * den)) * dp0 * DM(P, l - 1, m - 1, mp);
			if(hasp)
				v += sqrt((double) (l - m) * (l - m - 1) / (2.0 * den)) * dm0 * DM(P, l - 1, m + 1, mp);
			DM(D, l, m, mp) = v;
		}
		/* edges: the coefficient denominators use |m'| = l */
		const double den = (double) (2 * l) * (2 * l - 1);
		double complex vlo = 0.0, vhi = 0.0;
		if(has0) {
			const double c = sqrt(2.0 * (l + m) * (l - m) / den);
			vlo += c * d0m * DM(P, l - 1, m, -l + 1);
			vhi += c * d0p * DM(P, l - 1, m, l - 1);
		}
		if(hasm) {
			const double c = sqrt((double) (l + m) * (l + m - 1) / den);
			vlo += c * dpm * DM(P, l - 1, m - 1, -l + 1);
			vhi += c * dpp * DM(P, l - 1, m - 1, l - 1);
		}
		if(hasp) {
			const double c = sqrt((double) (l - m) * (l - m - 1) / den);
			vlo += c * dmm * DM(P, l - 1, m + 1, -l + 1);
			vhi += c * dmp * DM(P, l - 1, m + 1, l - 1);
		}
		DM(D, l, m, -l) = vlo;
		DM(D, l, m, l) = vhi;
	}
}

static struct sh_series_rotation_plan *plan_alloc(unsigned l_max)
{
	struct sh_series_rotation_plan *p = malloc(sizeof(*p));
	double complex **D = calloc(l_max + 1, sizeof(*D));
	double complex *base = l_max ? malloc(blocks_total(l_max) * sizeof(*base)) : NULL;
	if(!p || !D || (l_max && !base)) {
		free(p);
		free(D);
		free(base);
		return NULL;
	}
	for(unsigned l = 1; l <= l_max; l++)
		D[l] = base + block_offset(l) + (2 * l + 1) * l + l;
	p->l_max = l_max;
	p->D = D;
	return p;
}

struct sh_series_rotation_plan *sphr_rotation_plan_alloc(unsigned l_max)
{
	return plan_alloc(l_max);
}

static double complex *plan_base(const struct sh_series_rotation_plan *p)
{
	return p->l_max ? p->D[1] - 4 : NULL;
}

const double *sphr_rotation_blocks(const struct sh_series_rotation_plan *p)
{
	return (const double *) plan_base(p);
}

/* sh_series_rotation.c:386-422 */
struct sh_series_rotation_plan *sh_series_rotation_plan_new(const struct sh_series *series, const double *R)
{
	struct sh_series_rotation_plan *p = plan_alloc(series->l_max);
	if(!p)
		return NULL;
	if(series->l_max >= 1)
		fill_D1(p->D[1], R);
	for(unsigned l = 2; l <= series->l_max; l++)
		fill_DL(p->D[l], (int) l, p->D[l - 1], p->D[1]);
	return p;
}

void sh_series_rotation_plan_free(struct sh_series_rotation_plan *p)
{
	if(p) {
		free(plan_base(p));
		free(p->D);
	}
	free(p);
}

/* sh_series_rotation.c:448-464 */
struct sh_series_rotation_plan *sh_series_rotation_plan_copy(const struct sh_series_rotation_plan *p)
{
	struct sh_series_rotation_plan *q = plan_alloc(p->l_max);
	if(q && p->l_max)
		memcpy(plan_base(q), plan_base(p), blocks_total(p->l_max) * sizeof(double complex));
	return q;
}

/* sh_series_rotation.c:473-513: only reduction is supported; blocks above
 * l_max are released (the tail of the single allocation) */
struct sh_series_rotation_plan *sh_series_rotation_plan_set_l(struct sh_series_rotation_plan *p, unsigned l_max)
{
	if(l_max > p->l_max)
		return NULL;
	if(l_max == p->l_max)
		return p;
	double complex *base = plan_base(p);
	if(l_max == 0) {
		free(base);
		free(p->D);
		p->D = calloc(1, sizeof(*p->D));
		p->l_max = 0;
		return p;
	}
	double complex *nb = realloc(base, blocks_total(l_max) * sizeof(*nb));
	if(!nb)
		nb = base;
	double complex **D = realloc(p->D, (l_max + 1) * sizeof(*D));
	if(D)
		p->D = D;
	for(unsigned l = 1; l <= l_max; l++)
		p->D[l] = nb + block_offset(l) + (2 * l + 1) * l + l;
	p->l_max = l_max;
	return p;
}

/* sh_series_rotation.c:523-529 */
double complex sh_series_rotation_plan_wigner_D(const struct sh_series_rotation_plan *p, unsigned l, int m, int mp)
{
	if(l > p->l_max || (unsigned) abs(m) > l || (unsigned) abs(mp) > l)
		return NAN + I * NAN;
	return l == 0 ? 1.0 : DM(p->D[l], (int) l, m, mp);
}

/* sh_series_rotation.c:546-589 (GPU) */
struct sh_series *sh_series_rotate(struct sh_series *result, const struct sh_series *series, const struct sh_series_rotation_plan *p)
{
	if(result->l_max != series->l_max || result->polar || result == series || p->l_max < series->l_max)
		return NULL;
	if(sphr_gpu_rotate(series->l_max, series->polar, (const double *) series->coeff, (const double *) plan_base(p), (double *) result->coeff))
		return NULL;
	return result;
}

/* sh_series_rotation.c:600-625 (GPU for non-polar series) */
struct sh_series *sh_series_rotate_z(struct sh_series *result, const struct sh_series *series, double phi)
{
	if(result->l_max != series->l_max)
		return NULL;
	if(series->polar) {
		if(result != series)
			sh_series_assign(result, series);
		return result;
	}
	if(result->polar)
		return NULL;
	if(sphr_gpu_rotate_z(series->l_max, (const double *) series->coeff, phi, (double *) result->coeff))
		return NULL;
	return result;
}
