/*
 * Oracle (test infrastructure only): products of SH series, restating
 * sh_series_product.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* sh_series_product.c:93-105 */
static int op_cmp(const void *p1, const void *p2)
{
	const struct or_product_op *a = p1, *b = p2;
	int r = a->dest_offset - b->dest_offset;
	if(!r) {
		r = a->a_offset - b->a_offset;
		if(!r)
			r = a->b_offset - b->b_offset;
	}
	return r;
}

/* sh_series_product.c:123-162: Gaunt microcode
 * factor = (-1)^{d_m} sqrt((2d+1)(2a+1)(2b+1)/4pi) (a b d;0 0 0)(a b d;a_m b_m -d_m) */
static int product_plan(struct or_product_op *op0, int Dl, int Dp, int Al, int Ap, int Bl, int Bp)
{
	struct or_product_op *op = op0;
	for(int dl = 0; dl <= Dl; dl++) {
		int dm_max = Dp || (Ap && Bp) ? 0 : dl;
		for(int dm = -dm_max; dm <= dm_max; dm++) {
			int al_min = dl > Bl ? dl - Bl : 0;
			int al_max = Al < dl + Bl ? Al : dl + Bl;
			for(int al = al_min; al <= al_max; al++) {
				int bl_max = Bl < dl + al ? Bl : dl + al;
				int bl_min = abs(dl - al);
				bl_min += (al + bl_min + dl) & 1;
				for(int bl = bl_min; bl <= bl_max; bl += 2) {
					int am_min = -(al < bl - dm ? al : bl - dm);
					int am_max = al < bl + dm ? al : bl + dm;
					if(Ap) {
						if(am_min > 0 || am_max < 0)
							continue;
						am_min = am_max = 0;
					}
					for(int am = am_min; am <= am_max; am++) {
						int bm = dm - am;
						if(Bp && bm != 0)
							continue;
						op->dest_offset = (int) or_lmoffset(Dl, dl, dm);
						op->a_offset = (int) or_lmoffset(Al, al, am);
						op->b_offset = (int) or_lmoffset(Bl, bl, bm);
						op->factor = ((dm & 1) ? -1. : 1.) * sqrt((2. * dl + 1) * (2. * al + 1) * (2. * bl + 1) / (4 * M_PI)) * or_wigner_3j(al, bl, dl, 0, 0, 0) * or_wigner_3j(al, bl, dl, am, bm, -dm);
						op++;
					}
				}
			}
		}
	}
	return (int) (op - op0);
}

/* sh_series_product.c:174-225; microcode iff dest_l_max < 60 (line 199) */
struct or_product_plan *or_product_plan_new_forced(const struct or_series *dest, const struct or_series *a, const struct or_series *b, int microcode)
{
	struct or_product_plan *p = calloc(1, sizeof(*p));
	const unsigned dl = dest->l_max <= a->l_max + b->l_max ? dest->l_max : a->l_max + b->l_max;
	p->a_l_max = a->l_max;
	p->a_polar = a->polar;
	p->b_l_max = b->l_max;
	p->b_polar = b->polar;
	p->dest_l_max = dl;
	p->dest_polar = dest->polar;
	if(microcode) {
		size_t worst = or_length(dl, dest->polar) * or_length(a->l_max, a->polar) * or_length(b->l_max, b->polar);
		p->microcode = malloc(worst * sizeof(*p->microcode));
		p->plan_length = product_plan(p->microcode, dl, dest->polar, a->l_max, a->polar, b->l_max, b->polar);
		qsort(p->microcode, p->plan_length, sizeof(*p->microcode), op_cmp);
	}
	return p;
}

struct or_product_plan *or_product_plan_new(const struct or_series *dest, const struct or_series *a, const struct or_series *b)
{
	const unsigned dl = dest->l_max <= a->l_max + b->l_max ? dest->l_max : a->l_max + b->l_max;
	return or_product_plan_new_forced(dest, a, b, dl < 60);
}

void or_product_plan_free(struct or_product_plan *p)
{
	if(p)
		free(p->microcode);
	free(p);
}

/* sh_series_product.c:250-266 */
static struct or_series *frequency_domain_product(struct or_series *dest, const struct or_series *a, const struct or_series *b, const struct or_product_plan *p)
{
	or_series_zero(dest);
	for(int i = 0; i < p->plan_length; i++) {
		const struct or_product_op *op = &p->microcode[i];
		dest->coeff[op->dest_offset] += a->coeff[op->a_offset] * b->coeff[op->b_offset] * op->factor;
	}
	return dest;
}

/* sh_series_product.c:269-329 for two polar factors.  to_mesh of a polar
 * series has phi-constant rows (only the m = 0 column of the phi DFT
 * input is non-zero, so every output column equals it exactly), the
 * pointwise product is phi-constant, and from_mesh of a polar destination
 * reads only the phi-DFT's DC term: nphi copies of the row value, scaled
 * by w_i 2pi/nphi.  This evaluates exactly those quantities without the
 * redundant phi columns. */
static struct or_series *pixel_product_polar(struct or_series *dest, const struct or_series *a, const struct or_series *b)
{
	const unsigned L = a->l_max + b->l_max;
	/* the mesh nodes and Pbar_l0 table depend only on L: cache them
	 * across the per-frequency calls of the plan build */
	static unsigned cL = (unsigned) -1;
	static int nt, np;
	static double *x, *w, *Ptab;
	if(cL != L) {
		free(x);
		free(w);
		free(Ptab);
		x = malloc(2 * (L + 1) * sizeof(*x));
		w = malloc(2 * (L + 1) * sizeof(*w));
		or_mesh_nodes(L, &nt, &np, x, w);
		Ptab = malloc((size_t) nt * (L + 1) * sizeof(*Ptab));
		for(int i = 0; i < nt; i++)
			or_sphPlm_array(L, 0, x[i], Ptab + (size_t) i * (L + 1));
		cL = L;
	}
	const unsigned Lout = dest->l_max < L ? dest->l_max : L;
	double complex *c = calloc(L + 1, sizeof(*c));
	for(int i = 0; i < nt; i++) {
		const double *P = Ptab + (size_t) i * (L + 1);
		double complex A = 0., B = 0.;
		for(unsigned l = 0; l <= a->l_max; l++)
			A += a->coeff[l] * P[l];
		for(unsigned l = 0; l <= b->l_max; l++)
			B += b->coeff[l] * P[l];
		const double complex F0 = (A * B * (double) np) * (w[i] * (2. * M_PI / np));
		for(unsigned l = 0; l <= Lout; l++)
			c[l] += F0 * P[l];
	}
	struct or_series tmp = { Lout, 1, c };
	or_series_assign(dest, &tmp);
	free(c);
	return dest;
}

/* sh_series_product.c:269-329, general case */
static struct or_series *pixel_product_general(struct or_series *dest, const struct or_series *a, const struct or_series *b)
{
	const unsigned L = a->l_max + b->l_max;
	struct or_series *wa = or_series_copy(a), *wb = or_series_copy(b);
	or_series_resize(wa, L);
	or_series_resize(wb, L);
	int nt, np;
	double complex *am = or_to_mesh(wa, &nt, &np);
	double complex *bm = or_to_mesh(wb, NULL, NULL);
	for(size_t i = 0; i < (size_t) nt * np; i++)
		am[i] *= bm[i];
	struct or_series *wk = or_series_new(L, a->polar && b->polar);
	or_from_mesh(wk, am);
	or_series_resize(wk, dest->l_max);
	or_series_assign(dest, wk);
	or_series_free(wk);
	free(am);
	free(bm);
	or_series_free(wa);
	or_series_free(wb);
	return dest;
}

/* sh_series_product.c:338-349 */
struct or_series *or_product(struct or_series *dest, const struct or_series *a, const struct or_series *b, const struct or_product_plan *p)
{
	if(p->a_l_max != a->l_max || p->b_l_max != b->l_max || p->dest_l_max > dest->l_max || p->a_polar != a->polar || p->b_polar != b->polar || p->dest_polar != dest->polar)
		return NULL;
	if(p->microcode)
		return frequency_domain_product(dest, a, b, p);
	if(a->polar && b->polar && dest->polar)
		return pixel_product_polar(dest, a, b);
	return pixel_product_general(dest, a, b);
}
