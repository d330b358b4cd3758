/*
 * Products of SH series (sphradiometer src/lib/sh_series_product.c).
 * Below a product order of 60 the plan is the Gaunt "microcode" of the
 * reference (one multiply-add per coupling, sorted by destination, then
 * a, then b offsets); above it the product is formed on the pixel mesh.
 * The Wigner 3j symbols come from the Schulten-Gordon recurrence
 * (numerics.c).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sh_series.h>
#include "../sphr_internal.h"

/* sh_series_product.c:80-83 */
double sh_series_wigner_3j(int ja, int jb, int jc, int ma, int mb, int mc)
{
	if(ma + mb + mc != 0 || abs(ma) > ja || abs(mb) > jb || abs(mc) > jc)
		return 0.0;
	if(jc < abs(ja - jb) || jc > ja + jb)
		return 0.0;
	double *f = malloc((ja + jb + 2) * sizeof(*f));
	if(!f)
		return NAN;
	const int jmin = sphr_3j_range(ja, jb, ma, mb, f);
	const double v = (jmin < 0 || jc < jmin) ? 0.0 : f[jc - jmin];
	free(f);
	return v;
}

/* sh_series_product.c:93-105 */
static int op_cmp(const void *p, const void *q)
{
	const struct _sh_series_product_plan_op *a = p, *b = q;
	if(a->dest_offset != b->dest_offset)
		return a->dest_offset < b->dest_offset ? -1 : 1;
	if(a->a_offset != b->a_offset)
		return a->a_offset < b->a_offset ? -1 : 1;
	return a->b_offset < b->b_offset ? -1 : a->b_offset > b->b_offset;
}

/*
 * The couplings of sh_series_product.c:123-162: every (d, dm) <- (a, am)
 * x (b, bm) with dm = am + bm, |a - b| <= d <= a + b, a + b + d even,
 * d <= Dl, and the polar restrictions.  Enumerated per (a, am, b, bm) so
 * each pair of 3j ranges over d is computed once, then sorted into the
 * reference order.
 */
static int build_microcode(struct sh_series_product_plan *p, unsigned Dl, int Dp, unsigned Al, int Ap, unsigned Bl, int Bp)
{
	size_t cap = 1024, n = 0;
	struct _sh_series_product_plan_op *ops = malloc(cap * sizeof(*ops));
	double *f0 = malloc((Al + Bl + 2) * sizeof(*f0)), *fm = malloc((Al + Bl + 2) * sizeof(*fm));
	if(!ops || !f0 || !fm)
		goto fail;
	const int dm_zero = Dp || (Ap && Bp);
	for(int al = 0; al <= (int) Al; al++) {
		for(int bl = 0; bl <= (int) Bl; bl++) {
			const int j0 = sphr_3j_range(al, bl, 0, 0, f0);
			for(int am = Ap ? 0 : -al; am <= (Ap ? 0 : al); am++) {
				for(int bm = Bp ? 0 : -bl; bm <= (Bp ? 0 : bl); bm++) {
					const int dm = am + bm;
					if(dm_zero && dm != 0)
						continue;
					const int jm = sphr_3j_range(al, bl, am, bm, fm);
					if(jm < 0)
						continue;
					int dlo = abs(al - bl) > abs(dm) ? abs(al - bl) : abs(dm);
					if((al + bl + dlo) & 1)
						dlo++;
					for(int dl = dlo; dl <= al + bl && dl <= (int) Dl; dl += 2) {
						if(n == cap) {
							cap *= 2;
							struct _sh_series_product_plan_op *t = realloc(ops, cap * sizeof(*ops));
							if(!t)
								goto fail;
							ops = t;
						}
						const double g0 = f0[dl - j0], gm = fm[dl - jm];
						ops[n].dest_offset = (int) sh_series_params_lmoffset(Dl, dl, dm);
						ops[n].a_offset = (int) sh_series_params_lmoffset(Al, al, am);
						ops[n].b_offset = (int) sh_series_params_lmoffset(Bl, bl, bm);
						ops[n].factor = ((dm & 1) ? -1.0 : 1.0) * sqrt((2.0 * dl + 1) * (2.0 * al + 1) * (2.0 * bl + 1) / (4 * M_PI)) * g0 * gm;
						n++;
					}
				}
			}
		}
	}
	free(f0);
	free(fm);
	qsort(ops, n, sizeof(*ops), op_cmp);
	p->microcode = n ? ops : (free(ops), malloc(sizeof(*ops)));
	p->plan_length = (int) n;
	return 0;
fail:
	free(ops);
	free(f0);
	free(fm);
	return -1;
}

/* sh_series_product.c:174-225 */
struct sh_series_product_plan *sh_series_product_plan_new(const struct sh_series *dest, const struct sh_series *a, const struct sh_series *b)
{
	struct sh_series_product_plan *p = calloc(1, sizeof(*p));
	if(!p)
		return NULL;
	const unsigned dl = dest->l_max <= a->l_max + b->l_max ? dest->l_max : a->l_max + b->l_max;
	p->a_l_max = a->l_max;
	p->a_polar = a->polar;
	p->b_l_max = b->l_max;
	p->b_polar = b->polar;
	p->dest_l_max = dl;
	p->dest_polar = dest->polar;
	/* the reference's threshold between the two algorithms (line 199) */
	if(dl < 60 && build_microcode(p, dl, dest->polar, a->l_max, a->polar, b->l_max, b->polar)) {
		free(p);
		return NULL;
	}
	return p;
}

void sh_series_product_plan_free(struct sh_series_product_plan *p)
{
	if(p)
		free(p->microcode);
	free(p);
}

/* sh_series_product.c:250-266 */
static struct sh_series *microcode_product(struct sh_series *dest, const struct sh_series *a, const struct sh_series *b, const struct sh_series_product_plan *p)
{
	sh_series_zero(dest);
	for(int i = 0; i < p->plan_length; i++) {
		const struct _sh_series_product_plan_op *op = &p->microcode[i];
		dest->coeff[op->dest_offset] += a->coeff[op->a_offset] * b->coeff[op->b_offset] * op->factor;
	}
	return dest;
}

/*
 * sh_series_product.c:269-329 for polar factors and a polar destination:
 * on the mesh of order L = la + lb the factors are phi-constant, so the
 * product's analysis reduces to the Gauss-Legendre quadrature
 *   c_d = sum_j (2 pi w_j) A(x_j) B(x_j) Pbar_d0(x_j).
 */
static struct sh_series *polar_pixel_product(struct sh_series *dest, const struct sh_series *a, const struct sh_series *b)
{
	const unsigned L = a->l_max + b->l_max;
	const int n = 2 * ((int) L + 1);
	const unsigned Lo = dest->l_max < L ? dest->l_max : L;
	const double *x, *w;
	sphr_mesh_rule(L, &x, &w);
	double *P = malloc((L + 1) * sizeof(*P));
	double complex *c = calloc(L + 1, sizeof(*c));
	if(!P || !c) {
		free(P);
		free(c);
		return NULL;
	}
	for(int j = 0; j < n; j++) {
		sphr_plm(L, 0, x[j], P);
		double complex A = 0.0, B = 0.0;
		for(unsigned l = 0; l <= a->l_max; l++)
			A += a->coeff[l] * P[l];
		for(unsigned l = 0; l <= b->l_max; l++)
			B += b->coeff[l] * P[l];
		const double complex v = (2.0 * M_PI * w[j]) * (A * B);
		for(unsigned l = 0; l <= Lo; l++)
			c[l] += v * P[l];
	}
	struct sh_series tmp = { Lo, 1, c };
	sh_series_assign(dest, &tmp);
	free(P);
	free(c);
	return dest;
}

/* sh_series_product.c:269-329, general case: synthesis on the GPU,
 * analysis on the host */
static struct sh_series *general_pixel_product(struct sh_series *dest, const struct sh_series *a, const struct sh_series *b)
{
	const unsigned L = a->l_max + b->l_max;
	struct sh_series *wa = sh_series_copy(a), *wb = sh_series_copy(b), *wk = NULL;
	double complex *ma = NULL, *mb = NULL;
	struct sh_series *ret = NULL;
	if(!wa || !wb || !sh_series_resize(wa, L) || !sh_series_resize(wb, L))
		goto done;
	ma = sh_series_to_mesh(wa, NULL, NULL, NULL);
	mb = sh_series_to_mesh(wb, NULL, NULL, NULL);
	if(!ma || !mb)
		goto done;
	const size_t nn = (size_t) (2 * (L + 1)) * (2 * (L + 1));
	for(size_t i = 0; i < nn; i++)
		ma[i] *= mb[i];
	wk = sh_series_new(L, a->polar && b->polar);
	if(!wk || !sh_series_from_mesh(wk, ma) || !sh_series_resize(wk, dest->l_max))
		goto done;
	ret = sh_series_assign(dest, wk);
done:
	sh_series_free(wa);
	sh_series_free(wb);
	sh_series_free(wk);
	free(ma);
	free(mb);
	return ret;
}

/* sh_series_product.c:338-349 */
struct sh_series *sh_series_product(struct sh_series *dest, const struct sh_series *a, const struct sh_series *b, const struct sh_series_product_plan *p)
{
	if(p->a_l_max != a->l_max || p->b_l_max != b->l_max || p->dest_l_max > dest->l_max || p->a_polar != a->polar || p->b_polar != b->polar || p->dest_polar != dest->polar)
		return NULL;
	if(p->microcode)
		return microcode_product(dest, a, b, p);
	if(a->polar && b->polar && dest->polar)
		return polar_pixel_product(dest, a, b);
	return general_pixel_product(dest, a, b);
}
