/*
 * Oracle (test infrastructure only): sh_series / sh_series_array objects.
 * Restates sh_series.h:172-227, sh_series.c and sh_series_array.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* sh_series.h:172-188 (sh_series_moffset): m-major blocks, m >= 0 first,
 * then m = -L..-1. */
size_t or_moffset(unsigned l_max, int m)
{
	const long L1 = (long) l_max + 1;
	const long x = (long) m * (m - 1);
	if(m >= 0)
		return (size_t) (L1 * m - x / 2);
	return (size_t) (L1 * (L1 + m) + x / 2);
}

/* sh_series.h:197-201 */
size_t or_lmoffset(unsigned l_max, unsigned l, int m)
{
	return or_moffset(l_max, m) + l - (size_t) abs(m);
}

/* sh_series.h:218-226 */
size_t or_length(unsigned l_max, int polar)
{
	return polar ? (size_t) l_max + 1 : ((size_t) l_max + 1) * ((size_t) l_max + 1);
}

/* sh_series.c:51-67 */
struct or_series *or_series_new(unsigned l_max, int polar)
{
	struct or_series *s = malloc(sizeof(*s));
	s->l_max = l_max;
	s->polar = polar ? 1 : 0;
	s->coeff = malloc(or_length(l_max, s->polar) * sizeof(*s->coeff));
	return s;
}

/* sh_series.c:77-93 */
struct or_series *or_series_new_zero(unsigned l_max, int polar)
{
	struct or_series *s = or_series_new(l_max, polar);
	memset(s->coeff, 0, or_length(l_max, s->polar) * sizeof(*s->coeff));
	return s;
}

/* sh_series.c:101-111 */
struct or_series *or_series_copy(const struct or_series *src)
{
	struct or_series *s = or_series_new(src->l_max, src->polar);
	memcpy(s->coeff, src->coeff, or_length(src->l_max, src->polar) * sizeof(*s->coeff));
	return s;
}

void or_series_free(struct or_series *s)
{
	if(s)
		free(s->coeff);
	free(s);
}

/* sh_series.c:241-246 */
struct or_series *or_series_zero(struct or_series *s)
{
	memset(s->coeff, 0, or_length(s->l_max, s->polar) * sizeof(*s->coeff));
	return s;
}

double complex or_get(const struct or_series *s, unsigned l, int m)
{
	return s->coeff[or_lmoffset(s->l_max, l, m)];
}

void or_set(struct or_series *s, unsigned l, int m, double complex v)
{
	s->coeff[or_lmoffset(s->l_max, l, m)] = v;
}

/* sh_series.c:123-184 (sh_series_assign): copy every (l, m) both series
 * hold, zero the rest of dst.  Restated element-wise into a temporary so
 * that overlapping buffers behave like the reference's memmove order. */
struct or_series *or_series_assign(struct or_series *dst, const struct or_series *src)
{
	const size_t n = or_length(dst->l_max, dst->polar);
	double complex *tmp = calloc(n, sizeof(*tmp));
	const int dst_m_max = dst->polar ? 0 : (int) dst->l_max;
	const int src_m_max = src->polar ? 0 : (int) src->l_max;
	const unsigned l_top = dst->l_max < src->l_max ? dst->l_max : src->l_max;
	for(int m = -dst_m_max; m <= dst_m_max; m++) {
		if(abs(m) > src_m_max)
			continue;
		for(unsigned l = abs(m); l <= l_top; l++)
			tmp[or_lmoffset(dst->l_max, l, m)] = src->coeff[or_lmoffset(src->l_max, l, m)];
	}
	memcpy(dst->coeff, tmp, n * sizeof(*tmp));
	free(tmp);
	return dst;
}

/* sh_series.c:200-233 */
struct or_series *or_series_resize(struct or_series *s, unsigned l_max)
{
	struct or_series old = *s;
	struct or_series tmp = { l_max, s->polar, malloc(or_length(l_max, s->polar) * sizeof(double complex)) };
	or_series_assign(&tmp, &old);
	free(old.coeff);
	*s = tmp;
	return s;
}

/* sh_series.c:254-283 */
struct or_series *or_series_set_polar(struct or_series *s, int polar)
{
	polar = polar ? 1 : 0;
	if(s->polar == polar)
		return s;
	struct or_series old = *s;
	struct or_series tmp = { s->l_max, polar, malloc(or_length(s->l_max, polar) * sizeof(double complex)) };
	or_series_assign(&tmp, &old);
	free(old.coeff);
	*s = tmp;
	return s;
}

/* sh_series.c:366-392 */
struct or_series *or_series_add(struct or_series *a, double complex z, const struct or_series *b)
{
	const int m_max = b->polar ? 0 : (int) b->l_max;
	if(a->l_max == b->l_max && a->polar == b->polar) {
		const size_t n = or_length(b->l_max, b->polar);
		for(size_t i = 0; i < n; i++)
			a->coeff[i] += z * b->coeff[i];
		return a;
	}
	if(a->l_max < b->l_max || (a->polar && !b->polar))
		return NULL;
	for(int m = -m_max; m <= m_max; m++)
		for(unsigned l = abs(m); l <= b->l_max; l++)
			a->coeff[or_lmoffset(a->l_max, l, m)] += z * b->coeff[or_lmoffset(b->l_max, l, m)];
	return a;
}

/* sh_series.c:402-411 */
struct or_series *or_series_scale(struct or_series *a, double complex z)
{
	const size_t n = or_length(a->l_max, a->polar);
	for(size_t i = 0; i < n; i++)
		a->coeff[i] *= z;
	return a;
}

/* sh_series.c:441-475: a'_lm = (-1)^m conj(a_l,-m) */
struct or_series *or_series_conj(struct or_series *a)
{
	const int m_max = a->polar ? 0 : (int) a->l_max;
	for(unsigned l = 0; l <= a->l_max; l++)
		or_set(a, l, 0, conj(or_get(a, l, 0)));
	for(int m = 1; m <= m_max; m++)
		for(unsigned l = m; l <= a->l_max; l++) {
			double complex ap = or_get(a, l, m), an = or_get(a, l, -m);
			double sg = (m & 1) ? -1. : 1.;
			or_set(a, l, m, sg * conj(an));
			or_set(a, l, -m, sg * conj(ap));
		}
	return a;
}

/* sh_series.c:487-521 */
struct or_series *or_series_real(struct or_series *a)
{
	const int m_max = a->polar ? 0 : (int) a->l_max;
	for(unsigned l = 0; l <= a->l_max; l++)
		or_set(a, l, 0, creal(or_get(a, l, 0)));
	for(int m = 1; m <= m_max; m++)
		for(unsigned l = m; l <= a->l_max; l++) {
			double sg = (m & 1) ? -1. : 1.;
			double complex c = (or_get(a, l, m) + conj(or_get(a, l, -m)) * sg) * 0.5;
			or_set(a, l, m, c);
			or_set(a, l, -m, sg * conj(c));
		}
	return a;
}

/* sh_series.c:608-622 */
struct or_series *or_series_scale_power_l(struct or_series *a, double power, double overall)
{
	const int polar = a->polar;
	for(unsigned l = 0; l <= a->l_max; l++) {
		double f = overall * pow((double) l / a->l_max, power);
		int mm = polar ? 0 : (int) l;
		for(int m = -mm; m <= mm; m++)
			or_set(a, l, m, f * or_get(a, l, m));
	}
	return a;
}

/* diagnostics.c:145-175 with sh_series_dot (sh_series.c:645-676).  The
 * reference sorts the products with a comparator that never returns -1,
 * so its summation order is unspecified; a plain sum of |d|^2 is used. */
double or_rms_error(const struct or_series *measured, const struct or_series *exact)
{
	unsigned l = measured->l_max > exact->l_max ? measured->l_max : exact->l_max;
	struct or_series *m = or_series_copy(measured), *e = or_series_copy(exact);
	or_series_resize(m, l);
	or_series_resize(e, l);
	if(!(measured->polar && exact->polar)) {
		or_series_set_polar(m, 0);
		or_series_set_polar(e, 0);
	}
	or_series_add(e, -1., m);
	double acc = 0.;
	const size_t n = or_length(e->l_max, e->polar);
	for(size_t i = 0; i < n; i++)
		acc += creal(e->coeff[i] * conj(e->coeff[i]));
	or_series_free(m);
	or_series_free(e);
	return sqrt(acc / (4. * M_PI));
}

/*
 * sh_series_array (sh_series_array.c)
 */

/* sh_series_array.c:64-96 */
struct or_array *or_array_new(int n, unsigned l_max, int polar)
{
	struct or_array *a = malloc(sizeof(*a));
	a->l_max = l_max;
	a->polar = polar ? 1 : 0;
	a->n = n;
	a->stride = (int) or_length(l_max, a->polar);
	a->coeff = calloc((size_t) n * a->stride, sizeof(*a->coeff));
	return a;
}

void or_array_free(struct or_array *a)
{
	if(a)
		free(a->coeff);
	free(a);
}

void or_array_series(const struct or_array *a, int i, struct or_series *view)
{
	view->l_max = a->l_max;
	view->polar = a->polar;
	view->coeff = a->coeff + (size_t) i * a->stride;
}

/* sh_series_array.c:125-171: new trailing series are zero */
struct or_array *or_array_set_len(struct or_array *a, int n)
{
	if(n == a->n)
		return a;
	a->coeff = realloc(a->coeff, (size_t) n * a->stride * sizeof(*a->coeff));
	if(n > a->n)
		memset(a->coeff + (size_t) a->n * a->stride, 0, (size_t) (n - a->n) * a->stride * sizeof(*a->coeff));
	a->n = n;
	return a;
}

/* sh_series_array.c:261-331: every series is re-laid out at the new l_max */
struct or_array *or_array_set_l(struct or_array *a, unsigned l_max)
{
	if(l_max == a->l_max)
		return a;
	const int new_stride = (int) or_length(l_max, a->polar);
	double complex *c = malloc((size_t) a->n * new_stride * sizeof(*c));
	for(int i = 0; i < a->n; i++) {
		struct or_series src = { a->l_max, a->polar, a->coeff + (size_t) i * a->stride };
		struct or_series dst = { l_max, a->polar, c + (size_t) i * new_stride };
		or_series_assign(&dst, &src);
	}
	free(a->coeff);
	a->coeff = c;
	a->l_max = l_max;
	a->stride = new_stride;
	return a;
}

/* sh_series_array.c:379-387 */
struct or_array *or_array_scale(struct or_array *a, double complex z)
{
	const size_t n = (size_t) a->n * a->stride;
	for(size_t i = 0; i < n; i++)
		a->coeff[i] *= z;
	return a;
}

/* sh_series_array.c:437-456: result = sum_i array[i] * v[i], outer loop
 * over i, inner loop over the stride, sequential accumulation. */
struct or_series *or_array_dotc(struct or_series *r, const struct or_array *a, const double complex *v)
{
	if(r->l_max != a->l_max || r->polar != a->polar)
		return NULL;
	or_series_zero(r);
	const double complex *c = a->coeff;
	for(int i = 0; i < a->n; i++) {
		const double complex vi = v[i];
		for(int k = 0; k < a->stride; k++)
			r->coeff[k] += *c++ * vi;
	}
	return r;
}

/* sh_series_array.c:528-539: unnormalised forward DFT along the series
 * index, independently for every coefficient. */
struct or_array *or_array_forward_fft(struct or_array *a)
{
	double complex *col = malloc((size_t) a->n * sizeof(*col));
	for(int k = 0; k < a->stride; k++) {
		for(int i = 0; i < a->n; i++)
			col[i] = a->coeff[(size_t) i * a->stride + k];
		or_dft(col, a->n, -1);
		for(int i = 0; i < a->n; i++)
			a->coeff[(size_t) i * a->stride + k] = col[i];
	}
	free(col);
	return a;
}
