/*
 * sh_series objects and their element-wise arithmetic (host side).
 * Same semantics as sphradiometer's src/lib/sh_series.c; the coefficient
 * order is the one fixed by the public offset helpers in sh_series.h.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sh_series.h>

/* sh_series.c:51-67 */
struct sh_series *sh_series_new(unsigned int l_max, int polar)
{
	struct sh_series *s = malloc(sizeof(*s));
	double complex *c = malloc(sh_series_length(l_max, polar) * sizeof(*c));
	if(!s || !c) {
		free(s);
		free(c);
		return NULL;
	}
	s->l_max = l_max;
	s->polar = polar ? 1 : 0;
	s->coeff = c;
	return s;
}

/* sh_series.c:77-93 */
struct sh_series *sh_series_new_zero(unsigned int l_max, int polar)
{
	struct sh_series *s = malloc(sizeof(*s));
	double complex *c = calloc(sh_series_length(l_max, polar), sizeof(*c));
	if(!s || !c) {
		free(s);
		free(c);
		return NULL;
	}
	s->l_max = l_max;
	s->polar = polar ? 1 : 0;
	s->coeff = c;
	return s;
}

/* sh_series.c:101-111 */
struct sh_series *sh_series_copy(const struct sh_series *src)
{
	struct sh_series *s = sh_series_new(src->l_max, src->polar);
	if(s)
		memcpy(s->coeff, src->coeff, sh_series_length(src->l_max, src->polar) * sizeof(*s->coeff));
	return s;
}

/*
 * sh_series.c:123-184.  Copy every coefficient both series hold and zero
 * the rest of dst.  m blocks are moved one at a time in an order that is
 * safe when dst and src share a buffer and dst's blocks lie at or below
 * src's (shrinking) or at or above them (growing), which is how resize
 * and the array re-layouts use it.
 */
struct sh_series *sh_series_assign(struct sh_series *dst, const struct sh_series *src)
{
	const int dst_mm = dst->polar ? 0 : (int) dst->l_max;
	const int both = (src->polar || dst->polar) ? 0 : (int) (src->l_max < dst->l_max ? src->l_max : dst->l_max);
	const unsigned lc = src->l_max < dst->l_max ? src->l_max : dst->l_max;
	if(dst->l_max == src->l_max && (dst->polar || !src->polar)) {
		memmove(dst->coeff, src->coeff, sh_series_length(dst->l_max, dst->polar) * sizeof(*dst->coeff));
		return dst;
	}
	/* block of m in a series: start offset, length */
#define BLK(s, m) ((s)->coeff + sh_series_moffset((s)->l_max, (m)))
#define BLEN(L, m) ((L) + 1 - (unsigned) abs(m))
	if(dst->l_max <= src->l_max) {
		/* destination blocks are no longer than the source's and sit
		 * at or below them: ascending address order */
		for(int m = 0; m <= dst_mm; m++) {
			if(m <= both)
				memmove(BLK(dst, m), BLK(src, m), BLEN(lc, m) * sizeof(double complex));
			else
				memset(BLK(dst, m), 0, BLEN(dst->l_max, m) * sizeof(double complex));
			if(m <= both && dst->l_max > lc)
				memset(BLK(dst, m) + BLEN(lc, m), 0, (dst->l_max - lc) * sizeof(double complex));
		}
		for(int m = -dst_mm; m < 0; m++) {
			if(-m <= both)
				memmove(BLK(dst, m), BLK(src, m), BLEN(lc, m) * sizeof(double complex));
			else
				memset(BLK(dst, m), 0, BLEN(dst->l_max, m) * sizeof(double complex));
		}
	} else {
		/* growing: descending address order */
		for(int m = -1; m >= -dst_mm; m--) {
			if(-m <= both) {
				memmove(BLK(dst, m), BLK(src, m), BLEN(lc, m) * sizeof(double complex));
				memset(BLK(dst, m) + BLEN(lc, m), 0, (dst->l_max - lc) * sizeof(double complex));
			} else
				memset(BLK(dst, m), 0, BLEN(dst->l_max, m) * sizeof(double complex));
		}
		for(int m = dst_mm; m >= 0; m--) {
			if(m <= both) {
				memmove(BLK(dst, m), BLK(src, m), BLEN(lc, m) * sizeof(double complex));
				memset(BLK(dst, m) + BLEN(lc, m), 0, (dst->l_max - lc) * sizeof(double complex));
			} else
				memset(BLK(dst, m), 0, BLEN(dst->l_max, m) * sizeof(double complex));
		}
	}
#undef BLK
#undef BLEN
	return dst;
}

/* sh_series.c:200-233 */
struct sh_series *sh_series_resize(struct sh_series *s, unsigned int l_max)
{
	struct sh_series old = *s;
	if(l_max == s->l_max)
		return s;
	s->l_max = l_max;
	if(old.l_max > l_max) {
		sh_series_assign(s, &old);
		double complex *c = realloc(s->coeff, sh_series_length(l_max, s->polar) * sizeof(*c));
		if(c)
			s->coeff = c;
		return s;
	}
	double complex *c = realloc(s->coeff, sh_series_length(l_max, s->polar) * sizeof(*c));
	if(!c) {
		s->l_max = old.l_max;
		return NULL;
	}
	old.coeff = s->coeff = c;
	return sh_series_assign(s, &old);
}

/* sh_series.c:241-246 */
struct sh_series *sh_series_zero(struct sh_series *s)
{
	memset(s->coeff, 0, sh_series_length(s->l_max, s->polar) * sizeof(*s->coeff));
	return s;
}

/* sh_series.c:254-283 */
struct sh_series *sh_series_set_polar(struct sh_series *s, int polar)
{
	polar = polar ? 1 : 0;
	if(s->polar == polar)
		return s;
	const size_t n_old = sh_series_length(s->l_max, s->polar), n_new = sh_series_length(s->l_max, polar);
	double complex *c = realloc(s->coeff, n_new * sizeof(*c));
	if(!c)
		return NULL;
	s->coeff = c;
	if(!polar)
		memset(c + n_old, 0, (n_new - n_old) * sizeof(*c));
	s->polar = polar;
	return s;
}

void sh_series_free(struct sh_series *s)
{
	if(s)
		free(s->coeff);
	free(s);
}

double complex sh_series_get(const struct sh_series *s, unsigned int l, int m)
{
	return s->coeff[sh_series_lmoffset(s, l, m)];
}

double complex sh_series_set(struct sh_series *s, unsigned int l, int m, double complex v)
{
	return s->coeff[sh_series_lmoffset(s, l, m)] = v;
}

int sh_series_is_nan(const struct sh_series *s)
{
	const size_t n = sh_series_length(s->l_max, s->polar);
	for(size_t i = 0; i < n; i++)
		if(isnan(creal(s->coeff[i])) || isnan(cimag(s->coeff[i])))
			return 1;
	return 0;
}

int sh_series_is_inf(const struct sh_series *s)
{
	const size_t n = sh_series_length(s->l_max, s->polar);
	for(size_t i = 0; i < n; i++)
		if(isinf(creal(s->coeff[i])) || isinf(cimag(s->coeff[i])))
			return 1;
	return 0;
}

/* sh_series.c:366-392: a += z b; a must hold every coefficient of b */
struct sh_series *sh_series_add(struct sh_series *a, const double complex z, const struct sh_series *b)
{
	if(a->l_max == b->l_max && a->polar == b->polar) {
		const size_t n = sh_series_length(b->l_max, b->polar);
		for(size_t i = 0; i < n; i++)
			a->coeff[i] += z * b->coeff[i];
		return a;
	}
	if(a->l_max < b->l_max || (a->polar && !b->polar))
		return NULL;
	const int mm = b->polar ? 0 : (int) b->l_max;
	for(int m = -mm; m <= mm; m++) {
		double complex *d = a->coeff + sh_series_moffset(a->l_max, m);
		const double complex *s = b->coeff + sh_series_moffset(b->l_max, m);
		const unsigned len = b->l_max + 1 - abs(m);
		for(unsigned i = 0; i < len; i++)
			d[i] += z * s[i];
	}
	return a;
}

struct sh_series *sh_series_scale(struct sh_series *a, double complex z)
{
	const size_t n = sh_series_length(a->l_max, a->polar);
	for(size_t i = 0; i < n; i++)
		a->coeff[i] *= z;
	return a;
}

struct sh_series *sh_series_neg(struct sh_series *a)
{
	const size_t n = sh_series_length(a->l_max, a->polar);
	for(size_t i = 0; i < n; i++)
		a->coeff[i] = -a->coeff[i];
	return a;
}

/* sh_series.c:441-475: conj f has coefficients (-1)^m conj(a_{l,-m}) */
struct sh_series *sh_series_conj(struct sh_series *a)
{
	const int mm = a->polar ? 0 : (int) a->l_max;
	double complex *c0 = a->coeff;
	for(unsigned l = 0; l <= a->l_max; l++)
		c0[l] = conj(c0[l]);
	for(int m = 1; m <= mm; m++) {
		double complex *p = a->coeff + sh_series_moffset(a->l_max, m);
		double complex *n = a->coeff + sh_series_moffset(a->l_max, -m);
		const double sg = (m & 1) ? -1.0 : 1.0;
		for(unsigned i = 0; i <= a->l_max - m; i++) {
			const double complex vp = p[i], vn = n[i];
			p[i] = sg * conj(vn);
			n[i] = sg * conj(vp);
		}
	}
	return a;
}

/* sh_series.c:487-521 */
struct sh_series *sh_series_real(struct sh_series *a)
{
	const int mm = a->polar ? 0 : (int) a->l_max;
	for(unsigned l = 0; l <= a->l_max; l++)
		a->coeff[l] = creal(a->coeff[l]);
	for(int m = 1; m <= mm; m++) {
		double complex *p = a->coeff + sh_series_moffset(a->l_max, m);
		double complex *n = a->coeff + sh_series_moffset(a->l_max, -m);
		const double sg = (m & 1) ? -1.0 : 1.0;
		for(unsigned i = 0; i <= a->l_max - m; i++) {
			const double complex c = 0.5 * (p[i] + sg * conj(n[i]));
			p[i] = c;
			n[i] = sg * conj(c);
		}
	}
	return a;
}

/* sh_series.c:533-567 */
struct sh_series *sh_series_imag(struct sh_series *a)
{
	const int mm = a->polar ? 0 : (int) a->l_max;
	for(unsigned l = 0; l <= a->l_max; l++)
		a->coeff[l] = cimag(a->coeff[l]);
	for(int m = 1; m <= mm; m++) {
		double complex *p = a->coeff + sh_series_moffset(a->l_max, m);
		double complex *n = a->coeff + sh_series_moffset(a->l_max, -m);
		const double sg = (m & 1) ? -1.0 : 1.0;
		for(unsigned i = 0; i <= a->l_max - m; i++) {
			const double complex c = -0.5 * I * (p[i] - sg * conj(n[i]));
			p[i] = c;
			n[i] = sg * conj(c);
		}
	}
	return a;
}

/* sh_series.c:576-600 */
struct sh_series *sh_series_clip(struct sh_series *s, double epsilon)
{
	const size_t n = sh_series_length(s->l_max, s->polar);
	double thr = 0.0;
	for(size_t i = 0; i < n; i++)
		thr = fmax(thr, cabs(s->coeff[i]));
	if(thr == 0.0)
		return s;
	thr *= fabs(epsilon);
	for(size_t i = 0; i < n; i++)
		if(cabs(s->coeff[i]) < thr)
			s->coeff[i] = 0.0;
	return s;
}

/* sh_series.c:608-622: a_lm *= overall (l / l_max)^power */
struct sh_series *sh_series_scale_power_l(struct sh_series *a, double power, double overall)
{
	const int mm = a->polar ? 0 : (int) a->l_max;
	for(unsigned l = 0; l <= a->l_max; l++) {
		const double f = overall * pow((double) l / a->l_max, power);
		const int top = (int) l < mm ? (int) l : mm;
		for(int m = -top; m <= top; m++)
			a->coeff[sh_series_lmoffset(a, l, m)] *= f;
	}
	return a;
}

/*
 * sh_series.c:645-676: integral of a conj(b) over the sphere.  The
 * reference sums the products after a sort meant to order them by
 * magnitude; here they are summed in ascending-magnitude order too (a
 * well-defined comparator), which is the accuracy the sort aims for.
 */
static int cmp_abs(const void *p, const void *q)
{
	const double a = cabs(*(const double complex *) p), b = cabs(*(const double complex *) q);
	return a < b ? -1 : a > b ? 1 : 0;
}

double complex sh_series_dot(const struct sh_series *a, const struct sh_series *b)
{
	if(b->l_max != a->l_max || b->polar != a->polar)
		return NAN + I * NAN;
	const size_t n = sh_series_length(a->l_max, a->polar);
	double complex *t = malloc(n * sizeof(*t));
	if(!t)
		return NAN + I * NAN;
	for(size_t i = 0; i < n; i++)
		t[i] = a->coeff[i] * conj(b->coeff[i]);
	qsort(t, n, sizeof(*t), cmp_abs);
	double complex s = 0.0;
	for(size_t i = 0; i < n; i++)
		s += t[i];
	free(t);
	return s;
}
