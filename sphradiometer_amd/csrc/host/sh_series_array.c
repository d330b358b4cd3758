/*
 * sh_series_array: n series sharing one coefficient buffer with a fixed
 * stride (sphradiometer src/lib/sh_series_array.c).  The inner products
 * (sh_series_array_dot / _dotc) run on the GPU.
 */
#include <complex.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sh_series.h>
#include "../sphr_internal.h"

static void point_series(struct sh_series_array *a)
{
	for(int i = 0; i < a->n; i++)
		a->series[i] = (struct sh_series) { .l_max = a->l_max, .polar = a->polar, .coeff = a->coeff + (size_t) i * a->stride };
}

/* sh_series_array.c:64-96 */
struct sh_series_array *sh_series_array_new(int n, unsigned int l_max, int polar)
{
	struct sh_series_array *a = malloc(sizeof(*a));
	struct sh_series *s = malloc((n > 0 ? n : 1) * sizeof(*s));
	double complex *c = malloc(((size_t) (n > 0 ? n : 1)) * sh_series_length(l_max, polar) * sizeof(*c));
	if(!a || !s || !c || n < 0) {
		free(a);
		free(s);
		free(c);
		return NULL;
	}
	a->l_max = l_max;
	a->polar = polar ? 1 : 0;
	a->n = n;
	a->series = s;
	a->coeff = c;
	a->stride = (int) sh_series_length(l_max, polar);
	point_series(a);
	return a;
}

void sh_series_array_free(struct sh_series_array *a)
{
	if(a) {
		free(a->coeff);
		free(a->series);
	}
	free(a);
}

/* sh_series_array.c:125-171: extra series are zeroed */
struct sh_series_array *sh_series_array_set_len(struct sh_series_array *a, int n)
{
	if(n == a->n)
		return a;
	double complex *c = realloc(a->coeff, (size_t) (n > 0 ? n : 1) * a->stride * sizeof(*c));
	if(!c)
		return NULL;
	a->coeff = c;
	struct sh_series *s = realloc(a->series, (n > 0 ? n : 1) * sizeof(*s));
	if(!s) {
		point_series(a);
		return NULL;
	}
	a->series = s;
	if(n > a->n)
		memset(a->coeff + (size_t) a->n * a->stride, 0, (size_t) (n - a->n) * a->stride * sizeof(*c));
	a->n = n;
	point_series(a);
	return a;
}

/* re-lay every series out with a new (l_max, polar), through a fresh
 * buffer (sh_series_array.c:181-331 move in place; the result is the
 * same) */
static struct sh_series_array *relayout(struct sh_series_array *a, unsigned l_max, int polar)
{
	const int ns = (int) sh_series_length(l_max, polar);
	double complex *c = malloc((size_t) (a->n > 0 ? a->n : 1) * ns * sizeof(*c));
	if(!c)
		return NULL;
	for(int i = 0; i < a->n; i++) {
		struct sh_series src = { a->l_max, a->polar, a->coeff + (size_t) i * a->stride };
		struct sh_series dst = { l_max, polar, c + (size_t) i * ns };
		sh_series_assign(&dst, &src);
	}
	free(a->coeff);
	a->coeff = c;
	a->l_max = l_max;
	a->polar = polar;
	a->stride = ns;
	point_series(a);
	return a;
}

/* sh_series_array.c:181-253 */
struct sh_series_array *sh_series_array_set_polar(struct sh_series_array *a, int polar)
{
	polar = polar ? 1 : 0;
	if(a->polar == polar)
		return a;
	return relayout(a, a->l_max, polar);
}

/* sh_series_array.c:261-331 */
struct sh_series_array *sh_series_array_set_l(struct sh_series_array *a, unsigned int l_max)
{
	if(a->l_max == l_max)
		return a;
	return relayout(a, l_max, a->polar);
}

/* sh_series_array.c:339-351 */
struct sh_series_array *sh_series_array_copy(const struct sh_series_array *a)
{
	struct sh_series_array *b = sh_series_array_new(a->n, a->l_max, a->polar);
	if(b)
		memcpy(b->coeff, a->coeff, (size_t) a->n * a->stride * sizeof(*b->coeff));
	return b;
}

/* sh_series_array.c:359-371 */
struct sh_series_array *sh_series_array_assign(struct sh_series_array *dst, const struct sh_series_array *src)
{
	if(dst->n != src->n)
		return NULL;
	for(int i = 0; i < dst->n; i++)
		if(!sh_series_assign(&dst->series[i], &src->series[i]))
			return NULL;
	return dst;
}

/* sh_series_array.c:379-387 */
struct sh_series_array *sh_series_array_scale(struct sh_series_array *a, double complex z)
{
	const size_t n = (size_t) a->n * a->stride;
	for(size_t i = 0; i < n; i++)
		a->coeff[i] *= z;
	return a;
}

/* sh_series_array.c:407-426 (GPU) */
struct sh_series *sh_series_array_dot(struct sh_series *r, const struct sh_series_array *a, const double *v)
{
	if(r->l_max != a->l_max || r->polar != a->polar)
		return NULL;
	if(sphr_gpu_dotc((const double *) a->coeff, a->n, a->stride, v, 1, (double *) r->coeff))
		return NULL;
	return r;
}

/* sh_series_array.c:437-456 (GPU): r = sum_i array[i] v[i] */
struct sh_series *sh_series_array_dotc(struct sh_series *r, const struct sh_series_array *a, const double complex *v)
{
	if(r->l_max != a->l_max || r->polar != a->polar)
		return NULL;
	if(sphr_gpu_dotc((const double *) a->coeff, a->n, a->stride, (const double *) v, 0, (double *) r->coeff))
		return NULL;
	return r;
}

/* sh_series_array.c:465-502 */
struct sh_series_array *sh_series_array_window(struct sh_series_array *a, const double *w)
{
	for(int i = 0; i < a->n; i++)
		for(int k = 0; k < a->stride; k++)
			a->coeff[(size_t) i * a->stride + k] *= w[i];
	return a;
}

struct sh_series_array *sh_series_array_windowc(struct sh_series_array *a, const double complex *w)
{
	for(int i = 0; i < a->n; i++)
		for(int k = 0; k < a->stride; k++)
			a->coeff[(size_t) i * a->stride + k] *= w[i];
	return a;
}

/* sh_series_array.c:528-559: unnormalised DFT along the series index,
 * independently for each coefficient (plan-time) */
static struct sh_series_array *array_fft(struct sh_series_array *a, int sign)
{
	double *col = malloc(2 * (size_t) (a->n > 0 ? a->n : 1) * sizeof(*col));
	if(!col)
		return NULL;
	for(int k = 0; k < a->stride; k++) {
		for(int i = 0; i < a->n; i++) {
			col[2 * i] = creal(a->coeff[(size_t) i * a->stride + k]);
			col[2 * i + 1] = cimag(a->coeff[(size_t) i * a->stride + k]);
		}
		sphr_host_dft(col, a->n, sign);
		for(int i = 0; i < a->n; i++)
			a->coeff[(size_t) i * a->stride + k] = col[2 * i] + I * col[2 * i + 1];
	}
	free(col);
	return a;
}

struct sh_series_array *sh_series_array_forward_fft(struct sh_series_array *a)
{
	return array_fft(a, -1);
}

struct sh_series_array *sh_series_array_reverse_fft(struct sh_series_array *a)
{
	return array_fft(a, +1);
}
