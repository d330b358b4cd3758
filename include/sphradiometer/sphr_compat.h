/*
 * Third-party types that appear in sphradiometer's public structs and
 * signatures, supplied here so the headers compile without GSL or FFTW.
 *
 *  - gsl_vector: when <gsl/gsl_vector.h> is available it is used as is;
 *    otherwise a struct with GSL's field order (size, stride, data, block,
 *    owner) is declared, so code reading v->data / gsl_vector_get works
 *    either way.  Vectors this library allocates keep their data in the
 *    same allocation with owner = 0, so GSL's gsl_vector_free releases
 *    them correctly too.
 *  - fftw_plan: the FFT helpers of correlator.h return a plan object of
 *    this library (a GPU FFT plan), declared with FFTW's own typedef
 *    spelling so code that also includes <fftw3.h> still compiles.  It is
 *    released with correlator_fft_plan_free(), not fftw_destroy_plan().
 */
#ifndef SPHR_COMPAT_H
#define SPHR_COMPAT_H

#include <stddef.h>

#if defined(__has_include)
#if __has_include(<gsl/gsl_vector.h>)
#include <gsl/gsl_vector.h>
#define SPHR_HAVE_GSL 1
#endif
#endif

#ifndef SPHR_HAVE_GSL
#ifdef __cplusplus
extern "C" {
#endif
typedef struct {
	size_t size;
	size_t stride;
	double *data;
	void *block;
	int owner;
} gsl_vector;

static inline double gsl_vector_get(const gsl_vector *v, const size_t i)
{
	return v->data[i * v->stride];
}

static inline void gsl_vector_set(gsl_vector *v, const size_t i, double x)
{
	v->data[i * v->stride] = x;
}

gsl_vector *sphr_vector_alloc(size_t n);
void sphr_vector_free(gsl_vector *v);
#define gsl_vector_alloc sphr_vector_alloc
#define gsl_vector_free sphr_vector_free
#ifdef __cplusplus
}
#endif
#endif

typedef struct fftw_plan_s *fftw_plan;

#endif
