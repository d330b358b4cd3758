/*
 * projection.h -- SH expansion of the sinc delay kernel.  Drop-in for
 * sphradiometer's src/include/sphradiometer/projection.h (lines 58-79).
 */
#ifndef __RADIOMETER_PROJECTION_H__
#define __RADIOMETER_PROJECTION_H__

#include <sphradiometer/sphr_compat.h>
#include <sphradiometer/sh_series.h>

#ifdef __cplusplus
extern "C" {
#endif

struct projection_delay_element_data {
	const gsl_vector *r;
	double delta_t;
	int j, k;
};

int projection_matrix_n_elements(double, double);
unsigned int projection_matrix_l_max(double, double);

double projection_delay_element(double, double, void *);
struct sh_series_array *projection_matrix_delay(unsigned int, unsigned int, const gsl_vector *, double);

#ifdef __cplusplus
}
#endif

#endif
