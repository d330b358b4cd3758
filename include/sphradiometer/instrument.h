/*
 * instrument.h -- detector phase centres.  Drop-in for sphradiometer's
 * src/include/sphradiometer/instrument.h (structs at lines 59-80,
 * prototypes 91-116).  Positions are in light-seconds.  The LAL
 * constructors (instrument_new_from_LALDetector / _from_name) are not part
 * of this library: they need LAL's detector tables.
 */
#ifndef __RADIOMETER_INSTRUMENT_H__
#define __RADIOMETER_INSTRUMENT_H__

#include <sphradiometer/sphr_compat.h>

#ifdef __cplusplus
extern "C" {
#endif

struct instrument {
	gsl_vector *phase_centre;
	/* caller's data; never freed or read by the library */
	void *data;
};

struct instrument_array {
	struct instrument **instruments;
	int n;
};

double vector_magnitude(const gsl_vector *);
double vector_r_dot_s(const gsl_vector *, double, double);
void vector_direction(const gsl_vector *, double *, double *);

struct instrument *instrument_new(double, double, double, void *);
void instrument_free(struct instrument *);
struct instrument *instrument_copy(const struct instrument *);
struct instrument *instrument_new_from_r_theta_phi(double, double, double, void *);

struct instrument_array *instrument_array_new(int);
int instrument_array_len(const struct instrument_array *);
struct instrument *instrument_array_get(const struct instrument_array *, int);
struct instrument *instrument_array_set(struct instrument_array *, int, struct instrument *);
struct instrument *instrument_array_append(struct instrument_array *, struct instrument *);
void instrument_array_free(struct instrument_array *);
struct instrument_array *instrument_array_copy(const struct instrument_array *);

double instrument_r(const struct instrument *);
gsl_vector *instrument_baseline(const struct instrument *, const struct instrument *);

#ifdef __cplusplus
}
#endif

#endif
