/*
 * Detector phase centres and baselines (sphradiometer src/lib/
 * instrument.c, lines 60-336), positions in light-seconds.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/instrument.h>

#ifndef SPHR_HAVE_GSL
/* vector and data in one allocation with owner = 0, so both this free
 * and GSL's gsl_vector_free release it */
gsl_vector *sphr_vector_alloc(size_t n)
{
	gsl_vector *v = malloc(sizeof(*v) + n * sizeof(double));
	if(!v)
		return NULL;
	v->size = n;
	v->stride = 1;
	v->data = (double *) (v + 1);
	v->block = NULL;
	v->owner = 0;
	return v;
}

void sphr_vector_free(gsl_vector *v)
{
	free(v);
}
#endif

static gsl_vector *vec3(double x, double y, double z)
{
	gsl_vector *v = gsl_vector_alloc(3);
	if(v) {
		gsl_vector_set(v, 0, x);
		gsl_vector_set(v, 1, y);
		gsl_vector_set(v, 2, z);
	}
	return v;
}

/* instrument.c:60-67 */
double vector_magnitude(const gsl_vector *v)
{
	const double x = gsl_vector_get(v, 0), y = gsl_vector_get(v, 1), z = gsl_vector_get(v, 2);
	return sqrt(x * x + y * y + z * z);
}

/* instrument.c:77-80: r . s-hat(theta, phi) */
double vector_r_dot_s(const gsl_vector *v, double theta, double phi)
{
	return (gsl_vector_get(v, 0) * cos(phi) + gsl_vector_get(v, 1) * sin(phi)) * sin(theta) + gsl_vector_get(v, 2) * cos(theta);
}

/* instrument.c:90-100: theta in [0, pi], phi in [0, 2 pi) */
void vector_direction(const gsl_vector *v, double *theta, double *phi)
{
	const double x = gsl_vector_get(v, 0), y = gsl_vector_get(v, 1), z = gsl_vector_get(v, 2);
	*theta = atan2(sqrt(x * x + y * y), z);
	*phi = atan2(y, x);
	if(*phi < 0)
		*phi += 2. * M_PI;
}

struct instrument *instrument_new(double x, double y, double z, void *data)
{
	struct instrument *i = malloc(sizeof(*i));
	gsl_vector *pc = vec3(x, y, z);
	if(!i || !pc) {
		free(i);
		if(pc)
			gsl_vector_free(pc);
		return NULL;
	}
	i->phase_centre = pc;
	i->data = data;
	return i;
}

void instrument_free(struct instrument *i)
{
	if(i && i->phase_centre)
		gsl_vector_free(i->phase_centre);
	free(i);
}

struct instrument *instrument_copy(const struct instrument *i)
{
	return instrument_new(gsl_vector_get(i->phase_centre, 0), gsl_vector_get(i->phase_centre, 1), gsl_vector_get(i->phase_centre, 2), i->data);
}

/* instrument.c:164-172 */
struct instrument *instrument_new_from_r_theta_phi(double r, double theta, double phi, void *data)
{
	return instrument_new(r * sin(theta) * cos(phi), r * sin(theta) * sin(phi), r * cos(theta), data);
}

double instrument_r(const struct instrument *i)
{
	return vector_magnitude(i->phase_centre);
}

/* instrument.c:210-218: a - b */
gsl_vector *instrument_baseline(const struct instrument *a, const struct instrument *b)
{
	if(!a || !b)
		return NULL;
	return vec3(gsl_vector_get(a->phase_centre, 0) - gsl_vector_get(b->phase_centre, 0),
		gsl_vector_get(a->phase_centre, 1) - gsl_vector_get(b->phase_centre, 1),
		gsl_vector_get(a->phase_centre, 2) - gsl_vector_get(b->phase_centre, 2));
}

struct instrument_array *instrument_array_new(int n)
{
	struct instrument_array *a = malloc(sizeof(*a));
	struct instrument **v = calloc(n > 0 ? n : 1, sizeof(*v));
	if(!a || !v || n < 0) {
		free(a);
		free(v);
		return NULL;
	}
	a->instruments = v;
	a->n = n;
	return a;
}

int instrument_array_len(const struct instrument_array *a)
{
	return a->n;
}

/* borrowed reference */
struct instrument *instrument_array_get(const struct instrument_array *a, int k)
{
	return 0 <= k && k < a->n ? a->instruments[k] : NULL;
}

/* takes ownership; an out-of-range index frees the instrument and
 * returns NULL (the reference falls through to an out-of-range write
 * there, instrument.c:268-279) */
struct instrument *instrument_array_set(struct instrument_array *a, int k, struct instrument *i)
{
	if(k < 0 || k >= a->n) {
		instrument_free(i);
		return NULL;
	}
	instrument_free(a->instruments[k]);
	a->instruments[k] = i;
	return i;
}

struct instrument *instrument_array_append(struct instrument_array *a, struct instrument *i)
{
	struct instrument **v = realloc(a->instruments, (a->n + 1) * sizeof(*v));
	if(!v) {
		instrument_free(i);
		return NULL;
	}
	a->instruments = v;
	v[a->n++] = i;
	return i;
}

void instrument_array_free(struct instrument_array *a)
{
	if(a) {
		for(int k = 0; k < a->n; k++)
			instrument_free(a->instruments[k]);
		free(a->instruments);
	}
	free(a);
}

struct instrument_array *instrument_array_copy(const struct instrument_array *a)
{
	struct instrument_array *b = instrument_array_new(a->n);
	if(!b)
		return NULL;
	for(int k = 0; k < a->n; k++)
		b->instruments[k] = a->instruments[k] ? instrument_copy(a->instruments[k]) : NULL;
	return b;
}
