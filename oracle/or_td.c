/*
 * Oracle (test infrastructure only): the time-domain correlator,
 * restating correlator.c:325-400 (plan), 648-673 (baseline integrate),
 * 820-858 (network plan) and 967-984 (network integrate) of the reference
 * sample by sample: for each output sample t, the two delay-tap
 * projections dotted with the series (sh_series_array_dot,
 * sh_series_array.c:407-426), their SH product, added to power_1d with
 * weight window[t]; then rotated to the baseline direction and averaged
 * over the baselines.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

struct or_plan_td {
	double d[3], theta, phi, delta_t;
	int transient;
	struct or_array *proj_a, *proj_b;
	struct or_series *sample_a, *sample_b, *product, *power_1d;
	struct or_product_plan *pp;
	struct or_rotation_plan *rp;
};

/* correlator.c:325-385 */
struct or_plan_td *or_plan_td_new(const double ra[3], const double rb[3], double delta_t)
{
	struct or_plan_td *p = calloc(1, sizeof(*p));
	for(int i = 0; i < 3; i++)
		p->d[i] = ra[i] - rb[i];
	or_vector_direction(p->d, &p->theta, &p->phi);
	p->delta_t = delta_t;
	const double D = sqrt(p->d[0] * p->d[0] + p->d[1] * p->d[1] + p->d[2] * p->d[2]);
	const unsigned a_l = or_projection_l_max(D / 2, delta_t), b_l = a_l;
	unsigned P = or_power_l_max_naive(D, delta_t);
	if(P > a_l + b_l)
		P = a_l + b_l;
	const double rza[3] = { 0, 0, +D / 2. }, rzb[3] = { 0, 0, -D / 2. };
	p->proj_a = or_projection_matrix_delay(or_projection_n_elements(D / 2, delta_t), a_l, rza, delta_t);
	p->proj_b = or_projection_matrix_delay(or_projection_n_elements(D / 2, delta_t), b_l, rzb, delta_t);
	p->transient = ((p->proj_a->n > p->proj_b->n ? p->proj_a->n : p->proj_b->n) - 1) / 2;
	p->sample_a = or_series_new(a_l, 1);
	p->sample_b = or_series_new(b_l, 1);
	p->product = or_series_new(P, 1);
	p->power_1d = or_series_new(P, 1);
	p->pp = or_product_plan_new(p->power_1d, p->sample_a, p->sample_b);
	double R[9];
	or_rot_matrix(p->theta, p->phi, R);
	p->rp = or_rotation_plan_new(P, R);
	return p;
}

void or_plan_td_free(struct or_plan_td *p)
{
	if(!p)
		return;
	or_rotation_plan_free(p->rp);
	or_product_plan_free(p->pp);
	or_series_free(p->power_1d);
	or_series_free(p->product);
	or_series_free(p->sample_b);
	or_series_free(p->sample_a);
	or_array_free(p->proj_b);
	or_array_free(p->proj_a);
	free(p);
}

/* sh_series_array_dot (sh_series_array.c:407-426): sum_i array[i] v[i] */
static void array_dot(struct or_series *r, const struct or_array *a, const double *v)
{
	or_series_zero(r);
	const double complex *c = a->coeff;
	for(int i = 0; i < a->n; i++)
		for(int k = 0; k < a->stride; k++)
			r->coeff[k] += *c++ * v[i];
}

/* correlator.c:648-673 */
struct or_series *or_baseline_integrate_power_td(const double *ta, const double *tb, const double *window, int n, struct or_plan_td *p)
{
	n -= 2 * p->transient;
	if(n < 0)
		return NULL;
	or_series_zero(p->power_1d);
	while(n--) {
		array_dot(p->sample_a, p->proj_a, ta++);
		array_dot(p->sample_b, p->proj_b, tb++);
		if(!or_product(p->product, p->sample_a, p->sample_b, p->pp))
			return NULL;
		or_series_add(p->power_1d, *window++, p->product);
	}
	struct or_series *p2 = or_series_new(p->power_1d->l_max, 0);
	if(!or_rotate(p2, p->power_1d, p->rp)) {
		or_series_free(p2);
		return NULL;
	}
	return p2;
}

/* correlator.c:728-757 + 820-846: baseline k pairs (i, j), i = 1..n-1,
 * j < i, d = r_i - r_j */
struct or_network_td *or_network_td_new(int n_instruments, const double *pos, double delta_t)
{
	struct or_network_td *net = calloc(1, sizeof(*net));
	net->n_instruments = n_instruments;
	net->n_baselines = n_instruments * (n_instruments - 1) / 2;
	net->index_a = malloc(net->n_baselines * sizeof(int));
	net->index_b = malloc(net->n_baselines * sizeof(int));
	net->plans = malloc(net->n_baselines * sizeof(*net->plans));
	int k = 0;
	for(int i = 1; i < n_instruments; i++)
		for(int j = 0; j < i; j++, k++) {
			net->index_a[k] = i;
			net->index_b[k] = j;
			net->plans[k] = or_plan_td_new(pos + 3 * i, pos + 3 * j, delta_t);
		}
	return net;
}

void or_network_td_free(struct or_network_td *net)
{
	if(!net)
		return;
	for(int k = 0; k < net->n_baselines; k++)
		or_plan_td_free(net->plans[k]);
	free(net->plans);
	free(net->index_a);
	free(net->index_b);
	free(net);
}

/* correlator.c:967-984: sky = (1 / n_baselines) sum_k power_2d(k) */
struct or_series *or_network_integrate_power_td(struct or_series *sky, double **ts, int n, double **windows, struct or_network_td *net)
{
	or_series_zero(sky);
	for(int k = 0; k < net->n_baselines; k++) {
		struct or_series *p2 = or_baseline_integrate_power_td(ts[net->index_a[k]], ts[net->index_b[k]], windows[k], n, net->plans[k]);
		if(!p2)
			return NULL;
		or_series_add(sky, 1.0 / net->n_baselines, p2);
		or_series_free(p2);
	}
	return sky;
}

int or_plan_td_transient(const struct or_plan_td *p)
{
	return p->transient;
}

unsigned or_plan_td_l_max(const struct or_plan_td *p)
{
	return p->power_1d->l_max;
}

const double complex *or_plan_td_power_1d(const struct or_plan_td *p)
{
	return p->power_1d->coeff;
}
