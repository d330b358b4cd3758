/*
 * Oracle (test infrastructure only): geometry, delay projection and the
 * frequency-domain correlator, restating instrument.c, projection.c and
 * correlator.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* instrument.c:77-80 */
double or_vector_r_dot_s(const double v[3], double theta, double phi)
{
	return (v[0] * cos(phi) + v[1] * sin(phi)) * sin(theta) + v[2] * cos(theta);
}

/* instrument.c:90-100 */
void or_vector_direction(const double v[3], double *theta, double *phi)
{
	*theta = atan2(sqrt(v[0] * v[0] + v[1] * v[1]), v[2]);
	*phi = atan2(v[1], v[0]);
	if(*phi < 0)
		*phi += 2. * M_PI;
}

/* correlator.c:55-65 */
unsigned or_power_l_max_naive(double D, double delta_t)
{
	return (unsigned) floor(M_PI / (2. * asin(delta_t / (2. * D))) + 1.);
}

/* projection.c:75-87 */
unsigned or_projection_l_max(double r, double delta_t)
{
	const int l = M_PI * r / delta_t + 1;
	return (unsigned) l;
}

/* projection.c:56-66 (buffer = 60) */
int or_projection_n_elements(double r, double delta_t)
{
	return 2 * (int) (60 + r / delta_t + 1) + 1;
}

/* projection.c:99-107 with j = 0 */
struct delay_data {
	const double *r;
	double delta_t;
	int k;
};

static double delay_element(double theta, double phi, void *p)
{
	const struct delay_data *d = p;
	double pi_x = M_PI * (0 - d->k - or_vector_r_dot_s(d->r, theta, phi) / d->delta_t);
	return pi_x == 0.0 ? 1.0 : sin(pi_x) / pi_x;
}

/* projection.c:110-131: tap i (k = i - (n-1)/2) holds the SH expansion of
 * sinc(-k - r.s/dt).  For a phase centre on the z axis the series is
 * polar, the mesh sampled by from_realfunc (sh_series_trans.c:538-565,
 * 663-734) is phi-constant, and only the DC term of each ring's phi DFT
 * is read: nphi copies of the ring value.  That DC term is evaluated
 * directly here; or_from_realfunc is the general route and the tests
 * check the two agree. */
struct or_array *or_projection_matrix_delay(unsigned n, unsigned l_max, const double r[3], double delta_t)
{
	const int polar = r[0] == 0.0 && r[1] == 0.0;
	struct or_array *m = or_array_new((int) n, l_max, polar);
	struct delay_data dd = { r, delta_t, 0 };
	if(!polar) {
		for(unsigned i = 0; i < n; i++) {
			struct or_series v;
			or_array_series(m, (int) i, &v);
			dd.k = -((int) n - 1) / 2 + (int) i;
			or_from_realfunc(&v, delay_element, &dd);
		}
		return m;
	}
	int nt, np;
	double *x = malloc(2 * (l_max + 1) * sizeof(*x)), *w = malloc(2 * (l_max + 1) * sizeof(*w));
	or_mesh_nodes(l_max, &nt, &np, x, w);
	double *P = malloc((size_t) nt * (l_max + 1) * sizeof(*P));
	for(int i = 0; i < nt; i++)
		or_sphPlm_array(l_max, 0, x[i], P + (size_t) i * (l_max + 1));
	double *F = malloc(nt * sizeof(*F));
	for(unsigned t = 0; t < n; t++) {
		dd.k = -((int) n - 1) / 2 + (int) t;
		for(int i = 0; i < nt; i++) {
			const double f = delay_element(acos(x[i]), 0., &dd);
			F[i] = (f * (double) np) * (w[i] * (2. * M_PI / np));
		}
		double complex *c = m->coeff + (size_t) t * m->stride;
		for(unsigned l = 0; l <= l_max; l++) {
			double acc = 0.;
			for(int i = 0; i < nt; i++)
				acc += F[i] * P[(size_t) i * (l_max + 1) + l];
			c[l] = acc;
		}
	}
	free(F);
	free(P);
	free(x);
	free(w);
	return m;
}

/* correlator.c:325-400 (the parts the FD plan keeps) + 424-519.
 * target_l >= 0 builds the plan as it is after correlator_plan_fd_set_l
 * (correlator.c:571-591) to target_l: the product coefficients d <=
 * target_l and the Wigner-D blocks l <= target_l do not depend on the
 * truncation (the analysis is per-d, the D recursion runs upward in l),
 * so building only those is the same arithmetic.  The product branch is
 * still chosen from the natural l_max as the reference does. */
struct or_plan_fd *or_plan_fd_new(const double ra[3], const double rb[3], int n, double delta_t, int target_l)
{
	struct or_plan_fd *p = calloc(1, sizeof(*p));
	for(int i = 0; i < 3; i++)
		p->d[i] = ra[i] - rb[i];
	or_vector_direction(p->d, &p->theta, &p->phi);
	p->delta_t = delta_t;
	const double D = sqrt(p->d[0] * p->d[0] + p->d[1] * p->d[1] + p->d[2] * p->d[2]);
	const unsigned a_l = or_projection_l_max(D / 2, delta_t);
	const unsigned b_l = or_projection_l_max(D / 2, delta_t);
	unsigned P = (delta_t <= 0. || D <= 0.) ? (unsigned) -1 : or_power_l_max_naive(D, delta_t);
	if(P > a_l + b_l)
		P = a_l + b_l;
	const unsigned Pe = (target_l >= 0 && (unsigned) target_l < P) ? (unsigned) target_l : P;
	const int np = or_projection_n_elements(D / 2, delta_t);
	const double rza[3] = { 0, 0, +D / 2. }, rzb[3] = { 0, 0, -D / 2. };
	struct or_array *pa = or_projection_matrix_delay(np, a_l, rza, delta_t);
	struct or_array *pb = or_projection_matrix_delay(np, b_l, rzb, delta_t);
	p->transient = ((pa->n > pb->n ? pa->n : pb->n) - 1) / 2;
	struct or_series dummy_a = { a_l, 1, NULL }, dummy_b = { b_l, 1, NULL }, dest = { Pe, 1, NULL };
	const unsigned dl_nat = P <= a_l + b_l ? P : a_l + b_l;
	struct or_product_plan *pp = or_product_plan_new_forced(&dest, &dummy_a, &dummy_b, dl_nat < 60);
	double R[9];
	or_rot_matrix(p->theta, p->phi, R);
	p->rotation_plan = or_rotation_plan_new(Pe, R);
	p->power_1d = or_series_new(Pe, 1);
	p->delay_product = or_array_new(n, Pe, 1);
	p->fseries_product = malloc((size_t) n * sizeof(*p->fseries_product));
	or_array_set_len(pa, n);
	or_array_set_len(pb, n);
	or_array_forward_fft(pa);
	or_array_forward_fft(pb);
	/* phases are equal (same tap count) so the re-centring of
	 * correlator.c:444-475 cancels and is skipped, as there */
	for(int i = 0; i < n; i++) {
		struct or_series sa, sb, sd;
		or_array_series(pa, i, &sa);
		or_array_series(pb, i, &sb);
		or_array_series(p->delay_product, (n - i) % n, &sd);
		or_series_conj(&sa);
		or_product(&sd, &sa, &sb, pp);
	}
	or_array_scale(p->delay_product, 1.0 / n);
	or_product_plan_free(pp);
	or_array_free(pa);
	or_array_free(pb);
	return p;
}

void or_plan_fd_free(struct or_plan_fd *p)
{
	if(p) {
		free(p->fseries_product);
		or_array_free(p->delay_product);
		or_series_free(p->power_1d);
		or_rotation_plan_free(p->rotation_plan);
	}
	free(p);
}

/* correlator.c:571-591 */
struct or_plan_fd *or_plan_fd_set_l(struct or_plan_fd *p, unsigned l_max)
{
	if(p->delay_product->l_max != p->power_1d->l_max || p->rotation_plan->l_max != p->power_1d->l_max)
		return NULL;
	if(l_max > p->power_1d->l_max)
		return NULL;
	if(l_max == p->power_1d->l_max)
		return p;
	or_array_set_l(p->delay_product, l_max);
	or_series_resize(p->power_1d, l_max);
	or_rotation_plan_set_l(p->rotation_plan, l_max);
	return p;
}

/* correlator.c:681-711 */
struct or_series *or_baseline_integrate_power_fd(const double complex *fa, const double complex *fb, struct or_plan_fd *p)
{
	for(int i = 0; i < p->delay_product->n; i++)
		p->fseries_product[i] = fb[i] * conj(fa[i]);
	if(!or_array_dotc(p->power_1d, p->delay_product, p->fseries_product))
		return NULL;
	or_series_scale(p->power_1d, 1.0 / p->delay_product->n);
	struct or_series *p2 = or_series_new(p->power_1d->l_max, 0);
	if(!or_rotate(p2, p->power_1d, p->rotation_plan)) {
		or_series_free(p2);
		return NULL;
	}
	return p2;
}

/* correlator.c:728-757 + 866-893: baseline k pairs (i, j), i = 1..n-1,
 * j < i, d = r_i - r_j */
struct or_network *or_network_new(int n_instruments, const double *pos, int n, double delta_t, int target_l)
{
	struct or_network *net = calloc(1, sizeof(*net));
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
			net->plans[k] = or_plan_fd_new(pos + 3 * i, pos + 3 * j, n, delta_t, target_l);
		}
	return net;
}

void or_network_free(struct or_network *net)
{
	if(net) {
		for(int k = 0; k < net->n_baselines; k++)
			or_plan_fd_free(net->plans[k]);
		free(net->plans);
		free(net->index_a);
		free(net->index_b);
	}
	free(net);
}

/* correlator.c:949-959 */
struct or_network *or_network_set_l(struct or_network *net, unsigned l_max)
{
	for(int k = 0; k < net->n_baselines; k++) {
		if(net->plans[k]->power_1d->l_max <= l_max)
			continue;
		if(!or_plan_fd_set_l(net->plans[k], l_max))
			return NULL;
	}
	return net;
}

/* correlator.c:993-1010 */
struct or_series *or_network_integrate_power_fd(struct or_series *sky, double complex **fs, struct or_network *net)
{
	or_series_zero(sky);
	int k = 0;
	for(int i = 1; i < net->n_instruments; i++)
		for(int j = 0; j < i; j++, k++) {
			struct or_series *p2 = or_baseline_integrate_power_fd(fs[i], fs[j], net->plans[k]);
			if(!p2)
				return NULL;
			if(!or_series_add(sky, 1.0 / net->n_baselines, p2)) {
				or_series_free(p2);
				return NULL;
			}
			or_series_free(p2);
		}
	return sky;
}

/* correlator.c:216-231: r2c then fseries[n - i] = conj(fseries[i]) for
 * 1 <= i < n/2.  For odd n the reference leaves bin (n+1)/2 unset; here
 * it is the DFT's own value. */
void or_tseries_to_fseries(const double *ts, double complex *fs, int n)
{
	for(int i = 0; i < n; i++)
		fs[i] = ts[i];
	or_dft(fs, n, -1);
	for(int i = 1; i < n / 2; i++)
		fs[n - i] = conj(fs[i]);
}

/* correlator.c:234-243: c2c forward, unnormalised */
void or_ctseries_to_fseries(const double complex *ts, double complex *fs, int n)
{
	memcpy(fs, ts, (size_t) n * sizeof(*fs));
	or_dft(fs, n, -1);
}
