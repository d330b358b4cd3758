/*
 * Baselines, frequency-domain correlation plans and the cross-power
 * integrators (sphradiometer src/lib/correlator.c).
 *
 * Plan construction is host code (OpenMP).  The per-segment integrators
 * and the FFT helpers run on the GPU (dropin.hip).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/correlator.h>
#include <sphradiometer/instrument.h>
#include <sphradiometer/projection.h>
#include <sphradiometer/sh_series.h>
#include "../sphr_internal.h"

/* correlator.c:55-65 */
unsigned int correlator_baseline_power_l_max_naive(double D, double delta_t)
{
	return floor(M_PI / (2. * asin(delta_t / (2. * D))) + 1.);
}

/* correlator.c:68-76 */
unsigned int correlator_baseline_power_l_max(const struct correlator_baseline *b, double delta_t)
{
	const double D = vector_magnitude(b->d);
	if(delta_t <= 0. || D <= 0.)
		return (unsigned int) -1;
	return correlator_baseline_power_l_max_naive(D, delta_t);
}

/* correlator.c:88-98: one sidereal day is 86164 s */
double correlator_dump_interval(unsigned int power_l_max, unsigned int dumps_per_lobe)
{
	const double rad_per_s = 2 * M_PI / 86164.0;
	const double rad_per_lobe = 2 * M_PI / power_l_max;
	return rad_per_lobe / rad_per_s / dumps_per_lobe;
}

/* correlator.c:105-111 */
int correlator_transient(const struct sh_series_array *a, const struct sh_series_array *b)
{
	return ((a->n > b->n ? a->n : b->n) - 1) / 2;
}

/* correlator.c:134-150 */
double *correlator_square_window_new(const int n, const int zero_samples, const double amplitude)
{
	if(n <= 0 || zero_samples < 0 || 2 * zero_samples > n)
		return NULL;
	double *w = malloc(n * sizeof(*w));
	if(!w)
		return NULL;
	for(int i = 0; i < n; i++)
		w[i] = (i < zero_samples || i >= n - zero_samples) ? 0.0 : amplitude;
	return w;
}

/* correlator.c:183-204: sin^2 tapers, mirrored for symmetry */
double *correlator_tukey_window_new(const int n, const int zero_samples, const int taper_samples, const double amplitude)
{
	if(n <= 0 || zero_samples < 0 || 2 * (zero_samples + taper_samples) > n)
		return NULL;
	double *w = malloc(n * sizeof(*w));
	if(!w)
		return NULL;
	const double dth = M_PI_2 / taper_samples;
	for(int i = 0; i < n; i++)
		w[i] = (i < zero_samples || i >= n - zero_samples) ? 0.0 : amplitude;
	for(int i = 0; i < taper_samples; i++) {
		const double s = sin(dth * i);
		w[zero_samples + i] = w[n - 1 - zero_samples - i] = amplitude * s * s;
	}
	return w;
}

/*
 * FFT helpers (correlator.c:216-243).  The plan is a GPU FFT bound to the
 * two buffers, like an FFTW plan; executing it transforms the current
 * contents of the input buffer.  The real-input transform writes the full
 * spectrum, negative frequencies as conjugates (for odd n every bin is
 * set, including the one the reference leaves unwritten).
 */
fftw_plan correlator_tseries_to_fseries_plan(double *tseries, double complex *fseries, int n)
{
	if(!tseries || !fseries || n <= 0)
		return NULL;
	return (fftw_plan) sphr_fft_new(n, 1, tseries, (double *) fseries);
}

void correlator_tseries_to_fseries(double *tseries, double complex *fseries, int n, fftw_plan plan)
{
	(void) tseries;
	(void) fseries;
	(void) n;
	sphr_fft_exec((struct sphr_fft *) plan);
}

fftw_plan correlator_ctseries_to_fseries_plan(double complex *tseries, double complex *fseries, int n)
{
	if(!tseries || !fseries || n <= 0)
		return NULL;
	return (fftw_plan) sphr_fft_new(n, 0, (double *) tseries, (double *) fseries);
}

void correlator_ctseries_to_fseries(fftw_plan plan)
{
	sphr_fft_exec((struct sphr_fft *) plan);
}

void correlator_fft_plan_free(fftw_plan plan)
{
	sphr_fft_free((struct sphr_fft *) plan);
}

/* correlator.c:260-278 */
struct correlator_baseline *correlator_baseline_new(const struct instrument_array *instruments, int index_a, int index_b)
{
	struct correlator_baseline *b = malloc(sizeof(*b));
	gsl_vector *d = instrument_baseline(instrument_array_get(instruments, index_a), instrument_array_get(instruments, index_b));
	if(!b || !d) {
		if(d)
			gsl_vector_free(d);
		free(b);
		return NULL;
	}
	b->instruments = instruments;
	b->index_a = index_a;
	b->index_b = index_b;
	b->d = d;
	vector_direction(d, &b->theta, &b->phi);
	return b;
}

void correlator_baseline_free(struct correlator_baseline *b)
{
	if(b && b->d)
		gsl_vector_free(b->d);
	free(b);
}

/* the instrument array is shared, not copied (correlator.c:290-308) */
struct correlator_baseline *correlator_baseline_copy(const struct correlator_baseline *b)
{
	struct correlator_baseline *c = malloc(sizeof(*c));
	if(!c)
		return NULL;
	*c = *b;
	c->d = gsl_vector_alloc(3);
	if(!c->d) {
		free(c);
		return NULL;
	}
	for(int i = 0; i < 3; i++)
		gsl_vector_set(c->d, i, gsl_vector_get(b->d, i));
	return c;
}

/* correlator.c:728-757: baseline k is the pair (i, j), i = 1..n-1, j < i */
struct correlator_network_baselines *correlator_network_baselines_new(const struct instrument_array *instruments)
{
	const int n = instrument_array_len(instruments);
	const int nb = n * (n - 1) / 2;
	struct correlator_network_baselines *net = malloc(sizeof(*net));
	struct correlator_baseline **bl = calloc(nb > 0 ? nb : 1, sizeof(*bl));
	if(!net || !bl) {
		free(net);
		free(bl);
		return NULL;
	}
	int k = 0;
	for(int i = 1; i < n; i++)
		for(int j = 0; j < i; j++, k++) {
			bl[k] = correlator_baseline_new(instruments, i, j);
			if(!bl[k]) {
				while(--k >= 0)
					correlator_baseline_free(bl[k]);
				free(bl);
				free(net);
				return NULL;
			}
		}
	net->baselines = bl;
	net->n_baselines = nb;
	return net;
}

void correlator_network_baselines_free(struct correlator_network_baselines *net)
{
	if(net) {
		for(int k = 0; k < net->n_baselines; k++)
			correlator_baseline_free(net->baselines[k]);
		free(net->baselines);
	}
	free(net);
}

struct correlator_network_baselines *correlator_network_baselines_copy(const struct correlator_network_baselines *net)
{
	struct correlator_network_baselines *c = malloc(sizeof(*c));
	if(!c)
		return NULL;
	*c = *net;
	c->baselines = calloc(net->n_baselines > 0 ? net->n_baselines : 1, sizeof(*c->baselines));
	if(!c->baselines) {
		free(c);
		return NULL;
	}
	for(int k = 0; k < net->n_baselines; k++)
		if(!(c->baselines[k] = correlator_baseline_copy(net->baselines[k]))) {
			correlator_network_baselines_free(c);
			return NULL;
		}
	return c;
}

/* correlator.c:800-812 */
unsigned int correlator_network_l_max(struct correlator_network_baselines *net, double delta_t)
{
	unsigned int l_max = 0;
	for(int k = 0; k < net->n_baselines; k++) {
		const unsigned int l = correlator_baseline_power_l_max(net->baselines[k], delta_t);
		if(l > l_max)
			l_max = l;
	}
	return l_max;
}

/*
 * The delay product of correlator.c:482-492 in the pixel-domain branch of
 * sh_series_product (polar factors, order L = 2a on the mesh):
 *   dp[(n - f) mod n]_d = (1/n) sum_j (2 pi w_j) conj(A_f(x_j)) B_f(x_j) Pbar_d0(x_j),
 *   A_f(x) = sum_l Pa[f]_l Pbar_l0(x),  B_f(x) = sum_l Pb[f]_l Pbar_l0(x)
 * (conj(Pa[f]) synthesised is conj(A_f) because Pbar is real).  Evaluated
 * for all frequencies at once, OpenMP over f.
 */
static void delay_product_pixel(struct sh_series_array *dp, const struct sh_series_array *pa, const struct sh_series_array *pb, int n)
{
	const unsigned la = pa->l_max, lb = pb->l_max, L = la + lb, Lo = dp->l_max;
	const int ng = 2 * ((int) L + 1);
	const double *x, *w;
	sphr_mesh_rule(L, &x, &w);
	/* T[l][j] = Pbar_l0(x_j), j contiguous */
	double *T = malloc((size_t) (L + 1) * ng * sizeof(*T));
	double *col = malloc((L + 1) * sizeof(*col));
	for(int j = 0; j < ng; j++) {
		sphr_plm(L, 0, x[j], col);
		for(unsigned l = 0; l <= L; l++)
			T[(size_t) l * ng + j] = col[l];
	}
	free(col);
	const double inv_n = 1.0 / n;
	#pragma omp parallel
	{
		double complex *A = malloc(ng * sizeof(*A)), *B = malloc(ng * sizeof(*B));
		#pragma omp for schedule(static)
		for(int f = 0; f < n; f++) {
			const double complex *ca = pa->coeff + (size_t) f * pa->stride;
			const double complex *cb = pb->coeff + (size_t) f * pb->stride;
			for(int j = 0; j < ng; j++)
				A[j] = B[j] = 0.0;
			for(unsigned l = 0; l <= la; l++) {
				const double complex c = conj(ca[l]);
				const double *t = T + (size_t) l * ng;
				for(int j = 0; j < ng; j++)
					A[j] += c * t[j];
			}
			for(unsigned l = 0; l <= lb; l++) {
				const double complex c = cb[l];
				const double *t = T + (size_t) l * ng;
				for(int j = 0; j < ng; j++)
					B[j] += c * t[j];
			}
			for(int j = 0; j < ng; j++)
				A[j] = (2.0 * M_PI * w[j]) * A[j] * B[j];
			double complex *out = dp->coeff + (size_t) ((n - f) % n) * dp->stride;
			for(unsigned d = 0; d <= Lo; d++) {
				const double *t = T + (size_t) d * ng;
				double complex acc = 0.0;
				for(int j = 0; j < ng; j++)
					acc += A[j] * t[j];
				out[d] = acc * inv_n;
			}
		}
		free(A);
		free(B);
	}
	free(T);
}

/*
 * correlator.c:325-400 (the parts the FD plan keeps) and 424-519.  The
 * projection phases of the two phase centres are equal (same tap count),
 * so the re-centring of correlator.c:444-475 cancels and is not applied.
 */
struct correlator_plan_fd *correlator_plan_fd_new(const struct correlator_baseline *baseline, int n, double delta_t)
{
	if(delta_t <= 0. || n <= 0)
		return NULL;
	const double D = vector_magnitude(baseline->d);
	const unsigned a_l = projection_matrix_l_max(D / 2, delta_t);
	const unsigned b_l = projection_matrix_l_max(D / 2, delta_t);
	unsigned P = correlator_baseline_power_l_max(baseline, delta_t);
	if(P > a_l + b_l)
		P = a_l + b_l;
	const int np = projection_matrix_n_elements(D / 2, delta_t);
	struct correlator_plan_fd *p = calloc(1, sizeof(*p));
	gsl_vector *ra = gsl_vector_alloc(3), *rb = gsl_vector_alloc(3);
	struct sh_series_array *pa = NULL, *pb = NULL;
	double *R = sh_series_rot_matrix(baseline->theta, baseline->phi);
	if(!p || !ra || !rb || !R)
		goto fail;
	gsl_vector_set(ra, 0, 0);
	gsl_vector_set(ra, 1, 0);
	gsl_vector_set(ra, 2, +D / 2.);
	gsl_vector_set(rb, 0, 0);
	gsl_vector_set(rb, 1, 0);
	gsl_vector_set(rb, 2, -D / 2.);
	pa = projection_matrix_delay(np, a_l, ra, delta_t);
	pb = projection_matrix_delay(np, b_l, rb, delta_t);
	if(!pa || !pb)
		goto fail;
	p->baseline = baseline;
	p->delta_t = delta_t;
	p->transient = correlator_transient(pa, pb);
	struct sh_series power_shape = { P, 1, NULL };
	p->rotation_plan = sh_series_rotation_plan_new(&power_shape, R);
	p->power_1d = sh_series_new_zero(P, 1);
	p->delay_product = sh_series_array_new(n, P, 1);
	p->fseries_product = malloc((size_t) n * sizeof(*p->fseries_product));
	if(!p->rotation_plan || !p->power_1d || !p->delay_product || !p->fseries_product)
		goto fail;
	/* zero-pad (or cut) the taps to n and transform along the tap axis */
	if(!sh_series_array_set_len(pa, n) || !sh_series_array_set_len(pb, n))
		goto fail;
	sh_series_array_forward_fft(pa);
	sh_series_array_forward_fft(pb);
	if(P < 60) {
		/* microcode branch of sh_series_product (sh_series_product.c:199) */
		struct sh_series sa = { a_l, 1, NULL }, sb = { b_l, 1, NULL };
		struct sh_series_product_plan *pp = sh_series_product_plan_new(p->power_1d, &sa, &sb);
		if(!pp)
			goto fail;
		for(int f = 0; f < n; f++) {
			sh_series_conj(&pa->series[f]);
			sh_series_product(&p->delay_product->series[(n - f) % n], &pa->series[f], &pb->series[f], pp);
		}
		sh_series_product_plan_free(pp);
		sh_series_array_scale(p->delay_product, 1.0 / n);
	} else {
		delay_product_pixel(p->delay_product, pa, pb, n);
	}
	sh_series_array_free(pa);
	sh_series_array_free(pb);
	gsl_vector_free(ra);
	gsl_vector_free(rb);
	free(R);
	return p;
fail:
	sh_series_array_free(pa);
	sh_series_array_free(pb);
	if(ra)
		gsl_vector_free(ra);
	if(rb)
		gsl_vector_free(rb);
	free(R);
	correlator_plan_fd_free(p);
	return NULL;
}

void correlator_plan_fd_free(struct correlator_plan_fd *p)
{
	if(p) {
		sphr_gpu_plan_forget(p);
		free(p->fseries_product);
		sh_series_array_free(p->delay_product);
		sh_series_free(p->power_1d);
		sh_series_rotation_plan_free(p->rotation_plan);
	}
	free(p);
}

/* correlator.c:534-561 */
struct correlator_plan_fd *correlator_plan_fd_copy(const struct correlator_plan_fd *p)
{
	struct correlator_plan_fd *c = calloc(1, sizeof(*c));
	if(!c)
		return NULL;
	c->baseline = p->baseline;
	c->delta_t = p->delta_t;
	c->transient = p->transient;
	c->fseries_product = malloc((size_t) p->delay_product->n * sizeof(*c->fseries_product));
	c->delay_product = sh_series_array_copy(p->delay_product);
	c->power_1d = sh_series_copy(p->power_1d);
	c->rotation_plan = sh_series_rotation_plan_copy(p->rotation_plan);
	if(!c->fseries_product || !c->delay_product || !c->power_1d || !c->rotation_plan) {
		correlator_plan_fd_free(c);
		return NULL;
	}
	memcpy(c->fseries_product, p->fseries_product, (size_t) p->delay_product->n * sizeof(*c->fseries_product));
	return c;
}

/* correlator.c:571-591 */
struct correlator_plan_fd *correlator_plan_fd_set_l(struct correlator_plan_fd *p, unsigned l_max)
{
	if(p->delay_product->l_max != p->power_1d->l_max || p->rotation_plan->l_max != p->power_1d->l_max)
		return NULL;
	if(l_max > p->power_1d->l_max)
		return NULL;
	if(l_max == p->power_1d->l_max)
		return p;
	sphr_gpu_plan_forget(p);
	if(!sh_series_array_set_l(p->delay_product, l_max) || !sh_series_resize(p->power_1d, l_max) || !sh_series_rotation_plan_set_l(p->rotation_plan, l_max))
		return NULL;
	return p;
}

/* correlator.c:866-893 */
struct correlator_network_plan_fd *correlator_network_plan_fd_new(struct correlator_network_baselines *bl, int n, double delta_t)
{
	struct correlator_network_plan_fd *net = malloc(sizeof(*net));
	struct correlator_plan_fd **plans = calloc(bl->n_baselines > 0 ? bl->n_baselines : 1, sizeof(*plans));
	if(!net || !plans) {
		free(net);
		free(plans);
		return NULL;
	}
	for(int k = 0; k < bl->n_baselines; k++) {
		plans[k] = correlator_plan_fd_new(bl->baselines[k], n, delta_t);
		if(!plans[k]) {
			while(--k >= 0)
				correlator_plan_fd_free(plans[k]);
			free(plans);
			free(net);
			return NULL;
		}
	}
	net->baselines = bl;
	net->plans = plans;
	return net;
}

void correlator_network_plan_fd_free(struct correlator_network_plan_fd *net)
{
	if(net) {
		for(int k = 0; k < net->baselines->n_baselines; k++)
			correlator_plan_fd_free(net->plans[k]);
		free(net->plans);
	}
	free(net);
}

/* correlator.c:908-935: the baselines are shared */
struct correlator_network_plan_fd *correlator_network_plan_fd_copy(const struct correlator_network_plan_fd *net)
{
	struct correlator_network_plan_fd *c = malloc(sizeof(*c));
	if(!c)
		return NULL;
	*c = *net;
	c->plans = calloc(net->baselines->n_baselines > 0 ? net->baselines->n_baselines : 1, sizeof(*c->plans));
	if(!c->plans) {
		free(c);
		return NULL;
	}
	for(int k = 0; k < net->baselines->n_baselines; k++)
		if(!(c->plans[k] = correlator_plan_fd_copy(net->plans[k]))) {
			correlator_network_plan_fd_free(c);
			return NULL;
		}
	return c;
}

/* correlator.c:949-959: baselines already at or below l_max are kept */
struct correlator_network_plan_fd *correlator_network_plan_fd_set_l(struct correlator_network_plan_fd *net, unsigned l_max)
{
	for(int k = 0; k < net->baselines->n_baselines; k++) {
		if(net->plans[k]->power_1d->l_max <= l_max)
			continue;
		if(!correlator_plan_fd_set_l(net->plans[k], l_max))
			return NULL;
	}
	return net;
}

static void fill_baseline(struct sphr_fd_baseline *b, struct correlator_plan_fd *p, int ia, int ib)
{
	b->key = p;
	b->index_a = ia;
	b->index_b = ib;
	b->l_max = p->power_1d->l_max;
	b->polar = p->power_1d->polar;
	b->stride = p->delay_product->stride;
	b->dp = (const double *) p->delay_product->coeff;
	b->blocks = sphr_rotation_blocks(p->rotation_plan);
	b->power_1d = (double *) p->power_1d->coeff;
	b->fsp = (double *) p->fseries_product;
}

static int plan_consistent(const struct correlator_plan_fd *p)
{
	return p->delay_product->l_max == p->power_1d->l_max && p->delay_product->polar == p->power_1d->polar && p->rotation_plan->l_max >= p->power_1d->l_max;
}

/* correlator.c:681-711: returns a new non-polar series, the baseline's
 * power rotated into the Earth frame (GPU) */
struct sh_series *correlator_baseline_integrate_power_fd(const double complex *fa, const double complex *fb, struct correlator_plan_fd *p)
{
	if(!plan_consistent(p))
		return NULL;
	struct sh_series *out = sh_series_new(p->power_1d->l_max, 0);
	if(!out)
		return NULL;
	const double *fs[2] = { (const double *) fa, (const double *) fb };
	struct sphr_fd_baseline b;
	fill_baseline(&b, p, 0, 1);
	if(sphr_gpu_network_fd(2, fs, p->delay_product->n, 1, &b, out->l_max, 1.0, (double *) out->coeff)) {
		sh_series_free(out);
		return NULL;
	}
	return out;
}

void correlator_plan_fd_device_invalidate(struct correlator_plan_fd *p)
{
	sphr_gpu_plan_forget(p);
}

/* correlator.c:993-1010: sky = (1/nb) sum_k rotate(power_1d_k) (GPU) */
struct sh_series *correlator_network_integrate_power_fd(struct sh_series *sky, double complex **fseries, struct correlator_network_plan_fd *net)
{
	const int nb = net->baselines->n_baselines;
	sh_series_zero(sky);
	if(nb <= 0)
		return sky;
	const int ni = instrument_array_len(net->baselines->baselines[0]->instruments);
	if(ni * (ni - 1) / 2 < nb)
		return NULL;
	struct sphr_fd_baseline *b = malloc(nb * sizeof(*b));
	if(!b)
		return NULL;
	int k = 0;
	for(int i = 1; i < ni && k < nb; i++)
		for(int j = 0; j < i && k < nb; j++, k++) {
			struct correlator_plan_fd *p = net->plans[k];
			/* sh_series_add into sky must be lossless (sh_series.c:383) */
			if(!plan_consistent(p) || sky->polar || p->power_1d->l_max > sky->l_max || p->delay_product->n != net->plans[0]->delay_product->n) {
				free(b);
				return NULL;
			}
			fill_baseline(&b[k], p, i, j);
		}
	int rc = sphr_gpu_network_fd(ni, (const double *const *) fseries, net->plans[0]->delay_product->n, nb, b, sky->l_max, 1.0 / nb, (double *) sky->coeff);
	free(b);
	return rc ? NULL : sky;
}
