/*
 * A C caller of the library's ABI, written against include/sphradiometer/
 * only, along the frequency-domain flow of the reference's
 * test/correlator_validate.c:221-406: instruments -> network baselines ->
 * correlator_network_plan_fd_new -> FFT helpers -> network integrate ->
 * rotate_z, reading struct fields the way callers do (plans[k]->transient,
 * delay_product->n, baselines[k]->d / theta / phi).
 *
 *   fd_validate analytic
 *       the reference's accuracy test: a white-noise point source at
 *       (ra, dec) = (0, 0) in three sites at 512 Hz over one dump interval
 *       (correlator_dump_interval(l, 20)); the network sky must be within
 *       -21 dB RMS of the analytic cross power truncated to the sky's order
 *       (correlator_validate.c:62-114, 376), at GMST 0 and pi/4.
 *   fd_validate segment IN OUT
 *       the three-site workload's per-segment path: IN holds 3 x 1024
 *       float64 samples; OUT receives the l_max 32 sky (1089 complex) and
 *       its 66 x 66 mesh (complex, row-major [theta][phi]).
 *
 * Exit status 0 on success.  The site table is the one of the workload:
 * (46.45 N, 119.41 W), (30.56 N, 90.77 W), (43.63 N, 10.50 E) on a
 * 6371 km sphere.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <sphradiometer/instrument.h>
#include <sphradiometer/sh_series.h>
#include <sphradiometer/correlator.h>
#include <sphradiometer/diagnostics.h>

#define C_SI 299792458.0

static const double site_lat[3] = { 46.45, 30.56, 43.63 };
static const double site_lon[3] = { -119.41, -90.77, 10.50 };

static struct instrument_array *three_sites(void)
{
	struct instrument_array *a = instrument_array_new(0);
	for(int i = 0; i < 3; i++) {
		const double th = M_PI / 2 - site_lat[i] * M_PI / 180, ph = site_lon[i] * M_PI / 180;
		if(!instrument_array_append(a, instrument_new_from_r_theta_phi(6371e3 / C_SI, th, ph, NULL)))
			return NULL;
	}
	return a;
}

/* splitmix64 + Box-Muller: unit-variance white noise */
static uint64_t rng_state = 0x9e3779b97f4a7c15ull;

static double uniform01(void)
{
	uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	z ^= z >> 31;
	return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static double gaussian(void)
{
	return sqrt(-2 * log(uniform01())) * cos(2 * M_PI * uniform01());
}

/* correlator_validate.c:62-114 */
struct solution_data {
	double f0, delta_f, d, theta, phi, cos_gamma_s;
};

static double sinc(double x)
{
	x *= M_PI;
	return x == 0.0 ? 1.0 : sin(x) / x;
}

static double solution_f(double theta, double phi, void *vd)
{
	const struct solution_data *s = vd;
	const double cg = cos(theta) * cos(s->theta) + sin(theta) * sin(s->theta) * cos(phi - s->phi);
	const double zeta = s->d * (cg - s->cos_gamma_s);
	return sinc(zeta * s->delta_f) * cos(2 * M_PI * zeta * s->f0);
}

static struct sh_series *analytic(const struct correlator_baseline *b, unsigned l, double dt, double ra, double dec, double gmst)
{
	const double it = M_PI_2 - dec, ip = ra;
	struct solution_data s = { 0.25 / dt, 0.5 / dt, vector_magnitude(b->d), b->theta, b->phi + gmst, 0.0 };
	s.cos_gamma_s = cos(it) * cos(s.theta) + sin(it) * sin(s.theta) * cos(ip - s.phi);
	return sh_series_from_realfunc(sh_series_new(l, 0), solution_f, &s);
}

static int run_analytic(void)
{
	const double dt = 1.0 / 512;
	struct instrument_array *inst = three_sites();
	struct correlator_network_baselines *bl = inst ? correlator_network_baselines_new(inst) : NULL;
	if(!bl)
		return 1;
	const unsigned sky_l = correlator_network_l_max(bl, dt);
	const int n = (int) (correlator_dump_interval(sky_l, 20) / dt);
	struct correlator_network_plan_fd *plan = correlator_network_plan_fd_new(bl, n, dt);
	if(!plan)
		return 1;
	fprintf(stderr, "sky l_max %u, %d samples; baseline 0: transient %d, delay product %d x l_max %u\n", sky_l, n, plan->plans[0]->transient, plan->plans[0]->delay_product->n, plan->plans[0]->delay_product->l_max);
	double *h = malloc(n * sizeof(*h));
	double complex *H = malloc(n * sizeof(*H));
	double complex *fs[3];
	for(int i = 0; i < 3; i++)
		fs[i] = malloc(n * sizeof(*fs[i]));
	fftw_plan fp = correlator_tseries_to_fseries_plan(h, H, n);
	struct sh_series *sky = sh_series_new_zero(sky_l, 0);
	if(!fp || !sky)
		return 1;
	const double gmsts[2] = { 0.0, M_PI / 4 };
	int fail = 0;
	for(int g = 0; g < 2; g++) {
		const double gmst = gmsts[g], ra = 0.0, dec = 0.0;
		for(int t = 0; t < n; t++)
			h[t] = gaussian();
		correlator_tseries_to_fseries(h, H, n, fp);
		/* each site sees h(t + r . s), circularly (inject.c:119-152):
		 * the phase factor on the half spectrum, its conjugate on the
		 * other half; the Nyquist bin keeps its real part */
		const double th = M_PI_2 - dec, ph = ra - gmst;
		for(int i = 0; i < 3; i++) {
			const double rs = vector_r_dot_s(instrument_array_get(inst, i)->phase_centre, th, ph);
			for(int k = 0; k <= n / 2; k++) {
				double complex v = H[k] * cexp(2 * M_PI * I * rs * k / (dt * n));
				if(2 * k == n)
					v = creal(v);
				fs[i][k] = v;
				if(k > 0 && n - k != k)
					fs[i][n - k] = conj(v);
			}
		}
		if(!correlator_network_integrate_power_fd(sky, fs, plan))
			return 1;
		sh_series_rotate_z(sky, sky, gmst);
		struct sh_series *sol = analytic(bl->baselines[0], 2 * sky_l, dt, ra, dec, gmst);
		for(int k = 1; k < bl->n_baselines; k++) {
			struct sh_series *tmp = analytic(bl->baselines[k], 2 * sky_l, dt, ra, dec, gmst);
			sh_series_add(sol, 1.0, tmp);
			sh_series_free(tmp);
		}
		sh_series_scale(sol, 1.0 / bl->n_baselines);
		struct sh_series *lo = sh_series_resize(sh_series_copy(sol), sky_l);
		const double db = 10 * log10(fabs(diagnostics_rms_error(sky, lo)));
		printf("gmst %.6f: RMS (FD - order-reduced exact) = %.2f dB\n", gmst, db);
		if(!(db <= -21.0))
			fail = 1;
		sh_series_free(sol);
		sh_series_free(lo);
	}
	correlator_fft_plan_free(fp);
	sh_series_free(sky);
	correlator_network_plan_fd_free(plan);
	correlator_network_baselines_free(bl);
	instrument_array_free(inst);
	for(int i = 0; i < 3; i++)
		free(fs[i]);
	free(h);
	free(H);
	return fail;
}

static int run_segment(const char *in, const char *out)
{
	const int n = 1024;
	const unsigned L = 32;
	const double dt = 1.0 / 4096;
	double *ts[3];
	double complex *fs[3];
	fftw_plan fp[3];
	FILE *f = fopen(in, "rb");
	if(!f)
		return 1;
	for(int i = 0; i < 3; i++) {
		ts[i] = malloc(n * sizeof(*ts[i]));
		fs[i] = malloc(n * sizeof(*fs[i]));
		if(fread(ts[i], sizeof(double), n, f) != (size_t) n)
			return 1;
		fp[i] = correlator_tseries_to_fseries_plan(ts[i], fs[i], n);
		if(!fp[i])
			return 1;
	}
	fclose(f);
	struct instrument_array *inst = three_sites();
	struct correlator_network_baselines *bl = inst ? correlator_network_baselines_new(inst) : NULL;
	struct correlator_network_plan_fd *plan = bl ? correlator_network_plan_fd_new(bl, n, dt) : NULL;
	if(!plan || !correlator_network_plan_fd_set_l(plan, L))
		return 1;
	for(int k = 0; k < bl->n_baselines; k++)
		fprintf(stderr, "baseline %d: (%d, %d) transient %d, delay product %d x l_max %u\n", k, bl->baselines[k]->index_a, bl->baselines[k]->index_b, plan->plans[k]->transient, plan->plans[k]->delay_product->n, plan->plans[k]->delay_product->l_max);
	for(int i = 0; i < 3; i++)
		correlator_tseries_to_fseries(ts[i], fs[i], n, fp[i]);
	struct sh_series *sky = sh_series_new_zero(L, 0);
	if(!sky || !correlator_network_integrate_power_fd(sky, fs, plan))
		return 1;
	int nt = 0, np = 0;
	double complex *mesh = sh_series_to_mesh(sky, &nt, &np, NULL);
	if(!mesh || nt != 2 * (int) (L + 1) || np != 2 * (int) (L + 1))
		return 1;
	f = fopen(out, "wb");
	if(!f)
		return 1;
	fwrite(sky->coeff, sizeof(double complex), (L + 1) * (L + 1), f);
	fwrite(mesh, sizeof(double complex), (size_t) nt * np, f);
	fclose(f);
	free(mesh);
	sh_series_free(sky);
	for(int i = 0; i < 3; i++) {
		correlator_fft_plan_free(fp[i]);
		free(ts[i]);
		free(fs[i]);
	}
	correlator_network_plan_fd_free(plan);
	correlator_network_baselines_free(bl);
	instrument_array_free(inst);
	return 0;
}

int main(int argc, char **argv)
{
	if(argc == 2 && !strcmp(argv[1], "analytic"))
		return run_analytic();
	if(argc == 4 && !strcmp(argv[1], "segment"))
		return run_segment(argv[2], argv[3]);
	fprintf(stderr, "usage: %s analytic | segment IN OUT\n", argv[0]);
	return 2;
}
