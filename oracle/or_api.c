/*
 * Oracle (test infrastructure only): flat entry points for the Python
 * tests and the CPU baseline (plain pointers, row-major arrays,
 * complex = two doubles).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

void *orx_network_new(int n_inst, const double *pos, int n, double dt, int target_l)
{
	return or_network_new(n_inst, pos, n, dt, target_l);
}

void orx_network_free(void *net) { or_network_free(net); }

int orx_network_set_l(void *net, unsigned l) { return or_network_set_l(net, l) ? 0 : -1; }

int orx_network_n_baselines(void *net) { return ((struct or_network *) net)->n_baselines; }

/* info: power_1d l_max, dp n, dp stride, transient, dp polar, rotation l_max, index_a, index_b */
void orx_plan_info(void *vnet, int k, int *info)
{
	struct or_network *net = vnet;
	struct or_plan_fd *p = net->plans[k];
	info[0] = (int) p->power_1d->l_max;
	info[1] = p->delay_product->n;
	info[2] = p->delay_product->stride;
	info[3] = p->transient;
	info[4] = p->delay_product->polar;
	info[5] = (int) p->rotation_plan->l_max;
	info[6] = net->index_a[k];
	info[7] = net->index_b[k];
}

void orx_plan_geometry(void *vnet, int k, double *out)
{
	struct or_plan_fd *p = ((struct or_network *) vnet)->plans[k];
	out[0] = p->d[0];
	out[1] = p->d[1];
	out[2] = p->d[2];
	out[3] = p->theta;
	out[4] = p->phi;
}

void orx_plan_get_dp(void *vnet, int k, double complex *out)
{
	struct or_plan_fd *p = ((struct or_network *) vnet)->plans[k];
	memcpy(out, p->delay_product->coeff, (size_t) p->delay_product->n * p->delay_product->stride * sizeof(*out));
}

void orx_plan_get_power_1d(void *vnet, int k, double complex *out)
{
	struct or_plan_fd *p = ((struct or_network *) vnet)->plans[k];
	memcpy(out, p->power_1d->coeff, or_length(p->power_1d->l_max, 1) * sizeof(*out));
}

/* all D blocks l = 1..l_max of the plan, each (2l+1)^2 in storage order */
void orx_plan_get_D(void *vnet, int k, double complex *out)
{
	struct or_rotation_plan *r = ((struct or_network *) vnet)->plans[k]->rotation_plan;
	for(unsigned l = 1; l <= r->l_max; l++) {
		const int n = (2 * l + 1) * (2 * l + 1);
		memcpy(out, r->D[l] - (2 * l + 1) * l - l, n * sizeof(*out));
		out += n;
	}
}

/* sky (non-polar, sky_l) = network integrate of full spectra
 * fseries[n_inst][n] */
int orx_network_integrate(void *vnet, const double complex *fseries, unsigned sky_l, double complex *sky_out)
{
	struct or_network *net = vnet;
	const int n = net->plans[0]->delay_product->n;
	double complex **fs = malloc(net->n_instruments * sizeof(*fs));
	for(int i = 0; i < net->n_instruments; i++)
		fs[i] = (double complex *) fseries + (size_t) i * n;
	struct or_series sky = { sky_l, 0, sky_out };
	int rc = or_network_integrate_power_fd(&sky, fs, net) ? 0 : -1;
	free(fs);
	return rc;
}

/* the per-segment pipeline a caller runs: (optional window) -> FFT ->
 * network integrate -> rotate_z(gmst) -> to_mesh */
int orx_segment_pipeline(void *vnet, const double *tseries, const double *window, unsigned sky_l, double gmst, double complex *sky_out, double complex *mesh_out)
{
	struct or_network *net = vnet;
	const int n = net->plans[0]->delay_product->n;
	double complex *fsb = malloc((size_t) net->n_instruments * n * sizeof(*fsb));
	double *tmp = malloc((size_t) n * sizeof(*tmp));
	for(int i = 0; i < net->n_instruments; i++) {
		for(int t = 0; t < n; t++)
			tmp[t] = window ? tseries[(size_t) i * n + t] * window[t] : tseries[(size_t) i * n + t];
		or_tseries_to_fseries(tmp, fsb + (size_t) i * n, n);
	}
	int rc = orx_network_integrate(vnet, fsb, sky_l, sky_out);
	struct or_series sky = { sky_l, 0, sky_out };
	if(!rc && gmst != 0.)
		or_rotate_z(&sky, &sky, gmst);
	if(!rc && mesh_out) {
		int nt, np;
		double complex *m = or_to_mesh(&sky, &nt, &np);
		memcpy(mesh_out, m, (size_t) nt * np * sizeof(*m));
		free(m);
	}
	free(tmp);
	free(fsb);
	return rc;
}

void orx_tseries_to_fseries(const double *ts, double complex *fs, int n) { or_tseries_to_fseries(ts, fs, n); }
void orx_ctseries_to_fseries(const double complex *ts, double complex *fs, int n) { or_ctseries_to_fseries(ts, fs, n); }
void orx_dft(double complex *d, int n, int sign) { or_dft(d, n, sign); }
void orx_gl_nodes(int n, double *x, double *w) { or_gl_nodes(n, x, w); }
void orx_sphPlm_array(unsigned l_max, int m, double x, double *out) { or_sphPlm_array(l_max, m, x, out); }
double orx_wigner_3j(int a, int b, int c, int d, int e, int f) { return or_wigner_3j(a, b, c, d, e, f); }

void orx_mesh_nodes(unsigned l, double *cos_theta, double *w)
{
	int nt, np;
	or_mesh_nodes(l, &nt, &np, cos_theta, w);
}

void orx_Y(unsigned l, int m, double theta, double phi, double complex *out) { *out = or_Y(l, m, theta, phi); }

void orx_eval(unsigned l, int polar, const double complex *coeff, double theta, double phi, double complex *out)
{
	struct or_series s = { l, polar, (double complex *) coeff };
	*out = or_eval(&s, theta, phi);
}

void orx_to_mesh(unsigned l, int polar, const double complex *coeff, double complex *mesh)
{
	struct or_series s = { l, polar, (double complex *) coeff };
	int nt, np;
	double complex *m = or_to_mesh(&s, &nt, &np);
	memcpy(mesh, m, (size_t) nt * np * sizeof(*m));
	free(m);
}

void orx_from_mesh(unsigned l, int polar, const double complex *mesh, double complex *coeff)
{
	struct or_series s = { l, polar, coeff };
	or_from_mesh(&s, mesh);
}

void orx_from_realmesh(unsigned l, int polar, const double *mesh, double complex *coeff)
{
	struct or_series s = { l, polar, coeff };
	or_from_realmesh(&s, mesh);
}

void orx_impulse(unsigned l, double theta, double phi, double complex *coeff)
{
	struct or_series *s = or_impulse(l, theta, phi);
	memcpy(coeff, s->coeff, or_length(l, 0) * sizeof(*coeff));
	or_series_free(s);
}

void orx_euler_rotation_matrix(double g, double b, double a, double *R) { or_euler_rotation_matrix(g, b, a, R); }
void orx_rot_matrix(double t, double p, double *R) { or_rot_matrix(t, p, R); }
void orx_invrot_matrix(double t, double p, double *R) { or_invrot_matrix(t, p, R); }

/* all D blocks for l = 1..l_max, (2l+1)^2 each, storage order */
void orx_rotation_D(unsigned l_max, const double *R, double complex *out)
{
	struct or_rotation_plan *p = or_rotation_plan_new(l_max, R);
	for(unsigned l = 1; l <= l_max; l++) {
		const int n = (2 * l + 1) * (2 * l + 1);
		memcpy(out, p->D[l] - (2 * l + 1) * l - l, n * sizeof(*out));
		out += n;
	}
	or_rotation_plan_free(p);
}

int orx_rotate(unsigned l, int polar, const double complex *coeff, const double *R, double complex *out)
{
	struct or_rotation_plan *p = or_rotation_plan_new(l, R);
	struct or_series s = { l, polar, (double complex *) coeff }, r = { l, 0, out };
	int rc = or_rotate(&r, &s, p) ? 0 : -1;
	or_rotation_plan_free(p);
	return rc;
}

void orx_rotate_z(unsigned l, const double complex *coeff, double phi, double complex *out)
{
	struct or_series s = { l, 0, (double complex *) coeff }, r = { l, 0, out };
	or_rotate_z(&r, &s, phi);
}

/* mode: -1 = the reference's own choice, 0 = pixel domain, 1 = microcode */
int orx_product(unsigned dl, int dp, unsigned al, int ap, const double complex *a, unsigned bl, int bp, const double complex *b, int mode, double complex *out)
{
	struct or_series A = { al, ap, (double complex *) a }, B = { bl, bp, (double complex *) b }, Dd = { dl, dp, out };
	struct or_product_plan *p = mode < 0 ? or_product_plan_new(&Dd, &A, &B) : or_product_plan_new_forced(&Dd, &A, &B, mode);
	int rc = or_product(&Dd, &A, &B, p) ? 0 : -1;
	or_product_plan_free(p);
	return rc;
}

int orx_product_plan_length(unsigned dl, int dp, unsigned al, int ap, unsigned bl, int bp)
{
	struct or_series A = { al, ap, NULL }, B = { bl, bp, NULL }, Dd = { dl, dp, NULL };
	struct or_product_plan *p = or_product_plan_new_forced(&Dd, &A, &B, 1);
	int n = p->plan_length;
	or_product_plan_free(p);
	return n;
}

void orx_projection_matrix_delay(unsigned n, unsigned l_max, const double *r, double dt, double complex *out)
{
	struct or_array *m = or_projection_matrix_delay(n, l_max, r, dt);
	memcpy(out, m->coeff, (size_t) m->n * m->stride * sizeof(*out));
	or_array_free(m);
}

/* the general from_realfunc route for a polar delay kernel (checks the
 * phi-constant evaluation used above) */
struct dd2 { const double *r; double dt; int k; };
static double dfun(double theta, double phi, void *p)
{
	const struct dd2 *d = p;
	double pi_x = M_PI * (0 - d->k - or_vector_r_dot_s(d->r, theta, phi) / d->dt);
	return pi_x == 0.0 ? 1.0 : sin(pi_x) / pi_x;
}

void orx_projection_element_general(unsigned l_max, int polar, const double *r, double dt, int k, double complex *out)
{
	struct dd2 d = { r, dt, k };
	struct or_series s = { l_max, polar, out };
	or_from_realfunc(&s, dfun, &d);
}

void orx_params(double dt, double D, int *out)
{
	out[0] = (int) or_power_l_max_naive(D, dt);
	out[1] = (int) or_projection_l_max(D / 2, dt);
	out[2] = or_projection_n_elements(D / 2, dt);
}

double orx_rms_error(unsigned lm, int pm, const double complex *m, unsigned le, int pe, const double complex *e)
{
	struct or_series M = { lm, pm, (double complex *) m }, E = { le, pe, (double complex *) e };
	return or_rms_error(&M, &E);
}

void orx_series_real(unsigned l, double complex *coeff)
{
	struct or_series s = { l, 0, coeff };
	or_series_real(&s);
}

void orx_series_conj(unsigned l, int polar, double complex *coeff)
{
	struct or_series s = { l, polar, coeff };
	or_series_conj(&s);
}

void orx_scale_power_l(unsigned l, double complex *coeff, double power, double overall)
{
	struct or_series s = { l, 0, coeff };
	or_series_scale_power_l(&s, power, overall);
}

int orx_array_set_l(int n, unsigned l_old, int polar, const double complex *in, unsigned l_new, double complex *out)
{
	struct or_array *a = or_array_new(n, l_old, polar);
	memcpy(a->coeff, in, (size_t) n * a->stride * sizeof(*in));
	or_array_set_l(a, l_new);
	memcpy(out, a->coeff, (size_t) n * a->stride * sizeof(*out));
	or_array_free(a);
	return 0;
}

/* HEALPix output (or_healpix.c) */
int orx_healpix_nside_c(unsigned l)
{
	return or_healpix_nside_c(l);
}

int orx_healpix_nside_py(unsigned l)
{
	return or_healpix_nside_py(l);
}

void orx_healpix_pix2ang(int nside, long npix, double *theta, double *phi)
{
	for(long p = 0; p < npix; p++)
		or_healpix_pix2ang(nside, p, theta + p, phi + p);
}

void orx_healpix_map(unsigned l, int polar, const double complex *coeff, int nside, double *map)
{
	struct or_series s = { l, polar, (double complex *) coeff };
	or_healpix_map(&s, nside, map);
}

void orx_healpix_posterior(const double *map_p, const double *map_n, long npix, double *prob)
{
	or_healpix_posterior(map_p, map_n, npix, prob);
}

/* time-domain correlator (or_td.c) */
void *orx_plan_td_new(const double *ra, const double *rb, double delta_t)
{
	return or_plan_td_new(ra, rb, delta_t);
}

void orx_plan_td_free(void *p)
{
	or_plan_td_free(p);
}

/* info: [0] transient, [1] power l_max */
void orx_plan_td_info(const void *p, int *info)
{
	info[0] = or_plan_td_transient(p);
	info[1] = (int) or_plan_td_l_max(p);
}

/* power_2d (full layout of order l_max) and power_1d (l_max + 1) out */
int orx_baseline_integrate_power_td(const double *ta, const double *tb, const double *window, int n, void *p, double complex *power_2d, double complex *power_1d)
{
	struct or_series *r = or_baseline_integrate_power_td(ta, tb, window, n, p);
	if(!r)
		return -1;
	const unsigned L = r->l_max;
	memcpy(power_2d, r->coeff, (size_t) (L + 1) * (L + 1) * sizeof(*power_2d));
	memcpy(power_1d, or_plan_td_power_1d(p), (size_t) (L + 1) * sizeof(*power_1d));
	or_series_free(r);
	return 0;
}

/*
 * CPU baseline: the per-segment pipeline (FFT of each detector, network
 * integrate, to_mesh) over S segments, tseries [D][S][N], outputs
 * sky [S][(L+1)^2] and mesh [S][2B][2B].
 *   seg_threads = 1: segments in order on this thread, the ring loop of
 *     to_mesh on mesh_threads threads (the reference's own split:
 *     integrate on one core, to_mesh on 4, sh_series_trans.c:829);
 *   seg_threads > 1: segments split over that many threads, each with its
 *     own plan scratch (fseries_product, power_1d), as callers of the
 *     reference do with correlator_network_plan_fd_copy (correlator.c:
 *     908-935); to_mesh on one thread inside each.
 * The delay products and rotation plans are read-only and shared.
 */
int orx_segments_run(void *vnet, const double *ts, int S, unsigned sky_l, int seg_threads, int mesh_threads, double complex *sky_out, double complex *mesh_out)
{
	struct or_network *net = vnet;
	const int D = net->n_instruments, nb = net->n_baselines;
	const int n = net->plans[0]->delay_product->n;
	const size_t nsky = or_length(sky_l, 0);
	const size_t nmesh = (size_t) (2 * (sky_l + 1)) * (2 * (sky_l + 1));
	int rc = 0;
	if(seg_threads < 1)
		seg_threads = 1;
#pragma omp parallel num_threads(seg_threads) reduction(| : rc)
	{
		/* this thread's copy of the network: shared tables, own scratch */
		struct or_plan_fd *pl = malloc(nb * sizeof(*pl));
		struct or_plan_fd **pp = malloc(nb * sizeof(*pp));
		for(int k = 0; k < nb; k++) {
			pl[k] = *net->plans[k];
			pl[k].fseries_product = malloc((size_t) n * sizeof(double complex));
			pl[k].power_1d = or_series_new(net->plans[k]->power_1d->l_max, net->plans[k]->power_1d->polar);
			pp[k] = &pl[k];
		}
		struct or_network mine = *net;
		mine.plans = pp;
		double complex *fsb = malloc((size_t) D * n * sizeof(*fsb));
		double complex **fs = malloc(D * sizeof(*fs));
		double *tmp = malloc((size_t) n * sizeof(*tmp));
		const int mt = seg_threads > 1 ? 1 : mesh_threads;
#pragma omp for schedule(dynamic, 4)
		for(int s = 0; s < S; s++) {
			for(int i = 0; i < D; i++) {
				memcpy(tmp, ts + ((size_t) i * S + s) * n, (size_t) n * sizeof(*tmp));
				or_tseries_to_fseries(tmp, fsb + (size_t) i * n, n);
				fs[i] = fsb + (size_t) i * n;
			}
			struct or_series sky = { sky_l, 0, sky_out + (size_t) s * nsky };
			if(!or_network_integrate_power_fd(&sky, fs, &mine)) {
				rc |= 1;
				continue;
			}
			if(mesh_out) {
				double complex *m = or_to_mesh_threads(&sky, NULL, NULL, mt);
				memcpy(mesh_out + (size_t) s * nmesh, m, nmesh * sizeof(*m));
				free(m);
			}
		}
		for(int k = 0; k < nb; k++) {
			free(pl[k].fseries_product);
			or_series_free(pl[k].power_1d);
		}
		free(pl);
		free(pp);
		free(fsb);
		free(fs);
		free(tmp);
	}
	return rc ? -1 : 0;
}

void *orx_network_td_new(int n_instruments, const double *pos, double delta_t)
{
	return or_network_td_new(n_instruments, pos, delta_t);
}

void orx_network_td_free(void *net)
{
	or_network_td_free(net);
}

/* ts [n_instruments][n], windows [n_baselines][n - 2 transient_k] packed
 * with stride wstride; sky of order sky_l out (full layout) */
int orx_network_integrate_power_td(void *net, const double *ts, int n, const double *windows, int wstride, unsigned sky_l, double complex *sky)
{
	struct or_network_td *nt = net;
	double **t = malloc(nt->n_instruments * sizeof(*t)), **w = malloc(nt->n_baselines * sizeof(*w));
	for(int i = 0; i < nt->n_instruments; i++)
		t[i] = (double *) ts + (size_t) i * n;
	for(int k = 0; k < nt->n_baselines; k++)
		w[k] = (double *) windows + (size_t) k * wstride;
	struct or_series s = { sky_l, 0, sky };
	struct or_series *r = or_network_integrate_power_td(&s, t, n, w, nt);
	free(t);
	free(w);
	return r ? 0 : -1;
}
