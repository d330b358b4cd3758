/*
 * Oracle (test infrastructure only): the localization log-likelihood sky
 * of src/lib/inspiral_sky_map.c and RapidLocalization.py.
 *
 * LAL pieces are restated from their published definitions and are parity
 * unpinned (the reference has no test of them):
 *   - XLALComputeDetAMResponse: F+ = X.D.X - Y.D.Y, Fx = X.D.Y + Y.D.X with
 *     the polarisation basis X, Y of Anderson et al., PRD 63 042003 (2001),
 *     eq. (B4)-(B5), gha = gmst - ra.
 *   - XLALCreateTukeyREAL8Window: w = 1 for |y| <= 1 - beta, else
 *     (1 + cos(pi (|y| - 1 + beta) / beta)) / 2, y = (2i - (n-1)) / (n-1).
 *   - LAL_REARTH_SI = 6378137 m, LAL_C_SI = 299792458 m/s.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define OR_REARTH_SI 6378137.0
#define OR_C_SI 299792458.0

/* XLALComputeDetAMResponse restated; resp is the 3x3 response tensor */
void or_det_response(const double resp[9], double ra, double dec, double psi, double gmst, double *fp, double *fx)
{
	const double gha = gmst - ra;
	const double cg = cos(gha), sg = sin(gha), cd = cos(dec), sd = sin(dec), cp = cos(psi), sp = sin(psi);
	const double X[3] = { -cp * sg - sp * cg * sd, -cp * cg + sp * sg * sd, sp * cd };
	const double Y[3] = { sp * sg - cp * cg * sd, sp * cg + cp * sg * sd, cp * cd };
	double p = 0., x = 0.;
	for(int i = 0; i < 3; i++) {
		double dx = 0., dy = 0.;
		for(int j = 0; j < 3; j++) {
			dx += resp[3 * i + j] * X[j];
			dy += resp[3 * i + j] * Y[j];
		}
		p += X[i] * dx - Y[i] * dy;
		x += X[i] * dy + Y[i] * dx;
	}
	*fp = p;
	*fx = x;
}

struct cbc_data {
	int i, j, n;
	const double *resp;	/* n x 9 */
	double beta, psi;
	const double *psds;	/* n */
};

/* inspiral_sky_map.c:342-364 */
static double complex cbc_projection(double theta, double phi, void *vd)
{
	const struct cbc_data *d = vd;
	double fplus[16], fcross[16];
	double np2 = 0., nc2 = 0.;
	for(int k = 0; k < d->n; k++) {
		or_det_response(d->resp + 9 * k, phi, M_PI_2 - theta, d->psi, 0.0, &fplus[k], &fcross[k]);
		fplus[k] /= d->psds[k];
		fcross[k] /= d->psds[k];
		np2 += fplus[k] * fplus[k];
		nc2 += fcross[k] * fcross[k];
	}
	return (fplus[d->i] + I * d->beta * fcross[d->i]) * (fplus[d->j] - I * d->beta * fcross[d->j]) / (np2 + d->beta * d->beta * nc2) - (double) (d->i == d->j);
}

double complex or_cbc_projection(double theta, double phi, int i, int j, const double *resp, int n, double beta, double psi, const double *psds)
{
	struct cbc_data d = { i, j, n, resp, beta, psi, psds };
	return cbc_projection(theta, phi, &d);
}

/*
 * inspiral_sky_map.c:423-549 with correlator_plan_mult_by_projection
 * (389-420).  psd is [N][n_instruments].  A bin whose PSD row equals the
 * previous bin's gets the same projection series (the function is the
 * same, so is its expansion).
 */
int or_network_mult_by_projection(struct or_network *net, const double *resp, double beta, double psi, const double *psd)
{
	const int n = net->n_instruments;
	for(int b = 0; b < net->n_baselines; b++) {
		struct or_plan_fd *p = net->plans[b];
		const int N = p->delay_product->n;
		const unsigned L = p->delay_product->l_max;
		struct or_series **proj = calloc(N, sizeof(*proj));
		double R[9];
		or_invrot_matrix(p->theta, p->phi, R);
		struct or_rotation_plan *rot = or_rotation_plan_new(L, R);
		for(int f = 0; f < N; f++) {
			if(f > 0 && !memcmp(psd + (size_t) f * n, psd + (size_t) (f - 1) * n, n * sizeof(double))) {
				proj[f] = or_series_copy(proj[f - 1]);
				continue;
			}
			struct cbc_data d = { net->index_a[b], net->index_b[b], n, resp, beta, psi, psd + (size_t) f * n };
			struct or_series *s = or_series_new(L, 0);
			or_from_func(s, cbc_projection, &d);
			struct or_series *cpy = or_series_copy(s);
			or_rotate(s, cpy, rot);
			or_series_free(cpy);
			proj[f] = s;
		}
		or_rotation_plan_free(rot);
		/* correlator_plan_mult_by_projection: the result is polar only
		 * when both factors are; the projection never is */
		struct or_array *res = or_array_new(N, L, 0);
		struct or_series d0, a0;
		or_array_series(res, 0, &d0);
		or_array_series(p->delay_product, 0, &a0);
		struct or_product_plan *pp = or_product_plan_new(&d0, &a0, proj[0]);
		/* the bins are independent; the general products hold no
		 * shared state (pixel_product_polar's caches are not reached:
		 * the projection is never polar) */
#pragma omp parallel for schedule(dynamic)
		for(int f = 0; f < N; f++) {
			struct or_series dd, aa;
			or_array_series(res, f, &dd);
			or_array_series(p->delay_product, f, &aa);
			or_product(&dd, &aa, proj[f], pp);
		}
		or_product_plan_free(pp);
		or_array_free(p->delay_product);
		p->delay_product = res;
		for(int f = 0; f < N; f++)
			or_series_free(proj[f]);
		free(proj);
		/* power_1d becomes non-polar at the product's order, with a new
		 * rotation plan (inspiral_sky_map.c:518-532) */
		or_series_set_polar(p->power_1d, 0);
		or_series_resize(p->power_1d, L);
		or_rot_matrix(p->theta, p->phi, R);
		or_rotation_plan_free(p->rotation_plan);
		p->rotation_plan = or_rotation_plan_new(L, R);
	}
	return 0;
}

/* inspiral_sky_map.c:840-856 */
static double log_uniformsky_prior(double theta, double phi, void *data)
{
	(void) phi;
	(void) data;
	return log(sin(theta) / (4 * M_PI));
}

struct or_series *or_log_uniformsky_prior(unsigned l_max)
{
	struct or_series *s = or_series_new(l_max, 0);
	or_from_realfunc(s, log_uniformsky_prior, NULL);
	return s;
}

/* XLALCreateTukeyREAL8Window restated */
double or_tukey(int n, int i, double beta)
{
	if(n <= 1)
		return 1.0;
	const double y = (2.0 * i - (n - 1)) / (double) (n - 1);
	const double ay = fabs(y);
	if(beta <= 0. || ay <= 1. - beta)
		return 1.0;
	return 0.5 * (1. + cos(M_PI * (ay - 1. + beta) / beta));
}

/* inspiral_sky_map.c:125-128 */
long or_precalc_length(long length, double delta_t)
{
	return (long) (round(2 * OR_REARTH_SI / OR_C_SI / delta_t) + length);
}

/*
 * inspiral_sky_map.c:192-219 for series that share one epoch and length:
 * Tukey(0.1), scale 1/sqrt(len), zero-pad to the precalculated length
 * with round((len - padded) / 2.) as the first retained index
 * (time_series_pad, 137-163), scale sqrt(padded).  out is [n][padded].
 */
void or_preprocess_snr(const double complex *series, int n_series, int len, double delta_t, double complex *out)
{
	const long plen = or_precalc_length(len, delta_t);
	const long first = (long) round((len - plen) / 2.);	/* <= 0: leading zeros */
	for(int k = 0; k < n_series; k++) {
		double complex *o = out + (size_t) k * plen;
		for(long t = 0; t < plen; t++)
			o[t] = 0.;
		for(int t = 0; t < len; t++) {
			const long dst = t - first;
			if(dst >= 0 && dst < plen)
				o[dst] = series[(size_t) k * len + t] * or_tukey(len, t, 0.1) * sqrt(1. / len) * sqrt((double) plen);
		}
	}
}

/*
 * inspiral_sky_map.c:869-939 (generate_alm_sky) with the spectra already
 * formed (generate_alm_skys, 942-1101, c2c FFT per detector): integrate,
 * x2, x(nb 4), + logprior, rotate_z(gmst).
 */
struct or_series *or_generate_alm_sky(struct or_series *sky, struct or_network *net, double complex **fseries, double gmst, const struct or_series *logprior)
{
	if(!or_network_integrate_power_fd(sky, fseries, net))
		return NULL;
	or_series_scale(sky, 2.0);
	or_series_scale(sky, net->n_baselines * 4.0);
	or_series_add(sky, 1.0, logprior);
	or_rotate_z(sky, sky, gmst);
	return sky;
}

/* flat API for ctypes */

void orx_det_response(const double *resp, double ra, double dec, double psi, double gmst, double *out)
{
	or_det_response(resp, ra, dec, psi, gmst, out, out + 1);
}

void orx_cbc_projection(double theta, double phi, int i, int j, const double *resp, int n, double beta, double psi, const double *psds, double complex *out)
{
	*out = or_cbc_projection(theta, phi, i, j, resp, n, beta, psi, psds);
}

int orx_network_mult_by_projection(void *net, const double *resp, double beta, double psi, const double *psd)
{
	return or_network_mult_by_projection(net, resp, beta, psi, psd);
}

void orx_log_uniformsky_prior(unsigned l_max, double complex *out)
{
	struct or_series *s = or_log_uniformsky_prior(l_max);
	memcpy(out, s->coeff, or_length(l_max, 0) * sizeof(*out));
	or_series_free(s);
}

void orx_tukey(int n, double beta, double *out)
{
	for(int i = 0; i < n; i++)
		out[i] = or_tukey(n, i, beta);
}

long orx_precalc_length(long length, double delta_t)
{
	return or_precalc_length(length, delta_t);
}

void orx_preprocess_snr(const double complex *series, int n_series, int len, double delta_t, double complex *out)
{
	or_preprocess_snr(series, n_series, len, delta_t, out);
}

/* generate_alm_sky for one event: raw complex series [n][len] (before
 * preprocessing), gmst; sky at sky_l (non-polar) + optional mesh */
int orx_localize(void *vnet, const double complex *series, int len, double delta_t, double gmst, const double complex *logprior, unsigned sky_l, double power, double overall, double complex *sky_out)
{
	struct or_network *net = vnet;
	const int n = net->n_instruments;
	const long plen = or_precalc_length(len, delta_t);
	double complex *pre = malloc((size_t) n * plen * sizeof(*pre));
	double complex *fs = malloc((size_t) n * plen * sizeof(*fs));
	double complex **fp = malloc(n * sizeof(*fp));
	or_preprocess_snr(series, n, len, delta_t, pre);
	for(int k = 0; k < n; k++) {
		or_ctseries_to_fseries(pre + (size_t) k * plen, fs + (size_t) k * plen, (int) plen);
		fp[k] = fs + (size_t) k * plen;
	}
	struct or_series lp = { sky_l, 0, (double complex *) logprior };
	struct or_series *sky = or_series_new_zero(sky_l, 0);
	int rc = or_generate_alm_sky(sky, net, fp, gmst, &lp) ? 0 : -1;
	if(!rc) {
		or_series_scale_power_l(sky, power, overall);
		memcpy(sky_out, sky->coeff, or_length(sky_l, 0) * sizeof(*sky_out));
	}
	or_series_free(sky);
	free(fp);
	free(fs);
	free(pre);
	return rc;
}
