/*
 * The localization log-likelihood sky (sphradiometer
 * src/lib/inspiral_sky_map.c, RapidLocalization.py:246-298): the CBC
 * projection operator multiplied into a network FD plan, the log prior,
 * SNR-series preprocessing and the per-event sky.
 *
 * The projection multiply (inspiral_sky_map.c:389-549) forms, for every
 * frequency bin f of a baseline, the SH product of the polar delay
 * product dp_f with the rotated projection series p (non-polar), both at
 * the plan's order L, truncated at L.  The reference evaluates it per bin
 * as a pixel-domain product on a 2(2L+1)-ring mesh.  Here the product is
 * written as one quadrature, exact for these degrees:
 *
 *   out_f[l, m] = sum_j g_f(x_j) H[j][l, m],
 *   g_f(x_j)    = sum_l1 dp_f[l1] Pbar_l1,0(x_j),
 *   H[j][l, m]  = 2 pi w_j Q_m(x_j) Pbar_l,|m|(x_j),
 *   Q_m(x_j)    = sum_l2 p[l2, m] Pbar_l2,|m|(x_j),
 *
 * on the n_q = floor(3L/2) + 1 point Gauss-Legendre rule (the integrand
 * has degree <= 3L in cos theta), so all bins that share one projection
 * are one complex GEMM [(L+1)^2 x n_q] x [n_q x bins], run on the GPU.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/correlator.h>
#include <sphradiometer/inspiral_sky_map.h>
#include <sphradiometer/instrument.h>
#include <sphradiometer/sh_series.h>
#include <sphradiometer/sphradiometer_batch.h>
#include "../sphr_internal.h"

#define SPHR_REARTH_SI 6378137.0	/* LAL_REARTH_SI */
#define SPHR_C_SI 299792458.0		/* LAL_C_SI */

long precalculated_TimeSeries_length(long length, double deltaT)
{
	return (long) (round(2 * SPHR_REARTH_SI / SPHR_C_SI / deltaT) + length);
}

static double log_uniformsky_prior(double theta, double phi, void *data)
{
	(void) phi;
	(void) data;
	return log(sin(theta) / (4 * M_PI));
}

struct sh_series *sh_series_log_uniformsky_prior(int l_max)
{
	if(l_max < 0)
		return NULL;
	struct sh_series *s = sh_series_new(l_max, 0);
	if(!s)
		return NULL;
	if(!sh_series_from_realfunc(s, log_uniformsky_prior, NULL)) {
		sh_series_free(s);
		return NULL;
	}
	return s;
}

/* LAL's Tukey window: 1 on |y| <= 1 - beta, a raised cosine outside,
 * y = (2i - (n - 1)) / (n - 1) */
void sphr_tukey_window(int n, double beta, double *w)
{
	for(int i = 0; i < n; i++) {
		if(n <= 1) {
			w[i] = 1.0;
			continue;
		}
		const double y = fabs((2.0 * i - (n - 1)) / (double) (n - 1));
		w[i] = (beta <= 0. || y <= 1. - beta) ? 1.0 : 0.5 * (1. + cos(M_PI * (y - 1. + beta) / beta));
	}
}

int preprocess_SNR_arrays(double complex *const *series, int n_series, int length, double delta_t, double complex **out)
{
	if(n_series < 1 || length < 1 || delta_t <= 0.)
		return -1;
	const long plen = precalculated_TimeSeries_length(length, delta_t);
	/* time_series_pad (inspiral_sky_map.c:137-163): the first retained
	 * sample of the padded interval is round((length - plen) / 2.) */
	const long first = (long) round((length - plen) / 2.);
	double *w = malloc(length * sizeof(*w));
	if(!w)
		return -1;
	sphr_tukey_window(length, 0.1, w);
	const double s1 = sqrt(1. / length), s2 = sqrt((double) plen);
	for(int k = 0; k < n_series; k++) {
		for(long t = 0; t < plen; t++)
			out[k][t] = 0.;
		for(int t = 0; t < length; t++) {
			const long dst = t - first;
			if(dst >= 0 && dst < plen) {
				double complex v = series[k][t] * w[t];
				v *= s1;
				out[k][dst] = v * s2;
			}
		}
	}
	free(w);
	return 0;
}

/* F+ and Fx of response tensor D for a source at (ra, dec) with
 * polarisation angle psi at sidereal time gmst: the polarisation basis X,
 * Y of Anderson et al., PRD 63 042003 (2001), eq. (B4)-(B5) */
static void antenna_response(const float D[3][3], double ra, double dec, double psi, double gmst, double *fplus, double *fcross)
{
	const double gha = gmst - ra;
	const double cg = cos(gha), sg = sin(gha), cd = cos(dec), sd = sin(dec), cp = cos(psi), sp = sin(psi);
	const double X[3] = { -cp * sg - sp * cg * sd, -cp * cg + sp * sg * sd, sp * cd };
	const double Y[3] = { sp * sg - cp * cg * sd, sp * cg + cp * sg * sd, cp * cd };
	double fp = 0., fx = 0.;
	for(int i = 0; i < 3; i++) {
		double dx = 0., dy = 0.;
		for(int j = 0; j < 3; j++) {
			dx += D[i][j] * X[j];
			dy += D[i][j] * Y[j];
		}
		fp += X[i] * dx - Y[i] * dy;
		fx += X[i] * dy + Y[i] * dx;
	}
	*fplus = fp;
	*fcross = fx;
}

struct cbc_args {
	int i, j, n;
	const struct sphr_detector *const *det;
	double beta, psi;
	const double *psds;
};

/* inspiral_sky_map.c:342-364 */
static double complex cbc_projection(double theta, double phi, void *vargs)
{
	const struct cbc_args *a = vargs;
	double fp[64], fx[64], np2 = 0., nc2 = 0.;
	for(int k = 0; k < a->n; k++) {
		antenna_response(a->det[k]->response, phi, M_PI_2 - theta, a->psi, 0.0, &fp[k], &fx[k]);
		fp[k] /= a->psds[k];
		fx[k] /= a->psds[k];
		np2 += fp[k] * fp[k];
		nc2 += fx[k] * fx[k];
	}
	return (fp[a->i] + I * a->beta * fx[a->i]) * (fp[a->j] - I * a->beta * fx[a->j]) / (np2 + a->beta * a->beta * nc2) - (double) (a->i == a->j);
}

/* out[f] = P_L(dp[f] * proj) for the bins listed in bins[0..nb), dp polar
 * at order L, proj non-polar at order L; out rows are (L+1)^2 complex */
static int polar_times_series(const struct sh_series_array *dp, const struct sh_series *proj, const int *bins, int nbins, struct sh_series_array *out)
{
	const int L = (int) dp->l_max;
	const int nq = 3 * L / 2 + 1;
	const size_t ncoef = (size_t) (L + 1) * (L + 1);
	double *x = malloc(nq * sizeof(*x)), *w = malloc(nq * sizeof(*w));
	double complex *HT = malloc(ncoef * nq * sizeof(*HT));		/* [(l,m)][j] */
	double complex *GT = malloc((size_t) nq * nbins * sizeof(*GT));	/* [j][bin] */
	double complex *CT = malloc(ncoef * nbins * sizeof(*CT));		/* [(l,m)][bin] */
	double *P0 = malloc((size_t) nq * (L + 1) * sizeof(*P0));		/* [j][l] */
	int rc = -1, fail = 0;
	if(!x || !w || !HT || !GT || !CT || !P0)
		goto done;
	sphr_gl_rule(nq, x, w);
	for(int j = 0; j < nq; j++)
		sphr_plm(L, 0, x[j], P0 + (size_t) j * (L + 1));
	/* H^T, one m per iteration */
#pragma omp parallel for schedule(dynamic)
	for(int mi = 0; mi < 2 * L + 1; mi++) {
		const int m = mi - L, am = abs(m);
		double *P = malloc((size_t) (L + 1 - am) * sizeof(*P));
		if(!P) {
#pragma omp atomic write
			fail = 1;
			continue;
		}
		for(int j = 0; j < nq; j++) {
			sphr_plm(L, am, x[j], P);
			double complex q = 0.;
			for(int l2 = am; l2 <= L; l2++)
				q += proj->coeff[sh_series_params_lmoffset(L, l2, m)] * P[l2 - am];
			const double complex h = 2. * M_PI * w[j] * q;
			for(int l = am; l <= L; l++)
				HT[sh_series_params_lmoffset(L, l, m) * nq + j] = h * P[l - am];
		}
		free(P);
	}
	if(fail)
		goto done;
	/* G^T: the polar delay product on the quadrature rings */
#pragma omp parallel for
	for(int j = 0; j < nq; j++)
		for(int b = 0; b < nbins; b++) {
			const double complex *c = dp->coeff + (size_t) bins[b] * dp->stride;
			double complex g = 0.;
			for(int l1 = 0; l1 <= L; l1++)
				g += c[l1] * P0[(size_t) j * (L + 1) + l1];
			GT[(size_t) j * nbins + b] = g;
		}
	if(sphr_gpu_zgemm((int) ncoef, nbins, nq, (const double *) HT, (const double *) GT, (double *) CT))
		goto done;
#pragma omp parallel for
	for(int b = 0; b < nbins; b++) {
		double complex *o = out->coeff + (size_t) bins[b] * out->stride;
		for(size_t c = 0; c < ncoef; c++)
			o[c] = CT[c * nbins + b];
	}
	rc = 0;
done:
	free(x);
	free(w);
	free(HT);
	free(GT);
	free(CT);
	free(P0);
	return rc;
}

int correlator_network_plan_mult_by_projection(struct correlator_network_plan_fd *plan, double beta, double psi, double **psd)
{
	if(!plan || !psd || plan->baselines->n_baselines < 1)
		return -1;
	const struct instrument_array *inst = plan->baselines->baselines[0]->instruments;
	const int n = instrument_array_len(inst);
	if(n > 64)
		return -1;
	const struct sphr_detector *det[64];
	for(int k = 0; k < n; k++) {
		det[k] = instrument_array_get(inst, k)->data;
		if(!det[k]) {
			sphr_set_error("instrument %d has no detector response (data == NULL)", k);
			return -1;
		}
	}
	for(int b = 0; b < plan->baselines->n_baselines; b++) {
		struct correlator_plan_fd *p = plan->plans[b];
		const struct correlator_baseline *bl = plan->baselines->baselines[b];
		struct sh_series_array *dp = p->delay_product;
		const int N = dp->n;
		const unsigned L = dp->l_max;
		struct sh_series_array *res = sh_series_array_new(N, L, 0);
		int *bins = malloc(N * sizeof(*bins));
		struct sh_series *proj = sh_series_new(L, 0), *cpy = sh_series_new(L, 0);
		double *R = sh_series_invrot_matrix(bl->theta, bl->phi);
		struct sh_series_rotation_plan *rot = R ? sh_series_rotation_plan_new(proj, R) : NULL;
		struct sh_series_product_plan *pp = NULL;
		int rc = -1;
		free(R);
		if(!res || !bins || !proj || !cpy || !rot)
			goto baseline_done;
		if(!dp->polar)
			pp = sh_series_product_plan_new(&res->series[0], &dp->series[0], proj);
		for(int f0 = 0; f0 < N;) {
			/* the run of bins sharing this bin's PSD row */
			int f1 = f0 + 1;
			while(f1 < N && !memcmp(psd[f1], psd[f0], n * sizeof(double)))
				f1++;
			struct cbc_args args = { bl->index_a, bl->index_b, n, det, beta, psi, psd[f0] };
			if(!sh_series_from_func(cpy, cbc_projection, &args) || !sh_series_rotate(proj, cpy, rot))
				goto baseline_done;
			if(dp->polar) {
				for(int f = f0; f < f1; f++)
					bins[f - f0] = f;
				if(polar_times_series(dp, proj, bins, f1 - f0, res))
					goto baseline_done;
			} else {
				/* a plan multiplied before: general products, per bin */
				if(!pp)
					goto baseline_done;
				for(int f = f0; f < f1; f++)
					if(!sh_series_product(&res->series[f], &dp->series[f], proj, pp))
						goto baseline_done;
			}
			f0 = f1;
		}
		rc = 0;
baseline_done:
		sh_series_product_plan_free(pp);
		sh_series_rotation_plan_free(rot);
		sh_series_free(proj);
		sh_series_free(cpy);
		free(bins);
		if(rc) {
			sh_series_array_free(res);
			return -1;
		}
		sphr_gpu_plan_forget(p);
		sh_series_array_free(p->delay_product);
		p->delay_product = res;
		/* inspiral_sky_map.c:518-532 */
		if(!sh_series_set_polar(p->power_1d, 0) || !sh_series_resize(p->power_1d, L))
			return -1;
		R = sh_series_rot_matrix(bl->theta, bl->phi);
		if(!R)
			return -1;
		sh_series_rotation_plan_free(p->rotation_plan);
		p->rotation_plan = sh_series_rotation_plan_new(p->power_1d, R);
		free(R);
		if(!p->rotation_plan)
			return -1;
	}
	return 0;
}

/* inspiral_sky_map.c:869-939 */
static int generate_alm_sky(struct sh_series *sky, struct correlator_network_plan_fd *plans, double complex **fseries, double gmst, const struct sh_series *logprior)
{
	if(!correlator_network_integrate_power_fd(sky, fseries, plans))
		return -1;
	sh_series_scale(sky, 2.0);
	sh_series_scale(sky, plans->baselines->n_baselines * 4.0);
	if(!sh_series_add(sky, 1.0, logprior))
		return -1;
	if(!sh_series_rotate_z(sky, sky, gmst))
		return -1;
	return 0;
}

int generate_alm_skys_arrays(struct sh_series **skyp, struct sh_series **skyn, struct correlator_network_plan_fd *plansp, struct correlator_network_plan_fd *plansn, double complex *const *series, int length, double gmst, const struct sh_series *logprior)
{
	*skyp = *skyn = NULL;
	if(!plansp || !plansn || !series || !logprior || length < 1)
		return -1;
	const int n = instrument_array_len(plansp->baselines->baselines[0]->instruments);
	double complex **fs = calloc(n, sizeof(*fs));
	double complex *buf = malloc((size_t) n * length * sizeof(*buf));
	double complex *in = malloc((size_t) length * sizeof(*in));
	int rc = -1;
	if(!fs || !buf || !in)
		goto done;
	for(int k = 0; k < n; k++) {
		fs[k] = buf + (size_t) k * length;
		memcpy(in, series[k], length * sizeof(*in));
		fftw_plan fp = correlator_ctseries_to_fseries_plan(in, fs[k], length);
		if(!fp)
			goto done;
		correlator_ctseries_to_fseries(fp);
		correlator_fft_plan_free(fp);
	}
	*skyp = sh_series_new_zero(logprior->l_max, 0);
	*skyn = sh_series_new_zero(logprior->l_max, 0);
	if(!*skyp || !*skyn)
		goto done;
	if(generate_alm_sky(*skyp, plansp, fs, gmst, logprior) || generate_alm_sky(*skyn, plansn, fs, gmst, logprior))
		goto done;
	rc = 0;
done:
	if(rc) {
		sh_series_free(*skyp);
		sh_series_free(*skyn);
		*skyp = *skyn = NULL;
	}
	free(fs);
	free(buf);
	free(in);
	return rc;
}

/*
 * Batched localization (sphradiometer_batch.h): the per-event chain of
 * generate_alm_skys + sh_series_scale_power_l folded into constant
 * tables.  Every step after the spectra is linear in the cross spectra
 * X_b[f] = F_ib[f] conj(F_ia[f]), so the two skies of an event are
 *   sky_beta[l, m] = e^{-i m gmst} (sum_{b, f} W_beta[b, f][l, m] X_b[f]
 *                    + s_l prior[l, m]),
 *   W_beta[b, f]   = s_l (1/N)(1/nb) 2 (4 nb) R_b dp'_beta,b[f],
 * with R_b the baseline's rotation (full Wigner-D), s_l the
 * scale_power_l factor overall (l/L)^power, and the SNR preprocessing +
 * c2c DFT one matrix [length][N].
 */
/* exp(sign 2 pi i k / n), k < n, from exact integer angles */
static void twiddles(int n, int sign, double *out)
{
	const long double two_pi = 6.283185307179586476925286766559005768L;
	for(int k = 0; k < n; k++) {
		const long double a = two_pi * (long double) k / (long double) n;
		out[2 * k] = (double) cosl(a);
		out[2 * k + 1] = (double) (sign < 0 ? -sinl(a) : sinl(a));
	}
}

static size_t rot_block_offset(int l)
{
	return ((size_t) 4 * l * l * l - l - 3) / 3;
}

static int loc_check_plans(const struct correlator_network_plan_fd *p, int N, unsigned L)
{
	for(int b = 0; b < p->baselines->n_baselines; b++) {
		const struct correlator_plan_fd *q = p->plans[b];
		if(q->delay_product->n != N || q->delay_product->polar || q->delay_product->l_max > L || q->power_1d->l_max != q->delay_product->l_max || q->rotation_plan->l_max < q->delay_product->l_max)
			return -1;
	}
	return 0;
}

struct sphr_loc *sphr_loc_new(const struct correlator_network_plan_fd *plansp, const struct correlator_network_plan_fd *plansn, const struct sh_series *logprior, unsigned l_max, double power, double overall, int length, double delta_t, int precision, int max_events)
{
	if(!plansp || !plansn || !logprior || logprior->l_max != l_max || logprior->polar || length < 1 || delta_t <= 0. || max_events < 1) {
		sphr_set_error("sphr_loc_new: bad arguments");
		return NULL;
	}
	const int nb = plansp->baselines->n_baselines;
	const int N = plansp->plans[0]->delay_product->n;
	const int D = instrument_array_len(plansp->baselines->baselines[0]->instruments);
	if(nb > 16 || plansn->baselines->n_baselines != nb || N != precalculated_TimeSeries_length(length, delta_t)
	   || loc_check_plans(plansp, N, l_max) || loc_check_plans(plansn, N, l_max)) {
		sphr_set_error("sphr_loc_new: the plans must be multiplied by the projection, share N = precalculated length (%ld) and have l_max <= %u", precalculated_TimeSeries_length(length, delta_t), l_max);
		return NULL;
	}
	const int L = (int) l_max;
	const size_t ncoef = (size_t) (L + 1) * (L + 1);
	double complex *W = calloc((size_t) 2 * nb * N * ncoef, sizeof(*W));
	double complex *prior = malloc(ncoef * sizeof(*prior));
	double complex *dft = malloc((size_t) length * N * sizeof(*dft));
	double *sl = malloc((L + 1) * sizeof(*sl)), *win = malloc(length * sizeof(*win)), *tw = malloc(2 * (size_t) N * sizeof(*tw));
	struct sphr_loc *res = NULL;
	if(!W || !prior || !dft || !sl || !win || !tw)
		goto done;
	for(int l = 0; l <= L; l++)
		sl[l] = overall * pow((double) l / L, power);
	for(int l = 0; l <= L; l++)
		for(int m = -l; m <= l; m++) {
			const size_t c = sh_series_params_lmoffset(L, l, m);
			prior[c] = sl[l] * logprior->coeff[c];
		}
	const double scale = (1.0 / N) * (1.0 / nb) * 2.0 * (nb * 4.0);
	for(int beta = 0; beta < 2; beta++) {
		const struct correlator_network_plan_fd *pl = beta ? plansn : plansp;
		for(int b = 0; b < nb; b++) {
			const struct correlator_plan_fd *q = pl->plans[b];
			const int Lb = (int) q->delay_product->l_max;
			const double complex *blk = (const double complex *) sphr_rotation_blocks(q->rotation_plan);
#pragma omp parallel for
			for(int f = 0; f < N; f++) {
				const double complex *s = q->delay_product->coeff + (size_t) f * q->delay_product->stride;
				double complex *w = W + (((size_t) beta * nb + b) * N + f) * ncoef;
				for(int l = 0; l <= Lb; l++)
					for(int mp = -l; mp <= l; mp++) {
						double complex acc = 0.;
						if(l == 0)
							acc = s[0];
						else {
							const double complex *Dr = blk + rot_block_offset(l) + (size_t) (2 * l + 1) * (mp + l);
							for(int m = -l; m <= l; m++)
								acc += Dr[m + l] * s[sh_series_params_lmoffset(Lb, l, m)];
						}
						w[sh_series_params_lmoffset(L, l, mp)] = scale * sl[l] * acc;
					}
			}
		}
	}
	/* preprocessing + c2c DFT: F[f] = sum_t x[t] w_t sqrt(1/len) sqrt(N)
	 * e^{-2 pi i f (t - first) / N} */
	sphr_tukey_window(length, 0.1, win);
	{
		const long first = (long) round((length - (long) N) / 2.);
		const double s1 = sqrt(1. / length), s2 = sqrt((double) N);
		twiddles(N, -1, tw);
		for(int t = 0; t < length; t++) {
			long dst = t - first;
			for(int f = 0; f < N; f++) {
				const long k = ((long) f * dst) % N;
				dft[(size_t) t * N + f] = (dst >= 0 && dst < N) ? win[t] * s1 * s2 * (tw[2 * k] + I * tw[2 * k + 1]) : 0.;
			}
		}
	}
	struct sphr_loc_desc d;
	memset(&d, 0, sizeof(d));
	d.precision = precision;
	d.D = D;
	d.nb = nb;
	d.N = N;
	d.len = length;
	d.L = L;
	d.max_events = max_events;
	for(int b = 0; b < nb; b++) {
		d.ia[b] = plansp->baselines->baselines[b]->index_a;
		d.ib[b] = plansp->baselines->baselines[b]->index_b;
	}
	d.W = (const double *) W;
	d.prior = (const double *) prior;
	d.dft = (const double *) dft;
	res = sphr_loc_create(&d);
done:
	free(W);
	free(prior);
	free(dft);
	free(sl);
	free(win);
	free(tw);
	return res;
}

/* ------------------------------------------------------------------ */
/* autocorrelator plans (the diagonal of the projection operator)      */
/* ------------------------------------------------------------------ */

/*
 * inspiral_sky_map.c:561-660 (diagonal_projections,
 * autocorrelator_network_plan_fd_new): projections[i][f] is the (i, i)
 * element of the projection operator of polarisation beta for the PSD
 * values psd[f], expanded in the Earth frame at order l_max.  The
 * detectors are the instruments' data pointers, as for
 * correlator_network_plan_mult_by_projection.
 */
struct autocorrelator_network_plan_fd *autocorrelator_network_plan_fd_new(const struct instrument_array *instruments, double beta, double psi, double **psd, int length, int l_max)
{
	if(!instruments || !psd || length < 0 || l_max < 0) {
		sphr_set_error("autocorrelator_network_plan_fd_new: bad arguments");
		return NULL;
	}
	const int n = instrument_array_len(instruments);
	if(n > 64) {
		sphr_set_error("autocorrelator_network_plan_fd_new: %d instruments exceed 64", n);
		return NULL;
	}
	const struct sphr_detector *det[64];
	for(int k = 0; k < n; k++) {
		det[k] = instrument_array_get(instruments, k)->data;
		if(!det[k]) {
			sphr_set_error("instrument %d has no detector response (data == NULL)", k);
			return NULL;
		}
	}
	struct autocorrelator_network_plan_fd *plan = calloc(1, sizeof(*plan));
	struct sh_series ***pr = calloc(n > 0 ? n : 1, sizeof(*pr));
	if(!plan || !pr) {
		free(plan);
		free(pr);
		sphr_set_error("autocorrelator_network_plan_fd_new: out of memory");
		return NULL;
	}
	plan->instruments = instruments;
	plan->length = length;
	plan->projections = pr;
	int fail = 0;
	for(int i = 0; i < n && !fail; i++) {
		pr[i] = calloc(length > 0 ? length : 1, sizeof(**pr));
		fail = !pr[i];
		for(int f = 0; f < length && !fail; f++)
			fail = !(pr[i][f] = sh_series_new(l_max, 0));
	}
	if(!fail) {
#pragma omp parallel for collapse(2) schedule(dynamic)
		for(int i = 0; i < n; i++)
			for(int f = 0; f < length; f++) {
				struct cbc_args args = { i, i, n, det, beta, psi, psd[f] };
				if(!sh_series_from_func(pr[i][f], cbc_projection, &args)) {
#pragma omp atomic write
					fail = 1;
				}
			}
	}
	if(fail) {
		sphr_set_error("autocorrelator_network_plan_fd_new: projection expansion failed");
		autocorrelator_network_plan_fd_free(plan);
		return NULL;
	}
	return plan;
}

/* inspiral_sky_map.c:663-675 */
struct autocorrelator_network_plan_fd *autocorrelator_network_plan_fd_set_l(struct autocorrelator_network_plan_fd *plan, unsigned l_max)
{
	for(int i = 0; i < instrument_array_len(plan->instruments); i++)
		for(int f = 0; f < plan->length; f++)
			if(!sh_series_resize(plan->projections[i][f], l_max))
				return NULL;
	return plan;
}

/* inspiral_sky_map.c:678-691; also releases the partial plans the
 * constructor and the precalc reader leave on failure */
void autocorrelator_network_plan_fd_free(struct autocorrelator_network_plan_fd *plan)
{
	if(!plan)
		return;
	if(plan->projections) {
		for(int i = 0; i < instrument_array_len(plan->instruments); i++) {
			if(plan->projections[i])
				for(int f = 0; f < plan->length; f++)
					sh_series_free(plan->projections[i][f]);
			free(plan->projections[i]);
		}
		free(plan->projections);
	}
	free(plan);
}
