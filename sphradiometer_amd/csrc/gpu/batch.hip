/*
 * Batched device pipeline (sphradiometer_batch.h): FFT -> cross spectrum x
 * projection -> sky coefficients -> mesh synthesis for many segments per
 * launch, in float64 (SPHR_FP64) or float32 (SPHR_MIXED).
 *
 * Plan tables built once per batch object from the host network plan:
 *   W     [N+2][CT] the delay product folded for real data: for baseline
 *         columns 2l, 2l+1 and frequency f <= N/2,
 *           rows 2f   : (dp[f] + dp[N-f]) / N        (dp[f] alone at 0, N/2)
 *           rows 2f+1 : i (dp[f] - dp[N-f]) / N      (i dp[f] at 0, N/2)
 *         so sum_{f<N} dp[f] X[f] / N = sum_k A[k] W[k] with
 *         A = (Re X[f], Im X[f]) over the half spectrum (X[N-f] = conj X[f]
 *         for real time series)
 *   R     [nb][(L+1)^2] (1/nb) D^l_{0 m'} of each baseline's rotation, in
 *         the sky coefficient layout (zero above the baseline's l_max)
 *   synthesis tables of order L and the FFT twiddles
 *   M     the lag-domain projection tables of xcorr.hip, where that route
 *         applies (two or three detectors, 1024 samples, or two detectors
 *         at 4096 in float32): the FFT, cross spectra and projection then
 *         run as one kernel and the spectra and K splits are not used
 */
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sphradiometer_batch.h>
#include "../sphr_internal.h"
#include "kernels.h"
#include "runtime.h"
#include "synth_tables.h"

struct ev_set {
	hipEvent_t e[5];
};

struct sphr_batch {
	int precision, D, nb, N, Nh, L, max_segments, nsplit, k_per_split;
	int tpw;			/* projection tiles per wave (autotuned) */
	int xms;			/* projection segments per workgroup, 16 or 32 (autotuned) */
	int fused;			/* use the one-kernel sky stage where compiled (default 1) */
	int sky_occ;			/* fused sky workgroups per CU, 1 to 3 (autotuned) */
	int sky_nsub;			/* fused sky segments per workgroup, 0 (default), 1, 2 or 4 (autotuned) */
	int ksplit_opt[4], nsplit_opt[4];	/* K-split candidates: 1, 2, 4 x CUs workgroups, 64 per split */
	int shape;			/* the projection shape index in use (sphr_batch_set_shape) */
	int lag_ok, lag;		/* lag-domain route built / in use */
	size_t m_bytes;			/* its table's bytes */
	int xc_forms, xc_form;		/* its kernel forms (bit f - 1 for form f), the one in use (autotuned) */
	struct xproj_geom g;
	struct xcorr_geom xg;
	struct synth_dev syn;
	void *W, *R, *tw, *spec, *partial, *F, *M, *twW;
	int *col_bl;
	int2 *lm;
	int profile;
	struct ev_set *ev;
	int ev_used, ev_cap;
};

/* the projection shapes the autotuner tries: one baseline per workgroup
 * (tpw 0), then tiles per wave x segments per workgroup; with the 4 K
 * splits, shape index = 7 x split + shape */
#define SPHR_TUNE_W 7
static const int tune_tpw[SPHR_TUNE_W] = { 0, 1, 2, 4, 1, 2, 4 }, tune_ms[SPHR_TUNE_W] = { 64, 16, 16, 16, 32, 32, 32 };
#define SPHR_BATCH_SHAPES (4 * SPHR_TUNE_W)

static size_t block_offset_h(int l)
{
	return ((size_t) 4 * l * l * l - l - 3) / 3;
}

template <typename T>
static int upload(void **dst, const T *src, size_t count)
{
	if(sphr_rt_check(hipMalloc(dst, count * sizeof(T) + 16), "hipMalloc"))
		return -1;
	return sphr_rt_check(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

template <typename T>
static int build_tables(struct sphr_batch *bt, const struct sphr_net_desc *d)
{
	const int N = bt->N, Nh = bt->Nh, CT = bt->g.CT, K2 = bt->g.K2, L = bt->L, nb = bt->nb;
	const size_t ncoef = (size_t) (L + 1) * (L + 1);
	T *W = (T *) calloc((size_t) K2 * CT, sizeof(T));
	T *R = (T *) calloc(2 * ncoef * nb, sizeof(T));
	T *tw = (T *) malloc(2 * (size_t) N * sizeof(T));
	double *twd = (double *) malloc(2 * (size_t) N * sizeof(double));
	int rc = -1;
	if(!W || !R || !tw || !twd)
		goto done;
	for(int b = 0; b < nb; b++) {
		const int Lb = d->Lb[b], c0 = bt->g.col_start[b];
		const double *dp = d->dp[b];		/* N x (Lb+1) complex */
		for(int f = 0; f < Nh; f++) {
			const int fn = (N - f) % N;
			const int self = (f == 0 || 2 * f == N);
			for(int l = 0; l <= Lb; l++) {
				const double ar = dp[2 * ((size_t) f * (Lb + 1) + l)], ai = dp[2 * ((size_t) f * (Lb + 1) + l) + 1];
				const double br = self ? 0.0 : dp[2 * ((size_t) fn * (Lb + 1) + l)];
				const double bi = self ? 0.0 : dp[2 * ((size_t) fn * (Lb + 1) + l) + 1];
				/* G = a + b, H = i (a - b) */
				const double gr = ar + br, gi = ai + bi;
				const double hr = -(ai - bi), hi = ar - br;
				W[(size_t) (2 * f) * CT + c0 + 2 * l] = (T) (gr / N);
				W[(size_t) (2 * f) * CT + c0 + 2 * l + 1] = (T) (gi / N);
				W[(size_t) (2 * f + 1) * CT + c0 + 2 * l] = (T) (hr / N);
				W[(size_t) (2 * f + 1) * CT + c0 + 2 * l + 1] = (T) (hi / N);
			}
		}
		/* rotation rows (sh_series_rotation.c:566-577), averaged over
		 * baselines (correlator.c:1004) */
		T *Rb = R + 2 * ncoef * b;
		Rb[0] = (T) (1.0 / nb);
		for(int l = 1; l <= Lb; l++)
			for(int mp = -l; mp <= l; mp++) {
				const size_t src = block_offset_h(l) + (size_t) (2 * l + 1) * (mp + l) + l;
				const int dst = lm_offset(L, l, mp);
				Rb[2 * dst] = (T) (d->blocks[b][2 * src] / nb);
				Rb[2 * dst + 1] = (T) (d->blocks[b][2 * src + 1] / nb);
			}
	}
	sphr_twiddles(N, -1, twd);
	for(int k = 0; k < 2 * N; k++)
		tw[k] = (T) twd[k];
	if(upload(&bt->W, W, (size_t) K2 * CT) || upload(&bt->R, R, 2 * ncoef * nb) || upload(&bt->tw, tw, 2 * (size_t) N))
		goto done;
	rc = 0;
done:
	free(W);
	free(R);
	free(tw);
	free(twd);
	return rc;
}

/*
 * The lag-domain tables (xcorr.hip): for each baseline,
 *   M[u][l] = (1/N) sum_f dp[f][l] exp(-2 pi i f u / N)
 * from its own delay product.  The route is taken only where the kernel is
 * compiled for the shape and the plan has the properties the kernel relies
 * on, each checked against the largest |M| of the baseline:
 *   - M vanishes past |u| = 2 transient (largest dropped value <= 1e-13),
 *   - M is real (imaginary parts <= 1e-12),
 *   - M[-u][l] = (-1)^l M[u][l] (to 1e-12).
 * The table holds Ms[u][l] / (4 N), Ms the symmetric part (the mean of
 * M[u][l] and (-1)^l M[-u][l]), as [chunk][row][l] with lag
 * u = chunk SPHR_XC_MR + row (zero past the last lag); 4 N undoes the
 * scale of the kernel's packed transforms.  0 with
 * the route off when it does not apply; -1 on an allocation failure.
 */
template <typename T>
static int build_lag(struct sphr_batch *bt, const struct sphr_net_desc *d)
{
	const int N = bt->N, nb = bt->nb;
	struct xcorr_geom *g = &bt->xg;
	bt->lag_ok = bt->lag = 0;
	bt->xc_forms = xcorr_forms<T>(N, bt->D, nb);
	bt->xc_form = 1;
	if(!bt->xc_forms)
		return 0;
	memset(g, 0, sizeof(*g));
	g->D = bt->D;
	g->nb = nb;
	g->CT = bt->g.CT;
	size_t total = 0;
	for(int b = 0; b < nb; b++) {
		const int nl = d->Lb[b] + 1, U = 2 * d->transient[b];
		if(nl > 256 || U < 0 || 2 * U + 1 > N)
			return 0;
		/* chunks of SPHR_XC_MR lags, one thread per (chunk, l) */
		const int rows = SPHR_XC_MR, nch = (U + 1 + rows - 1) / rows;
		if(nch * nl > 256 || nch * rows > N / 2)
			return 0;
		g->ia[b] = bt->g.ia[b];
		g->ib[b] = bt->g.ib[b];
		g->nl[b] = nl;
		g->nch[b] = nch;
		g->rows[b] = rows;
		g->U1[b] = U + 1;
		g->col_start[b] = bt->g.col_start[b];
		g->width[b] = (b + 1 < nb ? bt->g.col_start[b + 1] : bt->g.CT) - bt->g.col_start[b];
		g->moff[b] = (int) total;
		total += (size_t) nch * rows * nl;
	}
	double *col = (double *) malloc(2 * (size_t) N * sizeof(double));
	double *Mr = (double *) malloc((size_t) N * sizeof(double));
	T *tab = (T *) calloc(total, sizeof(T));
	int rc = -1, ok = 1;
	if(!col || !Mr || !tab)
		goto done;
	for(int b = 0; b < nb && ok; b++) {
		const int nl = g->nl[b], U = g->U1[b] - 1;
		const double *dp = d->dp[b];
		double mmax = 0.0, tail = 0.0, imag = 0.0, asym = 0.0;
		T *tb = tab + g->moff[b];
		/* the largest |M| of the baseline first, then the checks and
		 * the table, one coefficient at a time */
		for(int pass = 0; pass < 2; pass++)
			for(int l = 0; l < nl; l++) {
				for(int f = 0; f < N; f++) {
					col[2 * f] = dp[2 * ((size_t) f * nl + l)];
					col[2 * f + 1] = dp[2 * ((size_t) f * nl + l) + 1];
				}
				sphr_host_dft(col, N, -1);
				for(int u = 0; u < N; u++) {
					const double re = col[2 * u] / N, im = col[2 * u + 1] / N;
					const int uu = u < N - u ? u : N - u;	/* |u| */
					if(pass == 0) {
						const double a = sqrt(re * re + im * im);
						mmax = a > mmax ? a : mmax;
						continue;
					}
					Mr[u] = re;
					if(uu > U)
						tail = fmax(tail, sqrt(re * re + im * im));
					else
						imag = fmax(imag, fabs(im));
				}
				if(pass == 0)
					continue;
				const double sg = (l & 1) ? -1.0 : 1.0;
				for(int u = 0; u <= U; u++) {
					const double ms = u == 0 ? Mr[0] : 0.5 * (Mr[u] + sg * Mr[N - u]);
					if(u > 0)
						asym = fmax(asym, fabs(Mr[u] - sg * Mr[N - u]));
					const int ch = u / g->rows[b], j = u - ch * g->rows[b];
					tb[((size_t) ch * g->rows[b] + j) * nl + l] = (T) (ms / (4.0 * N));
				}
			}
		if(!(mmax > 0.0) || tail > 1e-13 * mmax || imag > 1e-12 * mmax || asym > 1e-12 * mmax)
			ok = 0;
	}
	if(ok) {
		if(upload(&bt->M, tab, total))
			goto done;
		bt->m_bytes = total * sizeof(T);
		if(bt->xc_forms & 4) {
			/* the wave-level form's twiddle tables (SPHR_XC_TWW) */
			T *twl = (T *) malloc(2 * (size_t) SPHR_XC_TWW * sizeof(T));
			double *tw64 = (double *) malloc(2 * (size_t) N * sizeof(double));
			int bad = !twl || !tw64;
			if(!bad) {
				sphr_twiddles(N, -1, tw64);
				for(int k = 0; k < 16; k++)
					for(int i = 0; i < 64; i++) {
						const int e1 = (i * k) % N, e2 = (16 * (i & 3) * k) % N;
						twl[2 * (k * 64 + i)] = (T) tw64[2 * e1];
						twl[2 * (k * 64 + i) + 1] = (T) tw64[2 * e1 + 1];
						twl[2 * (1024 + k * 64 + i)] = (T) tw64[2 * e2];
						twl[2 * (1024 + k * 64 + i) + 1] = (T) tw64[2 * e2 + 1];
					}
				bad = upload(&bt->twW, twl, 2 * SPHR_XC_TWW);
			}
			free(twl);
			free(tw64);
			if(bad)
				goto done;
		}
		bt->lag_ok = bt->lag = 1;
	}
	rc = 0;
done:
	free(col);
	free(Mr);
	free(tab);
	return rc;
}

extern "C" void sphr_batch_free(struct sphr_batch *bt)
{
	if(!bt)
		return;
	void *ptrs[] = { bt->W, bt->R, bt->tw, bt->spec, bt->partial, bt->col_bl, bt->lm, bt->F, bt->M, bt->twW };
	for(unsigned i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++)
		if(ptrs[i])
			hipFree(ptrs[i]);
	synth_tables_free(&bt->syn);
	for(int i = 0; i < bt->ev_cap; i++)
		for(int k = 0; k < 5; k++)
			hipEventDestroy(bt->ev[i].e[k]);
	free(bt->ev);
	free(bt);
}

extern "C" struct sphr_batch *sphr_batch_create(const struct sphr_net_desc *d, unsigned sky_l, int precision, int max_segments)
{
	if(sphr_gpu_ensure())
		return NULL;
	const int N = d->N;
	if(N < 64 || (N & (N - 1))) {
		sphr_set_error("batched path needs a power-of-two segment length >= 64 (got %d)", N);
		return NULL;
	}
	if(d->nb < 1 || d->nb > SPHR_MAX_BASELINES || d->n_det > 2 * SPHR_MAX_BASELINES || max_segments < 1) {
		sphr_set_error("unsupported network shape");
		return NULL;
	}
	const size_t esz = precision == SPHR_MIXED ? sizeof(float) : sizeof(double);
	if((size_t) N * 2 * esz > sphr_rt_lds_max()) {
		sphr_set_error("segment length %d does not fit the FFT's LDS tile", N);
		return NULL;
	}
	struct sphr_batch *bt = (struct sphr_batch *) calloc(1, sizeof(*bt));
	if(!bt)
		return NULL;
	bt->precision = precision;
	bt->D = d->n_det;
	bt->nb = d->nb;
	bt->N = N;
	bt->Nh = N / 2 + 1;
	bt->L = (int) sky_l;
	bt->max_segments = max_segments;
	bt->g.D = d->n_det;
	bt->g.nb = d->nb;
	bt->g.N = N;
	bt->g.Nh = bt->Nh;
	bt->g.K2 = 2 * bt->Nh;
	int CT = 0;
	for(int b = 0; b < d->nb; b++) {
		if(!d->polar[b] || d->Lb[b] > (int) sky_l) {
			sphr_set_error("baseline %d: the batched path needs polar powers with l_max <= the sky's", b);
			free(bt);
			return NULL;
		}
		bt->g.ia[b] = d->ia[b];
		bt->g.ib[b] = d->ib[b];
		bt->g.Lb[b] = d->Lb[b];
		bt->g.col_start[b] = CT;
		CT += (2 * (d->Lb[b] + 1) + 15) / 16 * 16;	/* whole MFMA tiles */
	}
	bt->g.CT = CT;
	/* split K so the projection launch has enough workgroups; three
	 * candidates (about 1, 2 and 4 workgroups per CU), the middle one by
	 * default, the autotuner may pick another */
	{
		const int base = ((max_segments + 15) / 16) * ((CT / 16 + 15) / 16);
		const int maxs = (bt->g.K2 + 63) / 64;
		for(int o = 0; o < 3; o++) {
			const int target = (1 << o) * sphr_rt_cus();
			int ns = (target + base - 1) / base;
			ns = ns < 1 ? 1 : ns > maxs ? maxs : ns;
			int kps = (bt->g.K2 + ns - 1) / ns;
			kps = (kps + 31) / 32 * 32;
			bt->ksplit_opt[o] = kps;
			bt->nsplit_opt[o] = (bt->g.K2 + kps - 1) / kps;
		}
		/* one 64-wide chunk of K per workgroup */
		bt->ksplit_opt[3] = 64;
		bt->nsplit_opt[3] = (bt->g.K2 + 63) / 64;
		bt->k_per_split = bt->ksplit_opt[1];
		bt->nsplit = bt->nsplit_opt[1];
		bt->tpw = 0;
		bt->xms = 64;
		bt->shape = SPHR_TUNE_W * 1 + 0;
		bt->fused = 1;
		bt->sky_occ = 1;
		bt->sky_nsub = 0;
	}
	int *colbl = (int *) malloc((CT / 16) * sizeof(int));
	const int ncoef = (bt->L + 1) * (bt->L + 1);
	int2 *lm = (int2 *) malloc(ncoef * sizeof(int2));
	if(!colbl || !lm)
		goto fail;
	for(int b = 0; b < d->nb; b++)
		for(int c = 0; c < 2 * (d->Lb[b] + 1); c += 16)
			colbl[(bt->g.col_start[b] + c) / 16] = b;
	for(int i = 0; i < ncoef; i++) {
		int l, m;
		lm_decode(bt->L, i, &l, &m);
		lm[i].x = l;
		lm[i].y = m;
	}
	if(upload((void **) &bt->col_bl, colbl, CT / 16) || upload((void **) &bt->lm, lm, ncoef))
		goto fail;
	free(colbl);
	free(lm);
	colbl = NULL;
	lm = NULL;
	if((precision == SPHR_MIXED ? build_tables<float>(bt, d) : build_tables<double>(bt, d)))
		goto fail;
	if((precision == SPHR_MIXED ? build_lag<float>(bt, d) : build_lag<double>(bt, d)))
		goto fail;
	if(synth_tables_build(bt->L, &bt->syn))
		goto fail;
	if(sphr_rt_check(hipMalloc(&bt->spec, (size_t) bt->D * max_segments * bt->Nh * 2 * esz), "hipMalloc spectra")
	   || sphr_rt_check(hipMalloc(&bt->partial, (size_t) bt->nsplit_opt[3] * max_segments * CT * esz), "hipMalloc partials")
	   || sphr_rt_check(hipMalloc(&bt->F, (size_t) max_segments * 2 * (bt->L + 1) * (2 * bt->L + 1) * 2 * esz), "hipMalloc ring sums"))
		goto fail;
	return bt;
fail:
	free(colbl);
	free(lm);
	sphr_batch_free(bt);
	return NULL;
}

/* one run on the given scratch: spectra, K-split partials and (for the
 * two-kernel sky form) ring sums; F may be NULL, then the sky stage uses
 * the fused kernel or the one-kernel form */
template <typename T>
static int run_t(struct sphr_batch *bt, void *spec, void *partial, void *F, const void *ts, int f32, int S, const double *window, const double *gmst, void *alm, void *map, hipStream_t st, hipEvent_t *ev)
{
	if(ev)
		hipEventRecord(ev[0], st);
	if(bt->lag) {
		/* FFT, cross spectra and projection in one kernel: stage 0 */
		if(launch_xcorr<T>(ts, f32, window, S, bt->N, &bt->xg, (const cplx<T> *) bt->tw, (const cplx<T> *) bt->twW, (const T *) bt->M, (T *) partial, bt->xc_form, st))
			return -1;
		if(ev) {
			hipEventRecord(ev[1], st);
			hipEventRecord(ev[2], st);
		}
	} else {
		if(launch_rfft_pairs<T>(ts, f32, window, bt->D * S, bt->N, (const cplx<T> *) bt->tw, (cplx<T> *) spec, st))
			return -1;
		if(ev)
			hipEventRecord(ev[1], st);
		if(launch_xproj<T>((const cplx<T> *) spec, S, &bt->g, (const T *) bt->W, bt->col_bl, bt->k_per_split, bt->nsplit, (T *) partial, st, bt->tpw, bt->xms))
			return -1;
		if(ev)
			hipEventRecord(ev[2], st);
	}
	const int nsplit = bt->lag ? 1 : bt->nsplit;
	/* the fused sky kernel at the orders it is compiled for; else a_lm
	 * and ring sums, then the phi DFT as a GEMM; the one-kernel form when
	 * the coefficients do not fit the LDS budget */
	if(bt->fused && launch_sky_fused<T>((const T *) partial, nsplit, S, &bt->g, (const cplx<T> *) bt->R, bt->lm, gmst, &bt->syn.g, &bt->syn.tb, (cplx<T> *) alm, (cplx<T> *) map, st, bt->sky_occ, bt->sky_nsub) == 0) {
		if(ev)
			hipEventRecord(ev[3], st);
	} else if((F || !map) && launch_sky_lf<T>((const T *) partial, nsplit, S, &bt->g, (const cplx<T> *) bt->R, bt->lm, gmst, &bt->syn.g, &bt->syn.tb, (cplx<T> *) alm, map ? (cplx<T> *) F : (cplx<T> *) NULL, st) == 0) {
		if(ev)
			hipEventRecord(ev[3], st);
		if(map && launch_phi<T>((const cplx<T> *) F, S, &bt->syn.g, &bt->syn.tb, (cplx<T> *) map, st))
			return -1;
	} else {
		if(ev)
			hipEventRecord(ev[3], st);
		if(launch_sky<T>((const T *) partial, nsplit, S, &bt->g, (const cplx<T> *) bt->R, bt->lm, gmst, &bt->syn.g, &bt->syn.tb, (cplx<T> *) alm, (cplx<T> *) map, st))
			return -1;
	}
	if(ev)
		hipEventRecord(ev[4], st);
	return 0;
}

extern "C" int sphr_batch_run(struct sphr_batch *bt, const void *ts, int f32, int S, const double *window, const double *gmst, void *alm, void *map, void *stream)
{
	if(!bt || S < 1 || S > bt->max_segments || !ts || (!alm && map) || !alm) {
		sphr_set_error("sphr_batch_run: bad arguments (segments %d of at most %d)", S, bt ? bt->max_segments : 0);
		return -1;
	}
	hipStream_t st = stream ? (hipStream_t) stream : sphr_rt_stream();
	hipEvent_t *ev = NULL;
	if(bt->profile) {
		if(bt->ev_used == bt->ev_cap) {
			const int cap = bt->ev_cap ? 2 * bt->ev_cap : 64;
			struct ev_set *n = (struct ev_set *) realloc(bt->ev, cap * sizeof(*n));
			if(!n)
				return -1;
			bt->ev = n;
			for(int i = bt->ev_cap; i < cap; i++)
				for(int k = 0; k < 5; k++)
					hipEventCreate(&bt->ev[i].e[k]);
			bt->ev_cap = cap;
		}
		ev = bt->ev[bt->ev_used++].e;
	}
	const int r = bt->precision == SPHR_MIXED ? run_t<float>(bt, bt->spec, bt->partial, bt->F, ts, f32, S, window, gmst, alm, map, st, ev) : run_t<double>(bt, bt->spec, bt->partial, bt->F, ts, f32, S, window, gmst, alm, map, st, ev);
	if(r)
		sphr_set_error("sphr_batch_run: kernel launch failed (%s)", hipGetErrorString(hipGetLastError()));
	return r;
}

/*
 * The allocation-free run.  Its scratch (spectra, K-split partials, ring
 * sums) lives in one file-scope device array instead of the batch
 * object's buffers.  Each GPU holds its own copy of the array, at its own
 * address, from the library's load until the process ends.  In host code
 * the array's name is only a stand-in: its device address comes from
 * hipGetSymbolAddress, looked up once per device and kept in
 * scratch_dev[].  No kernel reads scratch it did not write in the same
 * run, so nothing is cleared between calls.
 */
#define SPHR_SCRATCH_BYTES ((size_t) 64 << 20)
#define SPHR_SCRATCH_DEVICES 64
static __device__ __align__(256) unsigned char scratch[SPHR_SCRATCH_BYTES];
static void *scratch_dev[SPHR_SCRATCH_DEVICES];
static pthread_mutex_t scratch_lock = PTHREAD_MUTEX_INITIALIZER;

static void *scratch_address(void)
{
	int dev = -1;
	if(sphr_rt_check(hipGetDevice(&dev), "hipGetDevice"))
		return NULL;
	if(dev < 0 || dev >= SPHR_SCRATCH_DEVICES) {
		sphr_set_error("sphr_batch_run_static: device %d beyond the %d the library serves", dev, SPHR_SCRATCH_DEVICES);
		return NULL;
	}
	pthread_mutex_lock(&scratch_lock);
	if(!scratch_dev[dev]) {
		void *p = NULL;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess || !p)
			sphr_set_error("sphr_batch_run_static: hipGetSymbolAddress of the scratch array failed (%s)", hipGetErrorString(e));
		else
			scratch_dev[dev] = p;
	}
	void *p = scratch_dev[dev];
	pthread_mutex_unlock(&scratch_lock);
	return p;
}

static size_t align256(size_t b)
{
	return (b + 255) & ~(size_t) 255;
}

extern "C" int sphr_batch_run_static(struct sphr_batch *bt, const void *ts, int f32, int S, const double *window, const double *gmst, void *alm, void *map, void *stream)
{
	if(!bt || S < 1 || !ts || !alm) {
		sphr_set_error("sphr_batch_run_static: bad arguments");
		return -1;
	}
	if(sphr_gpu_ensure())
		return -1;
	const size_t esz = bt->precision == SPHR_MIXED ? sizeof(float) : sizeof(double);
	const int need_F = map && !(bt->fused && sky_fused_order(bt->L));
	const size_t b_spec = bt->lag ? 0 : align256((size_t) bt->D * S * bt->Nh * 2 * esz);
	const size_t b_part = align256((size_t) (bt->lag ? 1 : bt->nsplit) * S * bt->g.CT * esz);
	const size_t b_F = need_F ? align256((size_t) S * 2 * (bt->L + 1) * (2 * bt->L + 1) * 2 * esz) : 0;
	if(b_spec + b_part + b_F > SPHR_SCRATCH_BYTES) {
		sphr_set_error("sphr_batch_run_static: %d segments need %zu bytes of scratch, more than the static %zu; run fewer segments per call", S, b_spec + b_part + b_F, SPHR_SCRATCH_BYTES);
		return -1;
	}
	unsigned char *base = (unsigned char *) scratch_address();
	if(!base)
		return -1;
	void *spec = base, *partial = base + b_spec, *F = need_F ? (void *) (base + b_spec + b_part) : NULL;
	hipStream_t st = stream ? (hipStream_t) stream : sphr_rt_stream();
	const int r = bt->precision == SPHR_MIXED ? run_t<float>(bt, spec, partial, F, ts, f32, S, window, gmst, alm, map, st, NULL) : run_t<double>(bt, spec, partial, F, ts, f32, S, window, gmst, alm, map, st, NULL);
	if(r)
		sphr_set_error("sphr_batch_run_static: kernel launch failed (%s)", hipGetErrorString(hipGetLastError()));
	return r;
}

extern "C" int sphr_batch_profile(struct sphr_batch *bt, int enable)
{
	bt->profile = enable;
	bt->ev_used = 0;
	return 0;
}

extern "C" int sphr_batch_profile_read(struct sphr_batch *bt, double ms[4], int *runs)
{
	for(int k = 0; k < 4; k++)
		ms[k] = 0.0;
	for(int i = 0; i < bt->ev_used; i++) {
		if(sphr_rt_check(hipEventSynchronize(bt->ev[i].e[4]), "hipEventSynchronize"))
			return -1;
		for(int k = 0; k < 4; k++) {
			float t = 0.f;
			hipEventElapsedTime(&t, bt->ev[i].e[k], bt->ev[i].e[k + 1]);
			ms[k] += t;
		}
	}
	if(runs)
		*runs = bt->ev_used;
	bt->ev_used = 0;
	return 0;
}

extern "C" void sphr_batch_set_fused(struct sphr_batch *bt, int on)
{
	if(bt)
		bt->fused = on ? 1 : 0;
}

extern "C" int sphr_batch_set_sky_occupancy(struct sphr_batch *bt, int workgroups_per_cu)
{
	if(!bt || workgroups_per_cu < 1 || workgroups_per_cu > 3) {
		sphr_set_error("sphr_batch_set_sky_occupancy: 1, 2 or 3 workgroups per CU");
		return -1;
	}
	bt->sky_occ = workgroups_per_cu;
	return 0;
}

extern "C" int sphr_batch_set_sky_segments(struct sphr_batch *bt, int segments_per_workgroup)
{
	const int n = segments_per_workgroup;
	if(!bt || !(n == 0 || n == 1 || n == 2 || n == 4) || (n == 4 && bt->precision == SPHR_FP64)) {
		sphr_set_error("sphr_batch_set_sky_segments: 0, 1, 2 or (mixed precision) 4 segments per workgroup");
		return -1;
	}
	bt->sky_nsub = n;
	return 0;
}

extern "C" void sphr_batch_info(const struct sphr_batch *bt, int info[13])
{
	info[12] = bt->sky_nsub;
	info[11] = bt->sky_occ;
	info[8] = bt->fused && sky_fused_order(bt->L);
	info[9] = bt->tpw;
	info[10] = bt->xms;
	info[0] = bt->D;
	info[1] = bt->nb;
	info[2] = bt->N;
	info[3] = bt->L;
	info[4] = bt->g.CT;
	info[5] = bt->nsplit;
	info[6] = bt->precision;
	info[7] = bt->k_per_split;
}

extern "C" int sphr_batch_peaks(struct sphr_batch *bt, const void *alm, const void *map, int S, int find_max, double *theta, double *phi, double *value, void *stream)
{
	if(!bt || S < 1 || S > bt->max_segments || !alm || !theta || !phi || !value) {
		sphr_set_error("sphr_batch_peaks: bad arguments (segments %d of at most %d)", S, bt ? bt->max_segments : 0);
		return -1;
	}
	hipStream_t st = stream ? (hipStream_t) stream : sphr_rt_stream();
	const double *sx = map ? NULL : sphr_gpu_seed_nodes();
	if(!map && !sx)
		return -1;
	const int ns = 2 * (SPHR_L_SEED + 1);
	int r;
	if(bt->precision == SPHR_MIXED)
		r = launch_extremize<float>((const cplx<float> *) alm, S, bt->L, find_max, SPHR_L_SEED, (const cplx<float> *) map, bt->syn.x, bt->syn.cm, bt->syn.alpha, sx, ns, 100, theta, phi, value, st);
	else
		r = launch_extremize<double>((const cplx<double> *) alm, S, bt->L, find_max, SPHR_L_SEED, (const cplx<double> *) map, bt->syn.x, bt->syn.cm, bt->syn.alpha, sx, ns, 100, theta, phi, value, st);
	if(r)
		sphr_set_error("sphr_batch_peaks: kernel launch failed (%s)", hipGetErrorString(hipGetLastError()));
	return r;
}

/* autotuner: time the projection's tile width (1, 2, 4 tiles per wave)
 * and K split (about 1, 2, 4 workgroups per CU) on the caller's data and
 * keep the fastest whole run; returns the chosen candidate */

static void set_shape(struct sphr_batch *bt, int o, int w)
{
	bt->shape = SPHR_TUNE_W * o + w;
	bt->k_per_split = bt->ksplit_opt[o];
	bt->nsplit = bt->nsplit_opt[o];
	bt->tpw = tune_tpw[w];
	bt->xms = tune_ms[w];
}

extern "C" int sphr_batch_set_shape(struct sphr_batch *bt, int index)
{
	if(!bt || index < 0 || index >= SPHR_BATCH_SHAPES) {
		sphr_set_error("sphr_batch_set_shape: index %d outside 0..%d", index, SPHR_BATCH_SHAPES - 1);
		return -1;
	}
	set_shape(bt, index / SPHR_TUNE_W, index % SPHR_TUNE_W);
	return 0;
}

extern "C" int sphr_batch_tune(struct sphr_batch *bt, const void *ts, int f32, int S, const double *window, const double *gmst, void *alm, void *map, double *best_ms)
{
	if(!bt || S < 1 || S > bt->max_segments || !ts || !alm) {
		sphr_set_error("sphr_batch_tune: bad arguments");
		return -1;
	}
	hipStream_t st = sphr_rt_stream();
	const int kps0 = bt->k_per_split, ns0 = bt->nsplit, tpw0 = bt->tpw, xms0 = bt->xms, occ0 = bt->sky_occ, nsub0 = bt->sky_nsub;
	const int shape0 = bt->shape, xform0 = bt->xc_form, lag0 = bt->lag;
	const int prof = bt->profile;
	hipEvent_t e0 = NULL, e1 = NULL;
	float best = 1e30f;
	int bo = -1, bw = 0, bocc = 1, bnsub = 0;
	bt->profile = 0;
	/* time one candidate: a run, then the fastest of three repetitions
	 * of five timed runs, per three runs (the unit of best_ms); -1 on
	 * failure */
	auto timed = [&](float *ms) -> int {
		int r = sphr_batch_run(bt, ts, f32, S, window, gmst, alm, map, st);
		float best = 1e30f;
		for(int rep = 0; rep < 3 && !r; rep++) {
			r = sphr_rt_check(hipEventRecord(e0, st), "tune");
			for(int i = 0; i < 5 && !r; i++)
				r = sphr_batch_run(bt, ts, f32, S, window, gmst, alm, map, st);
			float t = 0.f;
			r = r || sphr_rt_check(hipEventRecord(e1, st), "tune") || sphr_rt_check(hipEventSynchronize(e1), "tune") || sphr_rt_check(hipEventElapsedTime(&t, e0, e1), "tune");
			if(!r && t < best)
				best = t;
		}
		*ms = best * 3.0f / 5.0f;
		return r ? -1 : 0;
	};
	if(sphr_rt_check(hipEventCreate(&e0), "event") || sphr_rt_check(hipEventCreate(&e1), "event"))
		goto fail;
	/* the projection's shapes and K splits with the fused sky kernel in
	 * its default form, then the sky kernel's other forms at the best */
	bt->sky_occ = 1;
	bt->sky_nsub = 0;
	{
		/* with the lag-domain route in use: its kernel forms, then the
		 * frequency-domain route's projection shapes, and the faster
		 * route stays (few segments leave the lag kernel, one segment
		 * pair per workgroup, with too few workgroups); with the
		 * frequency-domain route selected, its shapes only */
		int bx = xform0, lag_best = 0;
		if(lag0) {
			bt->lag = 1;
			for(int x = 1; x <= 4; x++) {
				if(!(bt->xc_forms & (1 << (x - 1))))
					continue;
				bt->xc_form = x;
				float ms = 0.f;
				if(timed(&ms))
					goto fail;
				if(ms < best) {
					best = ms;
					bx = x;
					lag_best = 1;
				}
			}
		}
		bt->xc_form = bx;
		bt->lag = 0;
		for(int o = 0; o < 4; o++)
			for(int w = 0; w < SPHR_TUNE_W; w++) {
				set_shape(bt, o, w);
				float ms = 0.f;
				if(timed(&ms))
					goto fail;
				if(ms < best) {
					best = ms;
					bo = o;
					bw = w;
					lag_best = 0;
				}
			}
		if(bo < 0) {
			bo = shape0 / SPHR_TUNE_W;
			bw = shape0 % SPHR_TUNE_W;
		}
		set_shape(bt, bo, bw);
		bt->lag = lag_best;
	}
	if(bt->fused && sky_fused_order(bt->L) && map) {
		/* (workgroups per CU, segments per workgroup) */
		static const int forms[3][2] = { { 2, 0 }, { 3, 1 }, { 1, 4 } };
		for(int f = 0; f < 3; f++) {
			if(forms[f][1] == 4 && bt->precision == SPHR_FP64)
				continue;
			bt->sky_occ = forms[f][0];
			bt->sky_nsub = forms[f][1];
			float ms = 0.f;
			if(timed(&ms))
				goto fail;
			if(ms < best) {
				best = ms;
				bocc = forms[f][0];
				bnsub = forms[f][1];
			}
		}
	}
	bt->sky_occ = bocc;
	bt->sky_nsub = bnsub;
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	bt->profile = prof;
	if(best_ms)
		*best_ms = best / 3.0;
	return SPHR_TUNE_W * bo + bw;
fail:
	if(e0)
		hipEventDestroy(e0);
	if(e1)
		hipEventDestroy(e1);
	bt->k_per_split = kps0;
	bt->nsplit = ns0;
	bt->tpw = tpw0;
	bt->xms = xms0;
	bt->sky_occ = occ0;
	bt->sky_nsub = nsub0;
	bt->shape = shape0;
	bt->xc_form = xform0;
	bt->lag = lag0;
	bt->profile = prof;
	return -1;
}

extern "C" int sphr_batch_set_lag(struct sphr_batch *bt, int on)
{
	if(!bt || (on && !bt->lag_ok)) {
		sphr_set_error("sphr_batch_set_lag: the lag-domain route does not apply to this batch");
		return -1;
	}
	bt->lag = on ? 1 : 0;
	return 0;
}

extern "C" int sphr_batch_get_lag(const struct sphr_batch *bt)
{
	return bt && bt->lag ? 1 : 0;
}

extern "C" size_t sphr_batch_table_bytes(const struct sphr_batch *bt)
{
	if(!bt)
		return 0;
	return bt->lag ? bt->m_bytes : (size_t) bt->g.K2 * bt->g.CT * (bt->precision == SPHR_MIXED ? sizeof(float) : sizeof(double));
}

extern "C" int sphr_batch_set_xcorr_form(struct sphr_batch *bt, int form)
{
	if(!bt || form < 1 || form > 4 || !(bt->xc_forms & (1 << (form - 1)))) {
		sphr_set_error("sphr_batch_set_xcorr_form: form %d is not compiled for this batch", form);
		return -1;
	}
	bt->xc_form = form;
	return 0;
}

extern "C" int sphr_batch_get_xcorr_form(const struct sphr_batch *bt)
{
	return bt && bt->lag ? bt->xc_form : 0;
}
