/*
 * GPU back ends of the synchronous drop-in entry points (sh_series_array_
 * dotc, sh_series_rotate(_z), sh_series_to_mesh, the FD integrators and
 * the FFT helpers).  Each call copies its operands to the device, runs the
 * kernels on the library stream and copies the results back before
 * returning, as the reference's functions return with results in host
 * memory.  Calls are serialised: they share the scratch buffers.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "../sphr_internal.h"
#include "kernels.h"
#include "runtime.h"
#include "synth_tables.h"

#define LOCK() pthread_mutex_lock(&sphr_dropin_lock)
#define UNLOCK() pthread_mutex_unlock(&sphr_dropin_lock)

enum { SL_A, SL_B, SL_C, SL_D, SL_E, SL_F, SL_G, SL_H };

void sphr_twiddles(int n, int sign, double *out)
{
	const long double two_pi = 6.283185307179586476925286766559005768L;
	for(int k = 0; k < n; k++) {
		const long double a = two_pi * (long double) k / (long double) n;
		out[2 * k] = (double) cosl(a);
		out[2 * k + 1] = (double) (sign < 0 ? -sinl(a) : sinl(a));
	}
}

int sphr_factor(int n, int *fac, int max)
{
	int k = 0;
	while(n % 4 == 0 && k < max) {
		fac[k++] = 4;
		n /= 4;
	}
	for(int p = 2; n > 1 && k < max; p++)
		while(n % p == 0 && k < max) {
			fac[k++] = p;
			n /= p;
		}
	return n == 1 ? k : -1;
}

/* ---- synthesis tables (host build, device copy) ---- */

int synth_tables_build(int L, struct synth_dev *t, int gemm_tables)
{
	memset(t, 0, sizeof(*t));
	const int B = L + 1, n = 2 * B;
	t->g.L = L;
	t->g.B = B;
	t->g.n = n;
	t->g.nfac = sphr_factor(n, t->g.fac, SPHR_MAX_FACTORS);
	if(t->g.nfac < 0) {
		sphr_set_error("cannot factor the phi transform length %d", n);
		return -1;
	}
	const double *x;
	sphr_mesh_rule(L, &x, NULL);
	/* every table in one device allocation, filled by one copy:
	 * x [B], Pbar_mm [B][L+1], alpha [nal], twiddles [2n], cm [L+1],
	 * the expanded phi-DFT matrix [2 (2L+1)][2 n] in float64 then float32,
	 * the ring sums' recurrence pairs [nal + 2] in float64 then float32 */
	const size_t nal = (size_t) (L + 1) * (L + 2) / 2;
	const int nm = 2 * L + 1;
	const size_t ne = gemm_tables ? (size_t) 2 * nm * 2 * n : 0;
	auto up = [](size_t b) { return (b + 255) & ~(size_t) 255; };
	const size_t o_x = 0, o_s = up(o_x + B * sizeof(double)), o_a = up(o_s + (size_t) B * (L + 1) * sizeof(double));
	const size_t o_t = up(o_a + nal * sizeof(double)), o_c = up(o_t + 2 * (size_t) n * sizeof(double));
	const size_t o_e = up(o_c + (L + 1) * sizeof(double)), o_f = up(o_e + ne * sizeof(double));
	const size_t o_r = up(o_f + ne * sizeof(float)), o_q = up(o_r + (nal + 2) * 2 * sizeof(double));
	const size_t total = o_q + (nal + 2) * 2 * sizeof(float);
	char *hb = (char *) calloc(total, 1);
	double *tmp = (double *) malloc((L + 1) * sizeof(double));
	int rc = -1;
	if(!hb || !tmp) {
		sphr_set_error("synthesis tables: out of memory");
		goto done;
	}
	{
		double *hx = (double *) (hb + o_x), *hs = (double *) (hb + o_s), *ha = (double *) (hb + o_a);
		double *ht = (double *) (hb + o_t), *hc = (double *) (hb + o_c), *e64 = (double *) (hb + o_e);
		float *e32 = (float *) (hb + o_f);
		for(int j = 0; j < B; j++) {
			hx[j] = x[j];
			for(int m = 0; m <= L; m++) {
				sphr_plm(m, m, x[j], tmp);
				hs[(size_t) j * (L + 1) + m] = tmp[0];
			}
		}
		/* alpha_lm = sqrt((4l^2 - 1) / (l^2 - m^2)) at the (l, m >= 0)
		 * slot of the coefficient layout */
		for(int m = 0; m <= L; m++)
			for(int l = m + 1; l <= L; l++)
				ha[lm_offset(L, l, m)] = sqrt((4.0 * l * l - 1.0) / ((double) (l - m) * (l + m)));
		sphr_twiddles(n, +1, ht);
		/* Pbar_mm at the equator is its sectoral constant */
		for(int m = 0; m <= L; m++) {
			double one;
			sphr_plm(m, m, 0.0, &one);
			hc[m] = one;
		}
		/* rows 2 i, 2 i + 1 of the phi DFT for column i (m = i for i <= L,
		 * i - (2L+1) above) */
		for(int i = 0; i < (gemm_tables ? nm : 0); i++) {
			const int m = i <= L ? i : i - nm;
			for(int k = 0; k < n; k++) {
				const long r = (((long) k * m) % n + n) % n;
				const double er = ht[2 * r], ei = ht[2 * r + 1];
				double *r0 = e64 + (size_t) (2 * i) * 2 * n + 2 * k, *r1 = r0 + 2 * n;
				r0[0] = er;
				r0[1] = ei;
				r1[0] = -ei;
				r1[1] = er;
			}
		}
		for(size_t q = 0; q < ne; q++)
			e32[q] = (float) e64[q];
		/* (alpha_l, alpha_l / alpha_{l-1}) per (l, m >= 0) slot, zero
		 * where l = m (and the second where l = m + 1) */
		double *hr = (double *) (hb + o_r);
		float *hq = (float *) (hb + o_q);
		for(int m = 0; m <= L; m++)
			for(int l = m + 1; l <= L; l++) {
				const size_t t = lm_offset(L, l, m);
				hr[2 * t] = ha[t];
				hr[2 * t + 1] = l > m + 1 ? ha[t] / ha[t - 1] : 0.0;
			}
		for(size_t q = 0; q < 2 * (nal + 2); q++)
			hq[q] = (float) hr[q];
	}
	if(sphr_rt_check(hipMalloc(&t->blob, total), "hipMalloc") || sphr_rt_check(hipMemcpy(t->blob, hb, total, hipMemcpyHostToDevice), "hipMemcpy"))
		goto done;
	{
		char *d = (char *) t->blob;
		t->x = (double *) (d + o_x);
		t->smm = (double *) (d + o_s);
		t->alpha = (double *) (d + o_a);
		t->twp = (double *) (d + o_t);
		t->cm = (double *) (d + o_c);
		t->E64 = gemm_tables ? (double *) (d + o_e) : NULL;
		t->E32 = gemm_tables ? (float *) (d + o_f) : NULL;
		t->tb.rc64 = (const cplx<double> *) (d + o_r);
		t->tb.rc32 = (const cplx<float> *) (d + o_q);
	}
	t->tb.x = t->x;
	t->tb.smm = t->smm;
	t->tb.alpha = t->alpha;
	t->tb.twp = t->twp;
	t->tb.E64 = t->E64;
	t->tb.E32 = t->E32;
	if(gemm_tables && sky_fused_order(L)) {
		if(sky_fused_tables(L, &t->tb, &t->fz))
			goto done;
		t->tb.bf64 = t->fz.bf64;
		t->tb.bf32 = t->fz.bf32;
		t->tb.sk64 = t->fz.sk64;
		t->tb.sk32 = t->fz.sk32;
	}
	rc = 0;
done:
	free(hb);
	free(tmp);
	if(rc)
		synth_tables_free(t);
	return rc;
}

void synth_tables_free(struct synth_dev *t)
{
	if(t->blob)
		hipFree(t->blob);
	if(t->fz.bf64)
		hipFree(t->fz.bf64);
	if(t->fz.bf32)
		hipFree(t->fz.bf32);
	if(t->fz.sk64)
		hipFree(t->fz.sk64);
	if(t->fz.sk32)
		hipFree(t->fz.sk32);
	memset(t, 0, sizeof(*t));
}

/* drop-in caches (accessed under the drop-in lock) */
struct synth_cache_entry {
	struct synth_dev t;
	struct synth_cache_entry *next;
};
static struct synth_cache_entry *synth_cache;

static struct synth_dev *synth_for(int L)
{
	for(struct synth_cache_entry *e = synth_cache; e; e = e->next)
		if(e->t.g.L == L)
			return &e->t;
	struct synth_cache_entry *e = (struct synth_cache_entry *) calloc(1, sizeof(*e));
	if(!e || synth_tables_build(L, &e->t)) {
		free(e);
		return NULL;
	}
	e->next = synth_cache;
	synth_cache = e;
	return &e->t;
}

struct tw_cache_entry {
	int n;
	cplx<double> *tw;
	struct tw_cache_entry *next;
};
static struct tw_cache_entry *tw_cache;

static const cplx<double> *twiddles_for(int n)
{
	for(struct tw_cache_entry *e = tw_cache; e; e = e->next)
		if(e->n == n)
			return e->tw;
	double *h = (double *) malloc(2 * (size_t) n * sizeof(double));
	struct tw_cache_entry *e = (struct tw_cache_entry *) calloc(1, sizeof(*e));
	if(!h || !e) {
		free(h);
		free(e);
		return NULL;
	}
	sphr_twiddles(n, -1, h);
	if(sphr_rt_check(hipMalloc(&e->tw, 2 * (size_t) n * sizeof(double)), "hipMalloc") || sphr_rt_check(hipMemcpy(e->tw, h, 2 * (size_t) n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy")) {
		free(h);
		free(e);
		return NULL;
	}
	free(h);
	e->n = n;
	e->next = tw_cache;
	tw_cache = e;
	return e->tw;
}

static size_t block_offset_host(int l)
{
	return ((size_t) 4 * l * l * l - l - 3) / 3;
}

/* D^l_{0 m'} rows of packed blocks, in the coefficient layout of order L */
void sphr_polar_rows(int L, const double *blocks, double *rows)
{
	rows[0] = 1.0;
	rows[1] = 0.0;
	for(int l = 1; l <= L; l++)
		for(int mp = -l; mp <= l; mp++) {
			const size_t src = block_offset_host(l) + (size_t) (2 * l + 1) * (mp + l) + l;
			const int dst = lm_offset(L, l, mp);
			rows[2 * dst] = blocks[2 * src];
			rows[2 * dst + 1] = blocks[2 * src + 1];
		}
}

#define CK(x, what) do { if(sphr_rt_check((x), what)) goto fail; } while(0)

extern "C" int sphr_gpu_dotc(const double *coeff, int n, int stride, const double *vec, int vec_is_real, double *result)
{
	if(sphr_gpu_ensure())
		return -1;
	if(n <= 0) {
		memset(result, 0, 2 * (size_t) stride * sizeof(double));
		return 0;
	}
	LOCK();
	hipStream_t st = sphr_rt_stream();
	const int nsplit = n >= 64 * 256 ? 256 : (n + 63) / 64;
	cplx<double> *dc = (cplx<double> *) sphr_rt_scratch(SL_A, (size_t) n * stride * sizeof(cplx<double>));
	void *dv = sphr_rt_scratch(SL_B, (size_t) n * sizeof(cplx<double>));
	cplx<double> *dp = (cplx<double> *) sphr_rt_scratch(SL_C, (size_t) nsplit * stride * sizeof(cplx<double>));
	cplx<double> *dr = (cplx<double> *) sphr_rt_scratch(SL_D, (size_t) stride * sizeof(cplx<double>));
	if(!dc || !dv || !dp || !dr)
		goto fail;
	CK(hipMemcpyAsync(dc, coeff, (size_t) n * stride * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "copy coefficients");
	CK(hipMemcpyAsync(dv, vec, (size_t) n * (vec_is_real ? 1 : 2) * sizeof(double), hipMemcpyHostToDevice, st), "copy vector");
	if(launch_dotc(dc, n, stride, vec_is_real ? NULL : (const cplx<double> *) dv, vec_is_real ? (const double *) dv : NULL, 1.0, dp, nsplit, dr, st))
		goto launch;
	CK(hipMemcpyAsync(result, dr, (size_t) stride * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "copy result");
	CK(hipStreamSynchronize(st), "synchronize");
	UNLOCK();
	return 0;
launch:
	sphr_set_error("dotc kernel launch failed");
fail:
	UNLOCK();
	return -1;
}

extern "C" int sphr_gpu_to_mesh(unsigned l_max, int polar, const double *coeff, double *mesh)
{
	if(sphr_gpu_ensure())
		return -1;
	LOCK();
	hipStream_t st = sphr_rt_stream();
	const int L = (int) l_max, n = 2 * (L + 1);
	const size_t nc = polar ? (size_t) L + 1 : (size_t) (L + 1) * (L + 1);
	struct synth_dev *t = synth_for(L);
	cplx<double> *da = (cplx<double> *) sphr_rt_scratch(SL_A, nc * sizeof(cplx<double>));
	cplx<double> *dm = (cplx<double> *) sphr_rt_scratch(SL_B, (size_t) n * n * sizeof(cplx<double>));
	/* pinned staging both ways: one DMA each, no driver-side bounce */
	cplx<double> *hin = (cplx<double> *) sphr_rt_pinned(2, nc * sizeof(cplx<double>));
	cplx<double> *hout = (cplx<double> *) sphr_rt_pinned(3, (size_t) n * n * sizeof(cplx<double>));
	if(!t || !da || !dm || !hin || !hout)
		goto fail;
	memcpy(hin, coeff, nc * sizeof(cplx<double>));
	CK(hipMemcpyAsync(da, hin, nc * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "copy coefficients");
	if(launch_synth<double, double, double>(da, 1, polar, &t->g, &t->tb, dm, st)) {
		sphr_set_error("synthesis launch failed");
		goto fail;
	}
	CK(hipMemcpyAsync(hout, dm, (size_t) n * n * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "copy mesh");
	CK(hipStreamSynchronize(st), "synchronize");
	memcpy(mesh, hout, (size_t) n * n * sizeof(cplx<double>));
	UNLOCK();
	return 0;
fail:
	UNLOCK();
	return -1;
}

extern "C" int sphr_gpu_rotate(unsigned l_max, int polar, const double *coeff, const double *blocks, double *result)
{
	if(sphr_gpu_ensure())
		return -1;
	const int L = (int) l_max;
	const size_t nc = (size_t) (L + 1) * (L + 1);
	double *rows = NULL;
	LOCK();
	hipStream_t st = sphr_rt_stream();
	cplx<double> *din = (cplx<double> *) sphr_rt_scratch(SL_A, (polar ? L + 1 : nc) * sizeof(cplx<double>));
	cplx<double> *dout = (cplx<double> *) sphr_rt_scratch(SL_B, nc * sizeof(cplx<double>));
	const size_t nbl = polar ? nc : block_offset_host(L + 1);
	cplx<double> *dbl = (cplx<double> *) sphr_rt_scratch(SL_C, (nbl ? nbl : 1) * sizeof(cplx<double>));
	if(!din || !dout || !dbl)
		goto fail;
	CK(hipMemcpyAsync(din, coeff, (polar ? L + 1 : nc) * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "copy series");
	if(polar) {
		rows = (double *) malloc(2 * nc * sizeof(double));
		if(!rows)
			goto fail;
		sphr_polar_rows(L, blocks, rows);
		CK(hipMemcpyAsync(dbl, rows, nc * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "copy rotation rows");
	} else if(nbl && L > 0) {
		CK(hipMemcpyAsync(dbl, blocks, nbl * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "copy rotation blocks");
	}
	{
		const cplx<double> *pw[1] = { din };
		const cplx<double> *bk[1] = { dbl };
		const int Lb[1] = { L }, pol[1] = { polar };
		if(launch_rotsum(1, pw, Lb, pol, bk, L, 1.0, dout, st)) {
			sphr_set_error("rotation launch failed");
			goto fail;
		}
	}
	CK(hipMemcpyAsync(result, dout, nc * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "copy result");
	CK(hipStreamSynchronize(st), "synchronize");
	UNLOCK();
	free(rows);
	return 0;
fail:
	UNLOCK();
	free(rows);
	return -1;
}

extern "C" int sphr_gpu_rotate_z(unsigned l_max, const double *coeff, double phi, double *result)
{
	if(sphr_gpu_ensure())
		return -1;
	const size_t nc = (size_t) (l_max + 1) * (l_max + 1);
	LOCK();
	hipStream_t st = sphr_rt_stream();
	cplx<double> *din = (cplx<double> *) sphr_rt_scratch(SL_A, nc * sizeof(cplx<double>));
	cplx<double> *dout = (cplx<double> *) sphr_rt_scratch(SL_B, nc * sizeof(cplx<double>));
	if(!din || !dout)
		goto fail;
	CK(hipMemcpyAsync(din, coeff, nc * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "copy series");
	if(launch_rotate_z((int) l_max, din, phi, dout, st)) {
		sphr_set_error("rotate_z launch failed");
		goto fail;
	}
	CK(hipMemcpyAsync(result, dout, nc * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "copy result");
	CK(hipStreamSynchronize(st), "synchronize");
	UNLOCK();
	return 0;
fail:
	UNLOCK();
	return -1;
}

/*
 * Device copies of FD plans, kept across calls: each plan's delay product
 * and its rotation table (the D^l_{0 m'} rows for a polar power, the packed
 * D blocks otherwise), keyed by the plan and checked against the host
 * arrays and shapes they were made from.  The library's plan mutators
 * (correlator_plan_fd_free, _set_l, mult_by_projection) drop a plan's
 * entry through sphr_gpu_plan_forget; a caller that writes into a plan's
 * delay_product or rotation_plan in place calls it too.  Callers hold the
 * drop-in lock.
 */
#define PLAN_CACHE 64
struct plan_dev {
	const void *key;
	const double *dp_host, *blocks_host;
	int n, stride, polar;
	unsigned l_max;
	cplx<double> *d_dp, *d_rot;
	unsigned long long used;
};
static struct plan_dev plan_cache[PLAN_CACHE];
static unsigned long long plan_tick;

static void plan_entry_clear(struct plan_dev *e)
{
	if(e->d_dp)
		hipFree(e->d_dp);
	if(e->d_rot)
		hipFree(e->d_rot);
	memset(e, 0, sizeof(*e));
}

static size_t rot_count(const struct sphr_fd_baseline *b)
{
	return b->polar ? (size_t) (b->l_max + 1) * (b->l_max + 1) : (b->l_max > 0 ? block_offset_host(b->l_max + 1) : 1);
}

static const struct plan_dev *plan_device(const struct sphr_fd_baseline *b, int n)
{
	struct plan_dev *e = NULL, *victim = &plan_cache[0];
	for(int i = 0; i < PLAN_CACHE; i++) {
		if(b->key && plan_cache[i].key == b->key) {
			e = &plan_cache[i];
			break;
		}
		if(plan_cache[i].used < victim->used)
			victim = &plan_cache[i];
	}
	if(e && e->dp_host == b->dp && e->blocks_host == b->blocks && e->n == n && e->stride == b->stride && e->polar == b->polar && e->l_max == b->l_max) {
		e->used = ++plan_tick;
		return e;
	}
	if(!e)
		e = victim;
	plan_entry_clear(e);
	const size_t ndp = (size_t) n * b->stride, nrot = rot_count(b);
	double *rows = NULL;
	if(sphr_rt_check(hipMalloc(&e->d_dp, ndp * sizeof(cplx<double>)), "hipMalloc plan") || sphr_rt_check(hipMalloc(&e->d_rot, nrot * sizeof(cplx<double>)), "hipMalloc plan"))
		goto fail;
	if(sphr_rt_check(hipMemcpy(e->d_dp, b->dp, ndp * sizeof(cplx<double>), hipMemcpyHostToDevice), "copy delay product"))
		goto fail;
	if(b->polar) {
		rows = (double *) malloc(2 * nrot * sizeof(double));
		if(!rows)
			goto fail;
		if(b->l_max > 0)
			sphr_polar_rows((int) b->l_max, b->blocks, rows);
		else {
			rows[0] = 1.0;
			rows[1] = 0.0;
		}
		if(sphr_rt_check(hipMemcpy(e->d_rot, rows, nrot * sizeof(cplx<double>), hipMemcpyHostToDevice), "copy rotation rows"))
			goto fail;
		free(rows);
		rows = NULL;
	} else if(b->l_max > 0 && sphr_rt_check(hipMemcpy(e->d_rot, b->blocks, nrot * sizeof(cplx<double>), hipMemcpyHostToDevice), "copy rotation blocks"))
		goto fail;
	e->key = b->key;
	e->dp_host = b->dp;
	e->blocks_host = b->blocks;
	e->n = n;
	e->stride = b->stride;
	e->polar = b->polar;
	e->l_max = b->l_max;
	e->used = ++plan_tick;
	return e;
fail:
	free(rows);
	plan_entry_clear(e);
	return NULL;
}

extern "C" void sphr_gpu_plan_forget(const void *key)
{
	if(!key)
		return;
	LOCK();
	for(int i = 0; i < PLAN_CACHE; i++)
		if(plan_cache[i].key == key)
			plan_entry_clear(&plan_cache[i]);
	UNLOCK();
}

extern "C" int sphr_gpu_network_fd(int n_inst, const double *const *fseries, int n, int nb, const struct sphr_fd_baseline *bl, unsigned sky_l, double scale, double *sky)
{
	if(sphr_gpu_ensure())
		return -1;
	if(nb > SPHR_MAX_BASELINES || n_inst > 2 * SPHR_MAX_BASELINES || n <= 0) {
		sphr_set_error("network of %d baselines / %d instruments / %d samples is outside the supported range", nb, n_inst, n);
		return -1;
	}
	const size_t nsky = (size_t) (sky_l + 1) * (sky_l + 1);
	size_t pw_total = 0, pw_off[SPHR_MAX_BASELINES];
	for(int b = 0; b < nb; b++) {
		pw_off[b] = pw_total;
		pw_total += bl[b].stride;
	}
	const int nsplit = n >= 64 * 256 ? 256 : (n + 63) / 64;
	int maxstride = 1;
	for(int b = 0; b < nb; b++)
		maxstride = bl[b].stride > maxstride ? bl[b].stride : maxstride;
	const struct plan_dev *pd[SPHR_MAX_BASELINES];
	LOCK();
	hipStream_t st = sphr_rt_stream();
	/* staging of the spectra in (pinned) and of the sky, powers and
	 * cross spectra out */
	const size_t in_bytes = (size_t) n_inst * n * sizeof(cplx<double>);
	const size_t out_bytes = (nsky + pw_total + (size_t) nb * n) * sizeof(cplx<double>);
	cplx<double> *dfs = (cplx<double> *) sphr_rt_scratch(SL_A, in_bytes);
	cplx<double> *dout = (cplx<double> *) sphr_rt_scratch(SL_B, out_bytes);
	cplx<double> *dpart = (cplx<double> *) sphr_rt_scratch(SL_F, (size_t) nb * nsplit * maxstride * sizeof(cplx<double>));
	cplx<double> *hin = (cplx<double> *) sphr_rt_pinned(0, in_bytes);
	cplx<double> *hout = (cplx<double> *) sphr_rt_pinned(1, out_bytes);
	cplx<double> *dsky = dout, *dpw = dout + nsky, *dfsp = dpw + pw_total;
	if(!dfs || !dout || !dpart || !hin || !hout)
		goto fail;
	for(int b = 0; b < nb; b++)
		if(!(pd[b] = plan_device(&bl[b], n)))
			goto fail;
	{
		const cplx<double> *fsd[2 * SPHR_MAX_BASELINES];
		int ia[SPHR_MAX_BASELINES], ib[SPHR_MAX_BASELINES];
		for(int i = 0; i < n_inst; i++) {
			memcpy(hin + (size_t) i * n, fseries[i], (size_t) n * sizeof(cplx<double>));
			fsd[i] = dfs + (size_t) i * n;
		}
		CK(hipMemcpyAsync(dfs, hin, in_bytes, hipMemcpyHostToDevice, st), "copy spectra");
		for(int b = 0; b < nb; b++) {
			ia[b] = bl[b].index_a;
			ib[b] = bl[b].index_b;
		}
		if(launch_xspec(fsd, n, nb, ia, ib, dfsp, st)) {
			sphr_set_error("cross-spectrum launch failed");
			goto fail;
		}
	}
	{
		/* power_1d = (1/n) sum_f dp[f] X[f] (correlator.c:690-703), all
		 * baselines in one pair of launches */
		struct dotc_net_args da;
		for(int b = 0; b < nb; b++) {
			da.coeff[b] = pd[b]->d_dp;
			da.vec[b] = dfsp + (size_t) b * n;
			da.partial[b] = dpart + (size_t) b * nsplit * maxstride;
			da.out[b] = dpw + pw_off[b];
			da.stride[b] = bl[b].stride;
		}
		if(launch_dotc_net(&da, nb, n, nsplit, 1.0 / n, st)) {
			sphr_set_error("projection launch failed");
			goto fail;
		}
	}
	{
		const cplx<double> *pw[SPHR_MAX_BASELINES], *bk[SPHR_MAX_BASELINES];
		int Lb[SPHR_MAX_BASELINES], pol[SPHR_MAX_BASELINES];
		for(int b = 0; b < nb; b++) {
			pw[b] = dpw + pw_off[b];
			bk[b] = pd[b]->d_rot;
			Lb[b] = (int) bl[b].l_max;
			pol[b] = bl[b].polar;
		}
		if(launch_rotsum(nb, pw, Lb, pol, bk, (int) sky_l, scale, dsky, st)) {
			sphr_set_error("rotation launch failed");
			goto fail;
		}
	}
	/* one copy back: sky, the baselines' powers and cross spectra */
	CK(hipMemcpyAsync(hout, dout, out_bytes, hipMemcpyDeviceToHost, st), "copy results");
	CK(hipStreamSynchronize(st), "synchronize");
	memcpy(sky, hout, nsky * sizeof(cplx<double>));
	for(int b = 0; b < nb; b++) {
		memcpy(bl[b].power_1d, hout + nsky + pw_off[b], (size_t) bl[b].stride * sizeof(cplx<double>));
		if(bl[b].fsp)
			memcpy(bl[b].fsp, hout + nsky + pw_total + (size_t) b * n, (size_t) n * sizeof(cplx<double>));
	}
	UNLOCK();
	return 0;
fail:
	UNLOCK();
	return -1;
}

/* ---- FFT helpers ---- */

struct sphr_fft {
	int n, real_input;
	double *in, *out;
};

extern "C" struct sphr_fft *sphr_fft_new(int n, int real_input, double *in, double *out)
{
	if(sphr_gpu_ensure())
		return NULL;
	struct sphr_fft *f = (struct sphr_fft *) malloc(sizeof(*f));
	if(!f)
		return NULL;
	f->n = n;
	f->real_input = real_input;
	f->in = in;
	f->out = out;
	return f;
}

extern "C" void sphr_fft_free(struct sphr_fft *f)
{
	free(f);
}

extern "C" int sphr_fft_exec(struct sphr_fft *f)
{
	if(!f || sphr_gpu_ensure())
		return -1;
	const int n = f->n;
	const size_t in_bytes = (size_t) n * (f->real_input ? 1 : 2) * sizeof(double);
	LOCK();
	hipStream_t st = sphr_rt_stream();
	const cplx<double> *tw = twiddles_for(n);
	double *din = (double *) sphr_rt_scratch(SL_A, in_bytes);
	cplx<double> *dout = (cplx<double> *) sphr_rt_scratch(SL_B, (size_t) n * sizeof(cplx<double>));
	/* pinned staging both ways */
	double *hin = (double *) sphr_rt_pinned(2, in_bytes);
	cplx<double> *hout = (cplx<double> *) sphr_rt_pinned(3, (size_t) n * sizeof(cplx<double>));
	if(!tw || !din || !dout || !hin || !hout)
		goto fail;
	memcpy(hin, f->in, in_bytes);
	CK(hipMemcpyAsync(din, hin, in_bytes, hipMemcpyHostToDevice, st), "copy series");
	{
		const int pow2 = !(n & (n - 1));
		const int r = (pow2 && (size_t) n * sizeof(cplx<double>) <= sphr_rt_lds_max()) ? launch_fft_single(din, f->real_input, n, tw, dout, st) : launch_dft_direct(din, f->real_input, n, tw, dout, st);
		if(r) {
			sphr_set_error("FFT launch failed");
			goto fail;
		}
	}
	CK(hipMemcpyAsync(hout, dout, (size_t) n * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "copy spectrum");
	CK(hipStreamSynchronize(st), "synchronize");
	memcpy(f->out, hout, (size_t) n * sizeof(cplx<double>));
	UNLOCK();
	return 0;
fail:
	UNLOCK();
	return -1;
}

/* ---- extremum (sh_series_real_minimum) ---- */

static double *seed_x_dev;

static const double *seed_nodes(void)
{
	if(!seed_x_dev) {
		const double *x;
		sphr_mesh_rule(SPHR_L_SEED, &x, NULL);
		const int n = 2 * (SPHR_L_SEED + 1);
		if(sphr_rt_check(hipMalloc(&seed_x_dev, n * sizeof(double)), "hipMalloc") || sphr_rt_check(hipMemcpy(seed_x_dev, x, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy"))
			return NULL;
	}
	return seed_x_dev;
}

extern "C" const double *sphr_gpu_seed_nodes(void)
{
	return seed_nodes();
}

extern "C" int sphr_gpu_extremize(unsigned l_max, const double *coeff, int find_max, double *theta, double *phi, double *value)
{
	if(sphr_gpu_ensure())
		return -1;
	const int L = (int) l_max;
	const size_t nc = (size_t) (L + 1) * (L + 1);
	LOCK();
	hipStream_t st = sphr_rt_stream();
	struct synth_dev *t = synth_for(L);
	const double *sx = seed_nodes();
	cplx<double> *da = (cplx<double> *) sphr_rt_scratch(SL_A, nc * sizeof(cplx<double>));
	double *dout = (double *) sphr_rt_scratch(SL_B, 3 * sizeof(double));
	double hout[3];
	if(!t || !sx || !da || !dout)
		goto fail;
	CK(hipMemcpyAsync(da, coeff, nc * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "copy series");
	if(launch_extremize<double>(da, 1, L, find_max, SPHR_L_SEED, NULL, NULL, t->cm, t->alpha, sx, 2 * (SPHR_L_SEED + 1), 100, dout, dout + 1, dout + 2, st)) {
		sphr_set_error("extremize launch failed");
		goto fail;
	}
	CK(hipMemcpyAsync(hout, dout, 3 * sizeof(double), hipMemcpyDeviceToHost, st), "copy result");
	CK(hipStreamSynchronize(st), "synchronize");
	UNLOCK();
	*theta = hout[0];
	*phi = hout[1];
	*value = hout[2];
	return 0;
fail:
	UNLOCK();
	return -1;
}

/* peak refinement from given seeds on device series (batched API) */
extern "C" int sphr_refine_peaks(const void *alm, int alm_f32, int n_series, int alm_div, int l_max, const double *seeds, int find_max, double *theta, double *phi, double *value, void *stream)
{
	if(sphr_gpu_ensure())
		return -1;
	if(!alm || !seeds || n_series < 1 || alm_div < 1 || !theta || !phi || !value) {
		sphr_set_error("sphr_refine_peaks: bad arguments");
		return -1;
	}
	LOCK();
	struct synth_dev *t = synth_for(l_max);
	UNLOCK();
	if(!t)
		return -1;
	hipStream_t st = stream ? (hipStream_t) stream : sphr_rt_stream();
	const int r = alm_f32 ? launch_extremize_seeded<float>((const cplx<float> *) alm, n_series, alm_div, l_max, find_max, seeds, t->cm, t->alpha, 100, theta, phi, value, st)
			      : launch_extremize_seeded<double>((const cplx<double> *) alm, n_series, alm_div, l_max, find_max, seeds, t->cm, t->alpha, 100, theta, phi, value, st);
	if(r)
		sphr_set_error("sphr_refine_peaks: launch failed");
	return r;
}
