/*
 * Time-domain correlator (correlator.c:325-400, 611-673, 820-858,
 * 967-984 of the reference) with the integration on the GPU.
 *
 * The reference integrates sample by sample: for each output sample t it
 * dots the two delay-tap projections with the series (sample_a(t) =
 * sum_k proj_a[k] a[t + k], sample_b likewise), multiplies the two polar
 * series (sh_series_product) and adds window[t] times the product to
 * power_1d.  As functions of x = cos theta, sample_a(t)(x) =
 * sum_k T^a_k(x) a[t + k] with T^a_k the synthesis of proj_a[k], and the
 * product's coefficient d is 2 pi int sample_a sample_b Pbar_d0 dx, which a
 * Gauss-Legendre rule of n_q = (a + b + P) / 2 + 1 nodes integrates
 * exactly.  Summed over t:
 *   power_1d[d] = sum_q 2 pi w_q Pbar_d0(x_q) U_q,
 *   U_q = sum_{k,k'} T^a_k(x_q) C[k][k'] T^b_k'(x_q),
 *   C[k][k'] = sum_t window[t] a[t + k] b[t + k'],
 * the same arithmetic up to rounding, with the per-sample work folded into
 * the windowed correlation matrix C (n_p x n_p, a GEMM over t with the
 * Hankel operands generated from the series in the tile loader, float64
 * MFMA), then Y = C T^b and its bilinear form with T^a in one pass
 * (k_td_yu) and the d sums (k_td_pw).  For a window that is flat over
 * most of the series, C comes from the samples where the window changes
 * (the same GEMM over that support gives the diagonal differences of C),
 * its first row and column, and a sum along each diagonal.
 * The rotation to the baseline direction and the network average run on
 * the device (k_rotsum) from the rows D^l_{0 m'} of the plan's rotation
 * plan, uploaded once when the plan is made; a call copies back only
 * power_1d and the rotated (or summed) series.
 *
 * The plan keeps the reference's struct (proj_a, proj_b, sample_a,
 * sample_b, product, power_1d, product_plan, rotation_plan) as the first
 * member of a larger allocation that holds the device tables.  After an
 * integration power_1d holds the axially symmetric power, as in the
 * reference; sample_a, sample_b and product are not updated.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/correlator.h>
#include <sphradiometer/projection.h>
#include <sphradiometer/sh_series.h>
#include <sphradiometer/sphradiometer_gpu.h>
#include "../sphr_internal.h"
#include "kernels.h"
#include "runtime.h"

struct td_plan {
	struct correlator_plan_td pub;	/* the reference layout, first */
	int np, nq, P;
	double *d_Ta, *d_Tb;		/* [np][nq]: tap projections at the product nodes */
	double *d_WP;			/* [nq][P + 1]: 2 pi w_q Pbar_d0(x_q) */
	/* per-call device buffers, grown as needed */
	size_t cap_n, cap_part;
	double *d_a, *d_b, *d_w, *d_part, *d_C;
	double *d_U;			/* [ceil(np / 16)][nq]: partial bilinear forms */
	double *d_powc;			/* power_1d as P + 1 complex (imaginary parts 0) */
	double *d_rows;			/* (P + 1)^2 complex: D^l_{0 m'} of the rotation plan */
	size_t cap_out, cap_sv, cap_edge;
	double *d_sv;			/* window differences on their support: values, then offsets */
	double *d_edge;			/* first row and column of C */
	double *d_out;			/* rotated series (or a network's partial sums) */
	/* a network's series and windows (kept on its first plan) */
	size_t cap_nts, cap_nws;
	double *d_nts, *d_nws;
};

namespace {

/*
 * partial[z][k][k'] = sum_{t in split z} w[t] a[s_t + k] b[s_t + k'],
 * k, k' < np, with s_t = t or, for a gathered support, s_t = sidx[t]
 * (reads at or past na, the series length, count as 0): 64 x 64 tiles of 4 waves (32 x 32 each) on
 * v_mfma_f64_16x16x4f64, t in chunks of TD_KC through LDS, the operands read
 * from the series (Hankel structure) with the next chunk's loads in flight
 * during the current chunk's MFMAs.  MFMA layout: A[i][k] in lane
 * i + 16 k, B[k][j] in lane j + 16 k, accumulator element e of lane
 * j + 16 q holds D[q + 4 e][j].
 */
#define TD_TILE 64
#define TD_KC 32
/* operand elements per thread per chunk */
#define TD_LR (TD_KC * TD_TILE / 256)
/* output samples per K split */
#define TD_TSPLIT 1024

__global__ void __launch_bounds__(256) k_td_corr(const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ w, const double *__restrict__ sidx, int na, int nt, int np, int tpz, int tiles, double *__restrict__ part)
{
	typedef double v4 __attribute__((ext_vector_type(4)));
	__shared__ double sA[TD_KC][TD_TILE + 4];	/* [t][k] */
	__shared__ double sB[TD_KC][TD_TILE + 4];	/* [t][k'] */
	const int tile = blockIdx.x % (tiles * tiles), z = blockIdx.x / (tiles * tiles);
	const int m0 = (tile % tiles) * TD_TILE, n0 = (tile / tiles) * TD_TILE;
	const int t0 = z * tpz, t1 = t0 + tpz < nt ? t0 + tpz : nt;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int li = lane & 15, lk = lane >> 4;
	const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
	double ra[TD_LR], rb[TD_LR];
	auto load = [&](int c0) {
#pragma unroll
		for(int r = 0; r < TD_LR; r++) {
			const int e = tid + 256 * r;
			const int t = c0 + e / TD_TILE, k = e % TD_TILE;
			const bool ok = t < t1;
			/* sample offset of term t: t itself, or the t-th entry of a
			 * gathered support */
			const int st = sidx && ok ? (int) sidx[t] : t;
			ra[r] = ok && m0 + k < np && st + m0 + k < na ? w[t] * a[st + m0 + k] : 0.0;
			rb[r] = ok && n0 + k < np && st + n0 + k < na ? b[st + n0 + k] : 0.0;
		}
	};
	v4 acc[2][2];
#pragma unroll
	for(int i = 0; i < 2; i++)
#pragma unroll
		for(int j = 0; j < 2; j++)
			acc[i][j] = (v4) 0;
	if(t0 < t1)
		load(t0);
	for(int c0 = t0; c0 < t1; c0 += TD_KC) {
#pragma unroll
		for(int r = 0; r < TD_LR; r++) {
			const int e = tid + 256 * r;
			sA[e / TD_TILE][e % TD_TILE] = ra[r];
			sB[e / TD_TILE][e % TD_TILE] = rb[r];
		}
		__syncthreads();
		if(c0 + TD_KC < t1)
			load(c0 + TD_KC);
#pragma unroll
		for(int ks = 0; ks < TD_KC; ks += 4) {
			double af[2], bf[2];
#pragma unroll
			for(int i = 0; i < 2; i++)
				af[i] = sA[ks + lk][wm + 16 * i + li];
#pragma unroll
			for(int j = 0; j < 2; j++)
				bf[j] = sB[ks + lk][wn + 16 * j + li];
#pragma unroll
			for(int i = 0; i < 2; i++)
#pragma unroll
				for(int j = 0; j < 2; j++)
					acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
		}
		__syncthreads();
	}
	double *P = part + (size_t) z * np * np;
#pragma unroll
	for(int i = 0; i < 2; i++)
#pragma unroll
		for(int j = 0; j < 2; j++) {
			const int c = n0 + wn + 16 * j + li;
			if(c >= np)
				continue;
#pragma unroll
			for(int e = 0; e < 4; e++) {
				const int r = m0 + wm + 16 * i + lk + 4 * e;
				if(r < np)
					P[(size_t) r * np + c] = acc[i][j][e];
			}
		}
}

/* C = sum_z partial[z] */
__global__ void k_td_reduce(const double *__restrict__ part, int nz, long n, double *__restrict__ C)
{
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n)
		return;
	double s = 0.0;
	for(int z = 0; z < nz; z++)
		s += part[(size_t) z * n + i];
	C[i] = s;
}

/* part[c][j] = sum_{t in chunk c} w[t] a[t] b[t + j] and part[c][np + j]
 * = sum_{t in chunk c} w[t] a[t + j] b[t], j < np (the first row and
 * column of C, summed over the chunks by k_td_reduce): TD_EC samples per
 * workgroup, staged in LDS with the np - 1 samples after them, each
 * thread one or more j of both */
#define TD_EC 1024
#define TD_EDGE_NP_MAX 1024
__global__ void __launch_bounds__(256) k_td_edge(const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ w, int nt, int n, int np, double *__restrict__ part)
{
	__shared__ double wa[TD_EC], wb[TD_EC], ya[TD_EC + TD_EDGE_NP_MAX], yb[TD_EC + TD_EDGE_NP_MAX];
	const int t0 = blockIdx.x * TD_EC;
	for(int t = threadIdx.x; t < TD_EC; t += 256) {
		const bool ok = t0 + t < nt;
		wa[t] = ok ? w[t0 + t] * a[t0 + t] : 0.0;
		wb[t] = ok ? w[t0 + t] * b[t0 + t] : 0.0;
	}
	for(int t = threadIdx.x; t < TD_EC + np - 1; t += 256) {
		const bool ok = t0 + t < n;
		ya[t] = ok ? a[t0 + t] : 0.0;
		yb[t] = ok ? b[t0 + t] : 0.0;
	}
	__syncthreads();
	for(int jj = threadIdx.x; jj < np; jj += 256) {
		double r = 0.0, c = 0.0;
		for(int t = 0; t < TD_EC; t++) {
			r = fma(wa[t], yb[t + jj], r);
			c = fma(wb[t], ya[t + jj], c);
		}
		part[(size_t) blockIdx.x * 2 * np + jj] = r;
		part[(size_t) blockIdx.x * 2 * np + np + jj] = c;
	}
}

/* w[t] = -sum_{s <= t} v[s], t < nt: the window rebuilt from its first
 * differences on their support (offsets sidx, running sums cum), so a
 * flat window need not be copied to the device */
__global__ void k_td_wfill(const double *__restrict__ sidx, const double *__restrict__ cum, int nnz, int nt, double *__restrict__ w)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if(t >= nt)
		return;
	int lo = 0, hi = nnz;
	while(lo < hi) {
		const int mid = (lo + hi) >> 1;
		if(sidx[mid] <= (double) t)
			lo = mid + 1;
		else
			hi = mid;
	}
	w[t] = lo ? -cum[lo - 1] : 0.0;
}

/* C from its first row and column (edge) and the diagonal differences D
 * (in X on entry, C on return): C[k + 1][k' + 1] = C[k][k'] + D[k][k'].
 * One wave per diagonal k' - k: each lane sums a segment of D, a scan
 * over the lanes gives each segment's starting value, and the lanes then
 * write their segments. */
__global__ void __launch_bounds__(256) k_td_diag(const double *__restrict__ edge, int np, double *__restrict__ X)
{
	const int lane = threadIdx.x & 63;
	const int d = (int) (blockIdx.x * 4 + (threadIdx.x >> 6)) - (np - 1);
	if(d > np - 1)
		return;
	const int k0 = d >= 0 ? 0 : -d, kp0 = d >= 0 ? d : 0;
	const int len = np - (d >= 0 ? d : -d);
	const int seg = (len + 63) / 64, b0 = lane * seg, b1 = b0 + seg < len ? b0 + seg : len;
	const size_t step = (size_t) np + 1, base = (size_t) k0 * np + kp0;
	double sum = 0.0;
	for(int e = b0; e < b1; e++)
		sum += X[base + e * step];
	/* exclusive scan of the segment sums over the wave */
	double incl = sum;
	for(int o = 1; o < 64; o <<= 1) {
		const double u = __shfl_up(incl, o);
		if(lane >= o)
			incl += u;
	}
	double c = (d >= 0 ? edge[d] : edge[np - d]) + (incl - sum);
	for(int e = b0; e < b1; e++) {
		const double x = X[base + e * step];
		X[base + e * step] = c;
		c += x;
	}
}

/* Upart[kb][q] = sum_{k in block kb} Ta[k][q] sum_k' C[k][k'] Tb[k'][q]:
 * Y = C Tb and its bilinear form with Ta in one pass, for 16 k (block kb)
 * by 64 q per workgroup; each wave takes 4 k, each lane one q.  The 16 C
 * rows go through LDS in chunks of TD_YU_KC (every lane of a wave reads
 * the same element: a broadcast), the Tb row reads are coalesced over q. */
#define TD_YU_KC 256
/* the product rule's node count that k_td_pw's LDS holds */
#define TD_NQ_MAX 7680
__global__ void __launch_bounds__(256) k_td_yu(const double *__restrict__ C, const double *__restrict__ Ta, const double *__restrict__ Tb, int np, int nq, double *__restrict__ Upart)
{
	__shared__ double Cs[16][TD_YU_KC];
	__shared__ double red[4][64];
	const int qi = threadIdx.x & 63, kw = threadIdx.x >> 6;
	const int q = blockIdx.x * 64 + qi, kb = blockIdx.y, k0 = kb * 16 + kw * 4;
	double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
	for(int c0 = 0; c0 < np; c0 += TD_YU_KC) {
		const int nc = np - c0 < TD_YU_KC ? np - c0 : TD_YU_KC;
		__syncthreads();
		for(int e = threadIdx.x; e < 16 * TD_YU_KC; e += 256) {
			const int r = e / TD_YU_KC, c = e % TD_YU_KC, k = kb * 16 + r;
			Cs[r][c] = (k < np && c < nc) ? C[(size_t) k * np + c0 + c] : 0.0;
		}
		__syncthreads();
		if(q < nq) {
			const double *tb = Tb + (size_t) c0 * nq + q;
#pragma unroll 4
			for(int c = 0; c < nc; c++) {
				const double b = tb[(size_t) c * nq];
				y0 = fma(Cs[kw * 4 + 0][c], b, y0);
				y1 = fma(Cs[kw * 4 + 1][c], b, y1);
				y2 = fma(Cs[kw * 4 + 2][c], b, y2);
				y3 = fma(Cs[kw * 4 + 3][c], b, y3);
			}
		}
	}
	double u = 0.0;
	if(q < nq) {
		const double y[4] = { y0, y1, y2, y3 };
		for(int j = 0; j < 4; j++)
			if(k0 + j < np)
				u = fma(Ta[(size_t) (k0 + j) * nq + q], y[j], u);
	}
	red[kw][qi] = u;
	__syncthreads();
	if(kw == 0 && q < nq)
		Upart[(size_t) kb * nq + q] = (red[0][qi] + red[1][qi]) + (red[2][qi] + red[3][qi]);
}

/* power[d] = sum_q U_q WP[q][d] with U_q = sum_kb Upart[kb][q] (stored as
 * the complex power[d] + 0 i): 16 d per workgroup, the q sum split 16
 * ways; every workgroup first sums U into LDS (nq doubles, dynamic) */
__global__ void __launch_bounds__(256) k_td_pw(const double *__restrict__ Upart, int nkb, const double *__restrict__ WP, int nq, int P, double *__restrict__ power)
{
	extern __shared__ double Us[];
	__shared__ double red[16][16];
	for(int q = threadIdx.x; q < nq; q += 256) {
		double u = 0.0;
		for(int kb = 0; kb < nkb; kb++)
			u += Upart[(size_t) kb * nq + q];
		Us[q] = u;
	}
	__syncthreads();
	const int di = threadIdx.x & 15, qs = threadIdx.x >> 4, d = blockIdx.x * 16 + di;
	double s = 0.0;
	if(d <= P)
		for(int q = qs; q < nq; q += 16)
			s = fma(Us[q], WP[(size_t) q * (P + 1) + d], s);
	red[qs][di] = s;
	__syncthreads();
	if(qs == 0 && d <= P) {
		double t = 0.0;
		for(int g = 0; g < 16; g++)
			t += red[g][di];
		power[2 * d] = t;
		power[2 * d + 1] = 0.0;
	}
}

}	/* namespace */

/* ------------------------------------------------------------- the plan */

static void td_dev_free(struct td_plan *p)
{
	double *b[] = { p->d_Ta, p->d_Tb, p->d_WP, p->d_a, p->d_b, p->d_w, p->d_part, p->d_C, p->d_U, p->d_powc, p->d_rows, p->d_out, p->d_sv, p->d_edge, p->d_nts, p->d_nws };
	for(unsigned i = 0; i < sizeof(b) / sizeof(b[0]); i++)
		if(b[i])
			(void) hipFree(b[i]);
}

/* correlator_plan_td_new (correlator.c:325-385) */
extern "C" struct correlator_plan_td *correlator_plan_td_new(const struct correlator_baseline *baseline, double delta_t)
{
	if(!baseline || delta_t <= 0. || sphr_gpu_ensure())
		return NULL;
	struct td_plan *p = (struct td_plan *) calloc(1, sizeof(*p));
	gsl_vector *dp = gsl_vector_alloc(3);
	double *R = sh_series_rot_matrix(baseline->theta, baseline->phi);
	double *Ta = NULL, *Tb = NULL, *WP = NULL, *rows = NULL;
	if(!p || !dp || !R)
		goto fail;
	{
		struct correlator_plan_td *t = &p->pub;
		const double D = vector_magnitude(baseline->d);
		const unsigned a_l = projection_matrix_l_max(D / 2, delta_t), b_l = projection_matrix_l_max(D / 2, delta_t);
		unsigned P = correlator_baseline_power_l_max(baseline, delta_t);
		if(P > a_l + b_l)
			P = a_l + b_l;
		t->sample_a = sh_series_new(a_l, 1);
		t->sample_b = sh_series_new(b_l, 1);
		t->product = sh_series_new(P, 1);
		t->power_1d = sh_series_new_zero(P, 1);
		if(!t->sample_a || !t->sample_b || !t->product || !t->power_1d)
			goto fail;
		/* d' = +D/2 z and -D/2 z */
		gsl_vector_set(dp, 0, 0);
		gsl_vector_set(dp, 1, 0);
		gsl_vector_set(dp, 2, +D / 2.);
		t->proj_a = projection_matrix_delay(projection_matrix_n_elements(vector_magnitude(dp), delta_t), a_l, dp, delta_t);
		gsl_vector_set(dp, 2, -D / 2.);
		t->proj_b = projection_matrix_delay(projection_matrix_n_elements(vector_magnitude(dp), delta_t), b_l, dp, delta_t);
		if(!t->proj_a || !t->proj_b)
			goto fail;
		t->product_plan = sh_series_product_plan_new(t->power_1d, t->sample_a, t->sample_b);
		t->rotation_plan = sh_series_rotation_plan_new(t->power_1d, R);
		if(!t->product_plan || !t->rotation_plan)
			goto fail;
		t->baseline = baseline;
		t->delta_t = delta_t;
		t->transient = correlator_transient(t->proj_a, t->proj_b);
		/* device tables on an exact product rule: degree a + b + P */
		/* the correlation kernel reads a[t + k] for t < n - 2 transient,
		 * k < n_p: in range when n_p = 2 transient + 1 (odd tap counts,
		 * projection.c's 2 (int) (60 + r / dt + 1) + 1) */
		if(t->proj_a->n != t->proj_b->n || t->proj_a->n != 2 * t->transient + 1) {
			sphr_set_error("correlator_plan_td_new: tap counts %d, %d with transient %d", t->proj_a->n, t->proj_b->n, t->transient);
			goto fail;
		}
		const int np = t->proj_a->n, nq = (int) ((a_l + b_l + P) / 2 + 1);
		/* k_td_pw holds the nq bilinear forms in LDS */
		if(nq > TD_NQ_MAX) {
			sphr_set_error("correlator_plan_td_new: %d product nodes exceed %d", nq, TD_NQ_MAX);
			goto fail;
		}
		p->np = np;
		p->nq = nq;
		p->P = (int) P;
		Ta = (double *) malloc((size_t) np * nq * sizeof(double));
		Tb = (double *) malloc((size_t) np * nq * sizeof(double));
		WP = (double *) malloc((size_t) nq * (P + 1) * sizeof(double));
		rows = (double *) malloc(2 * (size_t) (P + 1) * (P + 1) * sizeof(double));
		if(!Ta || !Tb || !WP || !rows || sphr_td_tables(t->proj_a, t->proj_b, nq, (int) P, Ta, Tb, WP))
			goto fail;
		sphr_polar_rows((int) P, sphr_rotation_blocks(t->rotation_plan), rows);
		if(sphr_rt_check(hipMalloc(&p->d_Ta, (size_t) np * nq * sizeof(double)), "hipMalloc")
		   || sphr_rt_check(hipMalloc(&p->d_Tb, (size_t) np * nq * sizeof(double)), "hipMalloc")
		   || sphr_rt_check(hipMalloc(&p->d_WP, (size_t) nq * (P + 1) * sizeof(double)), "hipMalloc")
		   || sphr_rt_check(hipMalloc(&p->d_C, (size_t) np * np * sizeof(double)), "hipMalloc")
		   || sphr_rt_check(hipMalloc(&p->d_U, (size_t) ((np + 15) / 16) * nq * sizeof(double)), "hipMalloc")
		   || sphr_rt_check(hipMalloc(&p->d_powc, 2 * (size_t) (P + 1) * sizeof(double)), "hipMalloc")
		   || sphr_rt_check(hipMalloc(&p->d_rows, 2 * (size_t) (P + 1) * (P + 1) * sizeof(double)), "hipMalloc")
		   || sphr_memcpy_htod(p->d_rows, rows, 2 * (size_t) (P + 1) * (P + 1) * sizeof(double))
		   || sphr_memcpy_htod(p->d_Ta, Ta, (size_t) np * nq * sizeof(double))
		   || sphr_memcpy_htod(p->d_Tb, Tb, (size_t) np * nq * sizeof(double))
		   || sphr_memcpy_htod(p->d_WP, WP, (size_t) nq * (P + 1) * sizeof(double)))
			goto fail;
	}
	free(Ta);
	free(Tb);
	free(WP);
	free(rows);
	free(R);
	gsl_vector_free(dp);
	return &p->pub;
fail:
	free(Ta);
	free(Tb);
	free(WP);
	free(rows);
	free(R);
	if(dp)
		gsl_vector_free(dp);
	if(p)
		correlator_plan_td_free(&p->pub);
	return NULL;
}

extern "C" void correlator_plan_td_free(struct correlator_plan_td *plan)
{
	if(!plan)
		return;
	struct td_plan *p = (struct td_plan *) plan;
	sh_series_rotation_plan_free(plan->rotation_plan);
	sh_series_product_plan_free(plan->product_plan);
	sh_series_free(plan->power_1d);
	sh_series_free(plan->product);
	sh_series_free(plan->sample_b);
	sh_series_free(plan->sample_a);
	sh_series_array_free(plan->proj_b);
	sh_series_array_free(plan->proj_a);
	td_dev_free(p);
	free(p);
}

static int grow(double **buf, size_t *cap, size_t need)
{
	if(need <= *cap)
		return 0;
	if(*buf)
		(void) hipFree(*buf);
	*buf = NULL;
	if(sphr_rt_check(hipMalloc(buf, need * sizeof(double)), "hipMalloc"))
		return -1;
	*cap = need;
	return 0;
}

/* the plan's power (complex, in d_powc) from device copies of the series
 * a, b (n samples) and window w (n - 2 transient; hw: the same window in
 * host memory, or NULL); launches only.  nnz is td_window_nnz of the
 * window: when it is not -1 the window is not read from w but rebuilt
 * there (w is then a device buffer of nt doubles to fill).
 *
 * With v[s] = w[s - 1] - w[s] for s = 0..nt (w = 0 outside [0, nt)),
 * C[k + 1][k' + 1] - C[k][k'] = sum_s v[s] a[s + k] b[s + k'] exactly.  For
 * windows that are constant over most of the series (the square and
 * Tukey windows of the reference's stitching) v has few nonzeros: the
 * correlation kernel then runs over that support only, and C follows from
 * its first row and column along the diagonals. */
/* the number of nonzero first differences of the host window hw (nt
 * samples) when the flat-window route applies to a plan of np taps, else
 * -1 */
static int td_window_nnz(const double *hw, int nt, int np)
{
	if(!hw || nt < 2 || np < 2 || np > TD_EDGE_NP_MAX)
		return -1;
	int nnz = 0;
	for(int s = 0; s <= nt; s++)
		if((s > 0 ? hw[s - 1] : 0.0) - (s < nt ? hw[s] : 0.0) != 0.0 && 2 * ++nnz > nt)
			return -1;
	return nnz;
}

static int td_power_dev(const double *a, const double *b, double *w, const double *hw, int nnz, int n, struct td_plan *p)
{
	const int np = p->np, nq = p->nq, nt = n - 2 * p->pub.transient;
	hipStream_t st = sphr_rt_stream();
	const int sparse = nnz >= 0;
	const double *cw = w, *sidx = NULL;
	int nk = nt;
	if(sparse) {
		/* values, offsets and running sums of the window's differences */
		double *hv = (double *) malloc(3 * (size_t) (nnz > 0 ? nnz : 1) * sizeof(double));
		if(!hv)
			return -1;
		double run = 0.0;
		for(int s = 0, e = 0; s <= nt; s++) {
			const double v = (s > 0 ? hw[s - 1] : 0.0) - (s < nt ? hw[s] : 0.0);
			if(v != 0.0) {
				run += v;
				hv[e] = v;
				hv[nnz + e] = (double) s;
				hv[2 * nnz + e++] = run;
			}
		}
		const int rc = grow(&p->d_sv, &p->cap_sv, 3 * (size_t) (nnz > 0 ? nnz : 1)) || (nnz > 0 && sphr_memcpy_htod(p->d_sv, hv, 3 * (size_t) nnz * sizeof(double)));
		free(hv);
		if(rc || grow(&p->d_edge, &p->cap_edge, 2 * (size_t) np))
			return -1;
		cw = p->d_sv;
		sidx = p->d_sv + nnz;
		nk = nnz;
		/* the window itself, for the first row and column */
		hipLaunchKernelGGL(k_td_wfill, dim3((nt + 255) / 256), dim3(256), 0, st, sidx, p->d_sv + 2 * nnz, nnz, nt, w);
	}
	/* K splits of about TD_TSPLIT terms, enough workgroups on a short
	 * series; on a long one at most eight workgroups per CU's worth of
	 * splits, each then longer, so the partial sums (nz np^2 doubles)
	 * stay bounded whatever the stream's length */
	const int tiles = (np + TD_TILE - 1) / TD_TILE;
	int nz = (nk + TD_TSPLIT - 1) / TD_TSPLIT;
	const int nz_max = 8 * (sphr_rt_cus() > 0 ? sphr_rt_cus() : 256) / (tiles * tiles);
	if(nz > nz_max)
		nz = nz_max;
	if(nz < 1)
		nz = 1;
	const int tpz = ((nk + nz - 1) / nz + TD_KC - 1) / TD_KC * TD_KC;
	nz = nk > 0 ? (nk + tpz - 1) / tpz : 1;
	/* d_part holds the K-split partials of C, then (sparse windows) the
	 * chunk sums of its first row and column: sized once for both, before
	 * any launch */
	const int nch = (nt + TD_EC - 1) / TD_EC;
	const size_t need = (size_t) nz * np * np > (size_t) nch * 2 * np ? (size_t) nz * np * np : (size_t) nch * 2 * np;
	if(grow(&p->d_part, &p->cap_part, need))
		return -1;
	if(nk > 0)
		hipLaunchKernelGGL(k_td_corr, dim3(tiles * tiles * nz), dim3(256), 0, st, a, b, cw, sidx, n, nk, np, tpz, tiles, p->d_part);
	else
		(void) hipMemsetAsync(p->d_part, 0, (size_t) np * np * sizeof(double), st);
	const long nc = (long) np * np;
	hipLaunchKernelGGL(k_td_reduce, dim3((nc + 255) / 256), dim3(256), 0, st, p->d_part, nk > 0 ? nz : 1, nc, p->d_C);
	if(sparse) {
		hipLaunchKernelGGL(k_td_edge, dim3(nch), dim3(256), 0, st, a, b, w, nt, n, np, p->d_part);
		hipLaunchKernelGGL(k_td_reduce, dim3((2 * np + 255) / 256), dim3(256), 0, st, p->d_part, nch, (long) 2 * np, p->d_edge);
		hipLaunchKernelGGL(k_td_diag, dim3((2 * np - 1 + 3) / 4), dim3(256), 0, st, p->d_edge, np, p->d_C);
	}
	const int nkb = (np + 15) / 16;
	hipLaunchKernelGGL(k_td_yu, dim3((nq + 63) / 64, nkb), dim3(256), 0, st, p->d_C, p->d_Ta, p->d_Tb, np, nq, p->d_U);
	hipLaunchKernelGGL(k_td_pw, dim3((p->P + 16) / 16), dim3(256), (size_t) nq * sizeof(double), st, p->d_U, nkb, p->d_WP, nq, p->P, p->d_powc);
	return sphr_rt_check(hipGetLastError(), "time-domain correlator") ? -1 : 0;
}

/* the same from host series: copied into the plan's buffers first */
static int td_power_1d(const double *ta, const double *tb, const double *window, int n, struct td_plan *p)
{
	const int nt = n - 2 * p->pub.transient;
	size_t cap_a = p->cap_n, cap_b = p->cap_n, cap_w = p->cap_n;
	if(grow(&p->d_a, &cap_a, (size_t) n) || grow(&p->d_b, &cap_b, (size_t) n) || grow(&p->d_w, &cap_w, (size_t) (nt > 0 ? nt : 1)))
		return -1;
	p->cap_n = cap_a < cap_b ? (cap_a < cap_w ? cap_a : cap_w) : (cap_b < cap_w ? cap_b : cap_w);
	const int nnz = td_window_nnz(window, nt, p->np);
	if(sphr_memcpy_htod(p->d_a, ta, (size_t) n * sizeof(double)) || sphr_memcpy_htod(p->d_b, tb, (size_t) n * sizeof(double))
	   || (nt > 0 && nnz < 0 && sphr_memcpy_htod(p->d_w, window, (size_t) nt * sizeof(double))))
		return -1;
	return td_power_dev(p->d_a, p->d_b, p->d_w, window, nnz, n, p);
}

/* power_1d of the plan set from the device power; buf holds 2 (P + 1)
 * doubles */
static int td_fetch_power(struct td_plan *p, double *buf)
{
	if(sphr_memcpy_dtoh(buf, p->d_powc, 2 * (size_t) (p->P + 1) * sizeof(double)))
		return -1;
	for(int d = 0; d <= p->P; d++)
		sh_series_set(p->pub.power_1d, d, 0, buf[2 * d]);
	return 0;
}

/* correlator_baseline_integrate_power_td (correlator.c:648-673) */
extern "C" struct sh_series *correlator_baseline_integrate_power_td(const double *time_series_a, const double *time_series_b, const double *window, int n, struct correlator_plan_td *plan)
{
	if(!plan || !time_series_a || !time_series_b || !window)
		return NULL;
	if(n - 2 * plan->transient < 0)
		return NULL;
	struct td_plan *p = (struct td_plan *) plan;
	const int P = p->P;
	const size_t nsky = (size_t) (P + 1) * (P + 1);
	double *buf = (double *) malloc(2 * (size_t) (P + 1) * sizeof(double));
	struct sh_series *power_2d = sh_series_new(P, 0);
	int rc = !buf || !power_2d;
	pthread_mutex_lock(&sphr_dropin_lock);
	if(!rc)
		rc = td_power_1d(time_series_a, time_series_b, window, n, p) || grow(&p->d_out, &p->cap_out, 2 * nsky);
	if(!rc) {
		const cplx<double> *pw[1] = { (const cplx<double> *) p->d_powc }, *bk[1] = { (const cplx<double> *) p->d_rows };
		const int Lb[1] = { P }, pol[1] = { 1 };
		rc = launch_rotsum(1, pw, Lb, pol, bk, P, 1.0, (cplx<double> *) p->d_out, sphr_rt_stream());
		if(rc)
			sphr_set_error("rotation launch failed");
	}
	if(!rc)
		rc = td_fetch_power(p, buf) || sphr_memcpy_dtoh(power_2d->coeff, p->d_out, 2 * nsky * sizeof(double));
	pthread_mutex_unlock(&sphr_dropin_lock);
	free(buf);
	if(rc) {
		sh_series_free(power_2d);
		return NULL;
	}
	return power_2d;
}

/* correlator_network_plan_td_new / _free (correlator.c:820-858) */
extern "C" struct correlator_network_plan_td *correlator_network_plan_td_new(struct correlator_network_baselines *baselines, double delta_t)
{
	struct correlator_network_plan_td *n = (struct correlator_network_plan_td *) malloc(sizeof(*n));
	struct correlator_plan_td **plans = baselines ? (struct correlator_plan_td **) calloc(baselines->n_baselines, sizeof(*plans)) : NULL;
	if(!n || !plans) {
		free(n);
		free(plans);
		return NULL;
	}
	for(int i = 0; i < baselines->n_baselines; i++) {
		plans[i] = correlator_plan_td_new(baselines->baselines[i], delta_t);
		if(!plans[i]) {
			while(--i >= 0)
				correlator_plan_td_free(plans[i]);
			free(n);
			free(plans);
			return NULL;
		}
	}
	n->baselines = baselines;
	n->plans = plans;
	return n;
}

extern "C" void correlator_network_plan_td_free(struct correlator_network_plan_td *plan)
{
	if(plan) {
		for(int i = 0; i < plan->baselines->n_baselines; i++)
			correlator_plan_td_free(plan->plans[i]);
		free(plan->plans);
	}
	free(plan);
}

/* correlator_network_integrate_power_td (correlator.c:967-984).  The
 * series go to the device once for all baselines, in the first plan's
 * buffers; each baseline's power stays on the device, and one k_rotsum
 * launch per SPHR_MAX_BASELINES baselines rotates and averages them.  As
 * in the reference (sh_series_add), a sky of lower order than a baseline's
 * power, or a polar sky, fails after the powers are computed. */
extern "C" struct sh_series *correlator_network_integrate_power_td(struct sh_series *sky, double **tseries, int tseries_length, double **windows, struct correlator_network_plan_td *plan)
{
	if(!sky || !tseries || !windows || !plan || !plan->baselines->n_baselines)
		return NULL;
	const int nb = plan->baselines->n_baselines;
	const int n_inst = instrument_array_len(plan->baselines->baselines[0]->instruments);
	int tmax = 0, Pmax = 0;
	for(int k = 0; k < nb; k++) {
		const struct td_plan *p = (const struct td_plan *) plan->plans[k];
		tmax = p->pub.transient > tmax ? p->pub.transient : tmax;
		Pmax = p->P > Pmax ? p->P : Pmax;
	}
	sh_series_zero(sky);
	if(tseries_length - 2 * tmax < 0)
		return NULL;
	const int L = (int) sky->l_max;
	const size_t nsky = (size_t) (L + 1) * (L + 1);
	const int nchunk = (nb + SPHR_MAX_BASELINES - 1) / SPHR_MAX_BASELINES;
	double *buf = (double *) malloc(2 * (size_t) (Pmax + 1) * sizeof(double));
	double *part = nchunk > 1 ? (double *) malloc(2 * nsky * sizeof(double)) : NULL;
	struct td_plan *p0 = (struct td_plan *) plan->plans[0];
	int rc = !buf || (nchunk > 1 && !part);
	pthread_mutex_lock(&sphr_dropin_lock);
	if(!rc)
		rc = grow(&p0->d_nts, &p0->cap_nts, (size_t) n_inst * tseries_length) || grow(&p0->d_nws, &p0->cap_nws, (size_t) nb * tseries_length);
	double *dts = p0->d_nts, *dws = p0->d_nws;
	for(int i = 0; i < n_inst && !rc; i++)
		rc = sphr_memcpy_htod(dts + (size_t) i * tseries_length, tseries[i], (size_t) tseries_length * sizeof(double));
	/* windows flat over most of the series are rebuilt on the device from
	 * their differences; the others are copied */
	int nnzs[SPHR_MAX_BASELINES];
	for(int k = 0; k < nb && !rc; k++) {
		const int nt = tseries_length - 2 * plan->plans[k]->transient;
		const int nnz = td_window_nnz(windows[k], nt, ((const struct td_plan *) plan->plans[k])->np);
		if(k < SPHR_MAX_BASELINES)
			nnzs[k] = nnz;
		rc = nt > 0 && (k >= SPHR_MAX_BASELINES || nnz < 0) && sphr_memcpy_htod(dws + (size_t) k * tseries_length, windows[k], (size_t) nt * sizeof(double));
	}
	int k = 0;
	for(int i = 1; i < n_inst && !rc; i++)
		for(int j = 0; j < i && !rc; j++, k++)
			rc = td_power_dev(dts + (size_t) i * tseries_length, dts + (size_t) j * tseries_length, dws + (size_t) k * tseries_length, windows[k],
					  k < SPHR_MAX_BASELINES ? nnzs[k] : -1, tseries_length, (struct td_plan *) plan->plans[k]);
	for(k = 0; k < nb && !rc; k++)
		rc = td_fetch_power((struct td_plan *) plan->plans[k], buf);
	if(!rc && (sky->polar || L < Pmax))
		rc = -1;
	if(!rc)
		rc = grow(&p0->d_out, &p0->cap_out, 2 * nsky * nchunk);
	for(int c = 0; c < nchunk && !rc; c++) {
		const cplx<double> *pw[SPHR_MAX_BASELINES], *bk[SPHR_MAX_BASELINES];
		int Lb[SPHR_MAX_BASELINES], pol[SPHR_MAX_BASELINES], m = 0;
		for(k = c * SPHR_MAX_BASELINES; k < nb && m < SPHR_MAX_BASELINES; k++, m++) {
			const struct td_plan *p = (const struct td_plan *) plan->plans[k];
			pw[m] = (const cplx<double> *) p->d_powc;
			bk[m] = (const cplx<double> *) p->d_rows;
			Lb[m] = p->P;
			pol[m] = 1;
		}
		rc = launch_rotsum(m, pw, Lb, pol, bk, L, 1.0 / nb, (cplx<double> *) p0->d_out + nsky * c, sphr_rt_stream());
		if(rc)
			sphr_set_error("rotation launch failed");
	}
	for(int c = 0; c < nchunk && !rc; c++) {
		double *dst = c ? part : (double *) sky->coeff;
		rc = sphr_memcpy_dtoh(dst, p0->d_out + 2 * nsky * c, 2 * nsky * sizeof(double));
		for(size_t i = 0; c && !rc && i < 2 * nsky; i++)
			((double *) sky->coeff)[i] += part[i];
	}
	pthread_mutex_unlock(&sphr_dropin_lock);
	free(buf);
	free(part);
	if(rc) {
		sh_series_zero(sky);
		return NULL;
	}
	return sky;
}
