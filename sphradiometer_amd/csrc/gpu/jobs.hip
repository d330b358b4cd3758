/*
 * Batches of independent sky-map jobs, each with its own detector network,
 * segment length N and sky order L (the uneven workload): plan generation
 * on the GPU and a lag-domain data path, launched as one captured HIP
 * graph.
 *
 * The frequency-domain plan of a baseline (correlator.c:424-519, then
 * set_l to Lt = min(P, L)),
 *   dp[(N-f) mod N]_d = (2 pi / N) sum_q w_q conj(A_f(x_q)) B_f(x_q) Pbar_d0(x_q),
 *   A_f(x) = sum_i e^{-2 pi i f i / N} T_i(x),  T_i(x) = sum_{l<=a} C_il Pbar_l0(x),
 * with C the tap projections (projection.c:99-131), is used by
 * power_1d = (1/N) sum_f dp[f] X[f], X = F_b conj(F_a).  Summing over f
 * first turns this into lags of the circular cross-correlation
 *   c[u] = sum_f X[f] e^{2 pi i f u / N} = N sum_s x_a[s] x_b[(s+u) mod N],
 *   power_1d[d] = sum_{|u| < n_p} c[u mod N] M[u][d],
 *   M[u][d] = (2 pi / N^2) sum_q w_q Pbar_d0(x_q) sum_i T_i(x_q) T'_{i+u}(x_q),
 * where T' are the taps of the other phase centre, T'_i(x) = T_i(-x).  The
 * identity is exact (a reordering of finite sums); the plan per baseline
 * is (2 n_p - 1) x (Lt + 1) real numbers instead of N x (Lt + 1) complex
 * ones, and no FFT is needed.  q runs over an n_q-point Gauss-Legendre
 * rule with 2 n_q - 1 >= 2a + Lt, exact for the polynomial integrand, as
 * the reference's microcode and pixel-domain products are.
 *
 * Plan generation (float64, per baseline): Legendre tables on the tap and
 * product nodes; C (sinc samples x tables); T (C x tables); the lag
 * correlation K of T with its mirror; M (K x weighted tables); the polar
 * rotation rows D^l_{0m}(baseline) / nb.  Data path: lag correlations
 * (LDS-tiled), M and rotation per job into a_lm, synthesis per l_max
 * class with gathered offsets.
 */
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sphradiometer_batch.h>
#include "../sphr_internal.h"
#include "kernels.h"
#include "runtime.h"
#include "synth_tables.h"

struct jb_bl {
	int job, ia, ib;
	int a, np, nt, nq, Lt, lags, L;
	double r_dt, theta, phi, inv_nb, scale;	/* scale = 1 / N^2 */
	long off_nt, off_nq;		/* node arrays (x then w) in the chunk's node buffer */
	long off_Pt, off_Pq, off_C, off_T, off_K, off_Y, off_F;	/* chunk workspace (doubles) */
	long off_M, off_R, off_c, off_pw;	/* persistent: M (T), R (cplx T), c (T), power slot */
};

struct jb_job {
	int D, N, L, nb, first_bl;
	long off_in, off_alm, off_map;
};

struct jb_class {
	int L, count;
	long *d_aoff, *d_moff;
	struct synth_dev syn;
};

struct sphr_jobs {
	int precision, n_jobs, n_bl, n_classes;
	double delta_t;
	struct jb_job *jobs;
	struct jb_bl *bl;
	struct jb_job *d_jobs;
	struct jb_bl *d_bl;
	void *M, *R, *c;
	long n_in, n_alm, n_map, n_pw_max;
	struct jb_class *cls;
	struct synth_cls *d_scls;	/* the classes as one synthesis launch */
	long scls_blocks;
	size_t scls_lds;
	int lag_sc, lag_nt;		/* tuned lag-kernel tile */
	hipGraph_t graph;
	hipGraphExec_t exec;
	const void *g_in;
	void *g_alm, *g_map;
	hipStream_t g_st;
};

/* ---------------------------------------------------------------- plan */

__device__ inline double sinc_pi(double x)
{
	const double px = M_PI * x;
	return px == 0.0 ? 1.0 : sin(px) / px;
}

/*
 * Per baseline, over the first half of each (mirror-symmetric) rule:
 * Pbar_l0 at the tap nodes (Pt[j][l], j < nt/2, l <= a) and at the
 * product nodes (Pq[q][l], q < (nq+1)/2, l <= max(a, Lt)); the delay
 * table of the tap nodes: y_j = r_dt cos(acos x_j) (the delay of node j
 * in samples, projection.c:99-107), g_j = sin(pi y_j) nt (2 pi w_j / nt)
 * / pi and the weight 2 pi w_j / nt; the product-rule weights.
 * Y = [y | g | weight] (nt/2 each) then (nq+1)/2 product weights.
 */
#define PG_TL 8		/* orders per staged block of the table rows */
__global__ void __launch_bounds__(256) k_pg_tables(const struct jb_bl *__restrict__ bl, int b0, const double *__restrict__ nodes, const double *__restrict__ lc, double *__restrict__ ws)
{
	/* each wave's 64 rows, PG_TL orders at a time: a lane runs its node's
	 * normalised Legendre recurrence
	 *   Pbar_l0 = a_l x Pbar_{l-1,0} - (a_l / a_{l-1}) Pbar_{l-2,0},
	 *   a_l = sqrt((4 l^2 - 1) / l^2), lc[2l] = a_l, lc[2l+1] = a_l / a_{l-1},
	 * and the block goes out through LDS as runs of PG_TL consecutive
	 * orders per row (a row per lane written straight out put 64 rows'
	 * cache lines in every store) */
	__shared__ double tile[4][64][PG_TL + 1];
	const struct jb_bl &B = bl[b0 + blockIdx.x];
	const int lq = B.a > B.Lt ? B.a : B.Lt;
	const int nth = B.nt / 2, nqh = (B.nq + 1) / 2, nn = nth + nqh;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	for(int base = 0; base < nn; base += 256) {
		const int i = base + threadIdx.x;
		double x = 0.0;
		if(i < nth) {
			x = nodes[B.off_nt + i];
			const double y = B.r_dt * cos(acos(x));
			const double ww = nodes[B.off_nt + B.nt + i] * (2. * M_PI / B.nt);
			/* sin(pi y) = (-1)^n sin(pi (y - n)), n = rint(y): y - n is
			 * exact and sin of the reduced argument keeps its relative
			 * accuracy where y is close to an integer, which the taps
			 * divide by (kk + y) with kk = -n */
			const double n = rint(y), sr = sin(M_PI * (y - n));
			const double sy = ((long long) n & 1) ? -sr : sr;
			ws[B.off_Y + i] = y;
			ws[B.off_Y + nth + i] = sy / M_PI * (double) B.nt * ww;
			ws[B.off_Y + 2 * nth + i] = ww;
		} else if(i < nn) {
			const int q = i - nth;
			x = nodes[B.off_nq + q];
			ws[B.off_Y + 3 * nth + q] = nodes[B.off_nq + B.nq + q];
		}
		/* Pbar_l0(x) for l = 0..lq (every lane, the same bound) */
		double p0 = 0.28209479177387814347, p1 = lc[2] * x * p0;
		for(int l0 = 0; l0 <= lq; l0 += PG_TL) {
#pragma unroll
			for(int c = 0; c < PG_TL; c++) {
				const int l = l0 + c;
				double v = 0.0;
				if(l == 0)
					v = p0;
				else if(l == 1)
					v = p1;
				else if(l <= lq) {
					v = lc[2 * l] * x * p1 - lc[2 * l + 1] * p0;
					p0 = p1;
					p1 = v;
				}
				tile[wave][lane][c] = v;
			}
			__syncthreads();
#pragma unroll
			for(int k = 0; k < 8; k++) {
				const int r = (lane >> 3) + 8 * k, c = lane & 7, nd = base + 64 * wave + r, l = l0 + c;
				if(nd < nth) {
					if(l <= B.a)
						ws[B.off_Pt + (size_t) nd * (B.a + 1) + l] = tile[wave][r][c];
				} else if(nd < nn && l <= lq)
					ws[B.off_Pq + (size_t) (nd - nth) * (lq + 1) + l] = tile[wave][r][c];
			}
			__syncthreads();
		}
	}
}

/*
 * The pairs (F[i][j] + F[i][j'], F[i][j] - F[i][j']) of weighted delay
 * elements of tap kk = i - (np-1)/2 at tap nodes j and j' = nt-1-j
 * (y_j' = -y_j), [np][nt/2] pairs per baseline: the A operand of the taps
 * product, made once here rather than in each of its column tiles.
 * F = (nt w 2 pi / nt) sinc(-kk - y), sinc(-kk - y) = (-1)^kk sin(pi y)
 * / (pi (kk + y)), so with g = sin(pi y) nt (w 2 pi / nt) / pi,
 *   F + F' = (-1)^kk g (-2 y) / ((kk + y)(kk - y)),
 *   F - F' = (-1)^kk g (2 kk) / ((kk + y)(kk - y)): one division.
 * Grid (blocks of PG_FB pairs, baselines of the chunk); blocks past a
 * baseline's pairs return.
 */
#define PG_FB 1024
__global__ void __launch_bounds__(256) k_pg_fpairs(const struct jb_bl *__restrict__ bl, int b0, double *__restrict__ ws)
{
	const struct jb_bl &B = bl[b0 + blockIdx.y];
	const int nth = B.nt / 2;
	const long n = (long) B.np * nth, e0 = (long) blockIdx.x * PG_FB;
	if(e0 >= n)
		return;
	const double *Y = ws + B.off_Y;
	for(long e = e0 + threadIdx.x; e < n && e < e0 + PG_FB; e += blockDim.x) {
		const int i = (int) (e / nth), j = (int) (e - (long) i * nth);
		const int kk = -(B.np - 1) / 2 + i;
		const double y = Y[j], g = Y[nth + j];
		const double dp = (double) kk + y, dm = (double) kk - y;
		const double pr = dp * dm;
		double sum, dif;
		if(pr != 0.0) {
			const double r = ((kk & 1) ? -g : g) / pr;
			sum = r * (-2.0 * y);
			dif = r * (2.0 * (double) kk);
		} else {
			/* a delay on the sample grid: sinc(0) = 1 there */
			const double one = (double) B.nt * Y[2 * nth + j];
			const double f1 = dp == 0.0 ? one : ((kk & 1) ? -g : g) / dp;
			const double f2 = dm == 0.0 ? one : ((kk & 1) ? g : -g) / dm;
			sum = f1 + f2;
			dif = f1 - f2;
		}
		*reinterpret_cast<double2 *>(ws + B.off_F + 2 * e) = make_double2(sum, dif);
	}
}

/* K[u][q] = sum_i T[i][q] T[i+u][nq-1-q], u = uu - (np-1), q < (nq+1)/2,
 * one thread per entry: for taps too many for the tiled kernel's LDS */
__global__ void k_pg_lag(const struct jb_bl *__restrict__ bl, int b0, double *__restrict__ ws)
{
	const struct jb_bl &B = bl[b0 + blockIdx.y];
	const int nqh = (B.nq + 1) / 2;
	const long e = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if(e >= (long) B.lags * nqh)
		return;
	const int uu = (int) (e / nqh), q = (int) (e - (long) uu * nqh);
	const int u = uu - (B.np - 1);
	const double *T = ws + B.off_T;
	const int lo = u < 0 ? -u : 0, hi = u > 0 ? B.np - u : B.np;
	double acc = 0.0;
	for(int i = lo; i < hi; i++)
		acc += T[(size_t) i * B.nq + q] * T[(size_t) (i + u) * B.nq + (B.nq - 1 - q)];
	ws[B.off_K + e] = acc;
}

/*
 * The plan's three products as one grouped float64 GEMM on MFMA
 * (v_mfma_f64_16x16x4f64) over all baselines of a chunk: workgroup = a
 * 64 x 64 tile of one baseline's product (4 waves of 32 x 32), K in
 * chunks of 16 staged in LDS through registers (the next chunk's loads in
 * flight during this chunk's MFMAs).  The workgroup finds its baseline
 * in the chunk's tile prefix sums.
 *
 * Both quadrature rules are mirror-symmetric: node n-1-j is -x_j with the
 * same weight, and Pbar_l0(-x) = (-1)^l Pbar_l0(x).  Each product folds
 * over the node pairs, which halves its arithmetic:
 *   PG_TAPS:  C[i][l] = sum_{j < nt/2} (F[i][j] + (-1)^l F[i][j']) Pt[j][l],
 *             F the weighted sinc samples (generated in the loader); a
 *             tile's 64 columns are 32 even l (A = F + F') and 32 odd l
 *             (A = F - F'), l = 64 t + 2c (+1);
 *   PG_SYNTH: E, O = sum over even, odd l of C[i][l] Pq[q][l] for
 *             q < nqh = (nq+1)/2; T[i][q] = E + O, T[i][nq-1-q] = E - O;
 *             each K chunk of 16 holds 8 even l then 8 odd l;
 *   PG_M:     M[u][d] = scale sum_{q < nqh} (K[u][q] + (-1)^d K[u][q'])
 *             (2 pi w_q) Pq[q][d] (the middle node once), with
 *             K[u][q'] = K[-u][q]; columns as in PG_TAPS.
 * MFMA 16x16x4 f64 operand layout: A[i][k] in lane i + 16 k, B[k][j] in
 * lane j + 16 k, accumulator element e of lane j + 16 q holds D[q + 4 e][j].
 */
enum { PG_TAPS, PG_SYNTH, PG_M };
#define PG_TILE 64
#define PG_KC 16

template <int MODE, typename TO>
__global__ void __launch_bounds__(256) k_pg_gemm(const struct jb_bl *__restrict__ bl, int b0, int nb, const int *__restrict__ tile_start, double *__restrict__ ws, TO *__restrict__ Mout)
{
	typedef double v4 __attribute__((ext_vector_type(4)));
	constexpr int NA = MODE == PG_SYNTH ? 1 : 2;	/* A tiles: pair sum and difference */
	__shared__ double sA[NA][PG_KC][PG_TILE + 4];	/* [k][row] */
	__shared__ double sB[PG_KC][PG_TILE + 4];	/* [k][col] */
	/* this workgroup's baseline: the last with tile_start <= blockIdx.x */
	int lo = 0, hi = nb - 1;
	while(lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if(tile_start[mid] <= (int) blockIdx.x)
			lo = mid;
		else
			hi = mid - 1;
	}
	const struct jb_bl &B = bl[b0 + lo];
	const int nqh = (B.nq + 1) / 2;
	const int M = MODE == PG_M ? B.lags : B.np;
	const int K = MODE == PG_TAPS ? B.nt / 2 : MODE == PG_SYNTH ? B.a + 1 : nqh;
	const int ncol = MODE == PG_TAPS ? B.a + 1 : MODE == PG_SYNTH ? nqh : B.Lt + 1;
	const int tiles_m = (M + PG_TILE - 1) / PG_TILE;
	const int local = blockIdx.x - tile_start[lo];
	const int m0 = (local % tiles_m) * PG_TILE, n0 = (local / tiles_m) * PG_TILE;
	const int lq = B.a > B.Lt ? B.a : B.Lt;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int li = lane & 15, lk = lane >> 4;
	const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
	const int sel = NA == 2 && wn ? 1 : 0;	/* odd columns: the difference tile */
	/* tile column c -> output column */
	auto col = [&](int c) -> int { return MODE == PG_SYNTH ? n0 + c : n0 + (c < 32 ? 2 * c : 2 * (c - 32) + 1); };
	/* K chunk entry -> l (synthesis: 8 even l, then 8 odd l) */
	auto kl = [&](int k0, int k) -> int { return MODE == PG_SYNTH ? k0 + (k < 8 ? 2 * k : 2 * (k - 8) + 1) : k0 + k; };
	const int nth = B.nt / 2;
	const double *Y = ws + B.off_Y;
	double ra[NA][4], rb[4];
	auto load = [&](int k0) {
#pragma unroll
		for(int r = 0; r < 4; r++) {
			const int e = tid + 256 * r;
			const int i = m0 + e / PG_KC, k = kl(k0, e % PG_KC);
			double v0 = 0.0, v1 = 0.0;
			if(i < M && k < K) {
				if(MODE == PG_TAPS) {
					const double2 f = *reinterpret_cast<const double2 *>(ws + B.off_F + 2 * ((size_t) i * nth + k));
					v0 = f.x;
					v1 = f.y;
				}
				else if(MODE == PG_SYNTH)
					v0 = ws[B.off_C + (size_t) i * (B.a + 1) + k];
				else {
					const double kp = ws[B.off_K + (size_t) i * nqh + k];
					if(2 * k == B.nq - 1)
						v0 = kp;
					else {
						const double km = ws[B.off_K + (size_t) (B.lags - 1 - i) * nqh + k];
						v0 = kp + km;
						v1 = kp - km;
					}
				}
			}
			ra[0][r] = v0;
			if(NA == 2)
				ra[NA - 1][r] = v1;
			const int kb = kl(k0, e / PG_TILE), j = col(e % PG_TILE);
			double vb = 0.0;
			if(kb < K && j < ncol) {
				if(MODE == PG_TAPS)
					vb = ws[B.off_Pt + (size_t) kb * (B.a + 1) + j];
				else if(MODE == PG_SYNTH)
					vb = ws[B.off_Pq + (size_t) j * (lq + 1) + kb];
				else
					vb = (2. * M_PI * Y[3 * nth + kb]) * ws[B.off_Pq + (size_t) kb * (lq + 1) + j];
			}
			rb[r] = vb;
		}
	};
	auto store = [&]() {
#pragma unroll
		for(int r = 0; r < 4; r++) {
			const int e = tid + 256 * r;
#pragma unroll
			for(int h = 0; h < NA; h++)
				sA[h][e % PG_KC][e / PG_KC] = ra[h][r];
			sB[e / PG_TILE][e % PG_TILE] = rb[r];
		}
	};
	v4 acc[2][2][2];	/* [parity of l (synthesis)][row block][column block] */
#pragma unroll
	for(int p = 0; p < 2; p++)
#pragma unroll
		for(int a = 0; a < 2; a++)
#pragma unroll
			for(int b = 0; b < 2; b++)
				acc[p][a][b] = (v4) 0;
	load(0);
	for(int k0 = 0; k0 < K; k0 += PG_KC) {
		store();
		__syncthreads();
		if(k0 + PG_KC < K)
			load(k0 + PG_KC);
#pragma unroll
		for(int ks = 0; ks < PG_KC; ks += 4) {
			const int p = MODE == PG_SYNTH && ks >= 8 ? 1 : 0;
			double af[2], bf[2];
#pragma unroll
			for(int a = 0; a < 2; a++)
				af[a] = sA[sel][ks + lk][wm + 16 * a + li];
#pragma unroll
			for(int b = 0; b < 2; b++)
				bf[b] = sB[ks + lk][wn + 16 * b + li];
#pragma unroll
			for(int a = 0; a < 2; a++)
#pragma unroll
				for(int b = 0; b < 2; b++)
					acc[p][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[p][a][b], 0, 0, 0);
		}
		__syncthreads();
	}
#pragma unroll
	for(int a = 0; a < 2; a++)
#pragma unroll
		for(int b = 0; b < 2; b++) {
			const int j = col(wn + 16 * b + li);
			if(j >= ncol)
				continue;
#pragma unroll
			for(int e = 0; e < 4; e++) {
				const int i = m0 + wm + 16 * a + lk + 4 * e;
				if(i >= M)
					continue;
				const double v = acc[0][a][b][e];
				if(MODE == PG_TAPS)
					ws[B.off_C + (size_t) i * (B.a + 1) + j] = v;
				else if(MODE == PG_SYNTH) {
					const double o = acc[1][a][b][e];
					double *Ti = ws + B.off_T + (size_t) i * B.nq;
					Ti[j] = v + o;
					if(2 * j != B.nq - 1)
						Ti[B.nq - 1 - j] = v - o;
				} else
					Mout[B.off_M + (size_t) i * (B.Lt + 1) + j] = (TO) (v * B.scale);
			}
		}
}

/*
 * K[u][q] = sum_i T[i][q] T[i+u][nq-1-q] for q < nqh = (nq+1)/2 (the
 * other half is K[u][nq-1-q] = K[-u][q]), stored [lags][nqh].
 * Tiled: workgroup = one baseline
 * x lq columns q (lq a power of two <= 16); the column blocks of T and of
 * its mirror in LDS, zero-padded so that no bound is tested in the inner
 * loop (Ta rows 0..np+6, Tb rows -7..np+13); each thread runs PG_LR = 8
 * consecutive lags of one column over a register window of the mirror
 * column: 16 LDS reads per 64 multiply-adds.
 */
#define PG_LR 8
#define PG_LPA (PG_LR - 1)		/* Ta rows past np */
#define PG_LPB (2 * PG_LR - 2)		/* Tb rows past np (7 before) */

__host__ __device__ inline size_t pg_lag_lds(int np, int lq)
{
	return (size_t) ((np + PG_LPA) + (PG_LR - 1 + np + PG_LPB)) * lq * sizeof(double);
}

__global__ void __launch_bounds__(256) k_pg_lag2(const struct jb_bl *__restrict__ bl, int b0, int nb, const int *__restrict__ qblk_start, int lq, double *__restrict__ ws)
{
	extern __shared__ double lg_sm[];
	int lo = 0, hi = nb - 1;
	while(lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if(qblk_start[mid] <= (int) blockIdx.x)
			lo = mid;
		else
			hi = mid - 1;
	}
	const struct jb_bl &B = bl[b0 + lo];
	const int q0 = (blockIdx.x - qblk_start[lo]) * lq;
	const int np = B.np, nq = B.nq;
	const int na = np + PG_LPA, nbr = PG_LR - 1 + np + PG_LPB;
	double *Ta = lg_sm;			/* [na][lq]: T[i][q0 + c] */
	double *Tb = lg_sm + na * lq;		/* [nbr][lq]: T[i - 7][nq - 1 - (q0 + c)] */
	const double *T = ws + B.off_T;
	for(int e = threadIdx.x; e < nbr * lq; e += 256) {
		const int r = e / lq, c = e - r * lq;
		const int q = q0 + c;
		if(r < na)
			Ta[e] = q < nq && r < np ? T[(size_t) r * nq + q] : 0.0;
		const int rb = r - (PG_LR - 1);
		Tb[e] = q < nq && rb >= 0 && rb < np ? T[(size_t) rb * nq + (nq - 1 - q)] : 0.0;
	}
	__syncthreads();
	const int c = threadIdx.x & (lq - 1), g = threadIdx.x / lq, ng = 256 / lq;
	const int q = q0 + c, nqh = (nq + 1) / 2;
	if(q >= nqh)
		return;
	double *Ko = ws + B.off_K;
	const int nblk = (B.lags + PG_LR - 1) / PG_LR;
	for(int blk = g; blk < nblk; blk += ng) {
		const int uu0 = PG_LR * blk, u0 = uu0 - (np - 1);
		/* rows i in [0, np) with 0 <= i + u < np for some lag u of the block */
		int ilo = -u0 - (PG_LR - 1);
		ilo = ilo < 0 ? 0 : ilo;
		int ihi = np - u0;
		ihi = ihi > np ? np : ihi;
		double acc[PG_LR];
#pragma unroll
		for(int r = 0; r < PG_LR; r++)
			acc[r] = 0.0;
		/* win[t] = T'[ib + u0 + t], T' the mirror column; rows past the
		 * range read the zero padding */
		const double *tb = Tb + (size_t) (u0 + PG_LR - 1) * lq + c;
		const double *ta = Ta + c;
		double win[2 * PG_LR - 1];
#pragma unroll
		for(int t = 0; t < PG_LR - 1; t++)
			win[PG_LR + t] = ilo < ihi ? tb[(ilo + t) * lq] : 0.0;
		for(int ib = ilo; ib < ihi; ib += PG_LR) {
#pragma unroll
			for(int t = 0; t < PG_LR - 1; t++)
				win[t] = win[PG_LR + t];
#pragma unroll
			for(int t = 0; t < PG_LR; t++)
				win[PG_LR - 1 + t] = tb[(ib + PG_LR - 1 + t) * lq];
			double tv[PG_LR];
#pragma unroll
			for(int s = 0; s < PG_LR; s++)
				tv[s] = ta[(ib + s) * lq];
#pragma unroll
			for(int s = 0; s < PG_LR; s++)
#pragma unroll
				for(int r = 0; r < PG_LR; r++)
					acc[r] = fma(tv[s], win[s + r], acc[r]);
		}
#pragma unroll
		for(int r = 0; r < PG_LR; r++)
			if(uu0 + r < B.lags)
				Ko[(size_t) (uu0 + r) * nqh + q] = acc[r];
	}
}

/* R[(l, m)] = D^l_{0m} / nb = sqrt(4 pi / (2l+1)) conj(Y_lm(theta, phi)) / nb
 * for the sky order L of the job (zero above Lt is handled by the user) */
template <typename T>
__global__ void k_pg_rot(const struct jb_bl *__restrict__ bl, int b0, cplx<T> *__restrict__ R)
{
	const struct jb_bl &B = bl[b0 + blockIdx.x];
	const int L = B.L;
	const double x = cos(B.theta), s = sin(B.theta);
	for(int mi = threadIdx.x; mi < 2 * L + 1; mi += blockDim.x) {
		const int m = mi - L, am = m < 0 ? -m : m;
		/* Pbar_mm by the sectoral recurrence */
		double pmm = 0.28209479177387814347;
		for(int k = 1; k <= am; k++)
			pmm *= -sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
		double sn, cs;
		sincos(-(double) m * B.phi, &sn, &cs);
		const double sg = (m < 0 && (am & 1)) ? -1.0 : 1.0;
		double p0 = pmm, p1 = 0.0, am1 = 1.0;
		for(int l = am; l <= L; l++) {
			if(l > am) {
				const double al = sqrt((4.0 * l * l - 1.0) / ((double) l * l - (double) am * am));
				const double pn = l == am + 1 ? al * x * p0 : al * (x * p0 - p1 / am1);
				p1 = p0;
				p0 = pn;
				am1 = al;
			}
			const double f = sqrt(4 * M_PI / (2 * l + 1)) * sg * p0 * B.inv_nb;
			R[B.off_R + lm_offset(L, l, m)] = mk((T) (f * cs), (T) (f * sn));
		}
	}
}

/* ---------------------------------------------------------- data path */

/*
 * c[u] = N sum_s x_a[s] x_b[(s+u) mod N] for the baseline's lags, four
 * consecutive lags per thread: thread tid of the workgroup sums lags
 * uu0 + 4 tid + r, r < 4, over a register window of the shifted series,
 * through LDS tiles of sc samples (sc a multiple of four).  Per four
 * samples it reads four values of x_a (the same address in every lane)
 * and eight of x_b, as 16-byte LDS reads, for sixteen multiply-adds: a
 * tenth of the LDS instructions of one lag per thread (two reads per
 * multiply-add), which bounded that form (float32 4.24 -> 1.84 ms).
 * The x_b chunk starts at sample s0 + u0 (mod N): one modulo per chunk,
 * the wrap of each sample by subtraction.
 */
template <typename T>
struct lag4 {
	T x, y, z, w;
};

__device__ __forceinline__ lag4<float> lag_ld4(const float *p)
{
	const float4 v = *reinterpret_cast<const float4 *>(p);
	return { v.x, v.y, v.z, v.w };
}

__device__ __forceinline__ lag4<double> lag_ld4(const double *p)
{
	const double2 u = *reinterpret_cast<const double2 *>(p), v = *reinterpret_cast<const double2 *>(p + 2);
	return { u.x, u.y, v.x, v.y };
}

#define LC_R 4
template <typename T>
__global__ void k_lagcorr4(const struct jb_job *__restrict__ jobs, const struct jb_bl *__restrict__ bl, const float *__restrict__ in, int sc, T *__restrict__ c)
{
	extern __shared__ __align__(16) unsigned char lag4_sm_raw[];
	T *sa = reinterpret_cast<T *>(lag4_sm_raw);	/* [sc] */
	T *sb = sa + sc;				/* [sc + LC_R nt + LC_R] */
	const struct jb_bl &B = bl[blockIdx.y];
	const struct jb_job &J = jobs[B.job];
	const int nt = blockDim.x;
	const int uu0 = blockIdx.x * nt * LC_R;
	if(uu0 >= B.lags)
		return;
	const int N = J.N;
	const long u0 = uu0 - (B.np - 1);
	const float *xa = in + J.off_in + (size_t) B.ia * N, *xb = in + J.off_in + (size_t) B.ib * N;
	const int wlen = sc + LC_R * nt + LC_R;
	const T *bw = sb + LC_R * threadIdx.x;
	T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
	for(int s0 = 0; s0 < N; s0 += sc) {
		const int len = N - s0 < sc ? N - s0 : sc;
		for(int t = threadIdx.x; t < sc; t += nt)
			sa[t] = t < len ? (T) xa[s0 + t] : (T) 0;
		int base = (int) (((long) s0 + u0) % N);
		if(base < 0)
			base += N;
		for(int t = threadIdx.x; t < wlen; t += nt) {
			int idx = base + t;
			while(idx >= N)
				idx -= N;
			sb[t] = (T) xb[idx];
		}
		__syncthreads();
		/* sa is zero past len, so the last block of four needs no test */
		for(int t = 0; t < len; t += 4) {
			const lag4<T> a = lag_ld4(sa + t), w = lag_ld4(bw + t), v = lag_ld4(bw + t + 4);
			a0 += a.x * w.x + a.y * w.y + a.z * w.z + a.w * w.w;
			a1 += a.x * w.y + a.y * w.z + a.z * w.w + a.w * v.x;
			a2 += a.x * w.z + a.y * w.w + a.z * v.x + a.w * v.y;
			a3 += a.x * w.w + a.y * v.x + a.z * v.y + a.w * v.z;
		}
		__syncthreads();
	}
	const int uu = uu0 + LC_R * threadIdx.x;
	const T acc[LC_R] = { a0, a1, a2, a3 };
#pragma unroll
	for(int r = 0; r < LC_R; r++)
		if(uu + r < B.lags)
			c[B.off_c + uu + r] = acc[r] * (T) N;
}

/* per job: power_1d_b = c_b M_b, a_lm = sum_b R_b power_1d_b[l] */
template <typename T>
__global__ void k_job_alm(const struct jb_job *__restrict__ jobs, const struct jb_bl *__restrict__ bl, const T *__restrict__ c, const T *__restrict__ M, const cplx<T> *__restrict__ R, cplx<T> *__restrict__ alm)
{
	extern __shared__ unsigned char alm_sm_raw[];
	T *pw = reinterpret_cast<T *>(alm_sm_raw);
	const struct jb_job &J = jobs[blockIdx.x];
	int tot = 0;
	for(int b = 0; b < J.nb; b++)
		tot += bl[J.first_bl + b].Lt + 1;
	for(int e = threadIdx.x; e < tot; e += blockDim.x) {
		int b = 0, d = e;
		while(d > bl[J.first_bl + b].Lt) {
			d -= bl[J.first_bl + b].Lt + 1;
			b++;
		}
		const struct jb_bl &B = bl[J.first_bl + b];
		const T *cb = c + B.off_c, *Mb = M + B.off_M + d;
		const size_t ld = B.Lt + 1;
		/* four partial sums: four loads of the table column in flight */
		T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
		int uu = 0;
		for(; uu + 4 <= B.lags; uu += 4) {
			a0 += cb[uu] * Mb[uu * ld];
			a1 += cb[uu + 1] * Mb[(uu + 1) * ld];
			a2 += cb[uu + 2] * Mb[(uu + 2) * ld];
			a3 += cb[uu + 3] * Mb[(uu + 3) * ld];
		}
		for(; uu < B.lags; uu++)
			a0 += cb[uu] * Mb[uu * ld];
		pw[e] = (a0 + a1) + (a2 + a3);
	}
	__syncthreads();
	const int L = J.L, ncoef = (L + 1) * (L + 1);
	for(int idx = threadIdx.x; idx < ncoef; idx += blockDim.x) {
		int l, m;
		lm_decode(L, idx, &l, &m);
		cplx<T> acc = mk((T) 0, (T) 0);
		int off = 0;
		for(int b = 0; b < J.nb; b++) {
			const struct jb_bl &B = bl[J.first_bl + b];
			if(l <= B.Lt)
				cfma(acc, R[B.off_R + idx], pw[off + l]);
			off += B.Lt + 1;
		}
		alm[J.off_alm + idx] = acc;
	}
}

/* ---------------------------------------------------------------- host */

static unsigned pl_power_l(double D, double dt)
{
	return (unsigned) floor(M_PI / (2 * asin(dt / (2 * D))) + 1);
}

extern "C" void sphr_jobs_free(struct sphr_jobs *J)
{
	if(!J)
		return;
	if(J->exec)
		hipGraphExecDestroy(J->exec);
	if(J->graph)
		hipGraphDestroy(J->graph);
	void *p[] = { J->d_jobs, J->d_bl, J->M, J->R, J->c };
	for(unsigned i = 0; i < sizeof(p) / sizeof(p[0]); i++)
		if(p[i])
			hipFree(p[i]);
	for(int k = 0; k < J->n_classes; k++) {
		if(J->cls[k].d_aoff)
			hipFree(J->cls[k].d_aoff);
		if(J->cls[k].d_moff)
			hipFree(J->cls[k].d_moff);
		synth_tables_free(&J->cls[k].syn);
	}
	if(J->d_scls)
		hipFree(J->d_scls);
	free(J->cls);
	free(J->jobs);
	free(J->bl);
	free(J);
}

/*
 * All plans, in chunks of baselines that share one float64 workspace of at
 * most 4 GiB (fewer, larger launches than at 1 GiB: 12 chunks instead of
 * 46 for 20,430 baselines, 72 -> 68 ms; 16 GiB gains nothing more).
 * Everything the kernels read besides the workspace is laid
 * out and uploaded once: the baseline table with each chunk's workspace
 * offsets, every distinct quadrature rule once, and the per-chunk tile
 * prefix sums; the chunks then run back to back on the stream (each
 * reuses the workspace in stream order) with one synchronisation at the
 * end.
 */
template <typename T>
static int generate_plans(struct sphr_jobs *J, hipStream_t st, int (*host_work)(struct sphr_jobs *))
{
	const size_t budget = (size_t) 4 << 30;
	const int nbl = J->n_bl;
	double *ws = NULL, *nodes = NULL, *hnodes = NULL;
	int *d_ts = NULL, *h_ts = NULL;	/* per chunk: tile prefix sums of its four plan launches */
	long *rule_t = NULL, *rule_q = NULL;
	struct pg_chunk {
		int b0, b1, lq, lag_tiled, np_max, lags_nq_max;
		long f_max;	/* most taps pairs of a baseline */
		long tot[4];
	} *ch = NULL;
	int nch = 0, rc = -1;
	size_t wsz_max = 0, nsz = 0, lc_off = 0;
	int lc_n = 0;
	int a_max = 0, nq_max = 0;
	for(int b = 0; b < nbl; b++) {
		a_max = J->bl[b].a > a_max ? J->bl[b].a : a_max;
		nq_max = J->bl[b].nq > nq_max ? J->bl[b].nq : nq_max;
	}
	rule_t = (long *) malloc((a_max + 1) * sizeof(long));
	rule_q = (long *) malloc((nq_max + 1) * sizeof(long));
	ch = (struct pg_chunk *) calloc(nbl, sizeof(*ch));
	h_ts = (int *) malloc((size_t) 4 * nbl * sizeof(int));
	if(!rule_t || !rule_q || !ch || !h_ts) {
		sphr_set_error("sphr_jobs_new: out of memory");
		goto done;
	}
	/* distinct rules: the tap mesh of order a lists cos theta descending
	 * (sh_series_trans.c:149-196); the product rule is any exact rule */
	for(int i = 0; i <= a_max; i++)
		rule_t[i] = -1;
	for(int i = 0; i <= nq_max; i++)
		rule_q[i] = -1;
	for(int b = 0; b < nbl; b++) {
		struct jb_bl &B = J->bl[b];
		if(rule_t[B.a] < 0) {
			rule_t[B.a] = nsz;
			nsz += 2 * (size_t) B.nt;
		}
		if(rule_q[B.nq] < 0) {
			rule_q[B.nq] = nsz;
			nsz += 2 * (size_t) B.nq;
		}
		B.off_nt = rule_t[B.a];
		B.off_nq = rule_q[B.nq];
	}
	{
		int lq_max = 0;
		for(int b = 0; b < nbl; b++) {
			const int lq = J->bl[b].a > J->bl[b].Lt ? J->bl[b].a : J->bl[b].Lt;
			lq_max = lq > lq_max ? lq : lq_max;
		}
		lc_off = nsz;
		lc_n = lq_max + 1;
		nsz += 2 * (size_t) lc_n;
	}
	hnodes = (double *) malloc(nsz * sizeof(double));
	if(!hnodes) {
		sphr_set_error("sphr_jobs_new: out of memory");
		goto done;
	}
	{
		/* compute the rules not cached yet, in parallel */
		int *ord = (int *) malloc((a_max + nq_max + 2) * sizeof(int)), n_t = 0, n_q = 0;
		if(!ord) {
			sphr_set_error("sphr_jobs_new: out of memory");
			goto done;
		}
		for(int i = 0; i <= a_max; i++)
			if(rule_t[i] >= 0)
				ord[n_t++] = i;
		for(int i = 0; i <= nq_max; i++)
			if(rule_q[i] >= 0)
				ord[n_t + n_q++] = i;
		sphr_rules_prefetch(ord, n_t, ord + n_t, n_q);
		free(ord);
	}
	for(int i = 0; i <= a_max; i++)
		if(rule_t[i] >= 0) {
			const int nt = 2 * (i + 1);
			const double *x, *w;
			sphr_mesh_rule((unsigned) i, &x, &w);
			memcpy(hnodes + rule_t[i], x, nt * sizeof(double));
			memcpy(hnodes + rule_t[i] + nt, w, nt * sizeof(double));
		}
	for(int i = 0; i <= nq_max; i++)
		if(rule_q[i] >= 0) {
			const double *x, *w;
			sphr_gl_rule_cached(i, &x, &w);
			memcpy(hnodes + rule_q[i], x, i * sizeof(double));
			memcpy(hnodes + rule_q[i] + i, w, i * sizeof(double));
		}
	/* Legendre recurrence coefficients a_l, a_l / a_{l-1} */
	hnodes[lc_off] = hnodes[lc_off + 1] = 0.0;
	for(int l = 1; l < lc_n; l++) {
		const double al = sqrt((4.0 * l * l - 1.0) / ((double) l * l));
		const double am1 = l == 1 ? 1.0 : sqrt((4.0 * (l - 1) * (l - 1) - 1.0) / ((double) (l - 1) * (l - 1)));
		hnodes[lc_off + 2 * l] = al;
		hnodes[lc_off + 2 * l + 1] = al / am1;
	}
	/* chunks, their workspace layout and tile counts */
	for(int b0 = 0; b0 < nbl; nch++) {
		struct pg_chunk &C = ch[nch];
		size_t wsz = 0;
		int b1 = b0;
		C.np_max = C.lags_nq_max = 0;
		C.f_max = 0;
		while(b1 < nbl) {
			struct jb_bl &B = J->bl[b1];
			const int lq = B.a > B.Lt ? B.a : B.Lt;
			const int nth = B.nt / 2, nqh = (B.nq + 1) / 2;
			const size_t need = (size_t) nth * (B.a + 1) + (size_t) nqh * (lq + 1) + (size_t) B.np * (B.a + 1) + (size_t) B.np * B.nq + (size_t) B.lags * nqh + 3 * (size_t) nth + nqh
					    + 2 * (size_t) B.np * nth + 1;
			if(b1 > b0 && (wsz + need) * sizeof(double) > budget)
				break;
			B.off_Pt = wsz;
			B.off_Pq = B.off_Pt + (long) nth * (B.a + 1);
			B.off_C = B.off_Pq + (long) nqh * (lq + 1);
			B.off_T = B.off_C + (long) B.np * (B.a + 1);
			B.off_K = B.off_T + (long) B.np * B.nq;
			B.off_Y = B.off_K + (long) B.lags * nqh;
			/* the taps' pairs, 16-byte aligned (the workspace and every
			 * offset before them count doubles) */
			B.off_F = B.off_Y + 3L * nth + nqh;
			B.off_F += B.off_F & 1;
			wsz += need;
			C.np_max = B.np > C.np_max ? B.np : C.np_max;
			C.f_max = (long) B.np * nth > C.f_max ? (long) B.np * nth : C.f_max;
			C.lags_nq_max = B.lags * B.nq > C.lags_nq_max ? B.lags * B.nq : C.lags_nq_max;	/* bounds lags * nqh */
			b1++;
		}
		C.b0 = b0;
		C.b1 = b1;
		wsz_max = wsz > wsz_max ? wsz : wsz_max;
		/* lag tile: the widest column block whose two LDS column blocks
		 * stay within 48 KiB (four workgroups per CU); past 160 KiB at
		 * one column the untiled kernel runs */
		C.lq = 16;
		while(C.lq > 1 && pg_lag_lds(C.np_max, C.lq) > 48 * 1024)
			C.lq >>= 1;
		C.lag_tiled = pg_lag_lds(C.np_max, C.lq) <= 160 * 1024;
		const int nb = b1 - b0;
		for(int k = 0; k < 4; k++)
			C.tot[k] = 0;
		for(int b = b0; b < b1; b++) {
			const struct jb_bl &B = J->bl[b];
			const int t = PG_TILE;
			const long cnt[4] = {
				(long) ((B.np + t - 1) / t) * ((B.a + 1 + t - 1) / t),
				(long) ((B.np + t - 1) / t) * (((B.nq + 1) / 2 + t - 1) / t),
				(long) ((B.lags + t - 1) / t) * ((B.Lt + 1 + t - 1) / t),
				(long) (((B.nq + 1) / 2 + C.lq - 1) / C.lq),
			};
			for(int k = 0; k < 4; k++) {
				h_ts[4 * (size_t) b0 + k * nb + (b - b0)] = (int) C.tot[k];
				C.tot[k] += cnt[k];
			}
		}
		for(int k = 0; k < 4; k++)
			if(C.tot[k] > INT_MAX) {
				sphr_set_error("sphr_jobs_new: plan chunk too large");
				goto done;
			}
		b0 = b1;
	}
	if(sphr_rt_check(hipMalloc(&ws, wsz_max * sizeof(double)), "hipMalloc") || sphr_rt_check(hipMalloc(&nodes, nsz * sizeof(double)), "hipMalloc")
	   || sphr_rt_check(hipMalloc(&d_ts, (size_t) 4 * nbl * sizeof(int)), "hipMalloc")
	   || sphr_rt_check(hipMemcpyAsync(nodes, hnodes, nsz * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpy")
	   || sphr_rt_check(hipMemcpyAsync(J->d_bl, J->bl, (size_t) nbl * sizeof(struct jb_bl), hipMemcpyHostToDevice, st), "hipMemcpy")
	   || sphr_rt_check(hipMemcpyAsync(d_ts, h_ts, (size_t) 4 * nbl * sizeof(int), hipMemcpyHostToDevice, st), "hipMemcpy"))
		goto done;
	for(int c = 0; c < nch; c++) {
		const struct pg_chunk &C = ch[c];
		const int b0 = C.b0, nb = C.b1 - C.b0;
		const int *ts = d_ts + 4 * (size_t) b0;
		hipLaunchKernelGGL(k_pg_tables, dim3(nb), dim3(256), 0, st, J->d_bl, b0, nodes, nodes + lc_off, ws);
		hipLaunchKernelGGL(k_pg_fpairs, dim3((unsigned) ((C.f_max + PG_FB - 1) / PG_FB), nb), dim3(256), 0, st, J->d_bl, b0, ws);
		hipLaunchKernelGGL((k_pg_gemm<PG_TAPS, double>), dim3(C.tot[0]), dim3(256), 0, st, J->d_bl, b0, nb, ts, ws, (double *) NULL);
		hipLaunchKernelGGL((k_pg_gemm<PG_SYNTH, double>), dim3(C.tot[1]), dim3(256), 0, st, J->d_bl, b0, nb, ts + nb, ws, (double *) NULL);
		if(C.lag_tiled)
			hipLaunchKernelGGL(k_pg_lag2, dim3(C.tot[3]), dim3(256), pg_lag_lds(C.np_max, C.lq), st, J->d_bl, b0, nb, ts + 3 * nb, C.lq, ws);
		else
			hipLaunchKernelGGL(k_pg_lag, dim3((C.lags_nq_max + 255) / 256, nb), dim3(256), 0, st, J->d_bl, b0, ws);
		hipLaunchKernelGGL((k_pg_gemm<PG_M, T>), dim3(C.tot[2]), dim3(256), 0, st, J->d_bl, b0, nb, ts + 2 * nb, ws, (T *) J->M);
		hipLaunchKernelGGL((k_pg_rot<T>), dim3(nb), dim3(128), 0, st, J->d_bl, b0, (cplx<T> *) J->R);
		if(sphr_rt_check(hipGetLastError(), "plan kernels"))
			goto done;
	}
	/* host work that does not need the plans runs while the GPU makes
	 * them */
	if(host_work && host_work(J))
		goto done;
	if(sphr_rt_check(hipStreamSynchronize(st), "plan"))
		goto done;
	rc = 0;
done:
	if(rc)
		hipStreamSynchronize(st);
	if(ws)
		hipFree(ws);
	if(nodes)
		hipFree(nodes);
	if(d_ts)
		hipFree(d_ts);
	free(hnodes);
	free(h_ts);
	free(rule_t);
	free(rule_q);
	free(ch);
	return rc;
}

/* the l_max classes of the synthesis: their job offsets and synthesis
 * tables (needs the jobs only, not their plans) */
static int build_classes(struct sphr_jobs *J)
{
	const int n_jobs = J->n_jobs;
	J->cls = (struct jb_class *) calloc(n_jobs, sizeof(struct jb_class));
	if(!J->cls)
		return -1;
	for(int j = 0; j < n_jobs; j++) {
		const int L = J->jobs[j].L;
		int k = 0;
		while(k < J->n_classes && J->cls[k].L != L)
			k++;
		if(k == J->n_classes) {
			J->cls[k].L = L;
			J->n_classes++;
		}
		J->cls[k].count++;
	}
	for(int k = 0; k < J->n_classes; k++) {
		struct jb_class &C = J->cls[k];
		long *ha = (long *) malloc(C.count * sizeof(long)), *hm = (long *) malloc(C.count * sizeof(long));
		int n = 0;
		for(int j = 0; j < n_jobs && ha && hm; j++)
			if(J->jobs[j].L == C.L) {
				ha[n] = J->jobs[j].off_alm;
				hm[n] = J->jobs[j].off_map;
				n++;
			}
		const int bad = !ha || !hm || sphr_rt_check(hipMalloc(&C.d_aoff, C.count * sizeof(long)), "hipMalloc") || sphr_rt_check(hipMalloc(&C.d_moff, C.count * sizeof(long)), "hipMalloc")
				|| sphr_rt_check(hipMemcpy(C.d_aoff, ha, C.count * sizeof(long), hipMemcpyHostToDevice), "hipMemcpy")
				|| sphr_rt_check(hipMemcpy(C.d_moff, hm, C.count * sizeof(long), hipMemcpyHostToDevice), "hipMemcpy")
				|| synth_tables_build(C.L, &C.syn, 0);
		free(ha);
		free(hm);
		if(bad)
			return -1;
	}
	/* one synthesis launch over all classes, largest order first so the
	 * longest workgroups start first */
	struct synth_cls *h = (struct synth_cls *) calloc(J->n_classes, sizeof(struct synth_cls));
	int *ord = (int *) malloc(J->n_classes * sizeof(int));
	if(!h || !ord) {
		free(h);
		free(ord);
		return -1;
	}
	for(int k = 0; k < J->n_classes; k++) {
		int i = k;
		for(; i > 0 && J->cls[ord[i - 1]].L < J->cls[k].L; i--)
			ord[i] = ord[i - 1];
		ord[i] = k;
	}
	long b0 = 0;
	size_t lds = 0;
	for(int i = 0; i < J->n_classes; i++) {
		const struct jb_class &C = J->cls[ord[i]];
		struct synth_cls &d = h[i];
		d.g = C.syn.g;
		d.tb = C.syn.tb;
		d.aoff = C.d_aoff;
		d.moff = C.d_moff;
		d.b0 = b0;
		const size_t l = J->precision == SPHR_MIXED ? synth_cls_shape<float>(&d) : synth_cls_shape<double>(&d);
		lds = l > lds ? l : lds;
		b0 += (long) d.gx * C.count;
	}
	J->scls_blocks = b0;
	J->scls_lds = lds;
	const int bad = sphr_rt_check(hipMalloc(&J->d_scls, J->n_classes * sizeof(struct synth_cls)), "hipMalloc") || sphr_rt_check(hipMemcpy(J->d_scls, h, J->n_classes * sizeof(struct synth_cls), hipMemcpyHostToDevice), "hipMemcpy");
	free(h);
	free(ord);
	return bad ? -1 : 0;
}

extern "C" struct sphr_jobs *sphr_jobs_new(int n_jobs, const int *n_det, const double *positions, const int *N, const int *l_max, double delta_t, int precision)
{
	if(sphr_gpu_ensure())
		return NULL;
	if(n_jobs < 1 || !n_det || !positions || !N || !l_max || delta_t <= 0. || (precision != SPHR_FP64 && precision != SPHR_MIXED)) {
		sphr_set_error("sphr_jobs_new: bad arguments");
		return NULL;
	}
	struct sphr_jobs *J = (struct sphr_jobs *) calloc(1, sizeof(*J));
	if(!J)
		return NULL;
	J->precision = precision;
	J->n_jobs = n_jobs;
	J->delta_t = delta_t;
	J->lag_sc = 1024;
	J->lag_nt = 256;
	J->jobs = (struct jb_job *) calloc(n_jobs, sizeof(struct jb_job));
	int nbl = 0;
	for(int j = 0; j < n_jobs; j++) {
		if(n_det[j] < 2 || n_det[j] > 16 || N[j] < 2 || l_max[j] < 0) {
			sphr_set_error("sphr_jobs_new: job %d: %d detectors, N %d, l_max %d", j, n_det[j], N[j], l_max[j]);
			sphr_jobs_free(J);
			return NULL;
		}
		nbl += n_det[j] * (n_det[j] - 1) / 2;
	}
	J->n_bl = nbl;
	J->bl = (struct jb_bl *) calloc(nbl, sizeof(struct jb_bl));
	if(!J->jobs || !J->bl) {
		sphr_jobs_free(J);
		return NULL;
	}
	/* layout: inputs, outputs, baselines and their plan / lag slots */
	long in = 0, aoff = 0, moff = 0, moff_M = 0, off_R = 0, off_c = 0;
	int b = 0, pdet = 0;
	const size_t esz = precision == SPHR_MIXED ? 4 : 8;
	for(int j = 0; j < n_jobs; j++) {
		struct jb_job &Jb = J->jobs[j];
		Jb.D = n_det[j];
		Jb.N = N[j];
		Jb.L = l_max[j];
		Jb.nb = Jb.D * (Jb.D - 1) / 2;
		Jb.first_bl = b;
		Jb.off_in = in;
		Jb.off_alm = aoff;
		Jb.off_map = moff;
		in += (long) Jb.D * Jb.N;
		aoff += (long) (Jb.L + 1) * (Jb.L + 1);
		moff += 4L * (Jb.L + 1) * (Jb.L + 1);
		/* baselines in the reference order: i = 1..D-1, j < i, d = r_i - r_j
		 * (correlator.c:728-757) */
		for(int i = 1; i < Jb.D; i++)
			for(int k = 0; k < i; k++, b++) {
				struct jb_bl &B = J->bl[b];
				const double *ri = positions + 3 * (pdet + i), *rk = positions + 3 * (pdet + k);
				const double d[3] = { ri[0] - rk[0], ri[1] - rk[1], ri[2] - rk[2] };
				const double D = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
				if(D <= 0.) {
					sphr_set_error("sphr_jobs_new: job %d has coincident sites", j);
					sphr_jobs_free(J);
					return NULL;
				}
				B.job = j;
				B.ia = i;
				B.ib = k;
				B.a = (int) (M_PI * (D / 2) / delta_t + 1);
				B.np = 2 * (int) (60 + (D / 2) / delta_t + 1) + 1;
				if(B.np > Jb.N) {
					sphr_set_error("sphr_jobs_new: job %d: %d taps exceed N = %d", j, B.np, Jb.N);
					sphr_jobs_free(J);
					return NULL;
				}
				B.nt = 2 * (B.a + 1);
				unsigned P = pl_power_l(D, delta_t);
				if(P > 2u * B.a)
					P = 2 * B.a;
				B.Lt = (int) P < Jb.L ? (int) P : Jb.L;
				B.L = Jb.L;
				B.nq = (2 * B.a + B.Lt + 2) / 2 + 1;
				B.lags = 2 * B.np - 1;
				B.r_dt = (D / 2) / delta_t;
				/* vector_direction (instrument.c:90-100) */
				B.theta = atan2(sqrt(d[0] * d[0] + d[1] * d[1]), d[2]);
				B.phi = atan2(d[1], d[0]);
				if(B.phi < 0)
					B.phi += 2 * M_PI;
				B.inv_nb = 1.0 / Jb.nb;
				B.scale = 1.0 / ((double) Jb.N * Jb.N);
				B.off_M = moff_M;
				moff_M += (long) B.lags * (B.Lt + 1);
				B.off_R = off_R;
				off_R += (long) (Jb.L + 1) * (Jb.L + 1);
				B.off_c = off_c;
				off_c += B.lags;
			}
		pdet += Jb.D;
	}
	J->n_in = in;
	J->n_alm = aoff;
	J->n_map = moff;
	hipStream_t st = sphr_rt_stream();
	if(sphr_rt_check(hipMalloc(&J->d_jobs, n_jobs * sizeof(struct jb_job)), "hipMalloc") || sphr_rt_check(hipMalloc(&J->d_bl, nbl * sizeof(struct jb_bl)), "hipMalloc")
	   || sphr_rt_check(hipMalloc(&J->M, moff_M * esz), "hipMalloc") || sphr_rt_check(hipMalloc(&J->R, off_R * 2 * esz), "hipMalloc")
	   || sphr_rt_check(hipMalloc(&J->c, off_c * esz), "hipMalloc")
	   || sphr_rt_check(hipMemcpy(J->d_jobs, J->jobs, n_jobs * sizeof(struct jb_job), hipMemcpyHostToDevice), "hipMemcpy")) {
		sphr_jobs_free(J);
		return NULL;
	}
	if((precision == SPHR_MIXED ? generate_plans<float>(J, st, build_classes) : generate_plans<double>(J, st, build_classes))) {
		sphr_jobs_free(J);
		return NULL;
	}
	return J;
}

template <typename T>
static int launch_run(struct sphr_jobs *J, const float *in, void *alm, void *map, int sc, int nt, hipStream_t st)
{
	int lags_max = 0;
	for(int b = 0; b < J->n_bl; b++)
		lags_max = J->bl[b].lags > lags_max ? J->bl[b].lags : lags_max;
	hipLaunchKernelGGL((k_lagcorr4<T>), dim3((lags_max + LC_R * nt - 1) / (LC_R * nt), J->n_bl), dim3(nt), (size_t) (2 * sc + LC_R * nt + LC_R) * sizeof(T), st, J->d_jobs, J->d_bl, in, sc, (T *) J->c);
	size_t pw = 0;
	for(int j = 0; j < J->n_jobs; j++) {
		size_t t = 0;
		for(int b = 0; b < J->jobs[j].nb; b++)
			t += J->bl[J->jobs[j].first_bl + b].Lt + 1;
		pw = t > pw ? t : pw;
	}
	hipLaunchKernelGGL((k_job_alm<T>), dim3(J->n_jobs), dim3(256), pw * sizeof(T), st, J->d_jobs, J->d_bl, (const T *) J->c, (const T *) J->M, (const cplx<T> *) J->R, (cplx<T> *) alm);
	if(hipGetLastError() != hipSuccess)
		return -1;
	if(map && launch_synth_classes<T>((const cplx<T> *) alm, J->d_scls, J->n_classes, J->scls_blocks, J->scls_lds, (cplx<T> *) map, st))
		return -1;
	return 0;
}

static int run_direct(struct sphr_jobs *J, const float *in, void *alm, void *map, hipStream_t st)
{
	return J->precision == SPHR_MIXED ? launch_run<float>(J, in, alm, map, J->lag_sc, J->lag_nt, st) : launch_run<double>(J, in, alm, map, J->lag_sc, J->lag_nt, st);
}

extern "C" int sphr_jobs_run(struct sphr_jobs *J, const float *in, void *alm, void *map, void *stream)
{
	if(!J || !in || !alm) {
		sphr_set_error("sphr_jobs_run: bad arguments");
		return -1;
	}
	hipStream_t st = stream ? (hipStream_t) stream : sphr_rt_stream();
	/* replay the captured graph when the buffers and stream are the same */
	if(!J->exec || J->g_in != in || J->g_alm != alm || J->g_map != map || J->g_st != st) {
		if(J->exec) {
			hipGraphExecDestroy(J->exec);
			J->exec = NULL;
		}
		if(J->graph) {
			hipGraphDestroy(J->graph);
			J->graph = NULL;
		}
		if(sphr_rt_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture"))
			return -1;
		const int r = run_direct(J, in, alm, map, st);
		hipGraph_t g = NULL;
		const hipError_t e = hipStreamEndCapture(st, &g);
		if(r || sphr_rt_check(e, "end capture")) {
			if(g)
				hipGraphDestroy(g);
			return -1;
		}
		if(sphr_rt_check(hipGraphInstantiate(&J->exec, g, NULL, NULL, 0), "instantiate")) {
			hipGraphDestroy(g);
			return -1;
		}
		J->graph = g;
		J->g_in = in;
		J->g_alm = alm;
		J->g_map = map;
		J->g_st = st;
	}
	return sphr_rt_check(hipGraphLaunch(J->exec, st), "graph launch");
}

/* autotuner: time the lag-correlation tiles on the caller's data and keep
 * the fastest (the graph is re-captured on the next run).  On any failure
 * the tile in use before the call is restored and -1 returned. */
extern "C" int sphr_jobs_tune(struct sphr_jobs *J, const float *in, void *alm, double *best_ms)
{
	static const int cand[][2] = { { 512, 128 }, { 1024, 128 }, { 2048, 128 }, { 512, 256 }, { 1024, 256 }, { 2048, 256 } };
	if(!J || !in || !alm) {
		sphr_set_error("sphr_jobs_tune: bad arguments");
		return -1;
	}
	hipStream_t st = sphr_rt_stream();
	const int sc0 = J->lag_sc, nt0 = J->lag_nt;
	hipEvent_t e0 = NULL, e1 = NULL;
	float best = 1e30f;
	int bi = -1;
	if(sphr_rt_check(hipEventCreate(&e0), "event") || sphr_rt_check(hipEventCreate(&e1), "event"))
		goto done;
	for(unsigned k = 0; k < sizeof(cand) / sizeof(cand[0]); k++) {
		J->lag_sc = cand[k][0];
		J->lag_nt = cand[k][1];
		int r = run_direct(J, in, alm, NULL, st);
		r = r || sphr_rt_check(hipEventRecord(e0, st), "tune");
		for(int i = 0; i < 3 && !r; i++)
			r = run_direct(J, in, alm, NULL, st);
		r = r || sphr_rt_check(hipEventRecord(e1, st), "tune") || sphr_rt_check(hipEventSynchronize(e1), "tune");
		float ms = 0.f;
		r = r || sphr_rt_check(hipEventElapsedTime(&ms, e0, e1), "tune");
		if(r) {
			bi = -1;
			goto done;
		}
		if(ms < best) {
			best = ms;
			bi = (int) k;
		}
	}
done:
	if(e0)
		hipEventDestroy(e0);
	if(e1)
		hipEventDestroy(e1);
	if(bi < 0) {
		J->lag_sc = sc0;
		J->lag_nt = nt0;
		return -1;
	}
	J->lag_sc = cand[bi][0];
	J->lag_nt = cand[bi][1];
	if(J->exec) {
		hipGraphExecDestroy(J->exec);
		J->exec = NULL;
		hipGraphDestroy(J->graph);
		J->graph = NULL;
	}
	if(best_ms)
		*best_ms = best / 3.0;
	return bi;
}

/* choose a lag tile by hand: one of the tuner's candidates (samples per
 * chunk 512, 1024 or 2048; 128 or 256 threads).  The captured graph is
 * dropped so the next run uses it. */
extern "C" int sphr_jobs_set_lag_tile(struct sphr_jobs *J, int samples, int threads)
{
	if(!J || (samples != 512 && samples != 1024 && samples != 2048) || (threads != 128 && threads != 256)) {
		sphr_set_error("sphr_jobs_set_lag_tile: not a tuner tile");
		return -1;
	}
	J->lag_sc = samples;
	J->lag_nt = threads;
	if(J->exec) {
		hipGraphExecDestroy(J->exec);
		J->exec = NULL;
		hipGraphDestroy(J->graph);
		J->graph = NULL;
	}
	return 0;
}

/* [0] input floats, [1] a_lm complex elements, [2] map complex elements,
 * [3] baselines, [4] l_max classes, [5] lag tile (samples), [6] lag tile
 * (threads), [7] plan bytes */
extern "C" void sphr_jobs_info(const struct sphr_jobs *J, long info[8])
{
	long plan = 0;
	for(int b = 0; b < J->n_bl; b++)
		plan += (long) J->bl[b].lags * (J->bl[b].Lt + 1);
	const long esz = J->precision == SPHR_MIXED ? 4 : 8;
	long rot = 0;
	for(int b = 0; b < J->n_bl; b++)
		rot += (long) (J->bl[b].L + 1) * (J->bl[b].L + 1) * 2;
	info[0] = J->n_in;
	info[1] = J->n_alm;
	info[2] = J->n_map;
	info[3] = J->n_bl;
	info[4] = J->n_classes;
	info[5] = J->lag_sc;
	info[6] = J->lag_nt;
	info[7] = (plan + rot) * esz;
}

/* per-job offsets: in (floats), alm and map (complex elements) */
extern "C" void sphr_jobs_offsets(const struct sphr_jobs *J, long *in, long *alm, long *map)
{
	for(int j = 0; j < J->n_jobs; j++) {
		if(in)
			in[j] = J->jobs[j].off_in;
		if(alm)
			alm[j] = J->jobs[j].off_alm;
		if(map)
			map[j] = J->jobs[j].off_map;
	}
}
