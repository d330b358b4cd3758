/*
 * Launch wrappers of the HIP kernels (kernels.hip).  All launches are
 * asynchronous on the given stream; pointers are device pointers.
 */
#ifndef SPHR_GPU_KERNELS_H
#define SPHR_GPU_KERNELS_H

#include "common.h"

#define SPHR_MAX_BASELINES 16
#define SPHR_MAX_FACTORS 16
/* rings per hemisphere of the ring-sum kernel's row scales (l_max < 128) */
#define SPHR_MAX_MESH_B 128

/* per-plan constants of the cross-power projection (real time series);
 * CT counts the projection columns, each baseline padded to whole
 * 16-column tiles starting at col_start[b] */
struct xproj_geom {
	int D, nb, N, Nh, CT, K2;
	int ia[SPHR_MAX_BASELINES], ib[SPHR_MAX_BASELINES];
	int col_start[SPHR_MAX_BASELINES], Lb[SPHR_MAX_BASELINES];
};

/* phi-synthesis constants */
struct synth_geom {
	int L, B, n;		/* n = 2B points along phi */
	int nfac, fac[SPHR_MAX_FACTORS];
};

/* device tables of the mesh synthesis at order L */
struct synth_tables {
	const double *x;	/* B: cos(theta_j), j < B (descending) */
	const double *smm;	/* B x (L+1): Pbar_mm(x_j) */
	const double *alpha;	/* alpha_lm of the l recurrence at the (l, m >= 0)
				 * slot of the coefficient layout */
	const double *twp;	/* 2n: exp(+2 pi i t / n) */
	/* the ring sums' recurrence pairs (alpha_l, alpha_l / alpha_{l-1}) at
	 * the (l, m >= 0) slots, two zero pads: (nal + 2) complex values */
	const cplx<double> *rc64;
	const cplx<float> *rc32;
	const double *E64;	/* [2 (2L+1)][2 n]: the phi DFT, real-expanded */
	const float *E32;
	/* the fused sky kernel's tables at this order, NULL where it is not
	 * compiled: B fragments of its phi DFT ([k-step][column tile][lane])
	 * and the blob of small tables it stages into LDS */
	const double *bf64;
	const float *bf32;
	const unsigned char *sk64, *sk32;
};

/* device allocations behind the fused kernel's tables */
struct sky_fused_dev {
	double *bf64;
	float *bf32;
	unsigned char *sk64, *sk32;
};

/* batched real FFT of length N (power of two), two real series per
 * workgroup, half spectra out (N/2+1 bins per series).  in_f32 selects a
 * float input.  window may be NULL. */
template <typename T>
int launch_rfft_pairs(const void *in, int in_f32, const double *window, int nseries, int N, const cplx<T> *tw, cplx<T> *spec, hipStream_t st);

/* the lag-domain route of the projection (xcorr.hip): per baseline b the
 * folded lag table [nch][SPHR_XC_MR][nl] at moff[b] of the table array,
 * row (ch, j) holding lag u = ch SPHR_XC_MR + j for u < U1 (zero beyond) */
#define SPHR_XC_MAXB 3		/* baselines (and detectors) of the route */
#define SPHR_XC_MR 40		/* lags per chunk: the table rows a thread holds */
struct xcorr_geom {
	int D, nb, CT;
	int ia[SPHR_XC_MAXB], ib[SPHR_XC_MAXB];
	int nl[SPHR_XC_MAXB];		/* coefficients: l_b + 1 */
	int nch[SPHR_XC_MAXB];		/* lag chunks per coefficient */
	int rows[SPHR_XC_MAXB];		/* lags per chunk: SPHR_XC_MR */
	int U1[SPHR_XC_MAXB];		/* folded lags: largest |u| + 1 */
	int col_start[SPHR_XC_MAXB], width[SPHR_XC_MAXB];
	int moff[SPHR_XC_MAXB];
};

/* FFT + cross-correlation + lag-domain projection of every segment pair
 * into partial[S][CT] (one split: the power_1d of each baseline in its
 * columns, real parts at even columns, zero elsewhere); N is 1024 or 4096;
 * tw the N twiddles exp(-2 pi i k / N); twW the wave-level form's tables
 * (SPHR_XC_TWW complex values); form 1 to 4 (xcorr.hip); -1 for a form or
 * shape that is not compiled */
template <typename T>
int launch_xcorr(const void *in, int in_f32, const double *window, int S, int N, const struct xcorr_geom *g, const cplx<T> *tw, const cplx<T> *twW, const T *M, T *partial, int form, hipStream_t st);
/* the forms compiled for this shape (precision by T): bit f - 1 for form f */
template <typename T>
int xcorr_forms(int N, int D, int nb);
/* the wave-level form's twiddle tables, 2 x 16 x 64 complex values:
 * [k][n1] = exp(-2 pi i n1 k / 1024), then [c][lane] = exp(-2 pi i 16 (lane % 4) c / 1024) */
#define SPHR_XC_TWW (2 * 16 * 64)

/* projection GEMM: partial[split][S][CT]; tile_bl[t] is the baseline of
 * column tile t; k_per_split a multiple of the K chunk (32); tpw the
 * 16-column tiles per wave (1, 2 or 4) */
template <typename T>
int launch_xproj(const cplx<T> *spec, int S, const struct xproj_geom *g, const T *W, const int *tile_bl, int k_per_split, int nsplit, T *partial, hipStream_t st, int tpw, int ms);

/* sum partials, rotate, average baselines, rotate_z: alm[S][(L+1)^2] */
template <typename T, typename TO>
int launch_alm(const T *partial, int nsplit, int S, const struct xproj_geom *g, int L, const cplx<T> *R, const int2 *lm, const double *gmst, cplx<TO> *alm, hipStream_t st);

/* synthesis of S series of order L (polar flag) onto 2B x 2B meshes;
 * series s read from alm + aoff[s] and written to map + moff[s] (element
 * offsets, device arrays) when given, else packed */
template <typename TI, typename T, typename TO>
int launch_synth(const cplx<TI> *alm, int S, int polar, const struct synth_geom *g, const struct synth_tables *tb, cplx<TO> *map, hipStream_t st, const long *aoff = NULL, const long *moff = NULL);

/* one order class of a ragged synthesis over several orders: full-sphere
 * series (polar 0) of order g.L at alm + aoff[i], map + moff[i], i < count */
struct synth_cls {
	struct synth_geom g;
	struct synth_tables tb;
	const long *aoff, *moff;
	long b0;		/* the class's first workgroup */
	int rp, gx, a_in_lds, al_in_lds;
};

/* fills rp, gx and the LDS flags of d from d->g; returns the class's LDS
 * bytes per workgroup */
template <typename T>
size_t synth_cls_shape(struct synth_cls *d);

/* all classes in one launch of nblocks workgroups (sum of gx * count),
 * each given lds bytes (the largest class's); d_cls in device memory,
 * sorted by b0 */
template <typename T>
int launch_synth_classes(const cplx<T> *alm, const struct synth_cls *d_cls, int ncls, long nblocks, size_t lds, cplx<T> *map, hipStream_t st);

/* fused: partials -> a_lm -> mesh (map may be NULL), one workgroup per
 * segment; falls back to launch_alm + launch_synth when the coefficients
 * exceed the LDS budget */
template <typename T>
int launch_sky(const T *partial, int nsplit, int S, const struct xproj_geom *xg, const cplx<T> *R, const int2 *lm, const double *gmst, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *alm, cplx<T> *map, hipStream_t st);

/* the same stage as two kernels: a_lm and the Legendre stage into F
 * ([S][2(L+1)][2L+1] complex scratch; F NULL: a_lm only; -1 when the
 * coefficients exceed the LDS budget), then the phi DFT of F as an MFMA
 * GEMM into the meshes */
template <typename T>
int launch_sky_lf(const T *partial, int nsplit, int S, const struct xproj_geom *xg, const cplx<T> *R, const int2 *lm, const double *gmst, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *alm, cplx<T> *F, hipStream_t st);
/* the whole sky stage in one kernel (partials -> a_lm -> Legendre ring
 * sums in LDS -> phi DFT on MFMA -> mesh) for the orders it is compiled
 * for; -1 at other orders */
int sky_fused_order(int L);	/* 1 when launch_sky_fused is compiled for order L */
/* device tables of the fused kernel at order L, built from the synthesis
 * tables tb (x, smm, alpha, twp set); all NULL at orders it is not
 * compiled for; 0, or -1 with the error set */
int sky_fused_tables(int L, const struct synth_tables *tb, struct sky_fused_dev *d);
/* occ persistent workgroups per CU, nsub segments per workgroup (0: the
 * precision's default); -1 for a form whose LDS does not fit */
template <typename T>
int launch_sky_fused(const T *partial, int nsplit, int S, const struct xproj_geom *xg, const cplx<T> *R, const int2 *lm, const double *gmst, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *alm, cplx<T> *map, hipStream_t st, int occ = 1, int nsub = 0);
template <typename T>
int launch_phi(const cplx<T> *F, int S, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *map, hipStream_t st);
/* the a_lm of the localization skies formed where they are read: series
 * s = beta E + e is row (e, beta) of the sky product S [E][2][ncoef]
 * (complex) plus prior [ncoef], rotated by exp(-i m gmst[e]) (gmst may be
 * NULL), and written to alm [2E][ncoef] */
template <typename T> struct sky_epi {
	const T *S;
	const cplx<T> *prior;
	const double *gmst;
	int E;
	cplx<T> *alm;
};
/* the Legendre ring sums F [S][2(L+1)][2L+1] of S series (the first half
 * of launch_synth_gemm); with irs [S][2(L+1)], each ring's inverse split
 * scale for launch_gemm_h3a; with epi, the series are formed from the sky
 * product (alm unused) */
template <typename T>
int launch_ring_sums(const cplx<T> *alm, int S, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *F, hipStream_t st, float *irs, const struct sky_epi<T> *epi);
/* GEMM-form synthesis of S packed series (order L, non-polar) through the
 * ring-sum scratch F; -1 when a series exceeds the LDS budget */
template <typename T>
int launch_synth_gemm(const cplx<T> *alm, int S, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *F, cplx<T> *map, hipStream_t st);

/* extremum of sign * Re f for S series (one wavefront each): seed from
 * seed_map (S meshes of order L, rings from map_x) when given, else from
 * the l_seed low-pass on the n_seed-ring mesh seed_x; then Newton */
template <typename TI>
int launch_extremize(const cplx<TI> *alm, int S, int L, int find_max, int l_seed, const cplx<TI> *seed_map, const double *map_x, const double *cm, const double *alpha, const double *seed_x, int n_seed, int max_iter, double *theta, double *phi, double *value, hipStream_t st);

/* the same refinement from given seeds (theta, phi) [S][2]; series s
 * refines alm[s / alm_div] */
template <typename TI>
int launch_extremize_seeded(const cplx<TI> *alm, int S, int alm_div, int L, int find_max, const double *seeds, const double *cm, const double *alpha, int max_iter, double *theta, double *phi, double *value, hipStream_t st);

/* C[M][N] = A[M][K] B[K][N] on MFMA, row-major with leading dimensions;
 * T is the arithmetic type, TA the storage type of A (widened on load) */
template <typename T, typename TA>
int launch_gemm(int M, int N, int K, const TA *A, long lda, const T *B, long ldb, T *C, long ldc, hipStream_t st);
/* the same product (A in T) through rocBLAS */
template <typename T>
int launch_gemm_blas(int M, int N, int K, const T *A, long lda, const T *B, long ldb, T *C, long ldc, hipStream_t st);

/* split-FP16 GEMM (gemm_h3.hip): C = diag(ira) (Ah Bh^T + Ah Bl^T + Al Bh^T)
 * diag(icb), operands K-contiguous fp16, K % 64 == 0 */
template <typename TO>
int launch_gemm_h3(int M, int N, int K, const _Float16 *Ah, const _Float16 *Al, long lda, const float *ira, const _Float16 *Bh, const _Float16 *Bl, long ldb, const float *icb, TO *C, long ldc, hipStream_t st);
/* the same product with A float32 [M][lda] (K columns, K and lda even),
 * split while staged, inverse row scales ira given (powers of two that
 * put each row's largest element at 2^13..2^14); B [N][ldb] split with K
 * zero-padded to a multiple of 64 */
template <typename TO>
int launch_gemm_h3a(int M, int N, int K, const float *A, long lda, const float *ira, const _Float16 *Bh, const _Float16 *Bl, long ldb, const float *icb, TO *C, long ldc, hipStream_t st);
/* the meshes' phi DFT with the half-turn split (gemm_h3.hip): A [M][lda]
 * float32 ring sums of order L (complex columns m = 0..L, -L..-1) with
 * inverse row scales ira; Be, Bo [NH = 2 (L + 1)][ldb] the split DFT
 * matrices of the even and the odd orders (each class as m >= 0
 * ascending, then m < 0 ascending) to the first L + 1 points, with column
 * scales; C [M][2 NH] (rows with no gap) gets Ev + Od in columns 0..NH-1
 * and Ev - Od in NH..2NH-1; -1 for L + 1 > 64 */
int launch_phi_h3(int M, int L, const float *A, long lda, const float *ira, const _Float16 *Beh, const _Float16 *Bel, const float *icbe, const _Float16 *Boh, const _Float16 *Bol, const float *icbo, long ldb, float *C, hipStream_t st);
/* rows of X [rows][K] -> fp16 pairs scaled per row (columns K..Kp-1 zero) */
template <typename TI>
int launch_split_rows(int rows, int K, int Kp, const TI *X, long ldx, _Float16 *Xh, _Float16 *Xl, long ldh, float *irs, hipStream_t st);
/* drop-in helpers (FP64) */
int launch_xspec(const cplx<double> *const *fs_dev, int n, int nb, const int *ia, const int *ib, cplx<double> *fsp, hipStream_t st);
int launch_dotc(const cplx<double> *coeff, int n, int stride, const cplx<double> *vec, const double *vec_real, double scale, cplx<double> *partial, int nsplit, cplx<double> *out, hipStream_t st);
/* the projections of every baseline of a network at once: out[b] =
 * scale sum_i coeff[b][i][.] vec[b][i]; partial[b] holds nsplit x
 * stride[b] complex */
struct dotc_net_args {
	const cplx<double> *coeff[SPHR_MAX_BASELINES], *vec[SPHR_MAX_BASELINES];
	cplx<double> *partial[SPHR_MAX_BASELINES], *out[SPHR_MAX_BASELINES];
	int stride[SPHR_MAX_BASELINES];
};
int launch_dotc_net(const struct dotc_net_args *a, int nb, int n, int nsplit, double scale, hipStream_t st);
int launch_rotsum(int nb, const cplx<double> *const *power, const int *Lb, const int *polar, const cplx<double> *const *blocks, int L, double scale, cplx<double> *sky, hipStream_t st);
/* the rows D^l_{0 m'} of packed rotation blocks (l = 1 first, each
 * (2l+1)^2) in the coefficient layout of order L, as k_rotsum reads them
 * for a polar power; rows holds 2 (L + 1)^2 doubles */
void sphr_polar_rows(int L, const double *blocks, double *rows);
int launch_rotate_z(int L, const cplx<double> *in, double phi, cplx<double> *out, hipStream_t st);
int launch_fft_single(const double *in, int real_input, int N, const cplx<double> *tw, cplx<double> *out, hipStream_t st);
int launch_dft_direct(const double *in, int real_input, int N, const cplx<double> *tw, cplx<double> *out, hipStream_t st);

#endif
