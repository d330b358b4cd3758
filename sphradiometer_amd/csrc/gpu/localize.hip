/*
 * Batched localization skies on the GPU (sphradiometer_batch.h,
 * sphr_loc_*): for E events of D complex SNR series,
 *
 *   1. spectra  F[d][e] = series[d][e] . DFT    (MFMA GEMM, the SNR
 *      preprocessing folded into the [len][N] DFT matrix)
 *   2. cross spectra X[e][b N + f] = F_ib[e][f] conj(F_ia[e][f])
 *   3. skies S[e][beta][lm] = X[e] . W           (MFMA GEMM, K = nb N,
 *      both beta = +1 / -1 plans side by side in the columns)
 *   4. alm[beta][e][lm] = (S + s_l prior) e^{-i m gmst_e}
 *   5. optional meshes (synthesis) and peaks (extremizer).
 *
 * Complex operands of the GEMMs are real-expanded (see gemm.hip): the
 * right-hand matrices are stored expanded, the left-hand ones are the
 * interleaved complex arrays read as real.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sphradiometer_batch.h>
#include "../sphr_internal.h"
#include "kernels.h"
#include "runtime.h"
#include "synth_tables.h"

struct sphr_loc {
	int precision, D, nb, N, len, L, max_events;
	int ia[16], ib[16];
	void *Wx;	/* T [2 nb N][4 ncoef] */
	void *Dx;	/* T [2 len][2 N] */
	void *prior;	/* cplx<T> [ncoef] */
	void *F;	/* T [D max_events][2 N] */
	void *X;	/* T [max_events][2 nb N] */
	void *S;	/* T [max_events][4 ncoef] */
	void *Fr;	/* cplx<T> [2 max_events][2(L+1)][2L+1] ring sums */
	/* mixed precision: both products as split-FP16 GEMMs (gemm_h3.hip) */
	int Kw, Kd;		/* K of the sky and DFT products, padded to 64 */
	_Float16 *Wh, *Wl;	/* [4 ncoef][Kw]: W transposed, per-column scale */
	float *icw;		/* [4 ncoef] */
	_Float16 *Dh, *Dl;	/* [2 N][Kd]: DFT matrix transposed, per-column scale */
	float *icd;		/* [2 N] */
	_Float16 *Ah, *Al;	/* [D max_events][Kd]: the series, split per row */
	float *ira;		/* [D max_events] */
	_Float16 *Xh, *Xl;	/* [max_events][Kw]: cross spectra, split per row */
	float *irx;		/* [max_events] */
	int Kp;			/* K of the phi DFT, 2 (2L+1) padded to 64 */
	_Float16 *Ph, *Pl;	/* [2 n][Kp]: the phi DFT matrix transposed, per-column scale */
	float *icp;		/* [2 n] */
	float *irr;		/* [2 max_events n]: inverse split scales of the ring sums */
	/* the half-turn form of the phi DFT (L + 1 <= 64): the even and odd
	 * orders' DFT matrices to the first n/2 points, [n][Kph] each */
	int ht, Kph;
	_Float16 *Peh, *Pel, *Poh, *Pol;
	float *icpe, *icpo;
	struct synth_dev syn;
};

struct loc_pairs {
	int ia[16], ib[16];
};

template <typename T>
__global__ void k_loc_xspec(const cplx<T> *__restrict__ F, int E, int N, int nb, struct loc_pairs p, cplx<T> *__restrict__ X)
{
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	const long tot = (long) E * nb * N;
	if(i >= tot)
		return;
	const int e = (int) (i / ((long) nb * N));
	const int r = (int) (i - (long) e * nb * N);
	const int b = r / N, f = r - b * N;
	const cplx<T> Fa = F[((size_t) p.ia[b] * E + e) * N + f];
	const cplx<T> Fb = F[((size_t) p.ib[b] * E + e) * N + f];
	X[i] = Fb * conjg(Fa);
}

/* cross spectra of one event (one workgroup), split into the fp16 pair of
 * the sky product with a power-of-two row scale (gemm_h3.hip) */
__global__ void __launch_bounds__(256) k_loc_xspec_h3(const cplx<float> *__restrict__ F, int E, int N, int nb, struct loc_pairs p, int Kp, _Float16 *__restrict__ Xh, _Float16 *__restrict__ Xl, float *__restrict__ irx)
{
	__shared__ float red[4];
	const int e = blockIdx.x;
	const int nc = nb * N;
	float mx = 0.f;
	for(int i = threadIdx.x; i < nc; i += 256) {
		const int b = i / N, f = i - b * N;
		const cplx<float> X = F[((size_t) p.ib[b] * E + e) * N + f] * conjg(F[((size_t) p.ia[b] * E + e) * N + f]);
		mx = fmaxf(mx, fmaxf(fabsf(X.re), fabsf(X.im)));
	}
	for(int o = 32; o > 0; o >>= 1)
		mx = fmaxf(mx, __shfl_xor(mx, o));
	if((threadIdx.x & 63) == 0)
		red[threadIdx.x >> 6] = mx;
	__syncthreads();
	mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
	int ex = 0;
	if(mx > 0.f) {
		frexpf(mx, &ex);
		ex = 14 - ex;
	}
	const float s = ldexpf(1.0f, ex);
	_Float16 *xh = Xh + (size_t) e * Kp, *xl = Xl + (size_t) e * Kp;
	for(int i = threadIdx.x; i < nc; i += 256) {
		const int b = i / N, f = i - b * N;
		const cplx<float> X = F[((size_t) p.ib[b] * E + e) * N + f] * conjg(F[((size_t) p.ia[b] * E + e) * N + f]);
		const float vr = X.re * s, vi = X.im * s;
		const _Float16 hr = (_Float16) vr, hi = (_Float16) vi;
		xh[2 * i] = hr;
		xh[2 * i + 1] = hi;
		xl[2 * i] = (_Float16) (vr - (float) hr);
		xl[2 * i + 1] = (_Float16) (vi - (float) hi);
	}
	for(int i = 2 * nc + threadIdx.x; i < Kp; i += 256) {
		xh[i] = (_Float16) 0.f;
		xl[i] = (_Float16) 0.f;
	}
	if(threadIdx.x == 0)
		irx[e] = ldexpf(1.0f, -ex);
}

template <typename T>
__global__ void __launch_bounds__(256) k_loc_epilogue(const T *__restrict__ S, int E, int L, const cplx<T> *__restrict__ prior, const double *__restrict__ gmst, cplx<T> *__restrict__ alm)
{
	/* one workgroup per series s = beta E + e: the phases exp(-i m gmst_e)
	 * once per series, then the coefficients */
	extern __shared__ unsigned char epi_lds[];
	cplx<T> *rot = reinterpret_cast<cplx<T> *>(epi_lds);	/* [L + 1] */
	const int ncoef = (L + 1) * (L + 1);
	const int s = blockIdx.x, beta = s / E, e = s - beta * E;
	const double ge = gmst ? gmst[e] : 0.0;
	for(int m = threadIdx.x; m <= L; m += blockDim.x) {
		double sn = 0.0, cs = 1.0;
		if(m)
			sincos(-(double) m * ge, &sn, &cs);
		rot[m] = mk((T) cs, (T) sn);
	}
	__syncthreads();
	const T *src = S + (size_t) e * 4 * ncoef + (size_t) beta * 2 * ncoef;
	for(int c = threadIdx.x; c < ncoef; c += blockDim.x) {
		const int m = lm_order(L, c);
		alm[(size_t) s * ncoef + c] = (mk(src[2 * c], src[2 * c + 1]) + prior[c]) * (m >= 0 ? rot[m] : conjg(rot[-m]));
	}
}

template <typename T>
static int upload_as(void **dst, const double *src, size_t n)
{
	T *h = (T *) malloc(n * sizeof(T));
	if(!h)
		return -1;
	for(size_t i = 0; i < n; i++)
		h[i] = (T) src[i];
	int rc = sphr_rt_check(hipMalloc(dst, n * sizeof(T)), "hipMalloc") || sphr_rt_check(hipMemcpy(*dst, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
	free(h);
	return rc ? -1 : 0;
}

/* complex [K][N] -> real [2K][2N] expansion on the host, as T */
template <typename T>
static int upload_expanded(void **dst, const double *src, size_t K, size_t N)
{
	T *h = (T *) malloc(4 * K * N * sizeof(T));
	if(!h)
		return -1;
	for(size_t k = 0; k < K; k++)
		for(size_t n = 0; n < N; n++) {
			const double re = src[2 * (k * N + n)], im = src[2 * (k * N + n) + 1];
			T *r0 = h + (2 * k) * (2 * N) + 2 * n, *r1 = r0 + 2 * N;
			r0[0] = (T) re;
			r0[1] = (T) im;
			r1[0] = (T) -im;
			r1[1] = (T) re;
		}
	int rc = sphr_rt_check(hipMalloc(dst, 4 * K * N * sizeof(T)), "hipMalloc") || sphr_rt_check(hipMemcpy(*dst, h, 4 * K * N * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
	free(h);
	return rc ? -1 : 0;
}

/* a complex matrix [K][N] real-expanded to [2K][2N] (rows 2k = (re, im),
 * 2k+1 = (-im, re) per column pair), transposed to [2N][Kp] and split
 * into fp16 pairs with a power-of-two scale per column (gemm_h3.hip) */
static int upload_split_cols(const double *src, size_t K, size_t N, int Kp, _Float16 **dh, _Float16 **dl, float **dic)
{
	const size_t rows = 2 * N;
	_Float16 *h = (_Float16 *) calloc(rows * Kp, sizeof(_Float16)), *l = (_Float16 *) calloc(rows * Kp, sizeof(_Float16));
	float *ic = (float *) malloc(rows * sizeof(float));
	int rc = -1;
	if(!h || !l || !ic)
		goto done;
	for(size_t j = 0; j < rows; j++) {
		const size_t n = j / 2;
		const int im = (int) (j & 1);
		/* column j of the expanded matrix: row 2k holds re or im of
		 * src[k][n], row 2k+1 holds -im or re */
		double mx = 0.0;
		for(size_t k = 0; k < K; k++) {
			const double re = src[2 * (k * N + n)], ii = src[2 * (k * N + n) + 1];
			mx = fmax(mx, fmax(fabs(re), fabs(ii)));
		}
		int ex = 0;
		if(mx > 0.0) {
			frexp(mx, &ex);
			ex = 14 - ex;
		}
		const double s = ldexp(1.0, ex);
		for(size_t k = 0; k < K; k++) {
			const double re = src[2 * (k * N + n)], ii = src[2 * (k * N + n) + 1];
			const double v0 = (im ? ii : re) * s, v1 = (im ? re : -ii) * s;
			const _Float16 h0 = (_Float16) v0, h1 = (_Float16) v1;
			h[j * Kp + 2 * k] = h0;
			l[j * Kp + 2 * k] = (_Float16) (v0 - (double) h0);
			h[j * Kp + 2 * k + 1] = h1;
			l[j * Kp + 2 * k + 1] = (_Float16) (v1 - (double) h1);
		}
		ic[j] = (float) ldexp(1.0, -ex);
	}
	if(sphr_rt_check(hipMalloc((void **) dh, rows * Kp * sizeof(_Float16)), "hipMalloc") || sphr_rt_check(hipMalloc((void **) dl, rows * Kp * sizeof(_Float16)), "hipMalloc")
	   || sphr_rt_check(hipMalloc((void **) dic, rows * sizeof(float)), "hipMalloc")
	   || sphr_rt_check(hipMemcpy(*dh, h, rows * Kp * sizeof(_Float16), hipMemcpyHostToDevice), "hipMemcpy")
	   || sphr_rt_check(hipMemcpy(*dl, l, rows * Kp * sizeof(_Float16), hipMemcpyHostToDevice), "hipMemcpy")
	   || sphr_rt_check(hipMemcpy(*dic, ic, rows * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy"))
		goto done;
	rc = 0;
done:
	free(h);
	free(l);
	free(ic);
	return rc;
}

extern "C" void sphr_loc_free(struct sphr_loc *lc)
{
	if(!lc)
		return;
	void *p[] = { lc->Wx, lc->Dx, lc->prior, lc->F, lc->X, lc->S, lc->Fr, lc->Wh, lc->Wl, lc->icw, lc->Dh, lc->Dl, lc->icd, lc->Ah, lc->Al, lc->ira, lc->Xh, lc->Xl, lc->irx,
		      lc->Ph, lc->Pl, lc->icp, lc->irr, lc->Peh, lc->Pel, lc->Poh, lc->Pol, lc->icpe, lc->icpo };
	for(unsigned i = 0; i < sizeof(p) / sizeof(p[0]); i++)
		if(p[i])
			hipFree(p[i]);
	synth_tables_free(&lc->syn);
	free(lc);
}

template <typename T>
static int create_t(struct sphr_loc *lc, const struct sphr_loc_desc *d)
{
	const size_t ncoef = (size_t) (d->L + 1) * (d->L + 1);
	/* W: rows (b, f), columns (beta, lm): reorder [beta][b][f][lm] */
	const size_t K = (size_t) d->nb * d->N;
	double *Wr = (double *) malloc(2 * K * 2 * ncoef * sizeof(double));
	if(!Wr)
		return -1;
	for(size_t k = 0; k < K; k++)
		for(int beta = 0; beta < 2; beta++)
			memcpy(Wr + 2 * (k * 2 * ncoef + beta * ncoef), d->W + 2 * ((beta * K + k) * ncoef), 2 * ncoef * sizeof(double));
	const int split = sizeof(T) == 4;
	lc->Kw = (int) ((2 * K + 63) / 64 * 64);
	lc->Kd = (2 * d->len + 63) / 64 * 64;
	int rc = split ? upload_split_cols(Wr, K, 2 * ncoef, lc->Kw, &lc->Wh, &lc->Wl, &lc->icw) : upload_expanded<T>(&lc->Wx, Wr, K, 2 * ncoef);
	free(Wr);
	if(rc)
		return -1;
	rc = split ? upload_split_cols(d->dft, d->len, d->N, lc->Kd, &lc->Dh, &lc->Dl, &lc->icd) : upload_expanded<T>(&lc->Dx, d->dft, d->len, d->N);
	if(rc || upload_as<T>(&lc->prior, d->prior, 2 * ncoef))
		return -1;
	const size_t E = d->max_events;
	if(sphr_rt_check(hipMalloc(&lc->F, (size_t) d->D * E * 2 * d->N * sizeof(T)), "hipMalloc")
	   || sphr_rt_check(hipMalloc(&lc->S, E * 4 * ncoef * sizeof(T)), "hipMalloc")
	   || sphr_rt_check(hipMalloc(&lc->Fr, 2 * E * 2 * (d->L + 1) * (2 * d->L + 1) * 2 * sizeof(T)), "hipMalloc"))
		return -1;
	if(!split)
		return sphr_rt_check(hipMalloc(&lc->X, E * 2 * K * sizeof(T)), "hipMalloc") ? -1 : 0;
	if(sphr_rt_check(hipMalloc((void **) &lc->Ah, (size_t) d->D * E * lc->Kd * sizeof(_Float16)), "hipMalloc")
	   || sphr_rt_check(hipMalloc((void **) &lc->Al, (size_t) d->D * E * lc->Kd * sizeof(_Float16)), "hipMalloc")
	   || sphr_rt_check(hipMalloc((void **) &lc->ira, (size_t) d->D * E * sizeof(float)), "hipMalloc")
	   || sphr_rt_check(hipMalloc((void **) &lc->Xh, E * lc->Kw * sizeof(_Float16)), "hipMalloc")
	   || sphr_rt_check(hipMalloc((void **) &lc->Xl, E * lc->Kw * sizeof(_Float16)), "hipMalloc")
	   || sphr_rt_check(hipMalloc((void **) &lc->irx, E * sizeof(float)), "hipMalloc"))
		return -1;
	/* the phi DFT of the meshes: exp(i m phi_k), m = 0..L then -L..-1
	 * (the ring sums' column order), k < n */
	{
		const int L = d->L, n = 2 * (L + 1), nm = 2 * L + 1;
		double *ph = (double *) malloc((size_t) nm * n * 2 * sizeof(double)), *tw = (double *) malloc(2 * (size_t) n * sizeof(double));
		if(!ph || !tw) {
			free(ph);
			free(tw);
			return -1;
		}
		sphr_twiddles(n, +1, tw);
		for(int i = 0; i < nm; i++) {
			const int m = i <= L ? i : i - nm;
			for(int k = 0; k < n; k++) {
				const long r = (((long) k * m) % n + n) % n;
				ph[2 * ((size_t) i * n + k)] = tw[2 * r];
				ph[2 * ((size_t) i * n + k) + 1] = tw[2 * r + 1];
			}
		}
		lc->Kp = (2 * nm + 63) / 64 * 64;
		rc = upload_split_cols(ph, nm, n, lc->Kp, &lc->Ph, &lc->Pl, &lc->icp);
		/* the half-turn matrices: exp(i m phi_k) for k < n/2 over the even
		 * orders and over the odd ones, each class as m >= 0 ascending
		 * then m < 0 ascending (the order launch_phi_h3 gathers) */
		if(!rc && n <= 128) {
			const int ev = L / 2 * 2, om = (L & 1) ? L : L - 1, NE = ev + 1, NO = nm - NE;
			lc->Kph = ((2 * (NE > NO ? NE : NO)) + 63) / 64 * 64;
			for(int par = 0; par < 2 && !rc; par++) {
				const int cnt = par ? NO : NE;
				for(int i = 0; i < cnt; i++) {
					const int np = par ? (om + 1) / 2 : ev / 2 + 1;	/* orders >= 0 of the class */
					const int m = i < np ? (par ? 1 : 0) + 2 * i : (par ? -om : -ev) + 2 * (i - np);
					for(int k = 0; k < n / 2; k++) {
						const long r = (((long) k * m) % n + n) % n;
						ph[2 * ((size_t) i * (n / 2) + k)] = tw[2 * r];
						ph[2 * ((size_t) i * (n / 2) + k) + 1] = tw[2 * r + 1];
					}
				}
				rc = par ? upload_split_cols(ph, NO, n / 2, lc->Kph, &lc->Poh, &lc->Pol, &lc->icpo)
					 : upload_split_cols(ph, NE, n / 2, lc->Kph, &lc->Peh, &lc->Pel, &lc->icpe);
			}
			lc->ht = !rc;
		}
		free(ph);
		free(tw);
		if(rc || sphr_rt_check(hipMalloc((void **) &lc->irr, 2 * E * (size_t) n * sizeof(float)), "hipMalloc"))
			return -1;
	}
	return 0;
}

extern "C" struct sphr_loc *sphr_loc_create(const struct sphr_loc_desc *d)
{
	if(sphr_gpu_ensure())
		return NULL;
	if(d->nb > 16 || d->D > 16 || (d->precision != SPHR_FP64 && d->precision != SPHR_MIXED)) {
		sphr_set_error("sphr_loc_new: unsupported shape or precision");
		return NULL;
	}
	struct sphr_loc *lc = (struct sphr_loc *) calloc(1, sizeof(*lc));
	if(!lc)
		return NULL;
	lc->precision = d->precision;
	lc->D = d->D;
	lc->nb = d->nb;
	lc->N = d->N;
	lc->len = d->len;
	lc->L = d->L;
	lc->max_events = d->max_events;
	memcpy(lc->ia, d->ia, sizeof(lc->ia));
	memcpy(lc->ib, d->ib, sizeof(lc->ib));
	const int rc = d->precision == SPHR_MIXED ? create_t<float>(lc, d) : create_t<double>(lc, d);
	if(rc || synth_tables_build(d->L, &lc->syn)) {
		sphr_loc_free(lc);
		return NULL;
	}
	return lc;
}

/* the mixed-precision products: the series and the cross spectra split
 * per row, the DFT and sky matrices per column (at creation), three fp16
 * MFMA products each with float32 accumulation */
static int run_products_h3(struct sphr_loc *lc, const void *series, int series_f64, int E, hipStream_t st)
{
	const int N = lc->N, nb = lc->nb, L = lc->L, DE = lc->D * E;
	const int ncoef = (L + 1) * (L + 1);
	/* 1: spectra; 2: cross spectra; 3: skies */
	int r = series_f64 ? launch_split_rows<double>(DE, 2 * lc->len, lc->Kd, (const double *) series, 2L * lc->len, lc->Ah, lc->Al, lc->Kd, lc->ira, st)
			   : launch_split_rows<float>(DE, 2 * lc->len, lc->Kd, (const float *) series, 2L * lc->len, lc->Ah, lc->Al, lc->Kd, lc->ira, st);
	if(r || launch_gemm_h3<float>(DE, 2 * N, lc->Kd, lc->Ah, lc->Al, lc->Kd, lc->ira, lc->Dh, lc->Dl, lc->Kd, lc->icd, (float *) lc->F, 2L * N, st))
		return -1;
	struct loc_pairs p;
	memcpy(p.ia, lc->ia, sizeof(p.ia));
	memcpy(p.ib, lc->ib, sizeof(p.ib));
	hipLaunchKernelGGL(k_loc_xspec_h3, dim3(E), dim3(256), 0, st, (const cplx<float> *) lc->F, E, N, nb, p, lc->Kw, lc->Xh, lc->Xl, lc->irx);
	if(hipGetLastError() != hipSuccess)
		return -1;
	return launch_gemm_h3<float>(E, 4 * ncoef, lc->Kw, lc->Xh, lc->Xl, lc->Kw, lc->irx, lc->Wh, lc->Wl, lc->Kw, lc->icw, (float *) lc->S, 4L * ncoef, st);
}

template <typename T>
static int run_products_f64(struct sphr_loc *lc, const void *series, int series_f64, int E, hipStream_t st)
{
	const int N = lc->N, nb = lc->nb, L = lc->L;
	const int ncoef = (L + 1) * (L + 1);
	/* 1: spectra, [D E][2 len] x [2 len][2 N] (the series widened or
	 * narrowed on the way in when they are not in the working precision) */
	int r;
	if(series_f64)
		r = launch_gemm<T, double>(lc->D * E, 2 * N, 2 * lc->len, (const double *) series, 2L * lc->len, (const T *) lc->Dx, 2L * N, (T *) lc->F, 2L * N, st);
	else
		r = launch_gemm<T, float>(lc->D * E, 2 * N, 2 * lc->len, (const float *) series, 2L * lc->len, (const T *) lc->Dx, 2L * N, (T *) lc->F, 2L * N, st);
	if(r)
		return -1;
	/* 2: cross spectra */
	struct loc_pairs p;
	memcpy(p.ia, lc->ia, sizeof(p.ia));
	memcpy(p.ib, lc->ib, sizeof(p.ib));
	const long nx = (long) E * nb * N;
	hipLaunchKernelGGL((k_loc_xspec<T>), dim3((unsigned) ((nx + 255) / 256)), dim3(256), 0, st, (const cplx<T> *) lc->F, E, N, nb, p, (cplx<T> *) lc->X);
	if(hipGetLastError() != hipSuccess)
		return -1;
	/* 3: skies, [E][2 nb N] x [2 nb N][4 ncoef] */
	if(launch_gemm<T, T>(E, 4 * ncoef, 2 * nb * N, (const T *) lc->X, 2L * nb * N, (const T *) lc->Wx, 4L * ncoef, (T *) lc->S, 4L * ncoef, st))
		return -1;
	return 0;
}

template <typename T>
static int run_epilogue(struct sphr_loc *lc, int E, const double *gmst, void *alm, void *map, hipStream_t st)
{
	const int L = lc->L;
	/* 4: prior, gmst: with the meshes, formed by the ring-sum kernel as it
	 * reads them; without, by the epilogue kernel */
	struct sky_epi<T> epi = { (const T *) lc->S, (const cplx<T> *) lc->prior, gmst, E, (cplx<T> *) alm };
	auto epilogue = [&]() -> int {
		hipLaunchKernelGGL((k_loc_epilogue<T>), dim3(2 * E), dim3(256), (L + 1) * sizeof(cplx<T>), st, (const T *) lc->S, E, L, (const cplx<T> *) lc->prior, gmst, (cplx<T> *) alm);
		return hipGetLastError() == hipSuccess ? 0 : -1;
	};
	if(!map)
		return epilogue();
	/* 5: meshes of both skies: Legendre ring sums, then the phi DFT as a
	 * GEMM (the one-kernel synthesis when the order is too large for it);
	 * in mixed precision the GEMM is the split-FP16 product, each ring's
	 * split scale from the ring-sum kernel, the split done while staging */
	const int n = 2 * (L + 1), nm = 2 * L + 1;
	if(launch_ring_sums<T>(NULL, 2 * E, &lc->syn.g, &lc->syn.tb, (cplx<T> *) lc->Fr, st, sizeof(T) == 4 ? lc->irr : NULL, &epi) == 0) {
		/* mixed precision at orders whose half turn fits one tile: the
		 * half-turn phi DFT */
		if(sizeof(T) == 4 && lc->ht)
			return launch_phi_h3(2 * E * n, L, (const float *) lc->Fr, 2L * nm, lc->irr, lc->Peh, lc->Pel, lc->icpe, lc->Poh, lc->Pol, lc->icpo, lc->Kph, (float *) map, st);
		if(sizeof(T) == 4)
			return launch_gemm_h3a<float>(2 * E * n, 2 * n, 2 * nm, (const float *) lc->Fr, 2L * nm, lc->irr, lc->Ph, lc->Pl, lc->Kp, lc->icp, (float *) map, 2L * n, st);
		return launch_phi<T>((const cplx<T> *) lc->Fr, 2 * E, &lc->syn.g, &lc->syn.tb, (cplx<T> *) map, st);
	}
	if(epilogue())
		return -1;
	return launch_synth<T, T, T>((const cplx<T> *) alm, 2 * E, 0, &lc->syn.g, &lc->syn.tb, (cplx<T> *) map, st);
}

template <typename T>
static int run_t(struct sphr_loc *lc, const void *series, int series_f64, int E, const double *gmst, void *alm, void *map, hipStream_t st)
{
	if(sizeof(T) == 4) {
		if(run_products_h3(lc, series, series_f64, E, st))
			return -1;
	} else if(run_products_f64<T>(lc, series, series_f64, E, st))
		return -1;
	return run_epilogue<T>(lc, E, gmst, alm, map, st);
}

extern "C" int sphr_loc_run(struct sphr_loc *lc, const void *series, int series_f64, int n_events, const double *gmst, void *alm, void *map, void *stream)
{
	if(!lc || !series || !alm || n_events < 1 || n_events > lc->max_events) {
		sphr_set_error("sphr_loc_run: bad arguments (events %d of at most %d)", n_events, lc ? lc->max_events : 0);
		return -1;
	}
	hipStream_t st = stream ? (hipStream_t) stream : sphr_rt_stream();
	const int r = lc->precision == SPHR_MIXED ? run_t<float>(lc, series, series_f64, n_events, gmst, alm, map, st) : run_t<double>(lc, series, series_f64, n_events, gmst, alm, map, st);
	if(r)
		sphr_set_error("sphr_loc_run: kernel launch failed (%s)", hipGetErrorString(hipGetLastError()));
	return r;
}

extern "C" int sphr_loc_peaks(struct sphr_loc *lc, const void *alm, const void *map, int n_events, double *theta, double *phi, double *value, void *stream)
{
	if(!lc || !alm || n_events < 1 || n_events > lc->max_events || !theta || !phi || !value) {
		sphr_set_error("sphr_loc_peaks: bad arguments");
		return -1;
	}
	hipStream_t st = stream ? (hipStream_t) stream : sphr_rt_stream();
	const double *sx = map ? NULL : sphr_gpu_seed_nodes();
	if(!map && !sx)
		return -1;
	const int ns = 2 * (SPHR_L_SEED + 1), S = 2 * n_events;
	int r;
	if(lc->precision == SPHR_MIXED)
		r = launch_extremize<float>((const cplx<float> *) alm, S, lc->L, 1, SPHR_L_SEED, (const cplx<float> *) map, lc->syn.x, lc->syn.cm, lc->syn.alpha, sx, ns, 100, theta, phi, value, st);
	else
		r = launch_extremize<double>((const cplx<double> *) alm, S, lc->L, 1, SPHR_L_SEED, (const cplx<double> *) map, lc->syn.x, lc->syn.cm, lc->syn.alpha, sx, ns, 100, theta, phi, value, st);
	if(r)
		sphr_set_error("sphr_loc_peaks: kernel launch failed (%s)", hipGetErrorString(hipGetLastError()));
	return r;
}

extern "C" void sphr_loc_info(const struct sphr_loc *lc, int info[8])
{
	info[0] = lc->D;
	info[1] = lc->nb;
	info[2] = lc->N;
	info[3] = lc->len;
	info[4] = lc->L;
	info[5] = lc->precision;
	info[6] = lc->max_events;
	info[7] = 0;
}
