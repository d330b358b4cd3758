/*
 * HIP kernels of the cross-power path for gfx950 (64-wide wavefronts).
 *
 *   rfft_pairs   batched real FFTs: two real series packed as one complex
 *                series, in-place radix-4 decimation in time in LDS from a
 *                bit-reversed load, then split into two half spectra
 *   xproj        cross spectrum X = F_b conj(F_a) formed in LDS and
 *                contracted with the packed delay-product matrix (a real
 *                GEMM over K = N + 2), split over K into partial sums
 *   alm          partial sums -> per-baseline polar power -> rotation into
 *                the Earth frame (the D^l_{0 m'} row), 1/nb average and
 *                rotate_z, written in the reference coefficient layout
 *   synth        Legendre stage per (ring pair, m) with the l recurrence
 *                run in registers (equatorial symmetry gives both rings),
 *                then a Stockham mixed-radix DFT along phi in LDS
 * plus the FP64 helpers of the synchronous drop-in entry points.
 */
#include <hip/hip_runtime.h>
#include "kernels.h"

#define NT 256

/* ------------------------------------------------------------------ */
/* FFT                                                                 */
/* ------------------------------------------------------------------ */

/* in-place DIT FFT of z[0..N) (LDS), input already in bit-reversed
 * order; forward, tw[k] = exp(-2 pi i k / N) */
template <typename T>
__device__ void fft_dit(cplx<T> *z, int N, int logN, const cplx<T> *__restrict__ tw)
{
	int s = 1;
	if(logN & 1) {
		for(int q = threadIdx.x; q < N / 2; q += blockDim.x) {
			const cplx<T> a = z[2 * q], b = z[2 * q + 1];
			z[2 * q] = a + b;
			z[2 * q + 1] = a - b;
		}
		__syncthreads();
		s = 2;
	}
	for(; s < N; s *= 4) {
		const int stride = N / (4 * s);
		for(int q = threadIdx.x; q < N / 4; q += blockDim.x) {
			const int j = q % s;
			const int i0 = (q / s) * 4 * s + j;
			const cplx<T> w4 = tw[j * stride], w2 = tw[2 * j * stride];
			const cplx<T> z0 = z[i0], z1 = z[i0 + s], z2 = z[i0 + 2 * s], z3 = z[i0 + 3 * s];
			const cplx<T> t1 = w2 * z1, t3 = w2 * z3;
			const cplx<T> a0 = z0 + t1, a1 = z0 - t1, a2 = z2 + t3, a3 = z2 - t3;
			const cplx<T> u2 = w4 * a2;
			const cplx<T> v3 = w4 * a3;
			const cplx<T> u3 = mk(v3.im, -v3.re);	/* times -i */
			z[i0] = a0 + u2;
			z[i0 + 2 * s] = a0 - u2;
			z[i0 + s] = a1 + u3;
			z[i0 + 3 * s] = a1 - u3;
		}
		__syncthreads();
	}
}

__device__ inline int ilog2(int n)
{
	return 31 - __clz(n);
}

template <typename T, typename TI>
__global__ void __launch_bounds__(NT) k_rfft_pairs(const TI *__restrict__ in, const double *__restrict__ window, int nseries, int N, const cplx<T> *__restrict__ tw, cplx<T> *__restrict__ spec)
{
	extern __shared__ unsigned char smem_raw[];
	cplx<T> *z = reinterpret_cast<cplx<T> *>(smem_raw);
	const int logN = ilog2(N);
	const int sx = 2 * blockIdx.x, sy = sx + 1;
	const TI *x = in + (size_t) sx * N;
	const TI *y = in + (size_t) sy * N;
	const bool has_y = sy < nseries;
	for(int t = threadIdx.x; t < N; t += blockDim.x) {
		const T w = window ? (T) window[t] : (T) 1;
		const int r = __brev((unsigned) t) >> (32 - logN);
		z[r] = mk((T) x[t] * w, has_y ? (T) y[t] * w : (T) 0);
	}
	__syncthreads();
	fft_dit<T>(z, N, logN, tw);
	const int Nh = N / 2 + 1;
	cplx<T> *ox = spec + (size_t) sx * Nh;
	cplx<T> *oy = spec + (size_t) sy * Nh;
	for(int f = threadIdx.x; f < Nh; f += blockDim.x) {
		const cplx<T> zf = z[f & (N - 1)], zc = conjg(z[(N - f) & (N - 1)]);
		const cplx<T> sum = zf + zc, dif = zf - zc;
		ox[f] = mk(sum.re * (T) 0.5, sum.im * (T) 0.5);
		if(has_y)
			oy[f] = mk(dif.im * (T) 0.5, -dif.re * (T) 0.5);
	}
}

template <typename T>
int launch_rfft_pairs(const void *in, int in_f32, const double *window, int nseries, int N, const cplx<T> *tw, cplx<T> *spec, hipStream_t st)
{
	const size_t lds = (size_t) N * sizeof(cplx<T>);
	dim3 grid((nseries + 1) / 2);
	if(in_f32)
		hipLaunchKernelGGL((k_rfft_pairs<T, float>), grid, dim3(NT), lds, st, (const float *) in, window, nseries, N, tw, spec);
	else
		hipLaunchKernelGGL((k_rfft_pairs<T, double>), grid, dim3(NT), lds, st, (const double *) in, window, nseries, N, tw, spec);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_rfft_pairs<double>(const void *, int, const double *, int, int, const cplx<double> *, cplx<double> *, hipStream_t);
template int launch_rfft_pairs<float>(const void *, int, const double *, int, int, const cplx<float> *, cplx<float> *, hipStream_t);

/* single transform of length N (power of two): real input (full spectrum
 * out, negative frequencies conjugate) or complex input */
__global__ void __launch_bounds__(NT) k_fft_single(const double *__restrict__ in, int real_input, int N, const cplx<double> *__restrict__ tw, cplx<double> *__restrict__ out)
{
	extern __shared__ unsigned char smem_raw[];
	cplx<double> *z = reinterpret_cast<cplx<double> *>(smem_raw);
	const int logN = ilog2(N);
	for(int t = threadIdx.x; t < N; t += blockDim.x) {
		const int r = logN ? (int) (__brev((unsigned) t) >> (32 - logN)) : 0;
		z[r] = real_input ? mk(in[t], 0.0) : mk(in[2 * t], in[2 * t + 1]);
	}
	__syncthreads();
	if(N > 1)
		fft_dit<double>(z, N, logN, tw);
	for(int f = threadIdx.x; f < N; f += blockDim.x)
		out[f] = z[f];
}

int launch_fft_single(const double *in, int real_input, int N, const cplx<double> *tw, cplx<double> *out, hipStream_t st)
{
	hipLaunchKernelGGL(k_fft_single, dim3(1), dim3(NT), (size_t) N * sizeof(cplx<double>), st, in, real_input, N, tw, out);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* direct DFT for lengths that are not powers of two; exact twiddle index
 * (j k mod N) into the length-N table */
__global__ void k_dft_direct(const double *__restrict__ in, int real_input, int N, const cplx<double> *__restrict__ tw, cplx<double> *__restrict__ out)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if(k >= N)
		return;
	cplx<double> acc = mk(0.0, 0.0);
	long t = 0;
	for(int j = 0; j < N; j++) {
		const cplx<double> v = real_input ? mk(in[j], 0.0) : mk(in[2 * j], in[2 * j + 1]);
		cfma(acc, v, tw[t]);
		t += k;
		if(t >= N)
			t -= N;
	}
	out[k] = acc;
}

int launch_dft_direct(const double *in, int real_input, int N, const cplx<double> *tw, cplx<double> *out, hipStream_t st)
{
	hipLaunchKernelGGL(k_dft_direct, dim3((N + NT - 1) / NT), dim3(NT), 0, st, in, real_input, N, tw, out);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* ------------------------------------------------------------------ */
/* cross spectrum x delay-product projection                           */
/* ------------------------------------------------------------------ */

#define XP_TS 16	/* segments per workgroup */
#define XP_KC 16	/* frequencies per LDS chunk */
#define XP_UC 4		/* 64-column groups per workgroup (256 columns) */

template <typename T>
__global__ void __launch_bounds__(NT) k_xproj(const cplx<T> *__restrict__ spec, int S, struct xproj_geom g, const T *__restrict__ W, const int *__restrict__ col_bl, int k_per_split, T *__restrict__ partial)
{
	extern __shared__ unsigned char smem_raw[];
	cplx<T> *sF = reinterpret_cast<cplx<T> *>(smem_raw);		/* [D][TS][KC] */
	T *sA = reinterpret_cast<T *>(sF + g.D * XP_TS * XP_KC);	/* [nb][TS][2 KC] */
	const int s0 = blockIdx.x * XP_TS;
	const int c0 = blockIdx.y * 64 * XP_UC;
	const int kbeg = blockIdx.z * k_per_split;
	const int kend = min(g.K2, kbeg + k_per_split);
	const int fbeg = kbeg / 2, fend = kend / 2;
	const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
	int cidx[XP_UC], cb[XP_UC];
	bool cv[XP_UC];
	for(int u = 0; u < XP_UC; u++) {
		cidx[u] = c0 + tx + 64 * u;
		cv[u] = cidx[u] < g.CT;
		cb[u] = cv[u] ? col_bl[cidx[u]] : 0;
	}
	T acc[XP_UC][4];
	for(int u = 0; u < XP_UC; u++)
		for(int v = 0; v < 4; v++)
			acc[u][v] = (T) 0;
	for(int f0 = fbeg; f0 < fend; f0 += XP_KC) {
		const int nf = min(XP_KC, fend - f0);
		for(int i = threadIdx.x; i < g.D * XP_TS * XP_KC; i += NT) {
			const int d = i / (XP_TS * XP_KC), r = i % (XP_TS * XP_KC);
			const int s = r / XP_KC, ff = r % XP_KC;
			sF[i] = (s0 + s < S && ff < nf) ? spec[((size_t) d * S + s0 + s) * g.Nh + f0 + ff] : mk((T) 0, (T) 0);
		}
		__syncthreads();
		for(int i = threadIdx.x; i < g.nb * XP_TS * XP_KC; i += NT) {
			const int b = i / (XP_TS * XP_KC), r = i % (XP_TS * XP_KC);
			const int s = r / XP_KC, ff = r % XP_KC;
			const cplx<T> Fa = sF[(g.ia[b] * XP_TS + s) * XP_KC + ff];
			const cplx<T> Fb = sF[(g.ib[b] * XP_TS + s) * XP_KC + ff];
			const cplx<T> X = Fb * conjg(Fa);
			sA[(b * XP_TS + s) * 2 * XP_KC + 2 * ff] = X.re;
			sA[(b * XP_TS + s) * 2 * XP_KC + 2 * ff + 1] = X.im;
		}
		__syncthreads();
		for(int kk = 0; kk < 2 * nf; kk++) {
			const size_t row = (size_t) (2 * f0 + kk) * g.CT;
			T w[XP_UC];
			for(int u = 0; u < XP_UC; u++)
				w[u] = cv[u] ? W[row + cidx[u]] : (T) 0;
			for(int u = 0; u < XP_UC; u++) {
				const T *a = sA + (cb[u] * XP_TS + ty * 4) * 2 * XP_KC + kk;
				for(int v = 0; v < 4; v++)
					acc[u][v] = fma(a[v * 2 * XP_KC], w[u], acc[u][v]);
			}
		}
		__syncthreads();
	}
	for(int u = 0; u < XP_UC; u++)
		for(int v = 0; v < 4; v++) {
			const int s = s0 + ty * 4 + v;
			if(s < S && cv[u])
				partial[((size_t) blockIdx.z * S + s) * g.CT + cidx[u]] = acc[u][v];
		}
}

template <typename T>
int launch_xproj(const cplx<T> *spec, int S, const struct xproj_geom *g, const T *W, const int *col_bl, int k_per_split, int nsplit, T *partial, hipStream_t st)
{
	const size_t lds = (size_t) g->D * XP_TS * XP_KC * sizeof(cplx<T>) + (size_t) g->nb * XP_TS * 2 * XP_KC * sizeof(T);
	dim3 grid((S + XP_TS - 1) / XP_TS, (g->CT + 64 * XP_UC - 1) / (64 * XP_UC), nsplit);
	hipLaunchKernelGGL((k_xproj<T>), grid, dim3(NT), lds, st, spec, S, *g, W, col_bl, k_per_split, partial);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_xproj<double>(const cplx<double> *, int, const struct xproj_geom *, const double *, const int *, int, int, double *, hipStream_t);
template int launch_xproj<float>(const cplx<float> *, int, const struct xproj_geom *, const float *, const int *, int, int, float *, hipStream_t);

/* ------------------------------------------------------------------ */
/* per-baseline power -> sky coefficients                              */
/* ------------------------------------------------------------------ */

template <typename T, typename TO>
__global__ void __launch_bounds__(NT) k_alm(const T *__restrict__ partial, int nsplit, int S, struct xproj_geom g, int L, const cplx<T> *__restrict__ R, const int2 *__restrict__ lm, const double *__restrict__ gmst, cplx<TO> *__restrict__ alm)
{
	const int ncoef = (L + 1) * (L + 1);
	const int idx = blockIdx.x * blockDim.x + threadIdx.x;
	const int s = blockIdx.y;
	if(idx >= ncoef)
		return;
	const int2 q = lm[idx];
	const int l = q.x, m = q.y;
	cplx<T> acc = mk((T) 0, (T) 0);
	for(int b = 0; b < g.nb; b++) {
		if(l > g.Lb[b])
			continue;
		const int c = g.col_start[b] + 2 * l;
		T pr = 0, pi = 0;
		for(int sp = 0; sp < nsplit; sp++) {
			const T *pp = partial + ((size_t) sp * S + s) * g.CT + c;
			pr += pp[0];
			pi += pp[1];
		}
		cfma(acc, R[(size_t) b * ncoef + idx], mk(pr, pi));
	}
	if(gmst && m) {
		double sn, cs;
		sincos(-(double) m * gmst[s], &sn, &cs);
		acc = acc * mk((T) cs, (T) sn);
	}
	alm[(size_t) s * ncoef + idx] = mk((TO) acc.re, (TO) acc.im);
}

template <typename T, typename TO>
int launch_alm(const T *partial, int nsplit, int S, const struct xproj_geom *g, int L, const cplx<T> *R, const int2 *lm, const double *gmst, cplx<TO> *alm, hipStream_t st)
{
	const int ncoef = (L + 1) * (L + 1);
	dim3 grid((ncoef + NT - 1) / NT, S);
	hipLaunchKernelGGL((k_alm<T, TO>), grid, dim3(NT), 0, st, partial, nsplit, S, *g, L, R, lm, gmst, alm);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_alm<double, double>(const double *, int, int, const struct xproj_geom *, int, const cplx<double> *, const int2 *, const double *, cplx<double> *, hipStream_t);
template int launch_alm<float, float>(const float *, int, int, const struct xproj_geom *, int, const cplx<float> *, const int2 *, const double *, cplx<float> *, hipStream_t);

/* ------------------------------------------------------------------ */
/* synthesis onto the Gauss-Legendre x equispaced mesh                 */
/* ------------------------------------------------------------------ */

#define SY_RP 4		/* ring pairs per workgroup */

template <typename TI, typename T, typename TO>
__global__ void __launch_bounds__(NT) k_synth(const cplx<TI> *__restrict__ alm, int polar, struct synth_geom g, struct synth_tables tb, cplx<TO> *__restrict__ map)
{
	extern __shared__ unsigned char smem_raw[];
	const int n = g.n, L = g.L, B = g.B;
	cplx<T> *buf0 = reinterpret_cast<cplx<T> *>(smem_raw);		/* [2 RP][n] */
	cplx<T> *buf1 = buf0 + 2 * SY_RP * n;
	cplx<T> *twp = buf1 + 2 * SY_RP * n;				/* [n] */
	const int s = blockIdx.y;
	const int j0 = blockIdx.x * SY_RP;
	const int ncoef = polar ? L + 1 : (L + 1) * (L + 1);
	const cplx<TI> *a = alm + (size_t) s * ncoef;
	for(int t = threadIdx.x; t < n; t += blockDim.x)
		twp[t] = mk((T) tb.twp[2 * t], (T) tb.twp[2 * t + 1]);
	for(int t = threadIdx.x; t < 2 * SY_RP * n; t += blockDim.x)
		buf0[t] = mk((T) 0, (T) 0);
	__syncthreads();
	/* Legendre stage: item = (ring pair, m) */
	const int nm = polar ? 1 : 2 * L + 1;
	for(int it = threadIdx.x; it < SY_RP * nm; it += blockDim.x) {
		const int rp = it / nm, mi = it % nm;
		const int j = j0 + rp;
		if(j >= B)
			continue;
		const int m = polar ? 0 : mi - L;
		const int am = m < 0 ? -m : m;
		const T x = (T) tb.x[j];
		const T *al = nullptr;
		(void) al;
		const double *alpha = tb.alpha + (size_t) am * (L + 1);
		T pprev = 0, pcur = (T) tb.smm[(size_t) j * (L + 1) + am];
		cplx<T> E = mk((T) 0, (T) 0), O = mk((T) 0, (T) 0);
		const int base = polar ? 0 : lm_offset(L, am, m);	/* offset of (l = |m|, m) */
		T alpha_prev = 0;
		for(int l = am; l <= L; l++) {
			if(l == am + 1) {
				pprev = pcur;
				alpha_prev = (T) alpha[l];
				pcur = alpha_prev * x * pprev;
			} else if(l > am + 1) {
				const T al_l = (T) alpha[l];
				const T pn = al_l * (x * pcur - pprev / alpha_prev);
				pprev = pcur;
				pcur = pn;
				alpha_prev = al_l;
			}
			const cplx<TI> c = a[base + l - am];
			const cplx<T> cc = mk((T) c.re, (T) c.im);
			if((l + am) & 1)
				cfma(O, cc, pcur);
			else
				cfma(E, cc, pcur);
		}
		const int col = m >= 0 ? m : n + m;
		const T sg = (m < 0 && (am & 1)) ? (T) -1 : (T) 1;
		buf0[(2 * rp) * n + col] = (E + O) * sg;
		buf0[(2 * rp + 1) * n + col] = (E - O) * sg;
	}
	__syncthreads();
	/* backward DFT along phi, Stockham, one output per thread per stage */
	cplx<T> *src = buf0, *dst = buf1;
	int Ns = 1;
	for(int fi = 0; fi < g.nfac; fi++) {
		const int R = g.fac[fi];
		const int NsR = Ns * R, nR = n / R, tstep = n / NsR;
		for(int o = threadIdx.x; o < 2 * SY_RP * n; o += blockDim.x) {
			const int ring = o / n, oo = o % n;
			const int e = oo % NsR;
			const int jj = (oo / NsR) * Ns + (oo % Ns);
			const cplx<T> *sr = src + ring * n;
			cplx<T> acc = mk((T) 0, (T) 0);
			int t = 0;
			const int tinc = (e * tstep) % n;
			for(int r = 0; r < R; r++) {
				cfma(acc, sr[jj + r * nR], twp[t]);
				t += tinc;
				if(t >= n)
					t -= n;
			}
			dst[o] = acc;
		}
		__syncthreads();
		cplx<T> *tmp = src;
		src = dst;
		dst = tmp;
		Ns = NsR;
	}
	/* write the rows: slot 2rp -> ring j, slot 2rp+1 -> ring 2B-1-j */
	for(int o = threadIdx.x; o < 2 * SY_RP * n; o += blockDim.x) {
		const int slot = o / n, k = o % n;
		const int rp = slot >> 1;
		const int j = j0 + rp;
		if(j >= B)
			continue;
		const int ring = (slot & 1) ? 2 * B - 1 - j : j;
		const cplx<T> v = src[o];
		map[((size_t) s * n + ring) * n + k] = mk((TO) v.re, (TO) v.im);
	}
}

template <typename TI, typename T, typename TO>
int launch_synth(const cplx<TI> *alm, int S, int polar, const struct synth_geom *g, const struct synth_tables *tb, cplx<TO> *map, hipStream_t st)
{
	const size_t lds = (size_t) (2 * 2 * SY_RP * g->n + g->n) * sizeof(cplx<T>);
	dim3 grid((g->B + SY_RP - 1) / SY_RP, S);
	hipLaunchKernelGGL((k_synth<TI, T, TO>), grid, dim3(NT), lds, st, alm, polar, *g, *tb, map);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_synth<double, double, double>(const cplx<double> *, int, int, const struct synth_geom *, const struct synth_tables *, cplx<double> *, hipStream_t);
template int launch_synth<float, float, float>(const cplx<float> *, int, int, const struct synth_geom *, const struct synth_tables *, cplx<float> *, hipStream_t);

/* ------------------------------------------------------------------ */
/* drop-in helpers (FP64)                                              */
/* ------------------------------------------------------------------ */

struct xspec_args {
	const cplx<double> *fs[2 * SPHR_MAX_BASELINES];
	int ia[SPHR_MAX_BASELINES], ib[SPHR_MAX_BASELINES];
};

/* fsp[b][i] = F_{ib}[i] conj(F_{ia}[i]) (correlator.c:687-688) */
__global__ void k_xspec(struct xspec_args a, int n, int nb, cplx<double> *__restrict__ fsp)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const int b = blockIdx.y;
	if(i >= n || b >= nb)
		return;
	fsp[(size_t) b * n + i] = a.fs[a.ib[b]][i] * conjg(a.fs[a.ia[b]][i]);
}

int launch_xspec(const cplx<double> *const *fs_dev, int n, int nb, const int *ia, const int *ib, cplx<double> *fsp, hipStream_t st)
{
	struct xspec_args a;
	if(nb > SPHR_MAX_BASELINES)
		return -1;
	int nmax = 0;
	for(int b = 0; b < nb; b++) {
		a.ia[b] = ia[b];
		a.ib[b] = ib[b];
		nmax = max(nmax, max(ia[b], ib[b]) + 1);
	}
	if(nmax > 2 * SPHR_MAX_BASELINES)
		return -1;
	for(int i = 0; i < nmax; i++)
		a.fs[i] = fs_dev[i];
	hipLaunchKernelGGL(k_xspec, dim3((n + NT - 1) / NT, nb), dim3(NT), 0, st, a, n, nb, fsp);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* partial[split][c] = sum_{i in split} coeff[i][c] vec[i] */
__global__ void __launch_bounds__(NT) k_dotc_partial(const cplx<double> *__restrict__ coeff, int n, int stride, const cplx<double> *__restrict__ vec, const double *__restrict__ vec_real, int rows_per_split, cplx<double> *__restrict__ partial)
{
	__shared__ cplx<double> red[4][64];
	const int c = blockIdx.x * 64 + (threadIdx.x & 63);
	const int lane = threadIdx.x >> 6;
	const int i0 = blockIdx.y * rows_per_split;
	const int i1 = min(n, i0 + rows_per_split);
	cplx<double> acc = mk(0.0, 0.0);
	if(c < stride)
		for(int i = i0 + lane; i < i1; i += 4) {
			const cplx<double> v = vec_real ? mk(vec_real[i], 0.0) : vec[i];
			cfma(acc, coeff[(size_t) i * stride + c], v);
		}
	red[lane][threadIdx.x & 63] = acc;
	__syncthreads();
	if(lane == 0 && c < stride) {
		cplx<double> t = red[0][threadIdx.x];
		t = t + red[1][threadIdx.x];
		t = t + red[2][threadIdx.x];
		t = t + red[3][threadIdx.x];
		partial[(size_t) blockIdx.y * stride + c] = t;
	}
}

__global__ void k_dotc_reduce(const cplx<double> *__restrict__ partial, int nsplit, int stride, double scale, cplx<double> *__restrict__ out)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if(c >= stride)
		return;
	cplx<double> t = mk(0.0, 0.0);
	for(int sp = 0; sp < nsplit; sp++)
		t = t + partial[(size_t) sp * stride + c];
	out[c] = t * scale;
}

int launch_dotc(const cplx<double> *coeff, int n, int stride, const cplx<double> *vec, const double *vec_real, double scale, cplx<double> *partial, int nsplit, cplx<double> *out, hipStream_t st)
{
	const int rps = (n + nsplit - 1) / nsplit;
	hipLaunchKernelGGL(k_dotc_partial, dim3((stride + 63) / 64, nsplit), dim3(NT), 0, st, coeff, n, stride, vec, vec_real, rps, partial);
	hipLaunchKernelGGL(k_dotc_reduce, dim3((stride + NT - 1) / NT), dim3(NT), 0, st, partial, nsplit, stride, scale, out);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

struct rotsum_args {
	const cplx<double> *power[SPHR_MAX_BASELINES];
	const cplx<double> *blocks[SPHR_MAX_BASELINES];
	int Lb[SPHR_MAX_BASELINES], polar[SPHR_MAX_BASELINES];
};

__device__ inline size_t block_offset_dev(int l)
{
	return ((size_t) 4 * l * l * l - l - 3) / 3;
}

/* sky[l, m'] = scale sum_b sum_m D_b^l_{m m'} power_b[l, m]
 * (sh_series_rotation.c:546-577 then the network sum); for a polar power
 * blocks[b] holds only the rows D^l_{0 m'} in the sky coefficient order */
__global__ void k_rotsum(struct rotsum_args a, int nb, int L, double scale, cplx<double> *__restrict__ sky)
{
	const int idx = blockIdx.x * blockDim.x + threadIdx.x;
	if(idx >= (L + 1) * (L + 1))
		return;
	int l, mp;
	lm_decode(L, idx, &l, &mp);
	cplx<double> acc = mk(0.0, 0.0);
	for(int b = 0; b < nb; b++) {
		const int Lb = a.Lb[b];
		if(l > Lb)
			continue;
		if(a.polar[b]) {
			/* row table in the layout of order Lb */
			cfma(acc, a.blocks[b][lm_offset(Lb, l, mp)], a.power[b][l]);
		} else if(l == 0) {
			acc = acc + a.power[b][0];
		} else {
			const cplx<double> *D = a.blocks[b] + block_offset_dev(l) + (size_t) (2 * l + 1) * (mp + l);
			cplx<double> t = mk(0.0, 0.0);
			for(int m = -l; m <= l; m++)
				cfma(t, D[m + l], a.power[b][lm_offset(Lb, l, m)]);
			acc = acc + t;
		}
	}
	sky[idx] = acc * scale;
}

int launch_rotsum(int nb, const cplx<double> *const *power, const int *Lb, const int *polar, const cplx<double> *const *blocks, int L, double scale, cplx<double> *sky, hipStream_t st)
{
	if(nb > SPHR_MAX_BASELINES)
		return -1;
	struct rotsum_args a;
	for(int b = 0; b < nb; b++) {
		a.power[b] = power[b];
		a.blocks[b] = blocks[b];
		a.Lb[b] = Lb[b];
		a.polar[b] = polar[b];
	}
	const int nc = (L + 1) * (L + 1);
	hipLaunchKernelGGL(k_rotsum, dim3((nc + NT - 1) / NT), dim3(NT), 0, st, a, nb, L, scale, sky);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* sh_series_rotation.c:600-625: a_lm exp(-i m phi) */
__global__ void k_rotate_z(int L, const cplx<double> *__restrict__ in, double phi, cplx<double> *__restrict__ out)
{
	const int idx = blockIdx.x * blockDim.x + threadIdx.x;
	if(idx >= (L + 1) * (L + 1))
		return;
	int l, m;
	lm_decode(L, idx, &l, &m);
	double sn, cs;
	sincos(-(double) m * phi, &sn, &cs);
	out[idx] = in[idx] * mk(cs, sn);
}

int launch_rotate_z(int L, const cplx<double> *in, double phi, cplx<double> *out, hipStream_t st)
{
	const int nc = (L + 1) * (L + 1);
	hipLaunchKernelGGL(k_rotate_z, dim3((nc + NT - 1) / NT), dim3(NT), 0, st, L, in, phi, out);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}
