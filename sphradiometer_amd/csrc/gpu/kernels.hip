/*
 * HIP kernels of the cross-power path for gfx950 (64-wide wavefronts).
 *
 *   rfft_pairs   batched real FFTs: two real series packed as one complex
 *                series, in-place radix-4 decimation in time in LDS from a
 *                bit-reversed load, then split into two half spectra
 *   xproj        cross spectrum X = F_b conj(F_a) formed in LDS and
 *                contracted with the packed delay-product matrix: a real
 *                GEMM over K = N + 2 on MFMA 16x16x4 (f32 / f64), split
 *                over K into partial sums
 *   alm          partial sums -> per-baseline polar power -> rotation into
 *                the Earth frame (the D^l_{0 m'} row), 1/nb average and
 *                rotate_z, written in the reference coefficient layout
 *   synth        Legendre stage per (ring pair, m) with the l recurrence
 *                run in registers (equatorial symmetry gives both rings),
 *                then the phi DFT, two columns (k, k + n/2) per thread
 * plus the FP64 helpers of the synchronous drop-in entry points.
 */
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "runtime.h"

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

/*
 * The same transform for N = 1024 NQ (a power of 4): radix-4 Stockham
 * (autosort, natural order in and out) with 256 threads, NQ butterflies
 * per thread per stage.  Stage 0 reads the series straight from global
 * memory (all loads of a thread issued together); each later stage reads
 * the previous one's output from an LDS ping-pong buffer, so one barrier
 * per stage; the twiddles of all stages are loaded into registers first.
 * Stage with sub-transform size p: butterfly q, k = q mod p, inputs
 * u_r = a[q + r N/4] times exp(-2 pi i r k / 4p), radix-4 DFT, outputs to
 * (q / p) 4p + k + r p.
 */
template <typename T, typename TI, int NQ>
__global__ void __launch_bounds__(256) k_rfft4(const TI *__restrict__ in, const double *__restrict__ window, int nseries, const cplx<T> *__restrict__ tw, cplx<T> *__restrict__ spec)
{
	constexpr int N = 1024 * NQ, N4 = N / 4;
	constexpr int NS = NQ == 1 ? 5 : NQ == 4 ? 6 : 0;	/* stages */
	static_assert(NS > 0, "N must be 1024 or 4096");
	__shared__ cplx<T> buf[2][N];
	const int tid = threadIdx.x;
	const int npairs = (nseries + 1) / 2;
	/* twiddles of every stage: w_r = tw[r k N / 4p], k = q mod p */
	cplx<T> wt[NS][NQ][3];
#pragma unroll
	for(int st = 1; st < NS; st++) {
		const int p = 1 << (2 * st);
#pragma unroll
		for(int u = 0; u < NQ; u++) {
			const int k = (tid + 256 * u) & (p - 1);
#pragma unroll
			for(int r = 1; r < 4; r++)
				wt[st][u][r - 1] = tw[r * k * (N / (4 * p))];
		}
	}
	double wv[NQ][4];
#pragma unroll
	for(int u = 0; u < NQ; u++)
#pragma unroll
		for(int r = 0; r < 4; r++)
			wv[u][r] = window ? window[tid + 256 * u + r * N4] : 1.0;
	/* samples q + r N/4 of pair pp into registers */
	TI xv[NQ][4], yv[NQ][4];
	auto load = [&](int pp) {
		const int sx = 2 * pp, sy = sx + 1 < nseries ? sx + 1 : sx;
		const TI *x = in + (size_t) sx * N, *y = in + (size_t) sy * N;
#pragma unroll
		for(int u = 0; u < NQ; u++)
#pragma unroll
			for(int r = 0; r < 4; r++) {
				const int i = tid + 256 * u + r * N4;
				xv[u][r] = x[i];
				yv[u][r] = y[i];
			}
	};
	/* persistent over pairs; the next pair's samples are loaded while
	 * this pair's stages run */
	if(blockIdx.x < npairs)
		load(blockIdx.x);
	for(int pp = blockIdx.x; pp < npairs; pp += gridDim.x) {
		const int sx = 2 * pp, sy = sx + 1;
		const bool has_y = sy < nseries;
		cplx<T> a[NQ][4];
#pragma unroll
		for(int u = 0; u < NQ; u++)
#pragma unroll
			for(int r = 0; r < 4; r++) {
				const T w = (T) wv[u][r];
				a[u][r] = mk((T) xv[u][r] * w, has_y ? (T) yv[u][r] * w : (T) 0);
			}
		if(pp + (int) gridDim.x < npairs)
			load(pp + gridDim.x);
#pragma unroll
		for(int st = 0; st < NS; st++) {
			const int p = 1 << (2 * st);
			cplx<T> *o = buf[st & 1];
			if(st > 0) {
				const cplx<T> *iv = buf[(st - 1) & 1];
#pragma unroll
				for(int u = 0; u < NQ; u++)
#pragma unroll
					for(int r = 0; r < 4; r++)
						a[u][r] = iv[tid + 256 * u + r * N4];
			}
#pragma unroll
			for(int u = 0; u < NQ; u++) {
				const int q = tid + 256 * u, k = q & (p - 1);
				cplx<T> u0 = a[u][0], u1 = a[u][1], u2 = a[u][2], u3 = a[u][3];
				if(st > 0) {
					u1 = u1 * wt[st][u][0];
					u2 = u2 * wt[st][u][1];
					u3 = u3 * wt[st][u][2];
				}
				const cplx<T> s02 = u0 + u2, d02 = u0 - u2, s13 = u1 + u3, d13 = u1 - u3;
				const cplx<T> jd13 = mk(d13.im, -d13.re);	/* -i (u1 - u3) */
				const int base = (q - k) * 4 + k;
				o[base] = s02 + s13;
				o[base + p] = d02 + jd13;
				o[base + 2 * p] = s02 - s13;
				o[base + 3 * p] = d02 - jd13;
			}
			__syncthreads();
		}
		/* split the packed pair: X = (Z[f] + conj Z[N-f]) / 2,
		 * Y = (Z[f] - conj Z[N-f]) / 2i */
		const cplx<T> *z = buf[(NS - 1) & 1];
		constexpr int Nh = N / 2 + 1;
		cplx<T> *ox = spec + (size_t) sx * Nh;
		cplx<T> *oy = spec + (size_t) sy * Nh;
		for(int f = tid; f < Nh; f += 256) {
			const cplx<T> zf = z[f & (N - 1)], zc = conjg(z[(N - f) & (N - 1)]);
			const cplx<T> sum = zf + zc, dif = zf - zc;
			ox[f] = mk(sum.re * (T) 0.5, sum.im * (T) 0.5);
			if(has_y)
				oy[f] = mk(dif.im * (T) 0.5, -dif.re * (T) 0.5);
		}
		/* the last stage's buffer is read above; the next pair's first
		 * stage writes the other one, its second this one: a barrier
		 * before that second write */
		__syncthreads();
	}
}

/* persistent: up to 4 workgroups per CU, each looping over pairs with the
 * next pair's samples loaded during this pair's stages.  Measured at 1536
 * pairs (three sites, 1024 segments, float32): 1024 workgroups 10.4 us,
 * 768 with two pairs each 10.8 us, 1536 with one pair each (the kernel's
 * full occupancy) 11.2 us */
template <typename T, typename TI, int NQ>
static void launch_rfft4_at(const TI *in, const double *window, int nseries, const cplx<T> *tw, cplx<T> *spec, hipStream_t st)
{
	const int npairs = (nseries + 1) / 2, cap = 4 * sphr_rt_cus();
	hipLaunchKernelGGL((k_rfft4<T, TI, NQ>), dim3(npairs < cap ? npairs : cap), dim3(256), 0, st, in, window, nseries, tw, spec);
}

template <typename T>
int launch_rfft_pairs(const void *in, int in_f32, const double *window, int nseries, int N, const cplx<T> *tw, cplx<T> *spec, hipStream_t st)
{
	const size_t lds = (size_t) N * sizeof(cplx<T>);
	dim3 grid((nseries + 1) / 2);
	if(N == 1024 || (N == 4096 && 2 * N * sizeof(cplx<T>) <= 160 * 1024)) {
		if(N == 1024) {
			if(in_f32)
				launch_rfft4_at<T, float, 1>((const float *) in, window, nseries, tw, spec, st);
			else
				launch_rfft4_at<T, double, 1>((const double *) in, window, nseries, tw, spec, st);
		} else {
			if(in_f32)
				launch_rfft4_at<T, float, 4>((const float *) in, window, nseries, tw, spec, st);
			else
				launch_rfft4_at<T, double, 4>((const double *) in, window, nseries, tw, spec, st);
		}
		return hipGetLastError() == hipSuccess ? 0 : -1;
	}
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

/*
 * Real GEMM partial[z][s][c] = sum_{k in split z} A_b(c)[s][k] W[k][c] on
 * MFMA 16x16x4 (f32 or f64), A_b[s][2f] = Re X_b[s][f],
 * A_b[s][2f+1] = Im X_b[s][f], X_b = F_ib conj(F_ia).  Each baseline's
 * columns are padded to whole 16-column tiles so a tile reads one
 * baseline's A.  Workgroup: 16 segments x up to 16 tiles (4 per wave),
 * one K split; per K chunk the A strip (all baselines) and the W slab are
 * staged in LDS.
 *
 * MFMA 16x16x4 operand layout (64 lanes, measured on gfx950): A[i][k] in
 * lane i + 16 k, B[k][j] in lane j + 16 k; the 4 accumulator elements e of
 * lane j + 16 q hold D[i][j] with i = 4 q + e for f32 and i = q + 4 e for
 * f64.
 */
template <typename T> struct mfma16;
template <> struct mfma16<float> {
	typedef float v4 __attribute__((ext_vector_type(4)));
	static __device__ inline v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
	static __device__ inline int row(int q, int e) { return 4 * q + e; }
	enum { KC = 32 };
};
template <> struct mfma16<double> {
	typedef double v4 __attribute__((ext_vector_type(4)));
	static __device__ inline v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
	static __device__ inline int row(int q, int e) { return q + 4 * e; }
	enum { KC = 16 };
};

#define XP_NW 4		/* waves per workgroup */

/* 16-byte vector of T for the W slab loads */
template <typename T> struct vec16;
template <> struct vec16<float> {
	typedef float4 type;
	enum { N = 4 };
};
template <> struct vec16<double> {
	typedef double2 type;
	enum { N = 2 };
};

template <typename T, int XP_MS, int XP_TPW, int NBM>
__global__ void __launch_bounds__(NT) k_xproj(const cplx<T> *__restrict__ spec, int S, struct xproj_geom g, const T *__restrict__ W, const int *__restrict__ tile_bl, int k_per_split, T *__restrict__ partial)
{
	typedef typename mfma16<T>::v4 v4;
	typedef typename vec16<T>::type V;
	const int KC = mfma16<T>::KC, VN = vec16<T>::N;
	const int CTW = XP_NW * XP_TPW * 16;	/* columns per workgroup */
	/* per chunk: W slab KC x CTW in 16-byte vectors, and the A strip
	 * (all baselines) as complex products, staged through registers so
	 * the next chunk's loads are in flight during this chunk's MFMAs */
	const int WV = KC * CTW / VN / NT;	/* W vectors per thread */
	const int AP = (NBM * XP_MS * (KC / 2) + NT - 1) / NT;	/* NBM >= nb */
	extern __shared__ unsigned char smem_raw[];
	const int KA = KC + 1;					/* padded row: no bank conflicts */
	T *sA = reinterpret_cast<T *>(smem_raw);		/* [nb][MS][KA] */
	T *sW = sA + g.nb * XP_MS * KA;				/* [KC][CTW] */
	const int s0 = blockIdx.x * XP_MS;
	const int t0 = blockIdx.y * XP_NW * XP_TPW;		/* first tile */
	const int ntile = g.CT / 16;
	const int kbeg = blockIdx.z * k_per_split;
	const int kend = min(g.K2, kbeg + k_per_split);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int li = lane & 15, lk = lane >> 4;
	const int na = g.nb * XP_MS * (KC / 2);			/* A products per chunk */
	int tb[XP_TPW];
	for(int u = 0; u < XP_TPW; u++) {
		const int t = t0 + wave * XP_TPW + u;
		tb[u] = t < ntile ? tile_bl[t] : 0;
	}
	constexpr int RT = XP_MS / 16;				/* row tiles */
	v4 acc[RT][XP_TPW];
	for(int rt = 0; rt < RT; rt++)
		for(int u = 0; u < XP_TPW; u++)
			acc[rt][u] = (v4) 0;
	const int c0 = t0 * 16;
	V wv[WV];
	cplx<T> fa[AP], fb[AP];
	/* this thread's A products: spectrum offsets of the two detectors
	 * (frequency 0 of the chunk added per chunk), LDS slot, whether the
	 * segment exists; fixed over the K loop, so the per-lane lookup of the
	 * baseline's detectors happens once */
	size_t oa[AP], ob[AP];
	int sl_ff[AP];
#pragma unroll
	for(int p = 0; p < AP; p++) {
		const int i = p * NT + threadIdx.x;
		oa[p] = ob[p] = 0;
		sl_ff[p] = -1;
		if(i < na) {
			const int b = i / (XP_MS * (KC / 2)), r = i - b * XP_MS * (KC / 2);
			const int sl = r / (KC / 2), ff = r - sl * (KC / 2);
			const int s = s0 + sl;
			if(s < S) {
				oa[p] = ((size_t) g.ia[b] * S + s) * g.Nh + ff;
				ob[p] = ((size_t) g.ib[b] * S + s) * g.Nh + ff;
				sl_ff[p] = ff;
			}
		}
	}
	/* issue the global loads of the chunk at k0 into registers */
	auto load = [&](int k0) {
		const int f0 = k0 / 2;
#pragma unroll
		for(int v = 0; v < WV; v++) {
			const int i = (v * NT + threadIdx.x) * VN;
			const int kk = i / CTW, c = i - kk * CTW;
			const int k = k0 + kk, cg = c0 + c;
			if(k < kend && cg < g.CT)
				wv[v] = *reinterpret_cast<const V *>(W + (size_t) k * g.CT + cg);
			else
				wv[v] = V{};
		}
#pragma unroll
		for(int p = 0; p < AP; p++) {
			fa[p] = mk((T) 0, (T) 0);
			fb[p] = fa[p];
			if(sl_ff[p] >= 0 && 2 * (f0 + sl_ff[p]) < kend) {
				fa[p] = spec[oa[p] + f0];
				fb[p] = spec[ob[p] + f0];
			}
		}
	};
	/* write the staged chunk to LDS */
	auto store = [&]() {
#pragma unroll
		for(int v = 0; v < WV; v++)
			*reinterpret_cast<V *>(sW + (v * NT + threadIdx.x) * VN) = wv[v];
#pragma unroll
		for(int p = 0; p < AP; p++) {
			const int i = p * NT + threadIdx.x;
			if(i < na) {
				const int b = i / (XP_MS * (KC / 2)), r = i - b * XP_MS * (KC / 2);
				const int sl = r / (KC / 2), ff = r - sl * (KC / 2);
				const cplx<T> X = fb[p] * conjg(fa[p]);
				T *a = sA + (b * XP_MS + sl) * KA + 2 * ff;
				a[0] = X.re;
				a[1] = X.im;
			}
		}
	};
	if(kbeg < kend)
		load(kbeg);
	for(int k0 = kbeg; k0 < kend; k0 += KC) {
		store();
		__syncthreads();
		if(k0 + KC < kend)
			load(k0 + KC);
		/* all tiles unguarded: tiles past the last column multiply the
		 * zero-filled W slab (their baseline index is 0, a valid row) */
#pragma unroll
		for(int ks = 0; ks < KC; ks += 4) {
#pragma unroll
			for(int u = 0; u < XP_TPW; u++) {
				const int tl = wave * XP_TPW + u;
				const T b = sW[(ks + lk) * CTW + tl * 16 + li];
#pragma unroll
				for(int rt = 0; rt < RT; rt++) {
					const T a = sA[(tb[u] * XP_MS + 16 * rt + li) * KA + ks + lk];
					acc[rt][u] = mfma16<T>::mma(a, b, acc[rt][u]);
				}
			}
		}
		__syncthreads();
	}
	for(int u = 0; u < XP_TPW; u++) {
		const int tl = wave * XP_TPW + u;
		if(t0 + tl >= ntile)
			break;
		const int c = c0 + tl * 16 + li;
#pragma unroll
		for(int rt = 0; rt < RT; rt++)
			for(int e = 0; e < 4; e++) {
				const int s = s0 + 16 * rt + mfma16<T>::row(lk, e);
				if(s < S)
					partial[((size_t) blockIdx.z * S + s) * g.CT + c] = acc[rt][u][e];
			}
	}
}

/*
 * The projection with one baseline per workgroup: 64 segments (one 16-row
 * tile per wave) x the baseline's columns (up to NTB 16-column tiles) x a
 * K split.  The baseline's detectors and columns are uniform per
 * workgroup, so the A strip is that baseline's X alone and each A
 * fragment feeds all the baseline's column tiles.  Chunks of KC reals are
 * staged global -> registers -> LDS, the next chunk's loads in flight
 * during this chunk's MFMAs.
 */
template <typename T> struct xpb;
template <> struct xpb<float> {
	enum { KC = 64, KA = 68, CW = 132 };	/* row strides = 4 mod 64 words */
};
template <> struct xpb<double> {
	enum { KC = 32, KA = 36, CW = 132 };
};
#define XPB_MS 64

template <typename T, int NTB>
__global__ void __launch_bounds__(NT) k_xproj_bl(const cplx<T> *__restrict__ spec, int S, struct xproj_geom g, const T *__restrict__ W, int k_per_split, T *__restrict__ partial)
{
	typedef typename mfma16<T>::v4 v4;
	typedef typename vec16<T>::type V;
	const int KC = xpb<T>::KC, KA = xpb<T>::KA, CW = xpb<T>::CW, VN = vec16<T>::N;
	const int AP = XPB_MS * (KC / 2) / NT;		/* A products per thread */
	const int WV = (KC * NTB * 16 / VN + NT - 1) / NT;	/* W vectors per thread (at most) */
	__shared__ __align__(16) T sA[XPB_MS * KA];
	__shared__ __align__(16) T sW[KC * CW];
	const int b = blockIdx.y;
	const int s0 = blockIdx.x * XPB_MS;
	const int kbeg = blockIdx.z * k_per_split;
	const int kend = min(g.K2, kbeg + k_per_split);
	const int cb0 = g.col_start[b];
	const int ncol = (b + 1 < g.nb ? g.col_start[b + 1] : g.CT) - cb0;
	const int ncb = ncol / 16;
	const size_t da = (size_t) g.ia[b] * S, db = (size_t) g.ib[b] * S;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int li = lane & 15, lk = lane >> 4;
	v4 acc[NTB];
#pragma unroll
	for(int ct = 0; ct < NTB; ct++)
		acc[ct] = (v4) 0;
	V wv[WV];
	cplx<T> fa[AP], fb[AP];
	auto load = [&](int k0) {
		const int f0 = k0 / 2;
#pragma unroll
		for(int p = 0; p < AP; p++) {
			const int q = p * NT + threadIdx.x;
			const int sl = q / (KC / 2), ff = q - sl * (KC / 2);
			const int s = s0 + sl, f = f0 + ff;
			fa[p] = mk((T) 0, (T) 0);
			fb[p] = fa[p];
			if(s < S && 2 * f < kend) {
				fa[p] = spec[(da + s) * g.Nh + f];
				fb[p] = spec[(db + s) * g.Nh + f];
			}
		}
#pragma unroll
		for(int v = 0; v < WV; v++) {
			const int i = (v * NT + threadIdx.x) * VN;
			const int kk = i / (NTB * 16), c = i - kk * (NTB * 16);
			const int k = k0 + kk;
			if(kk < KC && k < kend && c < ncol)
				wv[v] = *reinterpret_cast<const V *>(W + (size_t) k * g.CT + cb0 + c);
			else
				wv[v] = V{};
		}
	};
	auto store = [&]() {
#pragma unroll
		for(int p = 0; p < AP; p++) {
			const int q = p * NT + threadIdx.x;
			const int sl = q / (KC / 2), ff = q - sl * (KC / 2);
			const cplx<T> X = fb[p] * conjg(fa[p]);
			sA[sl * KA + 2 * ff] = X.re;
			sA[sl * KA + 2 * ff + 1] = X.im;
		}
#pragma unroll
		for(int v = 0; v < WV; v++) {
			const int i = (v * NT + threadIdx.x) * VN;
			const int kk = i / (NTB * 16), c = i - kk * (NTB * 16);
			if(kk < KC)
				*reinterpret_cast<V *>(sW + kk * CW + c) = wv[v];
		}
	};
	if(kbeg < kend)
		load(kbeg);
	for(int k0 = kbeg; k0 < kend; k0 += KC) {
		store();
		__syncthreads();
		if(k0 + KC < kend)
			load(k0 + KC);
		/* all NTB tiles unguarded (tiles past this baseline's columns
		 * multiply the zero-filled W slab), operands of the next k-step
		 * read before this one's MFMAs */
		const T *ar = sA + (16 * wave + li) * KA + lk;
		const T *br = sW + lk * CW + li;
		T an = ar[0], bn[NTB];
#pragma unroll
		for(int ct = 0; ct < NTB; ct++)
			bn[ct] = br[ct * 16];
#pragma unroll
		for(int ks = 0; ks < KC; ks += 4) {
			const T a = an;
			T bb[NTB];
#pragma unroll
			for(int ct = 0; ct < NTB; ct++)
				bb[ct] = bn[ct];
			if(ks + 4 < KC) {
				an = ar[ks + 4];
#pragma unroll
				for(int ct = 0; ct < NTB; ct++)
					bn[ct] = br[(ks + 4) * CW + ct * 16];
			}
#pragma unroll
			for(int ct = 0; ct < NTB; ct++)
				acc[ct] = mfma16<T>::mma(a, bb[ct], acc[ct]);
		}
		__syncthreads();
	}
#pragma unroll
	for(int ct = 0; ct < NTB; ct++) {
		if(ct >= ncb)
			break;
		const int c = cb0 + ct * 16 + li;
		for(int e = 0; e < 4; e++) {
			const int s = s0 + 16 * wave + mfma16<T>::row(lk, e);
			if(s < S)
				partial[((size_t) blockIdx.z * S + s) * g.CT + c] = acc[ct][e];
		}
	}
}

template <typename T, int XP_MS, int XP_TPW>
static void launch_xproj_tpw(const cplx<T> *spec, int S, const struct xproj_geom *g, const T *W, const int *tile_bl, int k_per_split, int nsplit, T *partial, hipStream_t st)
{
	const int KC = mfma16<T>::KC;
	const size_t lds = ((size_t) g->nb * XP_MS * (KC + 1) + (size_t) KC * XP_NW * XP_TPW * 16) * sizeof(T);
	const int ntile = g->CT / 16;
	dim3 grid((S + XP_MS - 1) / XP_MS, (ntile + XP_NW * XP_TPW - 1) / (XP_NW * XP_TPW), nsplit);
	/* the register staging of the A strip is sized for up to 6 baselines
	 * (4 detectors); larger networks take the 16-baseline form */
	if(g->nb <= 6)
		hipLaunchKernelGGL((k_xproj<T, XP_MS, XP_TPW, 6>), grid, dim3(NT), lds, st, spec, S, *g, W, tile_bl, k_per_split, partial);
	else
		hipLaunchKernelGGL((k_xproj<T, XP_MS, XP_TPW, SPHR_MAX_BASELINES>), grid, dim3(NT), lds, st, spec, S, *g, W, tile_bl, k_per_split, partial);
}

/* tpw: 16-column tiles per wave (1, 2 or 4); ms: segments per workgroup
 * (16 or 32); tpw 0: one baseline per workgroup (64 segments, up to 8
 * column tiles per baseline); the shapes the batch's autotuner picks from */
template <typename T>
int launch_xproj(const cplx<T> *spec, int S, const struct xproj_geom *g, const T *W, const int *tile_bl, int k_per_split, int nsplit, T *partial, hipStream_t st, int tpw, int ms)
{
	int maxcol = 0;
	for(int b = 0; b < g->nb; b++) {
		const int nc = (b + 1 < g->nb ? g->col_start[b + 1] : g->CT) - g->col_start[b];
		maxcol = nc > maxcol ? nc : maxcol;
	}
	if(tpw == 0 && maxcol <= 8 * 16) {
		dim3 grid((S + XPB_MS - 1) / XPB_MS, g->nb, nsplit);
		if(maxcol <= 5 * 16)
			hipLaunchKernelGGL((k_xproj_bl<T, 5>), grid, dim3(NT), 0, st, spec, S, *g, W, k_per_split, partial);
		else
			hipLaunchKernelGGL((k_xproj_bl<T, 8>), grid, dim3(NT), 0, st, spec, S, *g, W, k_per_split, partial);
	} else if(ms == 32) {
		if(tpw == 1)
			launch_xproj_tpw<T, 32, 1>(spec, S, g, W, tile_bl, k_per_split, nsplit, partial, st);
		else if(tpw == 2)
			launch_xproj_tpw<T, 32, 2>(spec, S, g, W, tile_bl, k_per_split, nsplit, partial, st);
		else
			launch_xproj_tpw<T, 32, 4>(spec, S, g, W, tile_bl, k_per_split, nsplit, partial, st);
	} else {
		if(tpw == 1)
			launch_xproj_tpw<T, 16, 1>(spec, S, g, W, tile_bl, k_per_split, nsplit, partial, st);
		else if(tpw == 2)
			launch_xproj_tpw<T, 16, 2>(spec, S, g, W, tile_bl, k_per_split, nsplit, partial, st);
		else
			launch_xproj_tpw<T, 16, 4>(spec, S, g, W, tile_bl, k_per_split, nsplit, partial, st);
	}
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_xproj<double>(const cplx<double> *, int, const struct xproj_geom *, const double *, const int *, int, int, double *, hipStream_t, int, int);
template int launch_xproj<float>(const cplx<float> *, int, const struct xproj_geom *, const float *, const int *, int, int, float *, hipStream_t, int, int);

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

/*
 * Synthesis stages.  Legendre stage: one item per (ring pair, m) runs the
 * l recurrence in registers and accumulates the even-(l+|m|) and
 * odd-(l+|m|) sums E, O, giving F(x_j) = E + O and F(-x_j) = E - O.  phi
 * stage: each thread produces columns k and k + n/2 of a ring,
 *   map[k] = sum_m F_m w_k^m,  map[k + n/2] = sum_m (-1)^m F_m w_k^m,
 * w_k = exp(2 pi i k / n), with the powers of w_k formed by recurrence in
 * registers (m = 0 .. L up from 1, m = -L .. -1 up from the tabulated
 * w_k^{-L}), so the twiddle table is read once per thread.
 */
template <typename T>
struct synth_lds {
	const cplx<T> *tw;	/* [n] twiddles exp(+2 pi i t / n) */
	const cplx<T> *sa;	/* series coefficients in LDS, or NULL */
	const T *al;		/* recurrence coefficients in LDS, or NULL */
};

template <typename TI, typename T>
__device__ void legendre_stage(const cplx<TI> *__restrict__ ag, int polar, int rp, int j0, const struct synth_geom &g, const struct synth_tables &tb, const synth_lds<T> &ls, cplx<T> *F)
{
	const int n = g.n, L = g.L, B = g.B;
	const int nm = polar ? 1 : 2 * L + 1;
	for(int it = threadIdx.x; it < rp * nm; it += NT) {
		const int r = it / nm, mi = it - r * nm;
		const int j = j0 + r;
		if(j >= B)
			continue;
		const int m = polar ? 0 : mi - L;
		const int am = m < 0 ? -m : m;
		const T x = (T) tb.x[j];
		const int abase = polar ? 0 : lm_offset(L, am, m);	/* (l = |m|, m) */
		const int albase = polar ? 0 : lm_offset(L, am, am);	/* alpha at (l = |m|, |m|) */
		T pprev = 0, pcur = (T) tb.smm[(size_t) j * (L + 1) + am], aprev = 1;
		cplx<T> E = mk((T) 0, (T) 0), O = mk((T) 0, (T) 0);
		for(int l = am; l <= L; l++) {
			if(l > am) {
				const T a_l = ls.al ? ls.al[albase + l - am] : (T) tb.alpha[albase + l - am];
				const T pn = (l == am + 1) ? a_l * x * pcur : a_l * (x * pcur - pprev / aprev);
				pprev = pcur;
				pcur = pn;
				aprev = a_l;
			}
			cplx<T> c;
			if(ls.sa)
				c = ls.sa[abase + l - am];
			else {
				const cplx<TI> cg = ag[abase + l - am];
				c = mk((T) cg.re, (T) cg.im);
			}
			if((l + am) & 1)
				cfma(O, c, pcur);
			else
				cfma(E, c, pcur);
		}
		const int col = m >= 0 ? m : n + m;
		const T sg = (m < 0 && (am & 1)) ? (T) -1 : (T) 1;
		F[(2 * r) * n + col] = (E + O) * sg;
		F[(2 * r + 1) * n + col] = (E - O) * sg;
	}
}

template <typename T, typename TO>
__device__ void phi_stage(const cplx<T> *F, int polar, int rp, int j0, const struct synth_geom &g, const synth_lds<T> &ls, cplx<TO> *__restrict__ mapseg)
{
	const int n = g.n, B = g.B, half = g.n / 2;
	const int mmax = polar ? 0 : g.L;
	for(int o = threadIdx.x; o < 2 * rp * half; o += NT) {
		const int slot = o / half, k = o - slot * half;
		const int j = j0 + (slot >> 1);
		if(j >= B)
			continue;
		const cplx<T> *Fr = F + slot * n;
		const cplx<T> wk = ls.tw[k];
		cplx<T> a0 = mk((T) 0, (T) 0), a1 = mk((T) 0, (T) 0);
		cplx<T> w = mk((T) 1, (T) 0);
		for(int m = 0; m <= mmax; m++) {
			const cplx<T> e = Fr[m] * w;
			a0 = a0 + e;
			a1 = (m & 1) ? a1 - e : a1 + e;
			w = w * wk;
		}
		if(mmax > 0) {
			w = ls.tw[(int) ((n - (long) mmax * k % n) % n)];
			for(int m = -mmax; m < 0; m++) {
				const cplx<T> e = Fr[n + m] * w;
				a0 = a0 + e;
				a1 = (m & 1) ? a1 - e : a1 + e;
				w = w * wk;
			}
		}
		const int ring = (slot & 1) ? 2 * B - 1 - j : j;
		cplx<TO> *row = mapseg + (size_t) ring * n;
		row[k] = mk((TO) a0.re, (TO) a0.im);
		row[k + half] = mk((TO) a1.re, (TO) a1.im);
	}
}

/* synthesis of a batch of series: one workgroup per (series, group of
 * rp ring pairs); coefficients and recurrence table in LDS when they fit */
template <typename TI, typename T, typename TO>
__device__ __forceinline__ void synth_block(int s, int bx, const cplx<TI> *__restrict__ alm, int polar, int rp, int a_in_lds, int al_in_lds, const struct synth_geom &g, const struct synth_tables &tb, cplx<TO> *__restrict__ map, const long *__restrict__ aoff, const long *__restrict__ moff)
{
	extern __shared__ unsigned char smem_raw[];
	const int n = g.n, L = g.L;
	const int ncoef = polar ? L + 1 : (L + 1) * (L + 1);
	const int nal = polar ? L + 1 : (L + 1) * (L + 2) / 2;
	cplx<T> *F = reinterpret_cast<cplx<T> *>(smem_raw);	/* [2 rp][n] */
	cplx<T> *tw = F + 2 * rp * n;				/* [n] */
	cplx<T> *sa = tw + n;					/* [ncoef] when staged */
	T *al = reinterpret_cast<T *>(sa + (a_in_lds ? ncoef : 0));	/* [nal] when staged */
	const int j0 = bx * rp;
	/* series s at aoff[s] / moff[s] when given (ragged batches), else
	 * packed at uniform strides */
	const cplx<TI> *ag = alm + (aoff ? (size_t) aoff[s] : (size_t) s * ncoef);
	for(int t = threadIdx.x; t < n; t += NT)
		tw[t] = mk((T) tb.twp[2 * t], (T) tb.twp[2 * t + 1]);
	if(a_in_lds)
		for(int t = threadIdx.x; t < ncoef; t += NT) {
			const cplx<TI> c = ag[t];
			sa[t] = mk((T) c.re, (T) c.im);
		}
	if(al_in_lds)
		for(int t = threadIdx.x; t < nal; t += NT)
			al[t] = (T) tb.alpha[t];
	__syncthreads();
	synth_lds<T> ls;
	ls.tw = tw;
	ls.sa = a_in_lds ? sa : nullptr;
	ls.al = al_in_lds ? al : nullptr;
	legendre_stage<TI, T>(ag, polar, rp, j0, g, tb, ls, F);
	__syncthreads();
	phi_stage<T, TO>(F, polar, rp, j0, g, ls, map + (moff ? (size_t) moff[s] : (size_t) s * n * n));
}

template <typename TI, typename T, typename TO>
__global__ void __launch_bounds__(NT) k_synth(const cplx<TI> *__restrict__ alm, int polar, int rp, int a_in_lds, int al_in_lds, struct synth_geom g, struct synth_tables tb, cplx<TO> *__restrict__ map, const long *__restrict__ aoff, const long *__restrict__ moff)
{
	synth_block<TI, T, TO>(blockIdx.y, blockIdx.x, alm, polar, rp, a_in_lds, al_in_lds, g, tb, map, aoff, moff);
}

/* synthesis of ragged series of several orders in one launch: workgroup b
 * belongs to the class c with cls[c].b0 <= b < cls[c + 1].b0 (classes
 * sorted by b0), inside which it is series (b - b0) / gx, ring-pair
 * group (b - b0) % gx */
template <typename T>
__global__ void __launch_bounds__(NT) k_synth_cls(const cplx<T> *__restrict__ alm, const struct synth_cls *__restrict__ cls, int ncls, cplx<T> *__restrict__ map)
{
	const long b = blockIdx.x;
	int lo = 0, hi = ncls - 1;
	while(lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if(cls[mid].b0 <= b)
			lo = mid;
		else
			hi = mid - 1;
	}
	const struct synth_cls &d = cls[lo];
	const long loc = b - d.b0;
	synth_block<T, T, T>((int) (loc / d.gx), (int) (loc % d.gx), alm, 0, d.rp, d.a_in_lds, d.al_in_lds, d.g, d.tb, map, d.aoff, d.moff);
}

/*
 * Fused sky stage of the batched pipeline, one workgroup per segment:
 * sum the projection's K-split partials, rotate each baseline's polar
 * power into the Earth frame with the D^l_{0 m'} rows (pre-scaled by
 * 1/nb), apply rotate_z(gmst), write a_lm, then synthesise the mesh ring
 * group by ring group from the coefficients held in LDS.
 */
template <typename T>
__global__ void __launch_bounds__(NT) k_sky(const T *__restrict__ partial, int nsplit, int S, struct xproj_geom xg, const cplx<T> *__restrict__ R, const int2 *__restrict__ lm, const double *__restrict__ gmst, struct synth_geom g, struct synth_tables tb, int rp, cplx<T> *__restrict__ alm, cplx<T> *__restrict__ map)
{
	extern __shared__ unsigned char smem_raw[];
	const int n = g.n, L = g.L, ncoef = (L + 1) * (L + 1), nal = (L + 1) * (L + 2) / 2;
	cplx<T> *F = reinterpret_cast<cplx<T> *>(smem_raw);	/* [2 rp][n] */
	cplx<T> *tw = F + 2 * rp * n;				/* [n] */
	cplx<T> *sa = tw + n;					/* [ncoef] */
	T *al = reinterpret_cast<T *>(sa + ncoef);		/* [nal] */
	T *pw = al + nal;					/* [CT] */
	const int s = blockIdx.x;
	for(int c = threadIdx.x; c < xg.CT; c += NT) {
		T v = 0;
		for(int sp = 0; sp < nsplit; sp++)
			v += partial[((size_t) sp * S + s) * xg.CT + c];
		pw[c] = v;
	}
	for(int t = threadIdx.x; t < n; t += NT)
		tw[t] = mk((T) tb.twp[2 * t], (T) tb.twp[2 * t + 1]);
	for(int t = threadIdx.x; t < nal; t += NT)
		al[t] = (T) tb.alpha[t];
	__syncthreads();
	double gs = 0.0;
	if(gmst)
		gs = gmst[s];
	for(int idx = threadIdx.x; idx < ncoef; idx += NT) {
		const int2 q = lm[idx];
		const int l = q.x, m = q.y;
		cplx<T> acc = mk((T) 0, (T) 0);
		for(int b = 0; b < xg.nb; b++) {
			if(l > xg.Lb[b])
				continue;
			const T *p = pw + xg.col_start[b] + 2 * l;
			cfma(acc, R[(size_t) b * ncoef + idx], mk(p[0], p[1]));
		}
		if(gmst && m) {
			double sn, cs;
			sincos(-(double) m * gs, &sn, &cs);
			acc = acc * mk((T) cs, (T) sn);
		}
		sa[idx] = acc;
		alm[(size_t) s * ncoef + idx] = acc;
	}
	__syncthreads();
	if(!map)
		return;
	synth_lds<T> ls;
	ls.tw = tw;
	ls.sa = sa;
	ls.al = al;
	for(int j0 = 0; j0 < g.B; j0 += rp) {
		legendre_stage<T, T>(sa, 0, rp, j0, g, tb, ls, F);
		__syncthreads();
		phi_stage<T, T>(F, 0, rp, j0, g, ls, map + (size_t) s * n * n);
		__syncthreads();
	}
}

/*
 * Legendre ring sums of one series into F [n rings][2L+1] (ring rows in
 * map order, columns m = 0..L, -L..-1), a workgroup of NT threads:
 *   F[r][m] = sum_l a_lm Pbar_l|m|(x_r), times (-1)^m for odd m < 0,
 * for the ring pairs (x_j, -x_j) at once through the parity (-1)^(l+|m|).
 * Item (group of RG rings, a) runs the recurrence over l for |m| = a and
 * for |m| = L - a (L + 2 steps for every item), each step serving both
 * signs of m and all RG rings.  The recurrence
 *   Pbar_l = alpha_l x Pbar_{l-1} - (alpha_l / alpha_{l-1}) Pbar_{l-2}
 * reads the pairs (alpha_l, alpha_l / alpha_{l-1}) from cf (LDS, the
 * (l, m >= 0) slots of the coefficient layout, two zero pads).  With
 * rmax, each ring's largest real or imaginary part (as float bits) is
 * folded in by atomicMax.
 */
#define RS_RG 3
template <typename T>
__device__ inline void ring_sums(const cplx<T> *sa, const cplx<T> *cf, int L, int B, const double *__restrict__ xg, const double *__restrict__ smm, cplx<T> *__restrict__ F, unsigned *rmax)
{
	const int nm = 2 * L + 1, NA = L / 2 + 1, NG = (B + RS_RG - 1) / RS_RG;
	for(int it = threadIdx.x; it < NG * NA; it += NT) {
		const int tg = it % NG, a = it / NG;
		T x[RS_RG];
#pragma unroll
		for(int r = 0; r < RS_RG; r++) {
			const int j = RS_RG * tg + r;
			x[r] = j < B ? (T) xg[j] : (T) 0;
		}
		for(int h = 0; h < 2; h++) {
			const int am = h ? L - a : a;
			if(h && am == a)
				break;
			const int bp = lm_offset(L, am, am), bn = am ? lm_offset(L, am, -am) : bp;
			const cplx<T> *sp = sa + bp, *sn = sa + bn, *cc = cf + bp;
			const int cnt = L - am;
			T p2[RS_RG], p1[RS_RG];
			cplx<T> Ep[RS_RG], Op[RS_RG], En[RS_RG], On[RS_RG];
			{
				const cplx<T> ap0 = sp[0], an0 = sn[0], ap1 = sp[1], an1 = sn[1];
				const T c1 = cc[1].re;
#pragma unroll
				for(int r = 0; r < RS_RG; r++) {
					const int j = RS_RG * tg + r;
					p2[r] = j < B ? (T) smm[(size_t) j * (L + 1) + am] : (T) 0;
					Ep[r] = ap0 * p2[r];
					En[r] = an0 * p2[r];
					/* l = am + 1 (unused past the end: cnt = 0) */
					p1[r] = cnt >= 1 ? c1 * x[r] * p2[r] : (T) 0;
					Op[r] = cnt >= 1 ? ap1 * p1[r] : mk((T) 0, (T) 0);
					On[r] = cnt >= 1 ? an1 * p1[r] : mk((T) 0, (T) 0);
				}
			}
			/* two steps (even, odd l - am) per iteration, then one */
			int i = 2;
			for(; i + 1 <= cnt; i += 2) {
				const cplx<T> c0 = cc[i], c1 = cc[i + 1];
				const cplx<T> ap0 = sp[i], ap1 = sp[i + 1], an0 = sn[i], an1 = sn[i + 1];
#pragma unroll
				for(int r = 0; r < RS_RG; r++) {
					const T pe = c0.re * x[r] * p1[r] - c0.im * p2[r];
					const T po = c1.re * x[r] * pe - c1.im * p1[r];
					cfma(Ep[r], ap0, pe);
					cfma(En[r], an0, pe);
					cfma(Op[r], ap1, po);
					cfma(On[r], an1, po);
					p2[r] = pe;
					p1[r] = po;
				}
			}
			if(i == cnt) {
				const cplx<T> c0 = cc[i], ap0 = sp[i], an0 = sn[i];
#pragma unroll
				for(int r = 0; r < RS_RG; r++) {
					const T pe = c0.re * x[r] * p1[r] - c0.im * p2[r];
					cfma(Ep[r], ap0, pe);
					cfma(En[r], an0, pe);
				}
			}
			const T sg = (am & 1) ? (T) -1 : (T) 1;
			/* ring row, column, value */
			auto put = [&](int row, int col, cplx<T> v) {
				F[row * nm + col] = v;
				if(rmax)
					atomicMax(&rmax[row], __float_as_uint((float) fmax(fabs(v.re), fabs(v.im))));
			};
#pragma unroll
			for(int r = 0; r < RS_RG; r++) {
				const int j = RS_RG * tg + r;
				if(j >= B)
					continue;
				put(j, am, Ep[r] + Op[r]);
				put(2 * B - 1 - j, am, Ep[r] - Op[r]);
				if(am) {
					put(j, nm - am, (En[r] + On[r]) * sg);
					put(2 * B - 1 - j, nm - am, (En[r] - On[r]) * sg);
				}
			}
		}
	}
}

/* the recurrence pairs (alpha_l, alpha_l / alpha_{l-1}) of ring_sums at
 * the (l, m >= 0) slots, two zero pads, from the order's table into LDS */
template <typename T> struct ring_table;
template <> struct ring_table<double> {
	static __device__ inline const cplx<double> *get(const struct synth_tables &tb) { return tb.rc64; }
};
template <> struct ring_table<float> {
	static __device__ inline const cplx<float> *get(const struct synth_tables &tb) { return tb.rc32; }
};
template <typename T>
__device__ inline void ring_coefs(int L, const struct synth_tables &tb, cplx<T> *cf)
{
	const int n = (L + 1) * (L + 2) / 2 + 2;
	const cplx<T> *__restrict__ rc = ring_table<T>::get(tb);
	for(int t = threadIdx.x; t < n; t += NT)
		cf[t] = rc[t];
}

/*
 * Sky stage in two kernels (batched path).  k_sky_lf: one workgroup per
 * segment sums the K-split partials, forms a_lm (rotation rows, baseline
 * average, rotate_z) in LDS, writes a_lm, and runs the Legendre stage for
 * all ring pairs into F [S][n rings][2L+1] in HBM (ring rows in map order,
 * columns m = 0..L, -L..-1).  The phi DFT is then one GEMM on MFMA,
 * map [S n][n] = F [S n][2L+1] E [2L+1][n] (complex, real-expanded E).
 */
template <typename T>
__global__ void __launch_bounds__(NT) k_sky_lf(const T *__restrict__ partial, int nsplit, int S, struct xproj_geom xg, const cplx<T> *__restrict__ R, const int2 *__restrict__ lm, const double *__restrict__ gmst, struct synth_geom g, struct synth_tables tb, cplx<T> *__restrict__ alm, cplx<T> *__restrict__ Fg)
{
	extern __shared__ unsigned char smem_raw[];
	const int L = g.L, B = g.B, n = g.n, nm = 2 * L + 1;
	const int ncoef = (L + 1) * (L + 1), nal = (L + 1) * (L + 2) / 2;
	cplx<T> *sa = reinterpret_cast<cplx<T> *>(smem_raw);	/* [ncoef] */
	cplx<T> *cf = sa + ncoef;				/* [nal + 2] */
	T *pw = reinterpret_cast<T *>(cf + nal + 2);		/* [CT] */
	const int s = blockIdx.x;
	for(int c = threadIdx.x; c < xg.CT; c += NT) {
		T v = 0;
		for(int sp = 0; sp < nsplit; sp++)
			v += partial[((size_t) sp * S + s) * xg.CT + c];
		pw[c] = v;
	}
	ring_coefs<T>(L, tb, cf);
	__syncthreads();
	double gs = 0.0;
	if(gmst)
		gs = gmst[s];
	for(int idx = threadIdx.x; idx < ncoef; idx += NT) {
		const int2 q = lm[idx];
		const int l = q.x, m = q.y;
		cplx<T> acc = mk((T) 0, (T) 0);
		for(int b = 0; b < xg.nb; b++) {
			if(l > xg.Lb[b])
				continue;
			const T *p = pw + xg.col_start[b] + 2 * l;
			cfma(acc, R[(size_t) b * ncoef + idx], mk(p[0], p[1]));
		}
		if(gmst && m) {
			double sn, cs;
			sincos(-(double) m * gs, &sn, &cs);
			acc = acc * mk((T) cs, (T) sn);
		}
		sa[idx] = acc;
		alm[(size_t) s * ncoef + idx] = acc;
	}
	if(!Fg)
		return;
	__syncthreads();
	ring_sums<T>(sa, cf, L, B, tb.x, tb.smm, Fg + (size_t) s * n * nm, NULL);
}

/* Legendre ring sums of given series (one workgroup per series), the
 * first half of the GEMM-form synthesis; with irs, also each ring's
 * inverse split scale (the power of two 2^-e that puts the ring's largest
 * real or imaginary part at 2^13..2^14 when multiplied by 2^e), the row
 * scales of the split-FP16 phi DFT */
template <typename T>
__global__ void __launch_bounds__(NT) k_alm_lf(const cplx<T> *__restrict__ alm, struct synth_geom g, struct synth_tables tb, cplx<T> *__restrict__ Fg, float *__restrict__ irs, struct sky_epi<T> epi)
{
	extern __shared__ unsigned char smem_raw[];
	__shared__ unsigned rmax[2 * SPHR_MAX_MESH_B];
	__shared__ cplx<T> rot[SPHR_MAX_MESH_B];
	const int L = g.L, B = g.B, n = g.n, nm = 2 * L + 1;
	const int ncoef = (L + 1) * (L + 1), nal = (L + 1) * (L + 2) / 2;
	cplx<T> *sa = reinterpret_cast<cplx<T> *>(smem_raw);
	cplx<T> *cf = sa + ncoef;	/* [nal + 2] */
	const int s = blockIdx.x;
	if(epi.S) {
		/* series s = beta E + e: the product's (beta, e) row plus the
		 * prior, rotated by exp(-i m gmst_e); written out, kept in LDS */
		const int beta = s / epi.E, e = s - beta * epi.E;
		const T *src = epi.S + (size_t) e * 4 * ncoef + (size_t) beta * 2 * ncoef;
		const double ge = epi.gmst ? epi.gmst[e] : 0.0;
		for(int m = threadIdx.x; m <= L; m += NT) {
			double sn = 0.0, cs = 1.0;
			if(m)
				sincos(-(double) m * ge, &sn, &cs);
			rot[m] = mk((T) cs, (T) sn);
		}
		__syncthreads();
		for(int t = threadIdx.x; t < ncoef; t += NT) {
			const int m = lm_order(L, t);
			const cplx<T> v = (mk(src[2 * t], src[2 * t + 1]) + epi.prior[t]) * (m >= 0 ? rot[m] : conjg(rot[-m]));
			sa[t] = v;
			epi.alm[(size_t) s * ncoef + t] = v;
		}
	} else
		for(int t = threadIdx.x; t < ncoef; t += NT)
			sa[t] = alm[(size_t) s * ncoef + t];
	ring_coefs<T>(L, tb, cf);
	if(irs)
		for(int t = threadIdx.x; t < n; t += NT)
			rmax[t] = 0;
	__syncthreads();
	ring_sums<T>(sa, cf, L, B, tb.x, tb.smm, Fg + (size_t) s * n * nm, irs ? rmax : NULL);
	if(!irs)
		return;
	__syncthreads();
	for(int r = threadIdx.x; r < n; r += NT) {
		const float mx = __uint_as_float(rmax[r]);
		int ex = 0;
		if(mx > 0.f) {
			frexpf(mx, &ex);
			ex = 14 - ex;
		}
		irs[(size_t) s * n + r] = ldexpf(1.0f, -ex);
	}
}

template <typename T> struct phi_table;
template <> struct phi_table<double> {
	static const double *get(const struct synth_tables *tb) { return tb->E64; }
};
template <> struct phi_table<float> {
	static const float *get(const struct synth_tables *tb) { return tb->E32; }
};

template <typename T>
int launch_sky_lf(const T *partial, int nsplit, int S, const struct xproj_geom *xg, const cplx<T> *R, const int2 *lm, const double *gmst, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *alm, cplx<T> *F, hipStream_t st)
{
	const int L = g->L;
	const size_t lds = ((size_t) (L + 1) * (L + 1) + (size_t) (L + 1) * (L + 2) / 2 + 2) * sizeof(cplx<T>) + (size_t) xg->CT * sizeof(T);
	if(lds > 65536)
		return -1;
	hipLaunchKernelGGL((k_sky_lf<T>), dim3(S), dim3(NT), lds, st, partial, nsplit, S, *xg, R, lm, gmst, *g, *tb, alm, F);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T>
int launch_phi(const cplx<T> *F, int S, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *map, hipStream_t st)
{
	const int n = g->n, nm = 2 * g->L + 1;
	return launch_gemm<T, T>(S * n, 2 * n, 2 * nm, (const T *) F, 2L * nm, phi_table<T>::get(tb), 2L * n, (T *) map, 2L * n, st);
}

template int launch_sky_lf<double>(const double *, int, int, const struct xproj_geom *, const cplx<double> *, const int2 *, const double *, const struct synth_geom *, const struct synth_tables *, cplx<double> *, cplx<double> *, hipStream_t);
template int launch_sky_lf<float>(const float *, int, int, const struct xproj_geom *, const cplx<float> *, const int2 *, const double *, const struct synth_geom *, const struct synth_tables *, cplx<float> *, cplx<float> *, hipStream_t);
template <typename T>
int launch_ring_sums(const cplx<T> *alm, int S, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *F, hipStream_t st, float *irs, const struct sky_epi<T> *epi)
{
	const int L = g->L;
	const size_t lds = ((size_t) (L + 1) * (L + 1) + (size_t) (L + 1) * (L + 2) / 2 + 2) * sizeof(cplx<T>);
	if(lds > 65536 || g->B > SPHR_MAX_MESH_B)
		return -1;
	struct sky_epi<T> none = {};
	hipLaunchKernelGGL((k_alm_lf<T>), dim3(S), dim3(NT), lds, st, alm, *g, *tb, F, irs, epi ? *epi : none);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T>
int launch_synth_gemm(const cplx<T> *alm, int S, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *F, cplx<T> *map, hipStream_t st)
{
	if(launch_ring_sums<T>(alm, S, g, tb, F, st, NULL, NULL))
		return -1;
	return launch_phi<T>(F, S, g, tb, map, st);
}

template int launch_ring_sums<double>(const cplx<double> *, int, const struct synth_geom *, const struct synth_tables *, cplx<double> *, hipStream_t, float *, const struct sky_epi<double> *);
template int launch_ring_sums<float>(const cplx<float> *, int, const struct synth_geom *, const struct synth_tables *, cplx<float> *, hipStream_t, float *, const struct sky_epi<float> *);

template int launch_synth_gemm<double>(const cplx<double> *, int, const struct synth_geom *, const struct synth_tables *, cplx<double> *, cplx<double> *, hipStream_t);
template int launch_synth_gemm<float>(const cplx<float> *, int, const struct synth_geom *, const struct synth_tables *, cplx<float> *, cplx<float> *, hipStream_t);
template int launch_phi<double>(const cplx<double> *, int, const struct synth_geom *, const struct synth_tables *, cplx<double> *, hipStream_t);
template int launch_phi<float>(const cplx<float> *, int, const struct synth_geom *, const struct synth_tables *, cplx<float> *, hipStream_t);

/* ring pairs per pass of k_sky that fit the LDS budget, or 0 */
template <typename T>
static int sky_rp(const struct xproj_geom *xg, int L, size_t *lds)
{
	const size_t budget = 65536;
	const int n = 2 * (L + 1), B = L + 1;
	const size_t fixed = (size_t) n * sizeof(cplx<T>) + (size_t) (L + 1) * (L + 1) * sizeof(cplx<T>) + ((size_t) (L + 1) * (L + 2) / 2 + xg->CT) * sizeof(T);
	if(fixed + 2 * n * sizeof(cplx<T>) > budget)
		return 0;
	int rp = (int) ((budget - fixed) / (2 * n * sizeof(cplx<T>)));
	rp = rp > B ? B : rp;
	/* equal passes */
	const int passes = (B + rp - 1) / rp;
	rp = (B + passes - 1) / passes;
	*lds = fixed + (size_t) 2 * rp * n * sizeof(cplx<T>);
	return rp;
}

template <typename T>
int launch_sky(const T *partial, int nsplit, int S, const struct xproj_geom *xg, const cplx<T> *R, const int2 *lm, const double *gmst, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *alm, cplx<T> *map, hipStream_t st)
{
	size_t lds = 0;
	const int rp = sky_rp<T>(xg, g->L, &lds);
	if(!rp) {
		/* coefficients too large for one workgroup's LDS: two kernels */
		if(launch_alm<T, T>(partial, nsplit, S, xg, g->L, R, lm, gmst, alm, st))
			return -1;
		return map ? launch_synth<T, T, T>(alm, S, 0, g, tb, map, st) : 0;
	}
	hipLaunchKernelGGL((k_sky<T>), dim3(S), dim3(NT), lds, st, partial, nsplit, S, *xg, R, lm, gmst, *g, *tb, rp, alm, map);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename TI, typename T, typename TO>
int launch_synth(const cplx<TI> *alm, int S, int polar, const struct synth_geom *g, const struct synth_tables *tb, cplx<TO> *map, hipStream_t st, const long *aoff, const long *moff)
{
	const size_t budget = 65536;
	const int L = g->L, n = g->n;
	const size_t ncoef = polar ? L + 1 : (size_t) (L + 1) * (L + 1);
	const size_t nal = polar ? L + 1 : (size_t) (L + 1) * (L + 2) / 2;
	int rp = (int) (24576 / (2 * n * sizeof(cplx<T>)));
	rp = rp < 1 ? 1 : rp > g->B ? g->B : rp;
	/* few series (the drop-in's one): fewer ring pairs per workgroup, so
	 * the launch spreads over the CUs instead of a handful of them */
	const int cus = sphr_rt_cus() > 0 ? sphr_rt_cus() : 256;
	if((long) S * ((g->B + rp - 1) / rp) < 2L * cus) {
		const int want = (int) (((long) g->B * S + 2L * cus - 1) / (2L * cus));
		rp = want < 1 ? 1 : want < rp ? want : rp;
	}
	size_t lds = (size_t) (2 * rp + 1) * n * sizeof(cplx<T>);
	const int a_in_lds = lds + ncoef * sizeof(cplx<T>) <= budget - 4096;
	if(a_in_lds)
		lds += ncoef * sizeof(cplx<T>);
	const int al_in_lds = lds + nal * sizeof(T) <= budget;
	if(al_in_lds)
		lds += nal * sizeof(T);
	dim3 grid((g->B + rp - 1) / rp, S);
	hipLaunchKernelGGL((k_synth<TI, T, TO>), grid, dim3(NT), lds, st, alm, polar, rp, a_in_lds, al_in_lds, *g, *tb, map, aoff, moff);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_synth<double, double, double>(const cplx<double> *, int, int, const struct synth_geom *, const struct synth_tables *, cplx<double> *, hipStream_t, const long *, const long *);
template int launch_synth<float, float, float>(const cplx<float> *, int, int, const struct synth_geom *, const struct synth_tables *, cplx<float> *, hipStream_t, const long *, const long *);

template <typename T>
size_t synth_cls_shape(struct synth_cls *d)
{
	/* every workgroup of the launch gets the largest class's LDS: 40 KB
	 * keeps four per CU, the largest orders reading their coefficients
	 * from L2 (float32 4.37 -> 3.86 ms on 4096 jobs of l_max 16-64 against
	 * 64 KB; 48, 32 and 24 KB: 4.02, 3.94, 4.51 ms) */
	const size_t budget = 40960;
	const int L = d->g.L, n = d->g.n;
	const size_t ncoef = (size_t) (L + 1) * (L + 1), nal = (size_t) (L + 1) * (L + 2) / 2;
	int rp = (int) (24576 / (2 * n * sizeof(cplx<T>)));
	rp = rp < 1 ? 1 : rp > d->g.B ? d->g.B : rp;
	size_t lds = (size_t) (2 * rp + 1) * n * sizeof(cplx<T>);
	d->a_in_lds = lds + ncoef * sizeof(cplx<T>) <= budget;
	if(d->a_in_lds)
		lds += ncoef * sizeof(cplx<T>);
	d->al_in_lds = lds + nal * sizeof(T) <= budget;
	if(d->al_in_lds)
		lds += nal * sizeof(T);
	d->rp = rp;
	d->gx = (d->g.B + rp - 1) / rp;
	return lds;
}

template <typename T>
int launch_synth_classes(const cplx<T> *alm, const struct synth_cls *d_cls, int ncls, long nblocks, size_t lds, cplx<T> *map, hipStream_t st)
{
	if(ncls < 1 || nblocks < 1 || nblocks > 0x7fffffffL || lds > 65536)
		return -1;
	hipLaunchKernelGGL((k_synth_cls<T>), dim3((unsigned) nblocks), dim3(NT), lds, st, alm, d_cls, ncls, map);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template size_t synth_cls_shape<double>(struct synth_cls *);
template size_t synth_cls_shape<float>(struct synth_cls *);
template int launch_synth_classes<double>(const cplx<double> *, const struct synth_cls *, int, long, size_t, cplx<double> *, hipStream_t);
template int launch_synth_classes<float>(const cplx<float> *, const struct synth_cls *, int, long, size_t, cplx<float> *, hipStream_t);
template int launch_sky<double>(const double *, int, int, const struct xproj_geom *, const cplx<double> *, const int2 *, const double *, const struct synth_geom *, const struct synth_tables *, cplx<double> *, cplx<double> *, hipStream_t);
template int launch_sky<float>(const float *, int, int, const struct xproj_geom *, const cplx<float> *, const int2 *, const double *, const struct synth_geom *, const struct synth_tables *, cplx<float> *, cplx<float> *, hipStream_t);

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

/* the projections of all baselines of a network in two launches: grid
 * (column blocks, K splits, baselines), then (column blocks, baselines) */
__global__ void __launch_bounds__(NT) k_dotc_net_partial(struct dotc_net_args a, int n, int rows_per_split, int nsplit)
{
	__shared__ cplx<double> red[4][64];
	const int b = blockIdx.z;
	const int stride = a.stride[b];
	const int c = blockIdx.x * 64 + (threadIdx.x & 63);
	const int lane = threadIdx.x >> 6;
	const int i0 = blockIdx.y * rows_per_split;
	const int i1 = min(n, i0 + rows_per_split);
	const cplx<double> *coeff = a.coeff[b], *vec = a.vec[b];
	cplx<double> acc = mk(0.0, 0.0);
	if(c < stride)
		for(int i = i0 + lane; i < i1; i += 4)
			cfma(acc, coeff[(size_t) i * stride + c], vec[i]);
	red[lane][threadIdx.x & 63] = acc;
	__syncthreads();
	if(lane == 0 && c < stride) {
		cplx<double> t = red[0][threadIdx.x];
		t = t + red[1][threadIdx.x];
		t = t + red[2][threadIdx.x];
		t = t + red[3][threadIdx.x];
		a.partial[b][(size_t) blockIdx.y * stride + c] = t;
	}
}

__global__ void k_dotc_net_reduce(struct dotc_net_args a, int nsplit, double scale)
{
	const int b = blockIdx.y;
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if(c >= a.stride[b])
		return;
	cplx<double> t = mk(0.0, 0.0);
	for(int sp = 0; sp < nsplit; sp++)
		t = t + a.partial[b][(size_t) sp * a.stride[b] + c];
	a.out[b][c] = t * scale;
}

int launch_dotc_net(const struct dotc_net_args *a, int nb, int n, int nsplit, double scale, hipStream_t st)
{
	int maxs = 1;
	for(int b = 0; b < nb; b++)
		maxs = a->stride[b] > maxs ? a->stride[b] : maxs;
	const int rps = (n + nsplit - 1) / nsplit;
	hipLaunchKernelGGL(k_dotc_net_partial, dim3((maxs + 63) / 64, nsplit, nb), dim3(NT), 0, st, *a, n, rps, nsplit);
	hipLaunchKernelGGL(k_dotc_net_reduce, dim3((maxs + NT - 1) / NT, nb), dim3(NT), 0, st, *a, nsplit, scale);
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
