/*
 * The FFT, cross-spectrum and projection stages of the batched cross-power
 * path in one kernel, with the projection done in the lag domain.
 *
 * The frequency-domain integrator (correlator.c:681-711,
 * sh_series_array.c:437-456) forms, per baseline (a, b) and coefficient l,
 *     power_1d[l] = (1/N) sum_f dp[f][l] X[f],   X = F_b conj(F_a).
 * With X[f] = sum_u c[u] exp(-2 pi i f u / N), c[u] = sum_t x_a[t] x_b[t+u]
 * (circular, windowed series), this is
 *     power_1d[l] = sum_u c[u] M[u][l],
 *     M[u][l] = (1/N) sum_f dp[f][l] exp(-2 pi i f u / N).
 * The delay product is the product of two tap projections of
 * 2 transient + 1 taps each (projection.c:99-131), so M vanishes for
 * |u| > 2 transient; M is real (dp[N-f] = conj dp[f]) and
 * M[-u][l] = (-1)^l M[u][l] (the baseline frame is symmetric about its
 * centre).  The host builds M from the plan's own dp and checks all three
 * properties before it takes this route.  Then
 *     power_1d[l] = sum_{u=0}^{U} Ms[u][l] cs_l[u],
 *     cs_l[0] = c[0], cs_l[u] = c[u] + (-1)^l c[-u],
 * U + 1 terms (163, 233 and 231 for three sites at 1024 samples) in place
 * of the N + 2 of the frequency-domain product, and the spectra never
 * leave the chip.
 *
 * One workgroup of 256 G threads (G = the detector count, 2 or 3),
 * persistent over passes of PP segment pairs (s, s + 1) each: with PP = 1
 * each transform runs over 256 threads, with PP = 2 (two pairs at once)
 * over 128 threads, two butterflies per thread per stage:
 *  1. group d: radix-4 Stockham FFT of z = x_d,s + i x_d,s+1 (windowed),
 *     the series read straight from HBM into registers, the next pair's
 *     samples loaded while this pair is transformed;
 *  2. group b (baseline): the packed cross spectra
 *     Y = X_s + i X_s+1 formed from the detectors' packed spectra where
 *     the inverse transform's first stage reads them; the inverse runs as
 *     the forward transform of conj(Y), giving 4 N (c_s - i c_s+1);
 *  3. the lags folded into cs for both parities of l;
 *  4. thread (b, l, chunk): the dot product of its chunk of lags with the
 *     table rows it holds in registers (loaded once per workgroup), the
 *     chunks summed through LDS, power_1d written to the partial-sum
 *     layout of the sky kernel (real part, zero imaginary part).
 * Fourteen barriers per pass; LDS 2 PP G N complex values.
 */
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "runtime.h"

template <typename T, typename TI, int NQ, int G, bool MREG, int PP>
__global__ void __launch_bounds__(256 * G) k_xcorr(const TI *__restrict__ in, const double *__restrict__ window, int S, struct xcorr_geom g, const cplx<T> *__restrict__ tw, const T *__restrict__ M, T *__restrict__ partial)
{
	constexpr int N = 1024 * NQ, N4 = N / 4, NH = N / 2;
	constexpr int NS = NQ == 1 ? 5 : 6;	/* radix-4 stages */
	/* PP segment pairs per pass, each transform over TPT threads with
	 * NB butterflies per thread per stage */
	constexpr int TPT = 256 / PP, NB = NQ * PP;
	__shared__ cplx<T> buf[2][PP * G][N];
	const int fg = __builtin_amdgcn_readfirstlane(threadIdx.x / TPT);
	const int ps = fg / G, fd = fg - G * ps;	/* pair slot, detector / baseline */
	int ft = threadIdx.x & (TPT - 1);
	const int npg = ((S + 1) / 2 + PP - 1) / PP;	/* passes of PP pairs */
	/* twiddles of every stage: w_r = tw[r k N / 4p], k = q mod p (the
	 * inverse transform is a forward one too); where registers are short
	 * (float64, 4096 samples, two pairs per pass) only w_1 is held and
	 * w_2, w_3 are formed from it at each use */
	constexpr bool TWF = sizeof(T) == sizeof(float) && NB == 1;
	constexpr int NTW = TWF ? 3 : 1;
	cplx<T> wt[NS][NB][NTW];
#pragma unroll
	for(int st = 1; st < NS; st++) {
		const int p = 1 << (2 * st);
#pragma unroll
		for(int u = 0; u < NB; u++) {
			const int k = (ft + TPT * u) & (p - 1);
#pragma unroll
			for(int r = 1; r <= NTW; r++)
				wt[st][u][r - 1] = tw[r * k * (N / (4 * p))];
		}
	}
	T wv[NB][4];
#pragma unroll
	for(int u = 0; u < NB; u++)
#pragma unroll
		for(int r = 0; r < 4; r++)
			wv[u][r] = window ? (T) window[ft + TPT * u + r * N4] : (T) 1;
	const bool fw = fd < g.D, iv = fd < g.nb;
	/* this thread's projection: baseline pb, coefficient l, lag chunk ch
	 * (256 threads per baseline) */
	const int pb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
	const bool pv = pb < g.nb;
	const int nl = pv ? g.nl[pb] : 1;
	constexpr int R = SPHR_XC_MR;	/* lags per chunk */
	const int ch = t / nl, l = t - ch * nl;
	const bool act = pv && ch < g.nch[pb];
	/* the table rows: with MREG (float32 at 1024 samples), held in
	 * registers, loaded once per workgroup (one 768-thread workgroup per
	 * CU); otherwise read from L2 in the projection loop itself (two
	 * workgroups per CU in float32) */
	constexpr bool MPRE = MREG && sizeof(T) == sizeof(float) && NQ == 1;
	constexpr int MRR = MPRE ? SPHR_XC_MR : 1;
	T mr[MRR];
	const T *Mb = M + (pv ? g.moff[pb] : 0);
	/* samples q + r N/4 of detector fd, segments 2 pr and 2 pr + 1 of
	 * pair pr = PP pg + ps (a missing pair reads the last one) */
	TI xv[NB][4], yv[NB][4];
	auto load = [&](int pg) {
		int pr = PP * pg + ps;
		pr = 2 * pr < S ? pr : (S - 1) / 2;
		const int s0 = 2 * pr, s1 = s0 + 1 < S ? s0 + 1 : s0;
		const TI *x = in + ((size_t) fd * S + s0) * N, *y = in + ((size_t) fd * S + s1) * N;
#pragma unroll
		for(int u = 0; u < NB; u++)
#pragma unroll
			for(int r = 0; r < 4; r++) {
				const int i = ft + TPT * u + r * N4;
				xv[u][r] = x[i];
				yv[u][r] = y[i];
			}
	};
	/* one radix-4 Stockham stage of the transform in a[] into o (the
	 * stage with sub-transform size p = 4^st: butterfly q, k = q mod p,
	 * outputs to (q / p) 4p + k + r p) */
	auto stage = [&](cplx<T> (&a)[NB][4], int st, cplx<T> *o) {
		const int p = 1 << (2 * st);
#pragma unroll
		for(int u = 0; u < NB; u++) {
			const int q = ft + TPT * u, k = q & (p - 1);
			cplx<T> u0 = a[u][0], u1 = a[u][1], u2 = a[u][2], u3 = a[u][3];
			if(st > 0) {
				const cplx<T> w1 = wt[st][u][0];
				const cplx<T> w2 = TWF ? wt[st][u][TWF ? 1 : 0] : w1 * w1;
				const cplx<T> w3 = TWF ? wt[st][u][TWF ? 2 : 0] : w2 * w1;
				u1 = u1 * w1;
				u2 = u2 * w2;
				u3 = u3 * w3;
			}
			const cplx<T> s02 = u0 + u2, d02 = u0 - u2, s13 = u1 + u3, d13 = u1 - u3;
			const cplx<T> jd13 = mk(d13.im, -d13.re);	/* -i (u1 - u3) */
			const int base = (q - k) * 4 + k;
			o[base] = s02 + s13;
			o[base + p] = d02 + jd13;
			o[base + 2 * p] = s02 - s13;
			o[base + 3 * p] = d02 - jd13;
		}
	};
	if(fw && (int) blockIdx.x < npg)
		load(blockIdx.x);
	/* the table rows after the first pair's samples, with no branch
	 * around the loads (threads without a chunk read a valid row and never
	 * use it): the first transform waits for the samples only, the rows
	 * arrive during the transforms */
	if(MPRE) {
#pragma unroll
		for(int j = 0; j < MRR; j++)
			mr[j] = Mb[act ? (ch * R + j) * nl + l : 0];
	}
	for(int pg = blockIdx.x; pg < npg; pg += gridDim.x) {
		/* the thread index and the twiddles enter each pass through an
		 * empty asm, so per-thread index arithmetic and derived twiddles
		 * are formed where they are used instead of being lifted out of
		 * the loop and held across every phase */
		asm volatile("" : "+v"(ft));
		if(!TWF) {
#pragma unroll
			for(int st = 1; st < NS; st++)
#pragma unroll
				for(int u = 0; u < NB; u++)
					asm volatile("" : "+v"(wt[st][u][0].re), "+v"(wt[st][u][0].im));
		}
		const int s0 = 2 * (PP * pg + ps);
		const bool has1 = s0 + 1 < S;
		cplx<T> a[NB][4];
		/* 1. forward transforms, one detector per group */
		if(fw) {
#pragma unroll
			for(int u = 0; u < NB; u++)
#pragma unroll
				for(int r = 0; r < 4; r++) {
					const T w = wv[u][r];
					a[u][r] = mk((T) xv[u][r] * w, has1 ? (T) yv[u][r] * w : (T) 0);
				}
			if(pg + (int) gridDim.x < npg)
				load(pg + gridDim.x);
		}
#pragma unroll
		for(int st = 0; st < NS; st++) {
			if(fw) {
				if(st > 0) {
					const cplx<T> *iv_ = buf[(st - 1) & 1][fg];
#pragma unroll
					for(int u = 0; u < NB; u++)
#pragma unroll
						for(int r = 0; r < 4; r++)
							a[u][r] = iv_[ft + TPT * u + r * N4];
				}
				stage(a, st, buf[st & 1][fg]);
			}
			__syncthreads();
		}
		/* 2. packed cross spectra of baseline fd: with Z_d the packed
		 * spectrum of detector d, 2 F_s = Z[f] + conj Z[-f] and
		 * 2 i F_s+1 = Z[f] - conj Z[-f], so 4 X_s = sb conj(sa) and
		 * 4 X_s+1 = db conj(da); the transform input is conj(4 Y) */
		if(iv) {
			const cplx<T> *Za = buf[(NS - 1) & 1][G * ps + g.ia[fd]], *Zb = buf[(NS - 1) & 1][G * ps + g.ib[fd]];
#pragma unroll
			for(int u = 0; u < NB; u++)
#pragma unroll
				for(int r = 0; r < 4; r++) {
					const int f = ft + TPT * u + r * N4, fn = (N - f) & (N - 1);
					const cplx<T> za = Za[f], zac = conjg(Za[fn]), zb = Zb[f], zbc = conjg(Zb[fn]);
					const cplx<T> sa = za + zac, da = za - zac, sb = zb + zbc, db = zb - zbc;
					const cplx<T> x0 = sb * conjg(sa), x1 = db * conjg(da);
					a[u][r] = mk(x0.re - x1.im, -(x0.im + x1.re));
				}
		}
#pragma unroll
		for(int st = 0; st < NS; st++) {
			if(iv) {
				if(st > 0) {
					const cplx<T> *iv_ = buf[(NS + st - 1) & 1][fg];
#pragma unroll
					for(int u = 0; u < NB; u++)
#pragma unroll
						for(int r = 0; r < 4; r++)
							a[u][r] = iv_[ft + TPT * u + r * N4];
				}
				stage(a, st, buf[(NS + st) & 1][fg]);
			}
			__syncthreads();
		}
		/* 3. 4 N (c_s - i c_s+1) is in buf[1][G ps + b] ((2 NS - 1) is
		 * odd); the 256 threads of baseline pb fold its lags into
		 * buf[0][G ps + pb] for each pair: even l at [0, N/2), odd l at
		 * [N/2, N), zero past the last lag */
		if(pv) {
			const int up = g.nch[pb] * R, U1 = g.U1[pb];
#pragma unroll
			for(int pq = 0; pq < PP; pq++) {
				const cplx<T> *y = buf[1][G * pq + pb];
				cplx<T> *f0 = buf[0][G * pq + pb], *f1 = f0 + NH;
				for(int u = t; u < up; u += 256) {
					cplx<T> P = mk((T) 0, (T) 0), Q = P;
					if(u < U1) {
						const cplx<T> yu = y[u];
						if(u == 0) {
							P = yu;
						} else {
							const cplx<T> yn = y[N - u];
							P = yu + yn;
							Q = yu - yn;
						}
					}
					f0[u] = P;
					f1[u] = Q;
				}
			}
		}
		__syncthreads();
		/* 4. this thread's chunk of lags against its table rows, for each
		 * pair of the pass; the chunk sums into buf[1] (y is read) */
		T acc0[PP], acc1[PP];
#pragma unroll
		for(int pq = 0; pq < PP; pq++)
			acc0[pq] = acc1[pq] = (T) 0;
		if(act) {
			/* opaque per pass: otherwise the loads are lifted out of the
			 * pass loop and their values held across the transforms */
			int mo = ch * R * nl + l;
			asm volatile("" : "+v"(mo));
			const T *Mt = Mb + mo;
#pragma unroll
			for(int j = 0; j < R; j++) {
				const T m = MPRE ? mr[MPRE ? j : 0] : Mt[j * nl];
#pragma unroll
				for(int pq = 0; pq < PP; pq++) {
					const cplx<T> v = buf[0][G * pq + pb][((l & 1) ? NH : 0) + ch * R + j];
					acc0[pq] = fma(m, v.re, acc0[pq]);
					acc1[pq] = fma(m, v.im, acc1[pq]);
				}
			}
		}
		if(pv) {
#pragma unroll
			for(int pq = 0; pq < PP; pq++)
				buf[1][G * pq + pb][t] = mk(acc0[pq], acc1[pq]);
		}
		__syncthreads();
		if(pv) {
			const int c0 = g.col_start[pb];
#pragma unroll
			for(int pq = 0; pq < PP; pq++) {
				const int s = 2 * (PP * pg + pq);
				if(s >= S)
					break;
				const bool h1 = s + 1 < S;
				T *o0 = partial + (size_t) s * g.CT + c0, *o1 = o0 + g.CT;
				if(t < nl) {
					T v0 = (T) 0, v1 = (T) 0;
					for(int c = 0; c < g.nch[pb]; c++) {
						const cplx<T> v = buf[1][G * pq + pb][c * nl + t];
						v0 += v.re;
						v1 += v.im;
					}
					o0[2 * t] = v0;
					o0[2 * t + 1] = (T) 0;
					if(h1) {
						o1[2 * t] = -v1;
						o1[2 * t + 1] = (T) 0;
					}
				} else if(t < nl + g.width[pb] - 2 * nl) {
					/* the baseline's padding columns */
					const int c = nl + t;
					o0[c] = (T) 0;
					if(h1)
						o1[c] = (T) 0;
				}
			}
		}
	}
}

/*
 * The same stages with each 1024-point transform in one wavefront (float32,
 * 1024 samples, two or three detectors).  The Stockham form above runs
 * every radix-4 stage of a transform across 256 threads, a workgroup
 * barrier per stage; here a wave holds 16 points per lane and the
 * transform is
 *     1024 = 16 (in registers) x 16 (in registers) x 4 (in registers),
 * two transposes through the wave's own LDS tile in between, ordered by
 * the wave alone.  The workgroup (12 waves) takes two segment pairs at a
 * time: waves 3p + d transform detector d of pair p, then baseline d's
 * cross spectra; every thread holds its table rows in registers, as in
 * the Stockham form.  Four workgroup barriers per two pairs.
 *
 * Index maps of the transform X[k] = sum_n x[n] W^{nk}, W = exp(-2 pi i / 1024):
 *   n = n1 + 64 n2, k = k2 + 16 k1:  lane n1 holds x[n1 + 64 n2], n2 < 16;
 *     a 16-point DFT over n2, then W^{n1 k2} (table twA[k2][n1]);
 *   n1 = a + 4 b, k1 = c + 16 d:  lane 4 k2 + a holds the values of
 *     n1 = a + 4 b; a 16-point DFT over b, then W^{16 a c} (table
 *     twC[c][lane]);
 *   lane 4 k2 + cg holds the values of c = 4 cg + c' and a < 4; 4-point
 *     DFTs over a give X[k2 + 64 cg + 16 c' + 256 d] in slot c' + 4 d.
 * The LDS tiles are padded (rows of 68, and of 17) so that each transpose
 * touches every bank equally.
 */
#define XW_PAIRS 2		/* segment pairs per workgroup pass */
#define XW_TILE 1088		/* complex values in a wave's LDS tile */

__device__ inline void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* y_k = sum_a x_a W4^{ak} in place, W4 = -i */
__device__ inline void dft4(cplx<float> &x0, cplx<float> &x1, cplx<float> &x2, cplx<float> &x3)
{
	const cplx<float> s02 = x0 + x2, d02 = x0 - x2, s13 = x1 + x3, d13 = x1 - x3;
	const cplx<float> jd13 = mk(d13.im, -d13.re);	/* -i (x1 - x3) */
	x0 = s02 + s13;
	x2 = s02 - s13;
	x1 = d02 + jd13;
	x3 = d02 - jd13;
}

/* 16-point DFT in registers: with m = 4 a + b and k = c + 4 d,
 * V[k] = sum_b W4^{bd} W16^{bc} sum_a v[4a + b] W4^{ac}; V[c + 4 d] is left
 * in slot 4 c + d */
__device__ inline void dft16(cplx<float> (&v)[16])
{
	const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
#pragma unroll
	for(int b = 0; b < 4; b++)
		dft4(v[b], v[4 + b], v[8 + b], v[12 + b]);
	/* slot 4 c + b: times W16^{bc} */
	v[5] = v[5] * mk(c1, -s1);		/* b 1, c 1 */
	v[9] = v[9] * mk(h, -h);		/* b 1, c 2 */
	v[13] = v[13] * mk(s1, -c1);		/* b 1, c 3 */
	v[6] = v[6] * mk(h, -h);		/* b 2, c 1 */
	v[10] = mk(v[10].im, -v[10].re);	/* b 2, c 2: -i */
	v[14] = v[14] * mk(-h, -h);		/* b 2, c 3 */
	v[7] = v[7] * mk(s1, -c1);		/* b 3, c 1 */
	v[11] = v[11] * mk(-h, -h);		/* b 3, c 2 */
	v[15] = v[15] * mk(-c1, s1);		/* b 3, c 3 */
#pragma unroll
	for(int c = 0; c < 4; c++)
		dft4(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]);
}

/* the 1024-point forward transform of the wave's points (lane n1 holds
 * x[n1 + 64 n2] in v[n2]); on return lane 4 k2 + cg holds
 * X[k2 + 64 cg + 16 c' + 256 d] in v[c' + 4 d].  ex: the wave's LDS tile */
__device__ inline void wave_fft1024(cplx<float> (&v)[16], cplx<float> *ex, const cplx<float> *__restrict__ twA, const cplx<float> *__restrict__ twC, int lane)
{
	dft16(v);
	/* V[k2] (slot 4 (k2 % 4) + k2 / 4) times W^{n1 k2}, into row k2 */
#pragma unroll
	for(int k2 = 0; k2 < 16; k2++) {
		const int sl = 4 * (k2 & 3) + (k2 >> 2);
		ex[k2 * 68 + lane] = k2 ? v[sl] * twA[k2 * 64 + lane] : v[sl];
	}
	wave_sync();
	const int k2 = lane >> 2, a = lane & 3;
#pragma unroll
	for(int b = 0; b < 16; b++)
		v[b] = ex[k2 * 68 + a + 4 * b];
	wave_sync();
	dft16(v);
	/* U[c] (slot 4 (c % 4) + c / 4) times W^{16 a c}, into row 16 a + c
	 * (rows of 17, column k2) */
#pragma unroll
	for(int c = 0; c < 16; c++) {
		const int sl = 4 * (c & 3) + (c >> 2);
		ex[(16 * a + c) * 17 + k2] = c ? v[sl] * twC[c * 64 + lane] : v[sl];
	}
	wave_sync();
	const int cg = lane & 3;
#pragma unroll
	for(int aa = 0; aa < 4; aa++)
#pragma unroll
		for(int cp = 0; cp < 4; cp++)
			v[cp + 4 * aa] = ex[(16 * aa + 4 * cg + cp) * 17 + k2];
	wave_sync();
#pragma unroll
	for(int cp = 0; cp < 4; cp++)
		dft4(v[cp], v[cp + 4], v[cp + 8], v[cp + 12]);
}

template <typename TI>
__global__ void __launch_bounds__(768) k_xcorr_w(const TI *__restrict__ in, const double *__restrict__ window, int S, struct xcorr_geom g, const cplx<float> *__restrict__ twW, const float *__restrict__ M, float *__restrict__ partial)
{
	typedef cplx<float> C;
	constexpr int N = 1024, NH = 512, R = SPHR_XC_MR;
	__shared__ C ex[3 * XW_PAIRS][XW_TILE];	/* the FFT waves' tiles; y after the inverse */
	__shared__ C zb[XW_PAIRS][3][N];	/* packed spectra; then the folded lags */
	__shared__ C red[XW_PAIRS][3][256];	/* the projection's chunk sums */
	const C *twA = twW, *twC = twW + 16 * 64;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	/* FFT role: pair fp, detector / baseline fd */
	const int fp = wave / 3, fd = wave - 3 * fp;
	const bool fwave = wave < 3 * XW_PAIRS;
	/* projection role: baseline pb, coefficient l, lag chunk ch */
	const int pb = threadIdx.x >> 8, t = threadIdx.x & 255;
	const bool pv = pb < g.nb;
	const int nl = pv ? g.nl[pb] : 1;
	const int ch = t / nl, l = t - ch * nl;
	const bool act = pv && ch < g.nch[pb];
	float mr[R];
	{
		const float *Mb = M + (pv ? g.moff[pb] : 0);
#pragma unroll
		for(int j = 0; j < R; j++)
			mr[j] = act ? Mb[(ch * R + j) * nl + l] : 0.f;
	}
	const int nquads = (S + 2 * XW_PAIRS - 1) / (2 * XW_PAIRS);
	for(int q = blockIdx.x; q < nquads; q += gridDim.x) {
		int ln = lane;
		asm volatile("" : "+v"(ln));
		const int s0 = 2 * XW_PAIRS * q + 2 * fp;
		const bool fpair = fwave && s0 < S, has1 = s0 + 1 < S;
		C v[16];
		/* 1. forward transforms: wave (fp, fd) takes detector fd of its
		 * pair, z = x_s0 + i x_s1, windowed */
		if(fpair && fd < g.D) {
			const TI *x0 = in + ((size_t) fd * S + s0) * N;
			const TI *x1 = in + ((size_t) fd * S + (has1 ? s0 + 1 : s0)) * N;
#pragma unroll
			for(int m = 0; m < 16; m++) {
				const int i = ln + 64 * m;
				const float w = window ? (float) window[i] : 1.f;
				v[m] = mk((float) x0[i] * w, has1 ? (float) x1[i] * w : 0.f);
			}
			wave_fft1024(v, ex[wave], twA, twC, ln);
			C *Z = zb[fp][fd];
			const int k2 = ln >> 2, cg = ln & 3;
#pragma unroll
			for(int d = 0; d < 4; d++)
#pragma unroll
				for(int cp = 0; cp < 4; cp++)
					Z[k2 + 64 * cg + 16 * cp + 256 * d] = v[cp + 4 * d];
		}
		__syncthreads();
		/* 2. baseline fd of pair fp: conj(4 Y) from the packed spectra
		 * (as in the Stockham form), transformed, 4 N (c_s - i c_s+1) left
		 * in natural order in the wave's tile */
		if(fpair && fd < g.nb) {
			const C *Za = zb[fp][g.ia[fd]], *Zb = zb[fp][g.ib[fd]];
#pragma unroll
			for(int m = 0; m < 16; m++) {
				const int f = ln + 64 * m, fn = (N - f) & (N - 1);
				const C za = Za[f], zac = conjg(Za[fn]), zbv = Zb[f], zbc = conjg(Zb[fn]);
				const C sa = za + zac, da = za - zac, sb = zbv + zbc, db = zbv - zbc;
				const C x0 = sb * conjg(sa), x1 = db * conjg(da);
				v[m] = mk(x0.re - x1.im, -(x0.im + x1.re));
			}
			C *y = ex[wave];
			wave_fft1024(v, y, twA, twC, ln);
			const int k2 = ln >> 2, cg = ln & 3;
#pragma unroll
			for(int d = 0; d < 4; d++)
#pragma unroll
				for(int cp = 0; cp < 4; cp++)
					y[k2 + 64 * cg + 16 * cp + 256 * d] = v[cp + 4 * d];
		}
		__syncthreads();
		/* 3. fold the lags of every (pair, baseline) into its spectrum
		 * slot: even l at [0, N/2), odd l at [N/2, N), zero past the
		 * last lag */
		for(int pp = 0; pp < XW_PAIRS; pp++)
			for(int b = 0; b < g.nb; b++) {
				const int up = g.nch[b] * R, U1 = g.U1[b];
				const C *y = ex[3 * pp + b];
				C *f0 = zb[pp][b], *f1 = f0 + NH;
				for(int u = threadIdx.x; u < up; u += 768) {
					C P = mk(0.f, 0.f), Q = P;
					if(u < U1) {
						const C yu = y[u];
						if(u == 0) {
							P = yu;
						} else {
							const C yn = y[N - u];
							P = yu + yn;
							Q = yu - yn;
						}
					}
					f0[u] = P;
					f1[u] = Q;
				}
			}
		__syncthreads();
		/* 4. each thread's chunk of lags against its table rows, for
		 * every pair */
#pragma unroll
		for(int pp = 0; pp < XW_PAIRS; pp++) {
			float a0 = 0.f, a1 = 0.f;
			if(act) {
				const C *F = zb[pp][pb] + ((l & 1) ? NH : 0) + ch * R;
#pragma unroll
				for(int j = 0; j < R; j++) {
					const C vv = F[j];
					a0 = fmaf(mr[j], vv.re, a0);
					a1 = fmaf(mr[j], vv.im, a1);
				}
			}
			if(pv)
				red[pp][pb][t] = mk(a0, a1);
		}
		__syncthreads();
		if(pv) {
			const int c0 = g.col_start[pb];
#pragma unroll
			for(int pp = 0; pp < XW_PAIRS; pp++) {
				const int s = 2 * XW_PAIRS * q + 2 * pp;
				if(s >= S)
					break;
				const bool h1 = s + 1 < S;
				float *o0 = partial + (size_t) s * g.CT + c0, *o1 = o0 + g.CT;
				if(t < nl) {
					float v0 = 0.f, v1 = 0.f;
					for(int c = 0; c < g.nch[pb]; c++) {
						const C vv = red[pp][pb][c * nl + t];
						v0 += vv.re;
						v1 += vv.im;
					}
					o0[2 * t] = v0;
					o0[2 * t + 1] = 0.f;
					if(h1) {
						o1[2 * t] = -v1;
						o1[2 * t + 1] = 0.f;
					}
				} else if(t < nl + g.width[pb] - 2 * nl) {
					const int c = nl + t;
					o0[c] = 0.f;
					if(h1)
						o1[c] = 0.f;
				}
			}
		}
	}
}

/* form 1: the Stockham kernel with the table rows in registers, one
 * workgroup per CU; form 2: the same with the rows read from L2, two per
 * CU; form 3: the wave-level kernel (float32, 1024 samples); form 4: the
 * Stockham kernel with two pairs per pass, 128 threads per transform
 * (float32, 1024 samples) */
template <typename T, typename TI, int NQ, int G>
static int xc_at(const void *in, const double *window, int S, const struct xcorr_geom *g, const cplx<T> *tw, const cplx<T> *twW, const T *M, T *partial, int form, hipStream_t st)
{
	const int ncu = sphr_rt_cus();
	if(form == 3) {
		if constexpr(sizeof(T) == sizeof(float) && NQ == 1) {
			const int nq = (S + 2 * XW_PAIRS - 1) / (2 * XW_PAIRS);
			hipLaunchKernelGGL((k_xcorr_w<TI>), dim3(nq < ncu ? nq : ncu), dim3(768), 0, st, (const TI *) in, window, S, *g, twW, M, partial);
			return hipGetLastError() == hipSuccess ? 0 : -1;
		}
		return -1;
	}
	const int npairs = (S + 1) / 2;
	const dim3 block(256 * G);
	if(form == 4) {
		if constexpr(sizeof(T) == sizeof(float) && NQ == 1) {
			const int npg = (npairs + 1) / 2;
			hipLaunchKernelGGL((k_xcorr<T, TI, NQ, G, true, 2>), dim3(npg < ncu ? npg : ncu), block, 0, st, (const TI *) in, window, S, *g, tw, M, partial);
			return hipGetLastError() == hipSuccess ? 0 : -1;
		}
		return -1;
	}
	const int cap = (form == 1 ? 1 : 2) * ncu;
	const dim3 grid(npairs < cap ? npairs : cap);
	if(form == 1)
		hipLaunchKernelGGL((k_xcorr<T, TI, NQ, G, true, 1>), grid, block, 0, st, (const TI *) in, window, S, *g, tw, M, partial);
	else
		hipLaunchKernelGGL((k_xcorr<T, TI, NQ, G, false, 1>), grid, block, 0, st, (const TI *) in, window, S, *g, tw, M, partial);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T, int NQ, int G>
static int xc_in(const void *in, int in_f32, const double *window, int S, const struct xcorr_geom *g, const cplx<T> *tw, const cplx<T> *twW, const T *M, T *partial, int form, hipStream_t st)
{
	return in_f32 ? xc_at<T, float, NQ, G>(in, window, S, g, tw, twW, M, partial, form, st) : xc_at<T, double, NQ, G>(in, window, S, g, tw, twW, M, partial, form, st);
}

/* compiled shapes: 1024 samples at two or three detectors in both
 * precisions (the wave-level form in float32), 4096 samples at two
 * detectors in float32 (the others would need more than the 160 KB of
 * LDS); bit f - 1 set for form f */
template <typename T>
int xcorr_forms(int N, int D, int nb)
{
	if(nb < 1 || nb > D || D < 2 || D > SPHR_XC_MAXB)
		return 0;
	if(N == 1024)
		return sizeof(T) == sizeof(float) ? 15 : 3;
	return N == 4096 && D == 2 && sizeof(T) == sizeof(float) ? 3 : 0;
}

template <typename T>
int launch_xcorr(const void *in, int in_f32, const double *window, int S, int N, const struct xcorr_geom *g, const cplx<T> *tw, const cplx<T> *twW, const T *M, T *partial, int form, hipStream_t st)
{
	if(form < 1 || form > 4 || !(xcorr_forms<T>(N, g->D, g->nb) & (1 << (form - 1))))
		return -1;
	if(N == 1024)
		return g->D == 2 ? xc_in<T, 1, 2>(in, in_f32, window, S, g, tw, twW, M, partial, form, st) : xc_in<T, 1, 3>(in, in_f32, window, S, g, tw, twW, M, partial, form, st);
	if constexpr(sizeof(T) == sizeof(float))
		return xc_in<T, 4, 2>(in, in_f32, window, S, g, tw, twW, M, partial, form, st);
	return -1;
}

template int xcorr_forms<float>(int, int, int);
template int xcorr_forms<double>(int, int, int);
template int launch_xcorr<float>(const void *, int, const double *, int, int, const struct xcorr_geom *, const cplx<float> *, const cplx<float> *, const float *, float *, int, hipStream_t);
template int launch_xcorr<double>(const void *, int, const double *, int, int, const struct xcorr_geom *, const cplx<double> *, const cplx<double> *, const double *, double *, int, hipStream_t);
