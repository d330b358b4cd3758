/*
 * HEALPix output on the GPU (sh_series_healpix.cc, healpix.py,
 * RapidLocalization.py:498-530 of the reference).
 *
 * The reference writes the real part of a series (sh_series_real) through
 * healpix_cxx alm2map into a RING-ordered map at the centres of the
 * pixels; for the real part's coefficients that is Re f at the centres,
 * f the series itself, which is what is computed here, for many series at
 * once.  Pixel centres (Gorski et al. 2005, ApJ 622, 759): 4 nside - 1
 * rings of constant z; ring i (1-based from the north pole) has
 *   north cap,  i < nside:        z = 1 - i^2 / (3 nside^2), 4 i pixels,
 *                                 phi_j = (j + 1/2) 2 pi / (4 i);
 *   equator, nside <= i <= 3 nside: z = 4/3 - 2 i / (3 nside), 4 nside
 *                                 pixels, phi_j = (j + o) 2 pi / (4 nside),
 *                                 o = 0 when i + nside is odd, else 1/2;
 *   south cap: the mirror of the north cap,
 * pixels numbered ring by ring from the north pole, phi ascending.
 *
 * Kernel: workgroup = one series, its coefficients staged in LDS once;
 * the northern rings (through the equator) in chunks of HP_RC, each with
 * its mirror ring 4 nside - i (same nphi and phi_0, z -> -z).  Per chunk,
 * task (|m|, ring) - rings fastest, so that the lanes of a wave share m
 * and read the same coefficients - runs the Legendre recurrence at the
 * ring's z once for the ring and its mirror (even and odd l - m sums,
 * Pbar_lm(-z) = (-1)^(l+m) Pbar_lm(z)) (float64,
 * from Pbar_mm = cm[m] sin^m theta and the division-free coefficient
 * pairs of the plan), giving F_m and F_-m; with G_m = F_m exp(i m phi_0)
 * the ring's values are
 *   f(phi_0 + 2 pi j / nphi) = Re sum_m G_m exp(i m phi'),  phi' = 2 pi j / nphi,
 *     = G_0.re + sum_{m>0} A_m cos(m phi') - B_m sin(m phi'),
 *   A_m = Re (G_m + G_-m), B_m = Im (G_m - G_-m),
 * kept in LDS per (ring, m); then one thread per four pixels of a ring
 * (a quarter turn apart) sums over m with exp(i m phi') advanced by one
 * complex rotation per m.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sh_series.h>
#include <sphradiometer/sphradiometer_batch.h>
#include <sphradiometer/sphradiometer_gpu.h>
#include "../sphr_internal.h"
#include "kernels.h"
#include "runtime.h"

#define HP_NT 256
#define HP_NSIDE_MAX 4096	/* the per-ring form keeps 4 nside twiddles in LDS */

struct sphr_healpix {
	int L, nside, precision;
	double *d_cf;	/* [nal][2]: alpha_lm, alpha_lm / alpha_{l-1,m} at the (l, m >= 0) slots */
	double *d_cm;	/* [L + 1]: Pbar_mm(cos theta) = cm[m] sin^m theta */
};

/* ring geometry: z, sin theta (without cancellation near the poles), the
 * pixel count, the index of its first pixel and the phase offset phi_0 */
__device__ inline void hp_ring(int nside, int i, double *z, double *st, int *nphi, long *start, double *phi0)
{
	const long N = nside, npix = 12 * N * N, ncap = 2 * N * (N - 1);
	if(i < nside) {
		const double u = (double) i * i / (3. * N * N);	/* 1 - z */
		*z = 1. - u;
		*st = sqrt(u * (2. - u));
		*nphi = 4 * i;
		*start = 2L * i * (i - 1);
		*phi0 = M_PI / (4. * i);
	} else if(i <= 3 * nside) {
		*z = (4. / 3.) - 2. * i / (3. * N);
		*st = sqrt((1. - *z) * (1. + *z));
		*nphi = 4 * nside;
		*start = ncap + (long) (i - nside) * 4 * N;
		*phi0 = ((i + nside) & 1) ? 0. : M_PI / (4. * N);
	} else {
		const int k = 4 * nside - i;
		const double u = (double) k * k / (3. * N * N);
		*z = u - 1.;
		*st = sqrt(u * (2. - u));
		*nphi = 4 * k;
		*start = npix - 2L * k * (k + 1);
		*phi0 = M_PI / (4. * k);
	}
}

#define HP_RC 16		/* rings per chunk */
#define HP_SPW 4		/* series per workgroup of the per-ring form */

/* sin and cos of pi a / b in the map's precision */
__device__ inline void hp_sincospi(int a, int b, double *sn, double *cs)
{
	sincospi((double) a / b, sn, cs);
}
__device__ inline void hp_sincospi(int a, int b, float *sn, float *cs)
{
	sincospif((float) a / (float) b, sn, cs);
}

template <typename T>
__global__ void __launch_bounds__(HP_NT) k_healpix(const cplx<T> *__restrict__ alm, int n, int L, int nside, const double *__restrict__ cf, const double *__restrict__ cm, T *__restrict__ map)
{
	extern __shared__ __align__(16) unsigned char hp_sm[];
	const int s = blockIdx.x, L1 = L + 1, N2 = 2 * nside;
	const int ncoef = L1 * L1;
	const long npix = 12L * nside * nside;
	cplx<T> *a = reinterpret_cast<cplx<T> *>(hp_sm);		/* [ncoef] */
	T *AB = reinterpret_cast<T *>(a + ncoef);			/* [2 HP_RC][L + 1][2]: north rows, then their mirrors */
	/* the chunk's rings: first pixel, pixel count, AB row, prefix sums of
	 * the quarter rings */
	__shared__ long es[2 * HP_RC];
	__shared__ int en[2 * HP_RC], ea[2 * HP_RC], eq[2 * HP_RC + 1], ne_s;
	for(int t = threadIdx.x; t < ncoef; t += HP_NT)
		a[t] = alm[(size_t) s * ncoef + t];
	__syncthreads();
	T *out = map + (size_t) s * npix;
	for(int r0 = 1; r0 <= N2; r0 += HP_RC) {
		const int nr = N2 - r0 + 1 < HP_RC ? N2 - r0 + 1 : HP_RC;
		if(threadIdx.x == 0) {
			int ne = 0;
			eq[0] = 0;
			for(int rr = 0; rr < nr; rr++)
				for(int h = 0; h < 2; h++) {
					/* ring r0 + rr and its mirror 4 nside - r0 - rr (the
					 * equator ring is its own mirror) */
					const int ring = h ? 4 * nside - r0 - rr : r0 + rr;
					if(h && ring == r0 + rr)
						continue;
					double z, st, phi0;
					int nphi;
					long start;
					hp_ring(nside, ring, &z, &st, &nphi, &start, &phi0);
					es[ne] = start;
					en[ne] = nphi;
					ea[ne] = h ? HP_RC + rr : rr;
					eq[ne + 1] = eq[ne] + nphi / 4;
					ne++;
				}
			ne_s = ne;
		}
		/* Legendre sums: task (|m|, northern ring of the chunk), split by
		 * the parity of l - m: Pbar_lm(-z) = (-1)^(l+m) Pbar_lm(z), so the
		 * ring gets E + O and its mirror E - O */
		for(int t = threadIdx.x; t < HP_RC * L1; t += HP_NT) {
			const int m = t / HP_RC, rr = t - m * HP_RC;
			double arn = 0., bin = 0., ars = 0., bis = 0.;
			if(rr < nr) {
				double z, st, phi0;
				int nphi;
				long start;
				hp_ring(nside, r0 + rr, &z, &st, &nphi, &start, &phi0);
				/* Pbar_mm = cm[m] sin^m theta, the power by squaring */
				double p2 = cm[m], b = st;
				for(int e = m; e; e >>= 1) {
					if(e & 1)
						p2 *= b;
					b *= b;
				}
				const cplx<T> *ap = a + lm_offset(L, m, m);
				const cplx<T> *an = m ? a + lm_offset(L, m, -m) : ap;
				/* even (e) and odd (o) l - m sums of +m (p) and -m (n) */
				double epr = (double) ap[0].re * p2, epi = (double) ap[0].im * p2;
				double enr = (double) an[0].re * p2, eni = (double) an[0].im * p2;
				double opr = 0., opi = 0., onr = 0., oni = 0.;
				const double *c = cf + 2 * lm_offset(L, m, m);
				double p1 = 0.;
				if(m < L) {
					p1 = c[2] * z * p2;
					opr = (double) ap[1].re * p1;
					opi = (double) ap[1].im * p1;
					onr = (double) an[1].re * p1;
					oni = (double) an[1].im * p1;
				}
				int d = 2;
				for(; d + 1 <= L - m; d += 2) {
					const double pe = c[2 * d] * z * p1 - c[2 * d + 1] * p2;
					const double po = c[2 * d + 2] * z * pe - c[2 * d + 3] * p1;
					epr = fma((double) ap[d].re, pe, epr);
					epi = fma((double) ap[d].im, pe, epi);
					enr = fma((double) an[d].re, pe, enr);
					eni = fma((double) an[d].im, pe, eni);
					opr = fma((double) ap[d + 1].re, po, opr);
					opi = fma((double) ap[d + 1].im, po, opi);
					onr = fma((double) an[d + 1].re, po, onr);
					oni = fma((double) an[d + 1].im, po, oni);
					p2 = pe;
					p1 = po;
				}
				if(d <= L - m) {
					const double pe = c[2 * d] * z * p1 - c[2 * d + 1] * p2;
					epr = fma((double) ap[d].re, pe, epr);
					epi = fma((double) ap[d].im, pe, epi);
					enr = fma((double) an[d].re, pe, enr);
					eni = fma((double) an[d].im, pe, eni);
				}
				double sn, cs;
				sincos(m * phi0, &sn, &cs);
				/* the ring (h = 0) and its mirror (h = 1): F = E +- O;
				 * Pbar_l,-m = (-1)^m Pbar_lm; G_+-m = F_+-m exp(+-i m phi_0) */
				for(int h = 0; h < 2; h++) {
					const double sg = h ? -1. : 1.;
					const double fpr = epr + sg * opr, fpi = epi + sg * opi;
					double fnr = enr + sg * onr, fni = eni + sg * oni;
					if(m & 1) {
						fnr = -fnr;
						fni = -fni;
					}
					const double gpr = fpr * cs - fpi * sn, gpi = fpr * sn + fpi * cs;
					const double gnr = fnr * cs + fni * sn, gni = -fnr * sn + fni * cs;
					const double ar = m ? gpr + gnr : gpr, bi = m ? gpi - gni : 0.;
					if(h) {
						ars = ar;
						bis = bi;
					} else {
						arn = ar;
						bin = bi;
					}
				}
			}
			AB[2 * (rr * L1 + m)] = (T) arn;
			AB[2 * (rr * L1 + m) + 1] = (T) bin;
			AB[2 * ((HP_RC + rr) * L1 + m)] = (T) ars;
			AB[2 * ((HP_RC + rr) * L1 + m) + 1] = (T) bis;
		}
		__syncthreads();
		/* pixels of the chunk's rings, four at once: nphi is a multiple of
		 * 4, and the pixels j + q nphi / 4 (q = 0..3) sit at phi' + q pi / 2,
		 * where exp(i m (phi' + q pi / 2)) = i^(m q) exp(i m phi').  With
		 * C_m = A_m cos(m phi') - B_m sin(m phi'),
		 * D_m = A_m sin(m phi') + B_m cos(m phi') and S_k, T_k the sums
		 * of C_m, D_m over m = k mod 4:
		 *   q = 0: S0 + S1 + S2 + S3,   q = 1: S0 - T1 - S2 + T3,
		 *   q = 2: S0 - S1 + S2 - S3,   q = 3: S0 + T1 - S2 - T3 */
		const int ne = ne_s;
		/* u only grows, so each thread's ring search resumes where its
		 * last one stopped */
		int k = 0;
		for(int u = threadIdx.x; u < eq[ne]; u += HP_NT) {
			while(k + 1 < ne && u >= eq[k + 1])
				k++;
			const int j = u - eq[k], n4 = en[k] / 4;
			T sn, cs;
			hp_sincospi(2 * j, en[k], &sn, &cs);
			const T *ab = AB + 2 * ea[k] * L1;
			const T er = cs, ei = sn;
			T cr = er, ci = ei;	/* exp(i m phi') for m = 1 */
			T s0 = ab[0], s1 = 0, s2 = 0, s3 = 0, t1 = 0, t3 = 0;
			for(int m = 1; m <= L; m++) {
				const T A = ab[2 * m], B = ab[2 * m + 1];
				const T c = fma(A, cr, -B * ci);
				switch(m & 3) {
				case 0:
					s0 += c;
					break;
				case 1:
					s1 += c;
					t1 = fma(A, ci, fma(B, cr, t1));
					break;
				case 2:
					s2 += c;
					break;
				default:
					s3 += c;
					t3 = fma(A, ci, fma(B, cr, t3));
					break;
				}
				const T nr_ = cr * er - ci * ei;
				ci = fma(cr, ei, ci * er);
				cr = nr_;
			}
			T *o = out + es[k] + j;
			o[0] = (s0 + s2) + (s1 + s3);
			o[n4] = (s0 - s2) - (t1 - t3);
			o[2 * n4] = (s0 + s2) - (s1 + s3);
			o[3 * n4] = (s0 - s2) + (t1 - t3);
		}
		__syncthreads();
	}
}

/* orders whose coefficients do not fit in LDS: workgroup = one ring x
 * HP_SPW series, the coefficients read from global memory, the ring's
 * twiddles in LDS at exact integer indices (m j mod nphi) */
template <typename T>
__global__ void __launch_bounds__(HP_NT) k_healpix_ring(const cplx<T> *__restrict__ alm, int n, int L, int nside, const double *__restrict__ cf, const double *__restrict__ cm, T *__restrict__ map)
{
	extern __shared__ __align__(16) unsigned char hp_sm[];
	const int ring = blockIdx.x + 1, s0 = blockIdx.y * HP_SPW;
	double z, st, phi0;
	int nphi;
	long start;
	hp_ring(nside, ring, &z, &st, &nphi, &start, &phi0);
	const long npix = 12L * nside * nside;
	const int ncoef = (L + 1) * (L + 1);
	T *AB = reinterpret_cast<T *>(hp_sm);				/* [HP_SPW][L + 1][2] */
	T *tw = AB + 2 * HP_SPW * (L + 1);				/* [nphi][2]: cos, sin */
	for(int k = threadIdx.x; k < nphi; k += HP_NT) {
		double sn, cs;
		sincospi(2. * k / nphi, &sn, &cs);
		tw[2 * k] = (T) cs;
		tw[2 * k + 1] = (T) sn;
	}
	/* Legendre sums: thread (series, |m|) */
	for(int t = threadIdx.x; t < HP_SPW * (L + 1); t += HP_NT) {
		const int sl = t / (L + 1), m = t - sl * (L + 1), s = s0 + sl;
		double ar = 0., bi = 0.;
		if(s < n) {
			const cplx<T> *a = alm + (size_t) s * ncoef;
			/* Pbar_mm = cm[m] sin^m theta by repeated products */
			double p2 = cm[m];
			for(int k = 0; k < m; k++)
				p2 *= st;
			const cplx<T> *ap = a + lm_offset(L, m, m);
			const cplx<T> *an = m ? a + lm_offset(L, m, -m) : ap;
			double fpr = (double) ap[0].re * p2, fpi = (double) ap[0].im * p2;
			double fnr = (double) an[0].re * p2, fni = (double) an[0].im * p2;
			const double *c = cf + 2 * lm_offset(L, m, m);
			double p1 = 0.;
			if(m < L) {
				p1 = c[2] * z * p2;
				fpr = fma((double) ap[1].re, p1, fpr);
				fpi = fma((double) ap[1].im, p1, fpi);
				fnr = fma((double) an[1].re, p1, fnr);
				fni = fma((double) an[1].im, p1, fni);
			}
			for(int l = m + 2; l <= L; l++) {
				const int d = l - m;
				const double p = c[2 * d] * z * p1 - c[2 * d + 1] * p2;
				fpr = fma((double) ap[d].re, p, fpr);
				fpi = fma((double) ap[d].im, p, fpi);
				fnr = fma((double) an[d].re, p, fnr);
				fni = fma((double) an[d].im, p, fni);
				p2 = p1;
				p1 = p;
			}
			/* Pbar_l,-m = (-1)^m Pbar_lm; G_+-m = F_+-m exp(+-i m phi_0) */
			if(m & 1) {
				fnr = -fnr;
				fni = -fni;
			}
			double sn, cs;
			sincos(m * phi0, &sn, &cs);
			const double gpr = fpr * cs - fpi * sn, gpi = fpr * sn + fpi * cs;
			const double gnr = fnr * cs + fni * sn, gni = -fnr * sn + fni * cs;
			if(m) {
				ar = gpr + gnr;
				bi = gpi - gni;
			} else {
				ar = gpr;
				bi = 0.;
			}
		}
		AB[2 * t] = (T) ar;
		AB[2 * t + 1] = (T) bi;
	}
	__syncthreads();
	/* pixels: thread (series, j) */
	for(int t = threadIdx.x; t < HP_SPW * nphi; t += HP_NT) {
		const int sl = t / nphi, j = t - sl * nphi, s = s0 + sl;
		if(s >= n)
			continue;
		const T *ab = AB + 2 * sl * (L + 1);
		T v0 = ab[0], v1 = 0;
		int k = j;	/* m j mod nphi for m = 1 */
		for(int m = 1; m <= L; m++) {
			const T c = tw[2 * k], sn = tw[2 * k + 1];
			if(m & 1)
				v1 = fma(ab[2 * m], c, fma(-ab[2 * m + 1], sn, v1));
			else
				v0 = fma(ab[2 * m], c, fma(-ab[2 * m + 1], sn, v0));
			k += j;
			if(k >= nphi)
				k -= nphi;
		}
		map[(size_t) s * npix + start + j] = v0 + v1;
	}
}

/* the posterior of the two beta skies of each event (RapidLocalization.py:
 * 498-530): prob = (e_p / sum e_p + e_n / sum e_n) / 2, e = exp(map -
 * max map); maps [2][E][npix], one workgroup per event */
template <typename T>
__global__ void __launch_bounds__(HP_NT) k_healpix_posterior(const T *__restrict__ maps, int E, long npix, T *__restrict__ prob)
{
	/* exp(x - max) in the maps' precision (float32 maps: float32
	 * exponentials, float64 sums) */
	auto ex = [](T x, double m) -> double { return sizeof(T) == sizeof(float) ? (double) expf((float) ((double) x - m)) : exp((double) x - m); };
	__shared__ double red[2][HP_NT / 64];
	const int e = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const T *mp = maps + (size_t) e * npix, *mn = maps + ((size_t) E + e) * npix;
	double xp = -INFINITY, xn = -INFINITY;
	for(long p = threadIdx.x; p < npix; p += HP_NT) {
		xp = fmax(xp, (double) mp[p]);
		xn = fmax(xn, (double) mn[p]);
	}
	for(int o = 32; o; o >>= 1) {
		xp = fmax(xp, __shfl_xor(xp, o));
		xn = fmax(xn, __shfl_xor(xn, o));
	}
	if(!lane) {
		red[0][wave] = xp;
		red[1][wave] = xn;
	}
	__syncthreads();
	xp = red[0][0];
	xn = red[1][0];
	for(int w = 1; w < HP_NT / 64; w++) {
		xp = fmax(xp, red[0][w]);
		xn = fmax(xn, red[1][w]);
	}
	__syncthreads();
	double sp = 0., sn = 0.;
	for(long p = threadIdx.x; p < npix; p += HP_NT) {
		sp += ex(mp[p], xp);
		sn += ex(mn[p], xn);
	}
	for(int o = 32; o; o >>= 1) {
		sp += __shfl_xor(sp, o);
		sn += __shfl_xor(sn, o);
	}
	if(!lane) {
		red[0][wave] = sp;
		red[1][wave] = sn;
	}
	__syncthreads();
	sp = 0.;
	sn = 0.;
	for(int w = 0; w < HP_NT / 64; w++) {
		sp += red[0][w];
		sn += red[1][w];
	}
	const double ip = 0.5 / sp, in = 0.5 / sn;
	T *pr = prob + (size_t) e * npix;
	for(long p = threadIdx.x; p < npix; p += HP_NT)
		pr[p] = (T) (ex(mp[p], xp) * ip + ex(mn[p], xn) * in);
}

/* ---------------------------------------------------------------- C ABI */

extern "C" int sphr_healpix_nside(unsigned l_max)
{
	/* lmax2nside, sh_series_healpix.cc:196-238 */
	const double x = ceil(l_max / 2. + 0.75);
	for(int i = 1; i > 0; i <<= 1)
		if(i >= x)
			return i;
	return -1;
}

extern "C" struct sphr_healpix *sphr_healpix_new(unsigned l_max, int nside, int precision)
{
	if(sphr_gpu_ensure())
		return NULL;
	if(nside < 1 || nside > HP_NSIDE_MAX || (precision != SPHR_FP64 && precision != SPHR_MIXED) || l_max > 4096) {
		sphr_set_error("sphr_healpix_new: bad arguments (l_max %u, nside %d, precision %d)", l_max, nside, precision);
		return NULL;
	}
	const int L = (int) l_max, nal = (L + 1) * (L + 2) / 2;
	struct sphr_healpix *h = (struct sphr_healpix *) calloc(1, sizeof(*h));
	double *hc = (double *) calloc(2 * (size_t) nal + L + 1, sizeof(double));
	if(!h || !hc) {
		free(h);
		free(hc);
		sphr_set_error("sphr_healpix_new: out of memory");
		return NULL;
	}
	h->L = L;
	h->nside = nside;
	h->precision = precision;
	/* alpha_lm = sqrt((4 l^2 - 1) / (l^2 - m^2)) and its ratio to
	 * alpha_{l-1,m} at the (l, m) slot; Pbar_mm at the equator */
	for(int m = 0; m <= L; m++) {
		double prev = 1.;
		for(int l = m + 1; l <= L; l++) {
			const double a = sqrt((4. * l * l - 1.) / ((double) (l - m) * (l + m)));
			const int o = lm_offset(L, l, m);
			hc[2 * o] = a;
			hc[2 * o + 1] = l > m + 1 ? a / prev : 0.;
			prev = a;
		}
		double one;
		sphr_plm(m, m, 0.0, &one);
		hc[2 * (size_t) nal + m] = one;
	}
	if(sphr_rt_check(hipMalloc(&h->d_cf, (2 * (size_t) nal + L + 1) * sizeof(double)), "hipMalloc")
	   || sphr_rt_check(hipMemcpy(h->d_cf, hc, (2 * (size_t) nal + L + 1) * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy")) {
		free(hc);
		sphr_healpix_free(h);
		return NULL;
	}
	h->d_cm = h->d_cf + 2 * (size_t) nal;
	free(hc);
	return h;
}

extern "C" void sphr_healpix_free(struct sphr_healpix *h)
{
	if(!h)
		return;
	if(h->d_cf)
		hipFree(h->d_cf);
	free(h);
}

extern "C" long sphr_healpix_npix(const struct sphr_healpix *h)
{
	return h ? 12L * h->nside * h->nside : -1;
}

extern "C" int sphr_healpix_maps(const struct sphr_healpix *h, const void *alm, int n, void *map, void *stream)
{
	if(!h || !alm || !map || n < 0) {
		sphr_set_error("sphr_healpix_maps: bad arguments");
		return -1;
	}
	if(!n)
		return 0;
	hipStream_t st = stream ? (hipStream_t) stream : sphr_rt_stream();
	const size_t esz = h->precision == SPHR_MIXED ? sizeof(float) : sizeof(double);
	const size_t lds = (2 * (size_t) (h->L + 1) * (h->L + 1) + 4 * HP_RC * (size_t) (h->L + 1)) * esz;
	if(lds <= 150 * 1024) {
		/* one workgroup per series, its coefficients in LDS */
		if(h->precision == SPHR_MIXED)
			hipLaunchKernelGGL(k_healpix<float>, dim3(n), dim3(HP_NT), lds, st, (const cplx<float> *) alm, n, h->L, h->nside, h->d_cf, h->d_cm, (float *) map);
		else
			hipLaunchKernelGGL(k_healpix<double>, dim3(n), dim3(HP_NT), lds, st, (const cplx<double> *) alm, n, h->L, h->nside, h->d_cf, h->d_cm, (double *) map);
	} else {
		/* one workgroup per (ring, HP_SPW series) */
		const size_t lr = (2 * HP_SPW * (size_t) (h->L + 1) + 2 * 4 * (size_t) h->nside) * esz;
		if(lr > 160 * 1024) {
			sphr_set_error("sphr_healpix_maps: l_max %d with nside %d needs %zu bytes of LDS", h->L, h->nside, lr);
			return -1;
		}
		const dim3 grid(4 * h->nside - 1, (n + HP_SPW - 1) / HP_SPW);
		if(h->precision == SPHR_MIXED)
			hipLaunchKernelGGL(k_healpix_ring<float>, grid, dim3(HP_NT), lr, st, (const cplx<float> *) alm, n, h->L, h->nside, h->d_cf, h->d_cm, (float *) map);
		else
			hipLaunchKernelGGL(k_healpix_ring<double>, grid, dim3(HP_NT), lr, st, (const cplx<double> *) alm, n, h->L, h->nside, h->d_cf, h->d_cm, (double *) map);
	}
	return sphr_rt_check(hipGetLastError(), "k_healpix") ? -1 : 0;
}

extern "C" int sphr_healpix_posterior(const struct sphr_healpix *h, const void *maps, int n_events, void *prob, void *stream)
{
	if(!h || !maps || !prob || n_events < 0) {
		sphr_set_error("sphr_healpix_posterior: bad arguments");
		return -1;
	}
	if(!n_events)
		return 0;
	hipStream_t st = stream ? (hipStream_t) stream : sphr_rt_stream();
	const long npix = 12L * h->nside * h->nside;
	if(h->precision == SPHR_MIXED)
		hipLaunchKernelGGL(k_healpix_posterior<float>, dim3(n_events), dim3(HP_NT), 0, st, (const float *) maps, n_events, npix, (float *) prob);
	else
		hipLaunchKernelGGL(k_healpix_posterior<double>, dim3(n_events), dim3(HP_NT), 0, st, (const double *) maps, n_events, npix, (double *) prob);
	return sphr_rt_check(hipGetLastError(), "k_healpix_posterior") ? -1 : 0;
}

/* sh_series_write_healpix_map (sh_series_healpix.cc:247-286): the map of
 * Re f at nside = lmax2nside(l_max), RING order, to a FITS file */
extern "C" int sh_series_write_healpix_map(const struct sh_series *series, const char *filename)
{
	if(!series || !filename) {
		sphr_set_error("sh_series_write_healpix_map: bad arguments");
		return -1;
	}
	const int L = (int) series->l_max, nside = sphr_healpix_nside(series->l_max);
	if(nside < 1)
		return -1;
	const size_t ncoef = (size_t) (L + 1) * (L + 1);
	const long npix = 12L * nside * nside;
	double _Complex *full = (double _Complex *) calloc(ncoef, sizeof(*full));
	double *hmap = (double *) malloc(npix * sizeof(double));
	struct sphr_healpix *h = NULL;
	void *d_alm = NULL, *d_map = NULL;
	int rc = -1;
	if(!full || !hmap) {
		sphr_set_error("sh_series_write_healpix_map: out of memory");
		goto done;
	}
	/* the full coefficient layout (a polar series has m = 0 only) */
	for(int l = 0; l <= L; l++)
		for(int m = series->polar ? 0 : -l; m <= (series->polar ? 0 : l); m++)
			full[lm_offset(L, l, m)] = sh_series_get(series, l, m);
	h = sphr_healpix_new(series->l_max, nside, SPHR_FP64);
	if(!h || sphr_rt_check(hipMalloc(&d_alm, ncoef * sizeof(*full)), "hipMalloc") || sphr_rt_check(hipMalloc(&d_map, npix * sizeof(double)), "hipMalloc")
	   || sphr_memcpy_htod(d_alm, full, ncoef * sizeof(*full)) || sphr_healpix_maps(h, d_alm, 1, d_map, NULL)
	   || sphr_memcpy_dtoh(hmap, d_map, npix * sizeof(double)))
		goto done;
	rc = sphr_fits_write_healpix_map(filename, hmap, nside);
done:
	if(d_alm)
		hipFree(d_alm);
	if(d_map)
		hipFree(d_map);
	sphr_healpix_free(h);
	free(full);
	free(hmap);
	return rc;
}
