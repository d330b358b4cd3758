/*
 * Device-side helpers shared by the HIP sources: a small complex type
 * templated on the real type, and error plumbing.
 */
#ifndef SPHR_GPU_COMMON_H
#define SPHR_GPU_COMMON_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

template <typename T> struct cplx {
	T re, im;
};

template <typename T> __host__ __device__ inline cplx<T> mk(T r, T i)
{
	cplx<T> c;
	c.re = r;
	c.im = i;
	return c;
}
template <typename T> __host__ __device__ inline cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T> __host__ __device__ inline cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T> __host__ __device__ inline cplx<T> operator*(cplx<T> a, cplx<T> b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
template <typename T> __host__ __device__ inline cplx<T> operator*(cplx<T> a, T s) { return mk(a.re * s, a.im * s); }
template <typename T> __host__ __device__ inline cplx<T> conjg(cplx<T> a) { return mk(a.re, -a.im); }
/* a += b * c */
template <typename T> __host__ __device__ inline void cfma(cplx<T> &a, cplx<T> b, cplx<T> c)
{
	a.re = fma(b.re, c.re, fma(-b.im, c.im, a.re));
	a.im = fma(b.re, c.im, fma(b.im, c.re, a.im));
}
/* a += b * s, s real */
template <typename T> __host__ __device__ inline void cfma(cplx<T> &a, cplx<T> b, T s)
{
	a.re = fma(b.re, s, a.re);
	a.im = fma(b.im, s, a.im);
}

/* offset of (l, m) in a series of order L (the reference layout) */
__host__ __device__ inline int lm_offset(int L, int l, int m)
{
	const int L1 = L + 1;
	const int x = m * (m - 1);
	const int base = m >= 0 ? L1 * m - x / 2 : L1 * (L1 + m) + x / 2;
	return base + l - (m < 0 ? -m : m);
}

/* the order m of coefficient idx in a series of order L (closed form of
 * lm_decode's m: the block of idx from the quadratic of the block starts,
 * then corrected by whole blocks) */
__host__ __device__ inline int lm_order(int L, int idx)
{
	const int L1 = L + 1;
	const int npos = L1 * (L1 + 1) / 2;
	if(idx < npos) {
		/* m >= 0: block mm starts at L1 mm - mm (mm - 1) / 2 */
		const float b = 2.0f * L1 + 1.0f;
		int mm = (int) ((b - sqrtf(b * b - 8.0f * idx)) * 0.5f);
		mm = mm < 0 ? 0 : mm > L ? L : mm;
		while(mm < L && L1 * (mm + 1) - (mm + 1) * mm / 2 <= idx)
			mm++;
		while(mm > 0 && L1 * mm - mm * (mm - 1) / 2 > idx)
			mm--;
		return mm;
	}
	/* m = -L + k, k < L: block k starts k (k + 1) / 2 past npos */
	const int rem = idx - npos;
	int k = (int) ((sqrtf(8.0f * rem + 1.0f) - 1.0f) * 0.5f);
	k = k < 0 ? 0 : k > L - 1 ? L - 1 : k;
	while(k < L - 1 && (k + 1) * (k + 2) / 2 <= rem)
		k++;
	while(k > 0 && k * (k + 1) / 2 > rem)
		k--;
	return -L + k;
}

/* inverse of lm_offset: decode index -> (l, m) */
__host__ __device__ inline void lm_decode(int L, int idx, int *l, int *m)
{
	const int L1 = L + 1;
	const int npos = L1 * (L1 + 1) / 2;
	if(idx < npos) {
		int mm = 0;
		while(L1 * (mm + 1) - (mm + 1) * mm / 2 <= idx)
			mm++;
		*m = mm;
		*l = idx - (L1 * mm - mm * (mm - 1) / 2) + mm;
	} else {
		/* blocks m = -L .. -1 follow */
		int rem = idx - npos;
		int mm = -L;
		while(rem >= L1 + mm) {
			rem -= L1 + mm;
			mm++;
		}
		*m = mm;
		*l = rem - mm;
	}
}

#endif
