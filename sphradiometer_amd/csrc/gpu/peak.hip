/*
 * Extremum of a real function on the sphere given by its SH series
 * (sh_series_real_minimum, sh_series_extremize.c:273-360), batched: one
 * wavefront per series.
 *
 * Seed: either the reference's (sh_series_extremize.c:168-241): the
 * series low-passed to l = l_seed, sampled on its 2(l_seed+1)-ring
 * Gauss-Legendre mesh, the extreme pixel; or the extreme pixel of a mesh
 * the caller already synthesised at full order.
 *
 * Refinement: Newton steps in the tangent plane at the current point p,
 * basis (theta-hat, phi-hat), with the gradient and Hessian from central
 * differences of the exact series, a step length capped at
 * 0.05 rad, and a gradient step when the Hessian is not definite in the
 * wanted direction.  The reference refines with GSL's BFGS2; the fixed
 * point is the same stationary point, the iterates are not.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include "kernels.h"

struct peak_tables {
	const double *cm;	/* L+1: Pbar_mm = cm[m] sin^m theta */
	const double *alpha;	/* packed alpha_lm of the l recurrence (order L) */
	const double *seed_x;	/* seed mesh rings, cos theta descending */
	int n_seed;		/* seed mesh size (rings = columns) */
};

/* sum over the wavefront; every lane returns lane 0's total so control
 * flow that depends on it stays uniform */
__device__ inline double wave_sum(double v)
{
	for(int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off);
	return __shfl(v, 0);
}

/* Re f(theta, phi) for the series truncated at Lcut, lanes over m,
 * reduced across the wavefront */
template <typename TI>
__device__ double eval_wave(const cplx<TI> *a, int L, int Lcut, double theta, double phi, const struct peak_tables &t)
{
	const int lane = threadIdx.x & 63;
	const double x = cos(theta), s = sin(theta);
	double acc = 0.0;
	for(int m = lane; m <= Lcut; m += 64) {
		double pprev = 0.0, pcur = t.cm[m] * (m ? pow(s, (double) m) : 1.0), aprev = 1.0;
		double spr = 0, spi = 0, smr = 0, smi = 0;
		for(int l = m; l <= Lcut; l++) {
			if(l > m) {
				const double al = t.alpha[lm_offset(L, l, m)];
				const double pn = (l == m + 1) ? al * x * pcur : al * (x * pcur - pprev / aprev);
				pprev = pcur;
				pcur = pn;
				aprev = al;
			}
			const cplx<TI> cp = a[lm_offset(L, l, m)];
			spr += (double) cp.re * pcur;
			spi += (double) cp.im * pcur;
			if(m) {
				const cplx<TI> cn = a[lm_offset(L, l, -m)];
				smr += (double) cn.re * pcur;
				smi += (double) cn.im * pcur;
			}
		}
		double sn, cs;
		sincos((double) m * phi, &sn, &cs);
		acc += cs * spr - sn * spi;
		if(m)
			acc += ((m & 1) ? -1.0 : 1.0) * (cs * smr + sn * smi);
	}
	return wave_sum(acc);
}

__device__ inline void condition(double *theta, double *phi)
{
	/* theta into [0, pi] (flipping phi), phi into [0, 2 pi):
	 * sh_series_extremize.c:63-87 */
	double t = fmod(*theta, 2 * M_PI);
	if(t < 0)
		t += 2 * M_PI;
	double p = *phi;
	if(t > M_PI) {
		t = 2 * M_PI - t;
		p += M_PI;
	}
	p = fmod(p, 2 * M_PI);
	if(p < 0)
		p += 2 * M_PI;
	*theta = t;
	*phi = p;
}

/* point at tangent offset (d1, d2) from p in the (theta-hat, phi-hat)
 * basis, projected back onto the sphere */
__device__ inline void offset_point(double th, double ph, double d1, double d2, double *to, double *po)
{
	const double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
	const double x = st * cp + d1 * ct * cp - d2 * sp;
	const double y = st * sp + d1 * ct * sp + d2 * cp;
	const double z = ct - d1 * st;
	*to = atan2(sqrt(x * x + y * y), z);
	double p = atan2(y, x);
	if(p < 0)
		p += 2 * M_PI;
	*po = p;
}

template <typename TI>
__global__ void __launch_bounds__(64) k_extremize(const cplx<TI> *__restrict__ alm, int alm_div, int L, int sign, int l_seed, const cplx<TI> *__restrict__ seed_map, const double *__restrict__ map_x, const double *__restrict__ seed_dir, struct peak_tables t, int max_iter, double *__restrict__ theta_out, double *__restrict__ phi_out, double *__restrict__ value_out)
{
	const int s = blockIdx.x, lane = threadIdx.x;
	const int ncoef = (L + 1) * (L + 1);
	const cplx<TI> *a = alm + (size_t) (s / alm_div) * ncoef;
	double th, ph;
	if(seed_dir) {
		th = seed_dir[2 * (size_t) s];
		ph = seed_dir[2 * (size_t) s + 1];
	} else {
	/* seed: the extreme of sign * Re f over a mesh, first in row-major
	 * order on ties (sh_series_extremize.c:214-230) */
	double best = INFINITY;
	int bi = 0x7fffffff;
	const int lc = l_seed < L ? l_seed : L;
	const int n = seed_map ? 2 * (L + 1) : t.n_seed;
	for(int i = lane; i < n * n; i += 64) {
		double v;
		if(seed_map)
			v = sign * (double) seed_map[(size_t) s * n * n + i].re;
		else {
			/* per-lane scalar evaluation of the low-passed series */
			const int j = i / n, k = i - j * n;
			const double x = t.seed_x[j], st = sqrt((1.0 - x) * (1.0 + x)), ph = 2.0 * M_PI * k / n;
			double f = 0.0;
			for(int m = 0; m <= lc; m++) {
				double pprev = 0.0, pcur = t.cm[m] * (m ? pow(st, (double) m) : 1.0), aprev = 1.0;
				double spr = 0, spi = 0, smr = 0, smi = 0;
				for(int l = m; l <= lc; l++) {
					if(l > m) {
						const double al = t.alpha[lm_offset(L, l, m)];
						const double pn = (l == m + 1) ? al * x * pcur : al * (x * pcur - pprev / aprev);
						pprev = pcur;
						pcur = pn;
						aprev = al;
					}
					const cplx<TI> cp = a[lm_offset(L, l, m)];
					spr += (double) cp.re * pcur;
					spi += (double) cp.im * pcur;
					if(m) {
						const cplx<TI> cn = a[lm_offset(L, l, -m)];
						smr += (double) cn.re * pcur;
						smi += (double) cn.im * pcur;
					}
				}
				double sn, cs;
				sincos((double) m * ph, &sn, &cs);
				f += cs * spr - sn * spi;
				if(m)
					f += ((m & 1) ? -1.0 : 1.0) * (cs * smr + sn * smi);
			}
			v = sign * f;
		}
		if(v < best) {
			best = v;
			bi = i;
		}
	}
	/* arg-extreme across the wavefront, lowest pixel index on ties */
	for(int off = 32; off > 0; off >>= 1) {
		const double ob = __shfl_xor(best, off);
		const int oi = __shfl_xor(bi, off);
		if(ob < best || (ob == best && oi < bi)) {
			best = ob;
			bi = oi;
		}
	}
	bi = __shfl(bi, 0);
	{
		const int j = bi / n, k = bi - j * n;
		double x;
		if(seed_map) {
			const int B = L + 1;
			x = j < B ? map_x[j] : -map_x[2 * B - 1 - j];
		} else
			x = t.seed_x[j];
		th = acos(x);
		ph = 2.0 * M_PI * k / n;
	}
	}
	/* Newton in the tangent plane on g(p) = sign * Re f(p).  The
	 * gradient comes from central differences at step hg (truncation
	 * error hg^2 f'''/6 and rounding error eps |f| / hg, both near 1e-10
	 * of the function's scale), the Hessian from step h. */
	const double h = 1e-4, hg = 1e-6;
	const double off[12][2] = {
		{ h, 0 }, { -h, 0 }, { 0, h }, { 0, -h },
		{ h, h }, { h, -h }, { -h, h }, { -h, -h },
		{ hg, 0 }, { -hg, 0 }, { 0, hg }, { 0, -hg },
	};
	double f0 = sign * eval_wave(a, L, L, th, ph, t);
	for(int it = 0; it < max_iter; it++) {
		double fv[12];
		for(int q = 0; q < 12; q++) {
			double tt, pp;
			offset_point(th, ph, off[q][0], off[q][1], &tt, &pp);
			fv[q] = sign * eval_wave(a, L, L, tt, pp, t);
		}
		const double g1 = (fv[8] - fv[9]) / (2 * hg), g2 = (fv[10] - fv[11]) / (2 * hg);
		const double h11 = (fv[0] - 2 * f0 + fv[1]) / (h * h), h22 = (fv[2] - 2 * f0 + fv[3]) / (h * h);
		const double h12 = (fv[4] - fv[5] - fv[6] + fv[7]) / (4 * h * h);
		/* eigenvalues of the Hessian; the step solves (H + mu I) d = -g
		 * with mu lifting the smallest eigenvalue to 1e-6 of the largest,
		 * so flat directions (a ring of minima of an axisymmetric
		 * function) take no large steps and negative curvature is
		 * followed downhill */
		const double tr2 = 0.5 * (h11 + h22), dh = 0.5 * (h11 - h22);
		const double disc = sqrt(dh * dh + h12 * h12);
		const double lmin = tr2 - disc, lmax = tr2 + disc;
		const double gn = sqrt(g1 * g1 + g2 * g2);
		if(it > 0 && gn <= 1e-10 * (fabs(f0) + fabs(lmax)))
			break;
		double d1, d2;
		if(lmax > 0) {
			const double floor_ = 1e-6 * lmax;
			const double mu = lmin < floor_ ? floor_ - lmin : 0.0;
			const double a11 = h11 + mu, a22 = h22 + mu;
			const double det = a11 * a22 - h12 * h12;
			d1 = -(a22 * g1 - h12 * g2) / det;
			d2 = -(-h12 * g1 + a11 * g2) / det;
		} else {
			/* concave both ways: steepest descent */
			d1 = -0.02 * g1 / (gn + 1e-300);
			d2 = -0.02 * g2 / (gn + 1e-300);
		}
		const double dn = sqrt(d1 * d1 + d2 * d2);
		if(dn > 0.05) {
			d1 *= 0.05 / dn;
			d2 *= 0.05 / dn;
		}
		/* accept only steps that do not increase g; halve otherwise */
		double nth = th, nph = ph, nf = f0;
		double scale = 1.0;
		for(int bt = 0; bt < 8; bt++) {
			offset_point(th, ph, scale * d1, scale * d2, &nth, &nph);
			nf = sign * eval_wave(a, L, L, nth, nph, t);
			if(nf <= f0)
				break;
			scale *= 0.5;
		}
		if(nf > f0)
			break;
		th = nth;
		ph = nph;
		f0 = nf;
		if(scale * dn < 1e-11)
			break;
	}
	condition(&th, &ph);
	if(lane == 0) {
		theta_out[s] = th;
		phi_out[s] = ph;
		value_out[s] = sign * f0;
	}
}

template <typename TI>
int launch_extremize(const cplx<TI> *alm, int S, int L, int find_max, int l_seed, const cplx<TI> *seed_map, const double *map_x, const double *cm, const double *alpha, const double *seed_x, int n_seed, int max_iter, double *theta, double *phi, double *value, hipStream_t st)
{
	struct peak_tables t;
	t.cm = cm;
	t.alpha = alpha;
	t.seed_x = seed_x;
	t.n_seed = n_seed;
	hipLaunchKernelGGL((k_extremize<TI>), dim3(S), dim3(64), 0, st, alm, 1, L, find_max ? -1 : 1, l_seed, seed_map, map_x, (const double *) NULL, t, max_iter, theta, phi, value);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename TI>
int launch_extremize_seeded(const cplx<TI> *alm, int S, int alm_div, int L, int find_max, const double *seeds, const double *cm, const double *alpha, int max_iter, double *theta, double *phi, double *value, hipStream_t st)
{
	struct peak_tables t;
	t.cm = cm;
	t.alpha = alpha;
	t.seed_x = NULL;
	t.n_seed = 0;
	hipLaunchKernelGGL((k_extremize<TI>), dim3(S), dim3(64), 0, st, alm, alm_div, L, find_max ? -1 : 1, 0, (const cplx<TI> *) NULL, (const double *) NULL, seeds, t, max_iter, theta, phi, value);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_extremize<double>(const cplx<double> *, int, int, int, int, const cplx<double> *, const double *, const double *, const double *, const double *, int, int, double *, double *, double *, hipStream_t);
template int launch_extremize<float>(const cplx<float> *, int, int, int, int, const cplx<float> *, const double *, const double *, const double *, const double *, int, int, double *, double *, double *, hipStream_t);
template int launch_extremize_seeded<double>(const cplx<double> *, int, int, int, int, const double *, const double *, const double *, int, double *, double *, double *, hipStream_t);
template int launch_extremize_seeded<float>(const cplx<float> *, int, int, int, int, const double *, const double *, const double *, int, double *, double *, double *, hipStream_t);
