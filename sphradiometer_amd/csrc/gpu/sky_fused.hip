/*
 * Sky stage of the batched path in one kernel, for the orders the batched
 * cross-power skies run at (compile-time L): K-split partials -> a_lm
 * (rotation rows, baseline average, rotate_z) -> Legendre ring sums F in
 * LDS -> phi DFT on MFMA -> mesh.  F never leaves the chip.
 *
 * phi DFT with the half-turn split.  The mesh has n = 2(L+1) points
 * phi_k = 2 pi k / n, so phi_{k + n/2} = phi_k + pi and
 * exp(i m phi_{k+n/2}) = (-1)^m exp(i m phi_k).  With Ev[r][k] the sum over
 * even m and Od[r][k] over odd m of F[r][m] exp(i m phi_k), k < n/2:
 *     map[r][k] = Ev + Od,      map[r][k + n/2] = Ev - Od,
 * which halves the DFT work.  Ev and Od are real-expanded GEMMs
 * (rows r = rings, K = (re, im) of the modes of one parity, columns
 * (re, im) of the half-turn points) on v_mfma_{f32,f64}_16x16x4.  Each of
 * the 4 waves owns row tiles of 16 rings; the columns that do not fill a
 * 16-wide tile (one point for L = 32) and the rings beyond the last full
 * row tile (two for L = 32) are summed on the VALU.
 *
 * B operands (the DFT matrix) come from the n-entry twiddle table
 * exp(2 pi i t / n), t = m k mod n.  In float32 each wave keeps its B
 * fragments in registers for the life of the workgroup (the workgroups
 * are persistent over segments); in float64 they are formed in the loop,
 * hidden behind the longer f64 MFMA.
 *
 * Legendre stage.  F[r][m] = sum_l a_lm Pbar_l|m|(x_r) for ring pairs
 * (x_j, -x_j) at once through the parity (-1)^(l+|m|).  Item (j, a)
 * handles |m| = a and |m| = L - a, so every item runs L + 2 recurrence
 * steps (balanced lanes), each step serving m = +|m| and -|m|.  The
 * recurrence Pbar_l = alpha_l x Pbar_{l-1} - (alpha_l / alpha_{l-1})
 * Pbar_{l-2} uses a division-free coefficient table built once per
 * workgroup.
 *
 * Layout (reference sh_series / mesh conventions, sh_series.c:33-80,
 * sh_series_to_mesh in sh_series_trans.c): a_lm in the m = 0, 1..L,
 * -L..-1 block order; mesh [2B][2B] complex, rings in descending cos theta,
 * points phi_k = 2 pi k / n.
 */
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace {

#define SF_NT 256

template <typename T> struct sf_mf;
template <> struct sf_mf<float> {
	typedef float v4 __attribute__((ext_vector_type(4)));
	static __device__ inline v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
	static __device__ inline int row(int q, int e) { return 4 * q + e; }
	enum { BREG = 1 };
};
template <> struct sf_mf<double> {
	typedef double v4 __attribute__((ext_vector_type(4)));
	static __device__ inline v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
	static __device__ inline int row(int q, int e) { return q + 4 * e; }
	enum { BREG = 0 };
};

/* compile-time geometry of order L */
template <int L> struct sf_geom {
	enum {
		B = L + 1,		/* rings per hemisphere, points per half turn */
		n = 2 * B,		/* rings and points */
		nm = 2 * L + 1,		/* orders m = -L..L */
		ncoef = (L + 1) * (L + 1),
		nal = (L + 1) * (L + 2) / 2,
		NE = (L % 2 == 0) ? L + 1 : L,	/* even orders */
		NO = nm - NE,			/* odd orders */
		ME0 = (L % 2 == 0) ? -L : -L + 1,	/* smallest even m */
		MO0 = (L % 2 == 0) ? -L + 1 : -L,	/* smallest odd m */
		KSE = (2 * NE + 3) / 4,	/* MFMA k-steps, even / odd */
		KSO = (2 * NO + 3) / 4,
		KMAX = 4 * (KSE > KSO ? KSE : KSO),
		/* row stride of the F tiles in LDS: >= KMAX and = 4 mod 32
		 * words, so the 16 rows x 4 columns of an A fragment fall on
		 * distinct banks */
		FS = ((KMAX + 27) / 32) * 32 + 4,
		NRT = n / 16,		/* full row tiles */
		XR = n - 16 * NRT,	/* rings summed on the VALU */
		NCT = (2 * B) / 16,	/* full column tiles of the half turn */
		XP = B - 8 * NCT,	/* half-turn points summed on the VALU */
		NA = L / 2 + 1,		/* Legendre item pairs |m| = a, L - a */
	};
};

/* element (kr, c) of the real-expanded DFT matrix of one parity:
 * kr = 2 q + (0: re, 1: im of mode q), c = 2 k + (0: re, 1: im of point k) */
template <typename T, int L>
__device__ inline T sf_bval(const cplx<T> *tw, int m0, int nq, int kr, int c)
{
	typedef sf_geom<L> G;
	const int q = kr >> 1;
	if(q >= nq)
		return (T) 0;
	const int m = m0 + 2 * q, k = c >> 1;
	int t = (m * k) % G::n;
	if(t < 0)
		t += G::n;
	const cplx<T> w = tw[t];
	const int ci = kr & 1, d = c & 1;
	return ci == d ? w.re : (ci == 0 ? w.im : -w.im);
}

template <typename T, int L>
__global__ void __launch_bounds__(SF_NT) k_sky_fused(const T *__restrict__ partial, int nsplit, int S, struct xproj_geom xg, const cplx<T> *__restrict__ R, const int2 *__restrict__ lm, const double *__restrict__ gmst, struct synth_tables tb, cplx<T> *__restrict__ alm, cplx<T> *__restrict__ map)
{
	typedef sf_geom<L> G;
	typedef typename sf_mf<T>::v4 v4;
	__shared__ T Fe[G::n * G::FS];		/* even orders [ring][2 q + re/im] */
	__shared__ T Fo[G::n * G::FS];		/* odd orders */
	__shared__ cplx<T> sa[G::ncoef];	/* a_lm of the segment */
	__shared__ T ca[G::nal], cb[G::nal];	/* alpha_l, alpha_l / alpha_{l-1} */
	__shared__ cplx<T> tw[G::n];		/* exp(2 pi i t / n) */
	__shared__ cplx<T> rot[L + 1];		/* exp(-i m gmst) */
	extern __shared__ unsigned char sf_dyn[];
	T *pw = reinterpret_cast<T *>(sf_dyn);	/* [CT] partial sums */

	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int li = lane & 15, lk = lane >> 4;

	/* once per workgroup: tables and the zero padding of the F tiles */
	for(int t = tid; t < G::n * G::FS; t += SF_NT) {
		Fe[t] = (T) 0;
		Fo[t] = (T) 0;
	}
	for(int t = tid; t < G::nal; t += SF_NT) {
		int l, m;
		lm_decode(L, t, &l, &m);
		const double a = l > m ? tb.alpha[t] : 0.0;
		ca[t] = (T) a;
		cb[t] = (T) (l > m + 1 ? a / tb.alpha[t - 1] : 0.0);
	}
	for(int t = tid; t < G::n; t += SF_NT)
		tw[t] = mk((T) tb.twp[2 * t], (T) tb.twp[2 * t + 1]);
	__syncthreads();

	/* B fragments of this lane: column tile ct, k-step ks */
	T bfe[sf_mf<T>::BREG ? G::KSE : 1][G::NCT], bfo[sf_mf<T>::BREG ? G::KSO : 1][G::NCT];
	if(sf_mf<T>::BREG) {
#pragma unroll
		for(int ks = 0; ks < (sf_mf<T>::BREG ? G::KSE : 1); ks++)
#pragma unroll
			for(int ct = 0; ct < G::NCT; ct++)
				bfe[ks][ct] = sf_bval<T, L>(tw, G::ME0, G::NE, 4 * ks + lk, 16 * ct + li);
#pragma unroll
		for(int ks = 0; ks < (sf_mf<T>::BREG ? G::KSO : 1); ks++)
#pragma unroll
			for(int ct = 0; ct < G::NCT; ct++)
				bfo[ks][ct] = sf_bval<T, L>(tw, G::MO0, G::NO, 4 * ks + lk, 16 * ct + li);
	}

	for(int s = blockIdx.x; s < S; s += gridDim.x) {
		/* ---- partial sums over the K splits, rotate_z phases ---- */
		for(int c = tid; c < xg.CT; c += SF_NT) {
			T v = 0;
			for(int sp = 0; sp < nsplit; sp++)
				v += partial[((size_t) sp * S + s) * xg.CT + c];
			pw[c] = v;
		}
		if(tid <= L) {
			double sn = 0.0, cs = 1.0;
			if(gmst && tid)
				sincos(-(double) tid * gmst[s], &sn, &cs);
			rot[tid] = mk((T) cs, (T) sn);
		}
		__syncthreads();

		/* ---- a_lm: rotation rows over the baselines, rotate_z ---- */
		for(int idx = tid; idx < G::ncoef; idx += SF_NT) {
			const int2 q = lm[idx];
			const int l = q.x, m = q.y;
			cplx<T> acc = mk((T) 0, (T) 0);
			for(int b = 0; b < xg.nb; b++) {
				if(l > xg.Lb[b])
					continue;
				const T *p = pw + xg.col_start[b] + 2 * l;
				cfma(acc, R[(size_t) b * G::ncoef + idx], mk(p[0], p[1]));
			}
			if(gmst && m)
				acc = acc * (m > 0 ? rot[m] : conjg(rot[-m]));
			sa[idx] = acc;
			alm[(size_t) s * G::ncoef + idx] = acc;
		}
		__syncthreads();
		if(!map)
			continue;

		/* ---- Legendre ring sums into the F tiles ---- */
		for(int it = tid; it < G::B * G::NA; it += SF_NT) {
			const int j = it % G::B, a = it / G::B;
			const T x = (T) tb.x[j];
#pragma unroll
			for(int h = 0; h < 2; h++) {
				const int am = h ? L - a : a;
				if(h && am == a)
					break;
				const int abase = lm_offset(L, am, am);	/* = coefficient table slot */
				const int anbase = am ? lm_offset(L, am, -am) : 0;
				T pprev = 0, pcur = (T) tb.smm[(size_t) j * (L + 1) + am];
				cplx<T> Ep = mk((T) 0, (T) 0), Op = Ep, En = Ep, On = Ep;
				for(int l = am; l <= L; l++) {
					if(l > am) {
						const T pn = ca[abase + l - am] * x * pcur - cb[abase + l - am] * pprev;
						pprev = pcur;
						pcur = pn;
					}
					const cplx<T> cp = sa[abase + l - am];
					const cplx<T> cn = am ? sa[anbase + l - am] : cp;
					if((l - am) & 1) {
						cfma(Op, cp, pcur);
						cfma(On, cn, pcur);
					} else {
						cfma(Ep, cp, pcur);
						cfma(En, cn, pcur);
					}
				}
				/* m = +am, then m = -am (Pbar_l,-m = (-1)^m Pbar_lm) */
#pragma unroll
				for(int sg = 0; sg < 2; sg++) {
					if(sg && !am)
						break;
					const int m = sg ? -am : am;
					const T f = (sg && (am & 1)) ? (T) -1 : (T) 1;
					const cplx<T> N = (sg ? En + On : Ep + Op) * f, Sth = (sg ? En - On : Ep - Op) * f;
					T *Fp = (m & 1) ? Fo : Fe;
					const int col = 2 * ((m - ((m & 1) ? G::MO0 : G::ME0)) >> 1);
					Fp[j * G::FS + col] = N.re;
					Fp[j * G::FS + col + 1] = N.im;
					Fp[(G::n - 1 - j) * G::FS + col] = Sth.re;
					Fp[(G::n - 1 - j) * G::FS + col + 1] = Sth.im;
				}
			}
		}
		__syncthreads();

		/* ---- phi DFT of the full tiles on MFMA ---- */
		T *mr = reinterpret_cast<T *>(map + (size_t) s * G::n * G::n);	/* [n][2n] real */
		for(int rt = wave; rt < G::NRT; rt += SF_NT / 64) {
			v4 ae[G::NCT], ao[G::NCT];
#pragma unroll
			for(int ct = 0; ct < G::NCT; ct++) {
				ae[ct] = (v4) 0;
				ao[ct] = (v4) 0;
			}
			const T *fe = Fe + (16 * rt + li) * G::FS + lk, *fo = Fo + (16 * rt + li) * G::FS + lk;
#pragma unroll
			for(int ks = 0; ks < G::KSE; ks++) {
				const T av = fe[4 * ks];
#pragma unroll
				for(int ct = 0; ct < G::NCT; ct++) {
					const T bv = sf_mf<T>::BREG ? bfe[sf_mf<T>::BREG ? ks : 0][ct] : sf_bval<T, L>(tw, G::ME0, G::NE, 4 * ks + lk, 16 * ct + li);
					ae[ct] = sf_mf<T>::mma(av, bv, ae[ct]);
				}
			}
#pragma unroll
			for(int ks = 0; ks < G::KSO; ks++) {
				const T av = fo[4 * ks];
#pragma unroll
				for(int ct = 0; ct < G::NCT; ct++) {
					const T bv = sf_mf<T>::BREG ? bfo[sf_mf<T>::BREG ? ks : 0][ct] : sf_bval<T, L>(tw, G::MO0, G::NO, 4 * ks + lk, 16 * ct + li);
					ao[ct] = sf_mf<T>::mma(av, bv, ao[ct]);
				}
			}
#pragma unroll
			for(int ct = 0; ct < G::NCT; ct++)
#pragma unroll
				for(int e = 0; e < 4; e++) {
					const int r = 16 * rt + sf_mf<T>::row(lk, e), c = 16 * ct + li;
					mr[(size_t) r * 2 * G::n + c] = ae[ct][e] + ao[ct][e];
					mr[(size_t) r * 2 * G::n + G::n + c] = ae[ct][e] - ao[ct][e];
				}
		}

		/* ---- leftover rings (all points) and points (full tiles'
		 * rings) on the VALU: one lane per (ring, point, parity) ---- */
		{
			const int ntask = G::XR * G::B + 16 * G::NRT * G::XP;
			for(int u = tid; u < 2 * ntask; u += SF_NT) {
				const int task = u >> 1, par = u & 1;
				int r, k;
				if(task < G::XR * G::B) {
					r = 16 * G::NRT + task / G::B;
					k = task % G::B;
				} else {
					const int t2 = task - G::XR * G::B;
					r = t2 / (G::XP > 0 ? G::XP : 1);
					k = 8 * G::NCT + t2 % (G::XP > 0 ? G::XP : 1);
				}
				const T *Fr = (par ? Fo : Fe) + r * G::FS;
				const int nq = par ? G::NO : G::NE, m0 = par ? G::MO0 : G::ME0;
				cplx<T> acc = mk((T) 0, (T) 0);
				int t = (m0 * k) % G::n;
				if(t < 0)
					t += G::n;
				const int dt = (2 * k) % G::n;
				for(int q = 0; q < nq; q++) {
					cfma(acc, mk(Fr[2 * q], Fr[2 * q + 1]), tw[t]);
					t += dt;
					if(t >= G::n)
						t -= G::n;
				}
				/* the pair (Ev, Od) sits in lanes u, u ^ 1 of one wave */
				const T ore = __shfl_xor(acc.re, 1), oim = __shfl_xor(acc.im, 1);
				if(!par) {
					cplx<T> *row = map + (size_t) s * G::n * G::n + (size_t) r * G::n;
					row[k] = mk(acc.re + ore, acc.im + oim);
					row[k + G::B] = mk(acc.re - ore, acc.im - oim);
				}
			}
		}
		__syncthreads();
	}
}

}	/* namespace */

int sky_fused_order(int L)
{
	return L == 25 || L == 32;
}

/* the orders with a compiled fused kernel */
template <typename T, int L>
static int sky_fused_at(const T *partial, int nsplit, int S, const struct xproj_geom *xg, const cplx<T> *R, const int2 *lm, const double *gmst, const struct synth_tables *tb, cplx<T> *alm, cplx<T> *map, int grid, hipStream_t st)
{
	hipLaunchKernelGGL((k_sky_fused<T, L>), dim3(grid), dim3(SF_NT), xg->CT * sizeof(T), st, partial, nsplit, S, *xg, R, lm, gmst, *tb, alm, map);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T>
int launch_sky_fused(const T *partial, int nsplit, int S, const struct xproj_geom *xg, const cplx<T> *R, const int2 *lm, const double *gmst, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *alm, cplx<T> *map, hipStream_t st)
{
	static int ncu = 0;
	if(!ncu) {
		int dev = 0, v = 0;
		if(hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
			v = 256;
		ncu = v;
	}
	if(S <= 0)
		return 0;
	/* persistent workgroups: two per CU in float32, one in float64 (LDS) */
	const int per_cu = sizeof(T) == 4 ? 2 : 1;
	const int grid = S < per_cu * ncu ? S : per_cu * ncu;
	switch(g->L) {
	case 25:
		return sky_fused_at<T, 25>(partial, nsplit, S, xg, R, lm, gmst, tb, alm, map, grid, st);
	case 32:
		return sky_fused_at<T, 32>(partial, nsplit, S, xg, R, lm, gmst, tb, alm, map, grid, st);
	default:
		return -1;
	}
}

template int launch_sky_fused<double>(const double *, int, int, const struct xproj_geom *, const cplx<double> *, const int2 *, const double *, const struct synth_geom *, const struct synth_tables *, cplx<double> *, cplx<double> *, hipStream_t);
template int launch_sky_fused<float>(const float *, int, int, const struct xproj_geom *, const cplx<float> *, const int2 *, const double *, const struct synth_geom *, const struct synth_tables *, cplx<float> *, cplx<float> *, hipStream_t);
