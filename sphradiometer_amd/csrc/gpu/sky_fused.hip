/*
 * Sky stage of the batched path in one kernel, for the orders the batched
 * cross-power skies run at (compile-time L): K-split partials -> a_lm
 * (rotation rows, baseline average, rotate_z) -> Legendre ring sums F in
 * LDS -> phi DFT on MFMA -> mesh.  F never leaves the chip.
 *
 * Real skies.  The batched skies are cross powers of real time series, so
 * a_{l,-m} = (-1)^m conj(a_lm) and the ring sums obey F[r][-m] =
 * conj(F[r][m]) (F[r][m] = sum_l a_lm Pbar_l|m|(x_r), times (-1)^m for
 * m < 0): the map is real,
 *     map[r][k] = Re F[r][0] + 2 Re sum_{m > 0} F[r][m] exp(i m phi_k),
 * and only the m >= 0 half of F is formed.  The map is written in the
 * reference's complex layout with a zero imaginary part (the reference's
 * is round-off).  The a_lm are written in full.
 *
 * phi DFT with the half-turn split.  The mesh has n = 2(L+1) points
 * phi_k = 2 pi k / n, so phi_{k + n/2} = phi_k + pi and
 * exp(i m phi_{k+n/2}) = (-1)^m exp(i m phi_k).  With Ev[r][k] the part of
 * the sum over even m and Od[r][k] over odd m, k < n/2:
 *     map[r][k] = Ev + Od,      map[r][k + n/2] = Ev - Od,
 * which halves the DFT work.  Ev and Od are real GEMMs (rows r = rings,
 * K = (Re, Im) of the m >= 0 modes of one parity, columns the half-turn
 * points; entries w cos(m phi_k) and -w sin(m phi_k), w = 1 at m = 0 and
 * 2 otherwise) on v_mfma_{f32,f64}_16x16x4.  Each of the 4 waves owns row
 * tiles of 16 rings; the points that do not fill a 16-wide tile (one for
 * L = 32) and the rings beyond the last full row tile (two for L = 32)
 * are summed on the VALU.
 *
 * B operands (the real DFT matrix from exp(2 pi i t / n), t = m k mod n)
 * are laid out once per order in the lane order of the MFMA fragments
 * ([k-step][column tile][lane], coalesced).  In float32 each wave loads
 * its B fragments into registers for the life of the workgroup (the
 * workgroups are persistent over segments); in float64 they are read from
 * the table (L2-resident) in the loop, behind the longer f64 MFMA.
 *
 * Legendre stage.  F[r][m] = sum_l a_lm Pbar_l|m|(x_r) for ring pairs
 * (x_j, -x_j) at once through the parity (-1)^(l+|m|).  Item (ring
 * group, a) handles m = a and m = L - a for 2 or 3 rings (3 where
 * items of 2 would not fit one round of the 256 threads: 187 items instead
 * of 289 for L = 32), so every item runs L + 2 recurrence steps (balanced
 * lanes).  The
 * recurrence Pbar_l = alpha_l x Pbar_{l-1} - (alpha_l / alpha_{l-1})
 * Pbar_{l-2} uses a division-free coefficient table built once per
 * workgroup, and runs two steps (one even, one odd l - |m|) per
 * iteration so the parity needs no branch.
 *
 * Balance: the Legendre items that do not fill the last round land on
 * the first wave, the VALU leftovers of the phi DFT that do not fill one
 * round on the last wave.
 *
 * Layout (reference sh_series / mesh conventions, sh_series.c:33-80,
 * sh_series_to_mesh in sh_series_trans.c): a_lm in the m = 0, 1..L,
 * -L..-1 block order; mesh [2B][2B] complex, rings in descending cos theta,
 * points phi_k = 2 pi k / n.
 */
#include <hip/hip_runtime.h>
#include <string.h>
#include "kernels.h"
#include "runtime.h"

namespace {

#define SF_NT 256

template <typename T> struct sf_mf;
template <> struct sf_mf<float> {
	typedef float v4 __attribute__((ext_vector_type(4)));
	static __device__ inline v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
	static __device__ inline int row(int q, int e) { return 4 * q + e; }
	static __device__ inline const float *btab(const struct synth_tables &tb) { return tb.bf32; }
	static __device__ inline const unsigned char *blob(const struct synth_tables &tb) { return tb.sk32; }
	enum { BREG = 1 };
};
template <> struct sf_mf<double> {
	typedef double v4 __attribute__((ext_vector_type(4)));
	static __device__ inline v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
	static __device__ inline int row(int q, int e) { return q + 4 * e; }
	static __device__ inline const double *btab(const struct synth_tables &tb) { return tb.bf64; }
	static __device__ inline const unsigned char *blob(const struct synth_tables &tb) { return tb.sk64; }
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
		NM0 = L / 2 + 1,	/* even orders m >= 0 */
		NM1 = (L + 1) / 2,	/* odd orders m > 0 */
		KSE = (2 * NM0 + 3) / 4,	/* MFMA k-steps, even / odd */
		KSO = (2 * NM1 + 3) / 4,
		KMAX = 4 * (KSE > KSO ? KSE : KSO),
		/* row stride of the F tiles in LDS: >= KMAX and = 4 mod 32
		 * words, so the 16 rows x 4 columns of an A fragment fall on
		 * distinct banks */
		FS = ((KMAX + 27) / 32) * 32 + 4,
		NRT = n / 16,		/* full row tiles */
		XR = n - 16 * NRT,	/* rings summed on the VALU */
		NCT = B / 16,		/* full column tiles of the half turn */
		XP = B - 16 * NCT,	/* half-turn points summed on the VALU */
		NA = L / 2 + 1,		/* Legendre item pairs |m| = a, L - a */
	};
};

/* element (kr, c) of the real DFT matrix of parity par: kr = 2 j + (0: Re,
 * 1: Im of mode m = par + 2 j), c = point k; w cos(m phi_k) for Re and
 * -w sin(m phi_k) for Im, w = 1 at m = 0 and 2 otherwise */
template <int L>
__device__ inline double sf_bval(const double *twp, int par, int kr, int c)
{
	typedef sf_geom<L> G;
	const int j = kr >> 1;
	if(j >= (par ? G::NM1 : G::NM0))
		return 0.0;
	const int m = par + 2 * j;
	const int t = (m * c) % G::n;
	const double w = m ? 2.0 : 1.0;
	return (kr & 1) ? -w * twp[2 * t + 1] : w * twp[2 * t];
}

/* B-fragment table: entry ((kidx NCT + ct) 64 + lane), kidx < KSE the even
 * orders' k-steps, then the odd orders' */
template <typename T, int L>
__global__ void k_sky_bfrag(const double *__restrict__ twp, T *__restrict__ out)
{
	typedef sf_geom<L> G;
	const int idx = blockIdx.x * blockDim.x + threadIdx.x;
	if(idx >= (G::KSE + G::KSO) * G::NCT * 64)
		return;
	const int lane = idx & 63, ct = (idx >> 6) % G::NCT, kidx = (idx >> 6) / G::NCT;
	const int li = lane & 15, lk = lane >> 4;
	const int odd = kidx >= G::KSE, ks = odd ? kidx - G::KSE : kidx;
	out[idx] = (T) sf_bval<L>(twp, odd, 4 * ks + lk, 16 * ct + li);
}

/* rings per Legendre item: 2, or 3 where items of 2 rings would not fit
 * one round of the sub-group; recurrence steps per loaded batch (even) */
template <int L> struct sf_rg {
	enum { V = ((L + 2) / 2) * (L / 2 + 1) > 256 ? 3 : 2 };
};
#define SF_US 4
/* K splits the first segment's partial sums are prefetched for */
#define SF_PF 8

/* baselines handled with all rotation-row loads issued first; networks
 * with more take the general loop */
#define SF_NBF 4

/* a_lm of one segment: sum over baselines of rotation row x the baseline's
 * projected (l) column, accumulated in float64 whatever T (the network's
 * cross-baseline sum), then rotate_z; RP is an LDS or a global array of
 * the rotation rows [nb][ncoef] */
__device__ inline cplx<double> sf_wide(cplx<float> a) { return mk((double) a.re, (double) a.im); }
__device__ inline cplx<double> sf_wide(cplx<double> a) { return a; }

template <typename T, int L, typename RP>
__device__ inline void sf_alm(RP Rr, const T *pw, const short *slm, const cplx<T> *rot, bool rz, const struct xproj_geom &xg, const int *csb, const int *lbb, int tid, cplx<T> *sa, cplx<T> *alm)
{
	typedef sf_geom<L> G;
	const int nb = xg.nb;
#pragma unroll 2
	for(int idx = tid; idx < G::ncoef; idx += SF_NT) {
		const int q = slm[idx];
		const int l = q >> 8, m = (signed char) (q & 255);
		cplx<double> acc = mk(0.0, 0.0);
		if(nb <= SF_NBF) {
			cplx<T> r[SF_NBF];
#pragma unroll
			for(int b = 0; b < SF_NBF; b++)
				if(b < nb)
					r[b] = Rr[b * G::ncoef + idx];
#pragma unroll
			for(int b = 0; b < SF_NBF; b++)
				if(b < nb && l <= lbb[b]) {
					const T *p = pw + csb[b] + 2 * l;
					cfma(acc, sf_wide(r[b]), mk((double) p[0], (double) p[1]));
				}
		} else {
			for(int b = 0; b < nb; b++)
				if(l <= xg.Lb[b]) {
					const T *p = pw + xg.col_start[b] + 2 * l;
					cfma(acc, sf_wide(Rr[b * G::ncoef + idx]), mk((double) p[0], (double) p[1]));
				}
		}
		cplx<T> a = mk((T) acc.re, (T) acc.im);
		if(rz && m)
			a = a * (m > 0 ? rot[m] : conjg(rot[-m]));
		sa[idx] = a;
		alm[idx] = a;
	}
}

/* the kernel's read-only tables as one blob, staged into LDS in one pass:
 * recurrence coefficient pairs (alpha_l, alpha_l / alpha_{l-1}) with two
 * zero pads, twiddles exp(2 pi i t / n), cos theta of the northern rings,
 * Pbar_mm(x_j), and (l << 8) | (m & 255) per coefficient; each part
 * 16-byte aligned */
template <typename T, int L> struct sf_blob {
	typedef sf_geom<L> G;
	enum {
		CF = 0,
		TW = CF + (((G::nal + 2) * 2 * (int) sizeof(T) + 15) / 16) * 16,
		XS = TW + ((G::n * 2 * (int) sizeof(T) + 15) / 16) * 16,
		SM = XS + ((G::B * (int) sizeof(T) + 15) / 16) * 16,
		LM = SM + ((G::B * (L + 1) * (int) sizeof(T) + 15) / 16) * 16,
		BYTES = LM + ((G::ncoef * 2 + 15) / 16) * 16,
	};
};

template <typename T, int L>
__global__ void k_sky_blob(const double *__restrict__ alpha, const double *__restrict__ twp, const double *__restrict__ x, const double *__restrict__ sm, unsigned char *__restrict__ out)
{
	typedef sf_geom<L> G;
	typedef sf_blob<T, L> Bl;
	cplx<T> *cf = reinterpret_cast<cplx<T> *>(out + Bl::CF);
	cplx<T> *tw = reinterpret_cast<cplx<T> *>(out + Bl::TW);
	T *xs = reinterpret_cast<T *>(out + Bl::XS);
	T *smm = reinterpret_cast<T *>(out + Bl::SM);
	short *slm = reinterpret_cast<short *>(out + Bl::LM);
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if(t < G::nal) {
		int l, m;
		lm_decode(L, t, &l, &m);
		const double a = l > m ? alpha[t] : 0.0;
		cf[t] = mk((T) a, (T) (l > m + 1 ? a / alpha[t - 1] : 0.0));
	} else if(t < G::nal + 2)
		cf[t] = mk((T) 0, (T) 0);
	if(t < G::n)
		tw[t] = mk((T) twp[2 * t], (T) twp[2 * t + 1]);
	if(t < G::B)
		xs[t] = (T) x[t];
	if(t < G::B * (L + 1))
		smm[t] = (T) sm[t];
	if(t < G::ncoef) {
		int l, m;
		lm_decode(L, t, &l, &m);
		slm[t] = (short) ((l << 8) | (m & 255));
	}
}

template <typename T, int L, int NSUB>
__global__ void __launch_bounds__(SF_NT * NSUB) k_sky_fused(const T *__restrict__ partial, int nsplit, int S, struct xproj_geom xg, const cplx<T> *__restrict__ R, int rmode, const int2 *__restrict__ lm, const double *__restrict__ gmst, struct synth_tables tb, cplx<T> *__restrict__ alm, cplx<T> *__restrict__ map)
{
	typedef sf_geom<L> G;
	typedef typename sf_mf<T>::v4 v4;
	/* per sub-group (one segment each) */
	/* F tiles: [0] even orders, [1] odd orders, [sub][ring][2 q + re/im];
	 * with rmode 2 they hold the rotation rows from the partial-sum phase
	 * to the a_lm phase */
	__shared__ __align__(16) T Ft_[2][NSUB][G::n * G::FS];
	__shared__ cplx<T> sa_[NSUB][G::ncoef];	/* a_lm of the segment */
	__shared__ cplx<T> rot_[NSUB][L + 1];	/* exp(-i m gmst) */
	/* shared by the sub-groups: the table blob */
	typedef sf_blob<T, L> Bl;
	__shared__ __align__(16) unsigned char tab[Bl::BYTES];
	const cplx<T> *cf = reinterpret_cast<const cplx<T> *>(tab + Bl::CF);
	const cplx<T> *tw = reinterpret_cast<const cplx<T> *>(tab + Bl::TW);
	const T *xs = reinterpret_cast<const T *>(tab + Bl::XS);
	const T *smm = reinterpret_cast<const T *>(tab + Bl::SM);
	const short *slm = reinterpret_cast<const short *>(tab + Bl::LM);
	extern __shared__ __align__(16) unsigned char sf_dyn[];
	/* rotation rows [nb][ncoef]: rmode 0 global memory, 1 the dynamic LDS
	 * (then the [NSUB][CT] partial sums), 2 the F tiles */
	cplx<T> *sR = reinterpret_cast<cplx<T> *>(sf_dyn), *sRf = reinterpret_cast<cplx<T> *>(&Ft_[0][0][0]);

	const int gtid = threadIdx.x, sub = gtid / SF_NT, tid = gtid % SF_NT;
	const int wave = tid >> 6;
	T *Fe = Ft_[0][sub], *Fo = Ft_[1][sub];
	cplx<T> *sa = sa_[sub], *rot = rot_[sub];
	T *pw = reinterpret_cast<T *>(sf_dyn + (rmode == 1 ? (size_t) xg.nb * G::ncoef * sizeof(cplx<T>) : 0)) + (size_t) sub * xg.CT;
	/* the rotation rows to LDS at dst, all threads of the workgroup, eight
	 * loads in flight per thread */
	auto load_rows = [&](cplx<T> *dst) {
		constexpr int NT = SF_NT * NSUB;
		const int nr = xg.nb * G::ncoef;
		for(int t = gtid; t < nr; t += 8 * NT) {
			cplx<T> w[8];
#pragma unroll
			for(int u = 0; u < 8; u++)
				if(t + u * NT < nr)
					w[u] = R[t + u * NT];
#pragma unroll
			for(int u = 0; u < 8; u++)
				if(t + u * NT < nr)
					dst[t + u * NT] = w[u];
		}
	};

	/* the first segment's K-split partial sums, when each thread sums one
	 * column of at most SF_PF splits: loads issued before the table
	 * prologue's, so the two global round trips overlap */
	const bool pf = xg.CT <= SF_NT && nsplit <= SF_PF;
	T pv[SF_PF];
	{
		const int s = blockIdx.x * NSUB + sub;
		const size_t st = (size_t) S * xg.CT;
#pragma unroll
		for(int u = 0; u < SF_PF; u++)
			pv[u] = (pf && s < S && tid < xg.CT && u < nsplit) ? partial[(size_t) s * xg.CT + tid + u * st] : (T) 0;
	}

	/* once per workgroup: the table blob and the rotation rows to LDS
	 * (all loads of a thread issued before its stores), the zero padding
	 * of the F tiles */
	{
		constexpr int NT = SF_NT * NSUB, N16 = Bl::BYTES / 16, RB = (N16 + NT - 1) / NT;
		const uint4 *src = reinterpret_cast<const uint4 *>(sf_mf<T>::blob(tb));
		uint4 v[RB];
#pragma unroll
		for(int r = 0; r < RB; r++)
			if(gtid + r * NT < N16)
				v[r] = src[gtid + r * NT];
		if(rmode)
			load_rows(rmode == 1 ? sR : sRf);
#pragma unroll
		for(int r = 0; r < RB; r++)
			if(gtid + r * NT < N16)
				reinterpret_cast<uint4 *>(tab)[gtid + r * NT] = v[r];
		/* the F tiles' zero padding (with rmode 2 refreshed in each
		 * segment's Legendre phase instead) */
		if(rmode != 2)
			for(int t = gtid; t < 2 * NSUB * G::n * G::FS; t += NT)
				(&Ft_[0][0][0])[t] = (T) 0;
	}
	/* per-baseline columns and orders of the first baselines, uniform */
	int csb[SF_NBF], lbb[SF_NBF];
#pragma unroll
	for(int b = 0; b < SF_NBF; b++) {
		csb[b] = xg.col_start[b];
		lbb[b] = xg.Lb[b];
	}
	__syncthreads();

	/* B fragments of this lane, [k-step][column tile], from the table in
	 * MFMA lane order: float32 keeps them in registers for the life of the
	 * workgroup, float64 reads them in the loop */
	const T *__restrict__ btab = sf_mf<T>::btab(tb);
	T bf[sf_mf<T>::BREG ? G::KSE + G::KSO : 1][G::NCT];
	if(sf_mf<T>::BREG) {
		const int lane = tid & 63;
#pragma unroll
		for(int k = 0; k < (sf_mf<T>::BREG ? G::KSE + G::KSO : 1); k++)
#pragma unroll
			for(int ct = 0; ct < G::NCT; ct++)
				bf[k][ct] = btab[(k * G::NCT + ct) * 64 + lane];
	}

	for(int base = blockIdx.x * NSUB; base < S; base += gridDim.x * NSUB) {
		/* the thread index through an empty asm, so that the per-thread
		 * index arithmetic and table loads of the phases below stay inside
		 * the loop instead of being hoisted and held live across it */
		int otid = tid;
		asm volatile("" : "+v"(otid));
		const int olane = otid & 63, oli = olane & 15, olk = olane >> 4;
		const int s = base + sub;
		const bool act = s < S;
		/* ---- partial sums over the K splits, rotate_z phases ---- */
		if(act && pf && base == (int) blockIdx.x * NSUB) {
			/* the first segment's, loaded before the prologue */
			if(otid < xg.CT)
				pw[otid] = ((pv[0] + pv[1]) + (pv[2] + pv[3])) + ((pv[4] + pv[5]) + (pv[6] + pv[7]));
		} else if(act) {
			for(int c = otid; c < xg.CT; c += SF_NT) {
				/* eight independent loads in flight */
				const T *p = partial + (size_t) s * xg.CT + c;
				const size_t st = (size_t) S * xg.CT;
				T v[8];
#pragma unroll
				for(int u = 0; u < 8; u++)
					v[u] = 0;
				int sp = 0;
				for(; sp + 8 <= nsplit; sp += 8)
#pragma unroll
					for(int u = 0; u < 8; u++)
						v[u] += p[(sp + u) * st];
				for(; sp < nsplit; sp++)
					v[sp & 7] += p[sp * st];
				pw[c] = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
			}
		}
		if(act && otid <= L) {
			double sn = 0.0, cs = 1.0;
			if(gmst && otid)
				sincos(-(double) otid * gmst[s], &sn, &cs);
			rot[otid] = mk((T) cs, (T) sn);
		}
		/* rotation rows in the F tiles: reloaded after the previous
		 * segment's phi DFT has read them */
		if(rmode == 2 && base != (int) blockIdx.x * NSUB)
			load_rows(sRf);
		__syncthreads();

		/* ---- a_lm: rotation rows over the baselines, rotate_z ---- */
		if(act) {
			if(rmode)
				sf_alm<T, L>(rmode == 1 ? sR : sRf, pw, slm, rot, gmst != NULL, xg, csb, lbb, otid, sa, alm + (size_t) s * G::ncoef);
			else
				sf_alm<T, L>(R, pw, slm, rot, gmst != NULL, xg, csb, lbb, otid, sa, alm + (size_t) s * G::ncoef);
		}
		__syncthreads();
		if(!map)
			continue;

		/* ---- Legendre ring sums into the F tiles: item (a, ring group)
		 * runs the recurrences of SF_RG rings at once, so each
		 * coefficient and a_lm read from LDS serves SF_RG rings ---- */
		if(rmode == 2) {
			/* the zero padding the rotation rows overwrote: columns
			 * past the modes up to the last MFMA k-step, disjoint from
			 * the ring sums written below */
			constexpr int PE = 4 * G::KSE - 2 * G::NM0, PO = 4 * G::KSO - 2 * G::NM1;
			for(int t = otid; t < G::n * (PE + PO); t += SF_NT) {
				const int r = t / (PE + PO), c = t - r * (PE + PO);
				if(c < PE)
					Fe[r * G::FS + 2 * G::NM0 + c] = (T) 0;
				else
					Fo[r * G::FS + 2 * G::NM1 + c - PE] = (T) 0;
			}
		}
		if(act) {
			constexpr int SF_RG = sf_rg<L>::V;
			constexpr int NG = (G::B + SF_RG - 1) / SF_RG;
			for(int it = otid; it < NG * G::NA; it += SF_NT) {
				const int tg = it % NG, a = it / NG;
				T x[SF_RG];
#pragma unroll
				for(int r = 0; r < SF_RG; r++) {
					const int j = SF_RG * tg + r;
					x[r] = j < G::B ? xs[j] : (T) 0;
				}
#pragma unroll
				for(int h = 0; h < 2; h++) {
					const int am = h ? L - a : a;
					if(h && am == a)
						break;
					const int cbase = lm_offset(L, am, am);	/* coefficient table slot too */
					const cplx<T> *sp = sa + cbase, *cc = cf + cbase;
					const int cnt = L - am;	/* steps after l = am */
					T p2[SF_RG], p1[SF_RG];
					cplx<T> Ep[SF_RG], Op[SF_RG];
					{
						const cplx<T> s0 = sp[0];
						const cplx<T> c1 = cc[1], s1 = sp[1];
#pragma unroll
						for(int r = 0; r < SF_RG; r++) {
							const int j = SF_RG * tg + r;
							p2[r] = j < G::B ? smm[j * (L + 1) + am] : (T) 0;
							Ep[r] = s0 * p2[r];
							/* l = am + 1 (c1.im is 0; unused past the end) */
							p1[r] = cnt >= 1 ? c1.re * x[r] * p2[r] : (T) 0;
							Op[r] = s1 * p1[r];
						}
					}
					/* SF_US steps (E, O, ...) per iteration with all
					 * operands loaded first, then two, then one */
					int i = 2;
					for(; i + SF_US - 1 <= cnt; i += SF_US) {
						cplx<T> c[SF_US], a4[SF_US];
#pragma unroll
						for(int u = 0; u < SF_US; u++) {
							c[u] = cc[i + u];
							a4[u] = sp[i + u];
						}
#pragma unroll
						for(int r = 0; r < SF_RG; r++) {
#pragma unroll
							for(int u = 0; u < SF_US; u += 2) {
								const T pe = c[u].re * x[r] * p1[r] - c[u].im * p2[r];
								const T po = c[u + 1].re * x[r] * pe - c[u + 1].im * p1[r];
								cfma(Ep[r], a4[u], pe);
								cfma(Op[r], a4[u + 1], po);
								p2[r] = pe;
								p1[r] = po;
							}
						}
					}
					if(i + 1 <= cnt) {
						const cplx<T> c0 = cc[i], c1 = cc[i + 1], a0 = sp[i], a1 = sp[i + 1];
#pragma unroll
						for(int r = 0; r < SF_RG; r++) {
							const T pe = c0.re * x[r] * p1[r] - c0.im * p2[r];
							const T po = c1.re * x[r] * pe - c1.im * p1[r];
							cfma(Ep[r], a0, pe);
							cfma(Op[r], a1, po);
							p2[r] = pe;
							p1[r] = po;
						}
						i += 2;
					}
					if(i == cnt) {
						const cplx<T> cl = cc[i], a0 = sp[i];
#pragma unroll
						for(int r = 0; r < SF_RG; r++) {
							const T pe = cl.re * x[r] * p1[r] - cl.im * p2[r];
							cfma(Ep[r], a0, pe);
						}
					}
					/* F[r][am] of the northern ring and its mirror in the
					 * tile of am's parity, columns 2 j (Re), 2 j + 1 (Im) of
					 * mode am = parity + 2 j */
					{
						T *Fp = (am & 1) ? Fo : Fe;
						const int col = 2 * (am >> 1);
#pragma unroll
						for(int r = 0; r < SF_RG; r++) {
							const int j = SF_RG * tg + r;
							if(j >= G::B)
								continue;
							const cplx<T> N = Ep[r] + Op[r], Sth = Ep[r] - Op[r];
							Fp[j * G::FS + col] = N.re;
							Fp[j * G::FS + col + 1] = N.im;
							Fp[(G::n - 1 - j) * G::FS + col] = Sth.re;
							Fp[(G::n - 1 - j) * G::FS + col + 1] = Sth.im;
						}
					}
				}
			}
		}
		__syncthreads();

		if(act) {
			/* ---- phi DFT of the full tiles on MFMA ---- */
			cplx<T> *mr = map + (size_t) s * G::n * G::n;	/* [n][n] complex */
			for(int rt = wave; rt < G::NRT; rt += SF_NT / 64) {
				v4 ae[G::NCT], ao[G::NCT];
#pragma unroll
				for(int ct = 0; ct < G::NCT; ct++) {
					ae[ct] = (v4) 0;
					ao[ct] = (v4) 0;
				}
				const T *fe = Fe + (16 * rt + oli) * G::FS + olk, *fo = Fo + (16 * rt + oli) * G::FS + olk;
#pragma unroll
				for(int ks = 0; ks < G::KSE; ks++) {
					const T av = fe[4 * ks];
#pragma unroll
					for(int ct = 0; ct < G::NCT; ct++) {
						const T bv = sf_mf<T>::BREG ? bf[sf_mf<T>::BREG ? ks : 0][ct] : btab[(ks * G::NCT + ct) * 64 + olane];
						ae[ct] = sf_mf<T>::mma(av, bv, ae[ct]);
					}
				}
#pragma unroll
				for(int ks = 0; ks < G::KSO; ks++) {
					const T av = fo[4 * ks];
#pragma unroll
					for(int ct = 0; ct < G::NCT; ct++) {
						const T bv = sf_mf<T>::BREG ? bf[sf_mf<T>::BREG ? G::KSE + ks : 0][ct] : btab[((G::KSE + ks) * G::NCT + ct) * 64 + olane];
						ao[ct] = sf_mf<T>::mma(av, bv, ao[ct]);
					}
				}
#pragma unroll
				for(int ct = 0; ct < G::NCT; ct++)
#pragma unroll
					for(int e = 0; e < 4; e++) {
						const int r = 16 * rt + sf_mf<T>::row(olk, e), k = 16 * ct + oli;
						mr[(size_t) r * G::n + k] = mk(ae[ct][e] + ao[ct][e], (T) 0);
						mr[(size_t) r * G::n + G::B + k] = mk(ae[ct][e] - ao[ct][e], (T) 0);
					}
			}

			/* ---- leftover rings (all points) and points (full tiles'
			 * rings) on the VALU: one lane per (ring, point, parity),
			 * units in reverse thread order so a partial last round
			 * lands on the last wave ---- */
			const int ntask = G::XR * G::B + 16 * G::NRT * G::XP;
			for(int u = SF_NT - 1 - otid; u < 2 * ntask; u += SF_NT) {
				const int task = u >> 1, par = u & 1;
				int r, k;
				if(task < G::XR * G::B) {
					r = 16 * G::NRT + task / G::B;
					k = task % G::B;
				} else {
					const int t2 = task - G::XR * G::B;
					r = t2 / (G::XP > 0 ? G::XP : 1);
					k = 16 * G::NCT + t2 % (G::XP > 0 ? G::XP : 1);
				}
				/* sum_j w (Re F cos(m phi_k) - Im F sin(m phi_k)) over the
				 * modes m = par + 2 j of this parity */
				const T *Fr = (par ? Fo : Fe) + r * G::FS;
				const int nq = par ? G::NM1 : G::NM0;
				T a0 = 0, acc = 0, acc2 = 0;
				int t = par ? k : 0;
				const int dt = (2 * k) % G::n;
				int q = 0;
				if(!par) {
					/* m = 0: weight 1, the other modes weight 2 */
					a0 = Fr[0];
					q = 1;
					t = dt;
				}
				for(; q + 2 <= nq; q += 2) {
					const int t1 = t + dt >= G::n ? t + dt - G::n : t + dt;
					const cplx<T> w0 = tw[t], w1 = tw[t1];
					acc += Fr[2 * q] * w0.re - Fr[2 * q + 1] * w0.im;
					acc2 += Fr[2 * q + 2] * w1.re - Fr[2 * q + 3] * w1.im;
					t = t1 + dt >= G::n ? t1 + dt - G::n : t1 + dt;
				}
				if(q < nq)
					acc += Fr[2 * q] * tw[t].re - Fr[2 * q + 1] * tw[t].im;
				const T v = a0 + (T) 2 * (acc + acc2);
				/* the pair (Ev, Od) sits in lanes tid, tid ^ 1 of one
				 * wave (u and u ^ 1: both in range or both out) */
				const T ov = __shfl_xor(v, 1);
				if(!par) {
					cplx<T> *row = map + (size_t) s * G::n * G::n + (size_t) r * G::n;
					row[k] = mk(v + ov, (T) 0);
					row[k + G::B] = mk(v - ov, (T) 0);
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

template <int L>
static int sky_tables_at(const struct synth_tables *tb, struct sky_fused_dev *d)
{
	typedef sf_geom<L> G;
	const int ne = (G::KSE + G::KSO) * G::NCT * 64;
	const int nb64 = sf_blob<double, L>::BYTES, nb32 = sf_blob<float, L>::BYTES;
	if(sphr_rt_check(hipMalloc(&d->bf64, ne * sizeof(double)), "hipMalloc") || sphr_rt_check(hipMalloc(&d->bf32, ne * sizeof(float)), "hipMalloc")
	   || sphr_rt_check(hipMalloc(&d->sk64, nb64), "hipMalloc") || sphr_rt_check(hipMalloc(&d->sk32, nb32), "hipMalloc"))
		return -1;
	hipLaunchKernelGGL((k_sky_bfrag<double, L>), dim3((ne + 255) / 256), dim3(256), 0, 0, tb->twp, d->bf64);
	hipLaunchKernelGGL((k_sky_bfrag<float, L>), dim3((ne + 255) / 256), dim3(256), 0, 0, tb->twp, d->bf32);
	const int nt = G::ncoef > G::B * (L + 1) ? G::ncoef : G::B * (L + 1);
	hipLaunchKernelGGL((k_sky_blob<double, L>), dim3((nt + 255) / 256), dim3(256), 0, 0, tb->alpha, tb->twp, tb->x, tb->smm, d->sk64);
	hipLaunchKernelGGL((k_sky_blob<float, L>), dim3((nt + 255) / 256), dim3(256), 0, 0, tb->alpha, tb->twp, tb->x, tb->smm, d->sk32);
	return sphr_rt_check(hipGetLastError(), "sky tables") || sphr_rt_check(hipDeviceSynchronize(), "sky tables") ? -1 : 0;
}

int sky_fused_tables(int L, const struct synth_tables *tb, struct sky_fused_dev *d)
{
	memset(d, 0, sizeof(*d));
	switch(L) {
	case 25:
		return sky_tables_at<25>(tb, d);
	case 32:
		return sky_tables_at<32>(tb, d);
	default:
		return 0;
	}
}

/* sub-groups (segments in flight) per workgroup by default: float32 two,
 * sharing the tables; float64 one (its F tiles take most of the LDS) */
template <typename T> struct sf_sub {
	enum { N = sizeof(T) == 4 ? 2 : 1 };
};

/* static LDS of the kernel at NS sub-groups (F tiles, a_lm, tables), with
 * slack */
template <typename T, int L>
static size_t sky_fused_lds(int NS)
{
	typedef sf_geom<L> G;
	return NS * (2 * (size_t) G::n * G::FS * sizeof(T) + (size_t) (G::ncoef + L + 1) * sizeof(cplx<T>))
	       + (size_t) (G::nal + 2 + G::n) * sizeof(cplx<T>) + G::ncoef * sizeof(short) + (size_t) G::B * (L + 2) * sizeof(T) + 1024;
}

template <typename T, int L, int NS>
static int sky_fused_at(const T *partial, int nsplit, int S, const struct xproj_geom *xg, const cplx<T> *R, const int2 *lm, const double *gmst, const struct synth_tables *tb, cplx<T> *alm, cplx<T> *map, int ncu, int occ, hipStream_t st)
{
	typedef sf_geom<L> G;
	const size_t stat = sky_fused_lds<T, L>(NS);
	const size_t rbytes = (size_t) xg->nb * G::ncoef * sizeof(cplx<T>), pbytes = (size_t) NS * xg->CT * sizeof(T);
	/* persistent workgroups, occ per CU, as many as the LDS of all of
	 * them allows within the CU's 160 KiB; the rotation rows go to LDS
	 * when they fit beside the rest at that occupancy */
	const size_t cu_lds = 160 * 1024;
	if(stat + pbytes > cu_lds)
		return -1;
	while(occ > 1 && occ * (stat + pbytes) > cu_lds)
		occ--;
	/* rotation rows in the dynamic LDS when they fit beside the rest at
	 * this occupancy, else in the F tiles when those hold them, else read
	 * from global memory */
	const size_t ftiles = 2 * (size_t) NS * G::n * G::FS * sizeof(T);
	const int rmode = occ * (stat + rbytes + pbytes) <= cu_lds ? 1 : rbytes <= ftiles ? 2 : 0;
	const size_t dyn = (rmode == 1 ? rbytes : 0) + pbytes;
	const int nwg = (S + NS - 1) / NS;
	const int grid = nwg < occ * ncu ? nwg : occ * ncu;
	hipLaunchKernelGGL((k_sky_fused<T, L, NS>), dim3(grid), dim3(SF_NT * NS), dyn, st, partial, nsplit, S, *xg, R, rmode, lm, gmst, *tb, alm, map);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* nsub segments per workgroup (0: the default, float32 two sharing the
 * tables, float64 one; float32 four at one workgroup per CU share the
 * tables and the rotation rows), occ workgroups per CU where the LDS
 * allows */
template <typename T, int L>
static int sky_fused_form(const T *partial, int nsplit, int S, const struct xproj_geom *xg, const cplx<T> *R, const int2 *lm, const double *gmst, const struct synth_tables *tb, cplx<T> *alm, cplx<T> *map, int ncu, int occ, int nsub, hipStream_t st)
{
	if(!nsub)
		nsub = occ >= 3 ? 1 : sf_sub<T>::N;
	if(nsub == 1)
		return sky_fused_at<T, L, 1>(partial, nsplit, S, xg, R, lm, gmst, tb, alm, map, ncu, occ, st);
	if(nsub == 2)
		return sky_fused_at<T, L, 2>(partial, nsplit, S, xg, R, lm, gmst, tb, alm, map, ncu, occ, st);
	if constexpr(sizeof(T) == 4)
		if(nsub == 4)
			return sky_fused_at<T, L, 4>(partial, nsplit, S, xg, R, lm, gmst, tb, alm, map, ncu, occ, st);
	return -1;
}

template <typename T>
int launch_sky_fused(const T *partial, int nsplit, int S, const struct xproj_geom *xg, const cplx<T> *R, const int2 *lm, const double *gmst, const struct synth_geom *g, const struct synth_tables *tb, cplx<T> *alm, cplx<T> *map, hipStream_t st, int occ, int nsub)
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
	/* persistent workgroups, one per CU */
	switch(g->L) {
	case 25:
		return sky_fused_form<T, 25>(partial, nsplit, S, xg, R, lm, gmst, tb, alm, map, ncu, occ, nsub, st);
	case 32:
		return sky_fused_form<T, 32>(partial, nsplit, S, xg, R, lm, gmst, tb, alm, map, ncu, occ, nsub, st);
	default:
		return -1;
	}
}

template int launch_sky_fused<double>(const double *, int, int, const struct xproj_geom *, const cplx<double> *, const int2 *, const double *, const struct synth_geom *, const struct synth_tables *, cplx<double> *, cplx<double> *, hipStream_t, int, int);
template int launch_sky_fused<float>(const float *, int, int, const struct xproj_geom *, const cplx<float> *, const int2 *, const double *, const struct synth_geom *, const struct synth_tables *, cplx<float> *, cplx<float> *, hipStream_t, int, int);
