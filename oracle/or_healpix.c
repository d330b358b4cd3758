/*
 * Oracle (test infrastructure only): HEALPix output of an sh_series,
 * restating sh_series_healpix.cc and healpix.py of the reference.
 *
 * The reference hands the real part of the series (sh_series_real) to
 * healpix_cxx / healpy alm2map, which evaluates
 *     f(pix) = sum_l [a_l0 Y_l0 + 2 Re sum_{m>0} a_lm Y_lm]
 * at the centres of a RING-ordered map; for the coefficients of a real
 * function that is Re f at the pixel centres, f the original series.  The
 * pixel centres follow the published HEALPix definition (Gorski et al.
 * 2005, ApJ 622, 759, eqs. 4-9; the pix2ang_ring form of the HEALPix
 * library): the pixel index is inverted to (ring, position in ring).
 *
 * healpix_cxx and healpy are not part of the reference tree; parity is
 * pinned on closed forms (nside 1 centres, ring structure, Y_00, Y_10,
 * Y_11) in the tests.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include "oracle.h"

/* ceilpow2 and lmax2nside of sh_series_healpix.cc:196-238: the smallest
 * power of two not below ceil(l_max / 2 + 0.75) */
int or_healpix_nside_c(unsigned l_max)
{
	const double x = ceil(l_max / 2. + 0.75);
	for(int i = 1; i > 0; i <<= 1)
		if(i >= x)
			return i;
	return -1;
}

/* healpy_alm_to_map (healpix.py:163): nside = 2^ceil(log2(l_max / 2)) */
int or_healpix_nside_py(unsigned l_max)
{
	return (int) llround(pow(2., ceil(log2(l_max / 2.))));
}

static long isqrt_l(long v)
{
	long r = (long) sqrt((double) v);
	while(r * r > v)
		r--;
	while((r + 1) * (r + 1) <= v)
		r++;
	return r;
}

/* centre of RING pixel pix of an nside map */
void or_healpix_pix2ang(int nside, long pix, double *theta, double *phi)
{
	const long N = nside, npix = 12 * N * N, ncap = 2 * N * (N - 1);
	double z, ph;
	if(pix < ncap) {
		/* north polar cap: ring i (1-based) holds 4 i pixels */
		const long i = (1 + isqrt_l(1 + 2 * pix)) >> 1;
		const long j = pix + 1 - 2 * i * (i - 1);	/* 1 .. 4 i */
		z = 1. - (double) (i * i) / (3. * N * N);
		ph = (j - 0.5) * M_PI / (2. * i);
	} else if(pix < npix - ncap) {
		/* equatorial belt: rings N .. 3N of 4N pixels, alternate rings
		 * shifted by half a pixel */
		const long ip = pix - ncap;
		const long i = ip / (4 * N) + N;
		const long j = ip % (4 * N) + 1;
		const double fodd = ((i + N) & 1) ? 1. : 0.5;
		z = (4. / 3.) - 2. * i / (3. * N);
		ph = (j - fodd) * M_PI / (2. * N);
	} else {
		/* south polar cap, mirrored */
		const long ip = npix - pix;
		const long i = (1 + isqrt_l(2 * ip - 1)) >> 1;
		const long j = 4 * i + 1 - (ip - 2 * i * (i - 1));
		z = -1. + (double) (i * i) / (3. * N * N);
		ph = (j - 0.5) * M_PI / (2. * i);
	}
	*theta = acos(z);
	*phi = ph;
}

/* RING map of Re f at nside: map[12 nside^2] */
void or_healpix_map(const struct or_series *s, int nside, double *map)
{
	const long npix = 12L * nside * nside;
	for(long p = 0; p < npix; p++) {
		double th, ph;
		or_healpix_pix2ang(nside, p, &th, &ph);
		map[p] = creal(or_eval(s, th, ph));
	}
}

/* the posterior of the two beta skies (RapidLocalization.py:498-530):
 * each map exponentiated relative to its maximum and normalised to unit
 * sum, then the two averaged */
void or_healpix_posterior(const double *map_p, const double *map_n, long npix, double *prob)
{
	double mp = map_p[0], mn = map_n[0], sp = 0., sn = 0.;
	for(long p = 1; p < npix; p++) {
		mp = map_p[p] > mp ? map_p[p] : mp;
		mn = map_n[p] > mn ? map_n[p] : mn;
	}
	for(long p = 0; p < npix; p++) {
		sp += exp(map_p[p] - mp);
		sn += exp(map_n[p] - mn);
	}
	for(long p = 0; p < npix; p++)
		prob[p] = (exp(map_p[p] - mp) / sp + exp(map_n[p] - mn) / sn) / 2.;
}
