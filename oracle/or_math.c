/*
 * Oracle (test infrastructure only): the third-party numerics the
 * reference takes from GSL and FFTW, restated from their definitions.
 *
 *   or_gl_nodes      <- gsl_integration_glfixed_table_alloc/_point as used
 *                       by pixels_from_l_max (sh_series_trans.c:149-196)
 *   or_sphPlm_array  <- gsl_sf_legendre_sphPlm_array (sh_series_trans.c:93,
 *                       627, 715, 836, 847)
 *   or_wigner_3j     <- gsl_sf_coupling_3j (sh_series_product.c:80-83)
 *   or_dft           <- fftw c2c, FFTW_FORWARD = exp(-2 pi i jk/n),
 *                       unnormalised (correlator.c:216-243,
 *                       sh_series_array.c:528-559, sh_series_trans.c:601,
 *                       689, 860)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <quadmath.h>
#include "oracle.h"

/* Legendre P_n(x) and P_n'(x) by the three-term recurrence. */
static void legendre_pn(int n, long double x, long double *p, long double *dp)
{
	long double p0 = 1.L, p1 = x;
	if(n == 0) {
		*p = 1.L;
		*dp = 0.L;
		return;
	}
	for(int k = 2; k <= n; k++) {
		long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
		p0 = p1;
		p1 = p2;
	}
	*p = p1;
	*dp = n * (x * p1 - p0) / (x * x - 1.L);
}

/* n-point Gauss-Legendre rule on [-1, 1], nodes ascending (the order of
 * gsl_integration_glfixed_point), exactly antisymmetric nodes and
 * symmetric weights (GSL tabulates the positive half and mirrors it). */
void or_gl_nodes(int n, double *x, double *w)
{
	const int half = n / 2;
	for(int i = 0; i < half; i++) {
		/* i-th largest root; Tricomi initial estimate then Newton in
		 * long double */
		long double z = cosl(M_PI * (i + 0.75L) / (n + 0.5L));
		long double p, dp;
		for(int it = 0; it < 100; it++) {
			legendre_pn(n, z, &p, &dp);
			long double dz = p / dp;
			z -= dz;
			if(fabsl(dz) < 1e-19L)
				break;
		}
		legendre_pn(n, z, &p, &dp);
		long double wt = 2.L / ((1.L - z * z) * dp * dp);
		/* positive root goes to the top, negative to the bottom */
		x[n - 1 - i] = (double) z;
		x[i] = -(double) z;
		w[n - 1 - i] = w[i] = (double) wt;
	}
	if(n & 1) {
		long double p, dp;
		legendre_pn(n, 0.L, &p, &dp);
		x[half] = 0.;
		w[half] = (double) (2.L / (dp * dp));
	}
}

/* Orthonormal associated Legendre functions with the Condon-Shortley
 * phase, the normalisation of gsl_sf_legendre_sphPlm:
 *   Pbar_lm(x) = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m(x),
 *   P_l^m including (-1)^m.
 * out[l - m] for l = m .. l_max. */
void or_sphPlm_array(unsigned l_max, int m, double x, double *out)
{
	const double s = sqrt((1. - x) * (1. + x));
	/* Pbar_mm = (-1)^m sqrt((2m+1)/(4pi) (2m-1)!!/(2m)!!) s^m, built
	 * by Pbar_kk = -sqrt((2k+1)/(2k)) s Pbar_{k-1,k-1} */
	double pmm = 1. / sqrt(4. * M_PI);
	for(int k = 1; k <= m; k++)
		pmm *= -sqrt((2. * k + 1.) / (2. * k)) * s;
	if((unsigned) m > l_max)
		return;
	out[0] = pmm;
	if((unsigned) m == l_max)
		return;
	double pm1 = x * sqrt(2. * m + 3.) * pmm;
	out[1] = pm1;
	double p2 = pmm, p1 = pm1;
	for(unsigned l = m + 2; l <= l_max; l++) {
		const double ll = l, mm = m;
		const double a = sqrt((4. * ll * ll - 1.) / (ll * ll - mm * mm));
		const double b = sqrt(((ll - 1.) * (ll - 1.) - mm * mm) / (4. * (ll - 1.) * (ll - 1.) - 1.));
		const double p = a * (x * p1 - b * p2);
		out[l - m] = p;
		p2 = p1;
		p1 = p;
	}
}

double or_sphPlm(unsigned l, int m, double x)
{
	double *tmp = malloc((l - m + 1) * sizeof(*tmp));
	or_sphPlm_array(l, m, x, tmp);
	double v = tmp[l - m];
	free(tmp);
	return v;
}

/* Wigner 3j by Racah's closed form.  The alternating sum cancels
 * heavily for j ~ 50, so it is evaluated in binary128 (libquadmath). */
#define QFACT_MAX 1200
static __float128 qfact(int n)
{
	static __float128 tab[QFACT_MAX + 1];
	static int filled = 0;
	if(!filled) {
		tab[0] = 1.Q;
		for(int k = 1; k <= QFACT_MAX; k++)
			tab[k] = tab[k - 1] * k;
		filled = 1;
	}
	return tab[n];
}

double or_wigner_3j(int j1, int j2, int j3, int m1, int m2, int m3)
{
	if(m1 + m2 + m3 != 0)
		return 0.;
	if(abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3)
		return 0.;
	if(j3 > j1 + j2 || j3 < abs(j1 - j2))
		return 0.;
	if(m1 == 0 && m2 == 0 && m3 == 0 && ((j1 + j2 + j3) & 1))
		return 0.;
	__float128 delta = qfact(j1 + j2 - j3) * qfact(j1 - j2 + j3) * qfact(-j1 + j2 + j3) / qfact(j1 + j2 + j3 + 1);
	__float128 pre = sqrtq(delta * qfact(j1 + m1) * qfact(j1 - m1) * qfact(j2 + m2) * qfact(j2 - m2) * qfact(j3 + m3) * qfact(j3 - m3));
	int kmin = 0;
	if(j2 - j3 - m1 > kmin)
		kmin = j2 - j3 - m1;
	if(j1 - j3 + m2 > kmin)
		kmin = j1 - j3 + m2;
	int kmax = j1 + j2 - j3;
	if(j1 - m1 < kmax)
		kmax = j1 - m1;
	if(j2 + m2 < kmax)
		kmax = j2 + m2;
	__float128 sum = 0.Q;
	for(int k = kmin; k <= kmax; k++) {
		__float128 t = 1.Q / (qfact(k) * qfact(j3 - j2 + k + m1) * qfact(j3 - j1 + k - m2) * qfact(j1 + j2 - j3 - k) * qfact(j1 - k - m1) * qfact(j2 - k + m2));
		sum += (k & 1) ? -t : t;
	}
	int ph = j1 - j2 - m3;
	__float128 r = pre * sum;
	return (double) ((ph & 1) ? -r : r);
}

/* Unnormalised DFT, out[k] = sum_j in[j] exp(sign 2 pi i jk / n), by
 * recursive mixed-radix decimation in time; direct sums for prime
 * lengths.  Twiddle angles are reduced exactly (integer t mod n). */
static void dft_rec(const double complex *in, int stride, double complex *out, int n, const double complex *tw, int tw_n)
{
	/* tw[t] = exp(sign 2 pi i t / tw_n); this sub-transform uses
	 * exp(sign 2 pi i t / n) = tw[t * (tw_n / n)] */
	const int step = tw_n / n;
	int p = 2;
	while(p * p <= n && n % p)
		p++;
	if(n % p)
		p = n;
	if(p == n) {
		for(int k = 0; k < n; k++) {
			double complex acc = 0.;
			for(int j = 0; j < n; j++)
				acc += in[(size_t) j * stride] * tw[(size_t) ((long) j * k % n) * step];
			out[k] = acc;
		}
		return;
	}
	const int m = n / p;
	double complex *sub = malloc((size_t) n * sizeof(*sub));
	for(int r = 0; r < p; r++)
		dft_rec(in + (size_t) r * stride, stride * p, sub + (size_t) r * m, m, tw, tw_n);
	for(int q = 0; q < p; q++)
		for(int k = 0; k < m; k++) {
			const int kk = k + m * q;
			double complex acc = 0.;
			for(int r = 0; r < p; r++)
				acc += sub[(size_t) r * m + k] * tw[(size_t) ((long) r * kk % n) * step];
			out[kk] = acc;
		}
	free(sub);
}

/* Bluestein: a length-n DFT as a circular convolution of length m >= 2n-1
 * (a power of two) with the chirp exp(sign i pi k^2 / n), k^2 reduced
 * exactly mod 2n.  Used when n has a prime factor above 64, where the
 * direct sums would cost O(p n) (FFTW also transforms such lengths in
 * O(n log n)). */
static void dft_bluestein(double complex *data, int n, int sign)
{
	int m = 1;
	while(m < 2 * n - 1)
		m <<= 1;
	double complex *chirp = malloc((size_t) n * sizeof(*chirp));
	for(long k = 0; k < n; k++) {
		const long t = (k * k) % (2L * n);
		const double ang = M_PI * (double) t / n;
		chirp[k] = cos(ang) + I * (sign < 0 ? -sin(ang) : sin(ang));
	}
	double complex *a = calloc(m, sizeof(*a)), *b = calloc(m, sizeof(*b));
	for(int k = 0; k < n; k++)
		a[k] = data[k] * chirp[k];
	b[0] = conj(chirp[0]);
	for(int k = 1; k < n; k++)
		b[k] = b[m - k] = conj(chirp[k]);
	or_dft(a, m, -1);
	or_dft(b, m, -1);
	for(int k = 0; k < m; k++)
		a[k] *= b[k];
	or_dft(a, m, +1);
	for(int k = 0; k < n; k++)
		data[k] = a[k] * chirp[k] / m;
	free(a);
	free(b);
	free(chirp);
}

static int largest_prime_factor(int n)
{
	int p = 1;
	for(int q = 2; (long) q * q <= n; q++)
		while(n % q == 0) {
			p = q;
			n /= q;
		}
	return n > 1 ? n : p;
}

void or_dft(double complex *data, int n, int sign)
{
	if(n <= 1)
		return;
	if(largest_prime_factor(n) > 64) {
		dft_bluestein(data, n, sign);
		return;
	}
	/* the twiddles of the last length and sign, per thread: the
	 * pipelines call the same transform many times */
	static __thread double complex *tw;
	static __thread int tw_n, tw_sign;
	if(!tw || tw_n != n || tw_sign != sign) {
		free(tw);
		tw = malloc((size_t) n * sizeof(*tw));
		for(int t = 0; t < n; t++) {
			double ang = 2. * M_PI * (double) t / n;
			tw[t] = cos(ang) + I * (sign < 0 ? -sin(ang) : sin(ang));
		}
		tw_n = n;
		tw_sign = sign;
	}
	double complex *out = malloc((size_t) n * sizeof(*out));
	dft_rec(data, 1, out, n, tw, n);
	memcpy(data, out, (size_t) n * sizeof(*out));
	free(out);
}
