/*
 * Plan-time numerics of the host library: Gauss-Legendre rules, normalised
 * associated Legendre functions, Wigner 3j symbols and a host DFT.  These
 * stand in for the GSL / FFTW routines the reference calls at plan time
 * (sh_series_trans.c:149-196 glfixed, sphPlm; sh_series_product.c:80-83
 * coupling_3j; sh_series_array.c:528-559 FFTW).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "../sphr_internal.h"

/*
 * Gauss-Legendre.  Roots of P_n by Newton's method in long double, started
 * from the asymptotic estimate x ~ cos(pi (i - 1/4) / (n + 1/2)) (Tricomi);
 * the weight is 2 / ((1 - x^2) P_n'(x)^2).  The negative half mirrors the
 * positive half so nodes are exactly antisymmetric.
 */
static void pn_and_derivative(int n, long double x, long double *pn, long double *dpn)
{
	long double pkm1 = 1.L, pk = x;
	for(int k = 1; k < n; k++) {
		long double pkp1 = ((2 * k + 1) * x * pk - k * pkm1) / (k + 1);
		pkm1 = pk;
		pk = pkp1;
	}
	*pn = n ? pk : 1.L;
	*dpn = n ? n * (pkm1 - x * pk) / (1.L - x * x) : 0.L;
}

void sphr_gl_rule(int n, double *x, double *w)
{
	for(int i = 1; i <= n / 2; i++) {
		long double t = cosl(3.14159265358979323846264338327950288L * (i - 0.25L) / (n + 0.5L));
		long double p, dp;
		for(int it = 0; it < 64; it++) {
			pn_and_derivative(n, t, &p, &dp);
			const long double step = p / dp;
			t -= step;
			if(fabsl(step) <= 4e-20L * fabsl(t))
				break;
		}
		pn_and_derivative(n, t, &p, &dp);
		const double wt = (double) (2.L / ((1.L - t * t) * dp * dp));
		x[n - i] = (double) t;
		x[i - 1] = -(double) t;
		w[n - i] = w[i - 1] = wt;
	}
	if(n & 1) {
		long double p, dp;
		pn_and_derivative(n, 0.L, &p, &dp);
		x[n / 2] = 0.;
		w[n / 2] = (double) (2.L / (dp * dp));
	}
}

/* cache of mesh rules, theta ascending (cos theta descending) */
struct mesh_rule {
	unsigned l_max;
	double *x, *w;
	struct mesh_rule *next;
};
static struct mesh_rule *mesh_rules;
static pthread_mutex_t mesh_lock = PTHREAD_MUTEX_INITIALIZER;

static struct mesh_rule *gl_rules;	/* keyed by the node count in l_max */

/* find key in list under the lock; NULL if absent */
static struct mesh_rule *rule_find(struct mesh_rule *list, unsigned key)
{
	for(struct mesh_rule *r = list; r; r = r->next)
		if(r->l_max == key)
			return r;
	return NULL;
}

/* insert a rule computed outside the lock, unless another thread has
 * inserted the same key meanwhile (then the copy is dropped); returns the
 * cached entry */
static struct mesh_rule *rule_insert(struct mesh_rule **list, struct mesh_rule *mine)
{
	pthread_mutex_lock(&mesh_lock);
	struct mesh_rule *r = rule_find(*list, mine->l_max);
	if(!r) {
		mine->next = *list;
		*list = mine;
		r = mine;
	}
	pthread_mutex_unlock(&mesh_lock);
	if(r != mine) {
		free(mine->x);
		free(mine->w);
		free(mine);
	}
	return r;
}

void sphr_gl_rule_cached(int n, const double **x, const double **w)
{
	pthread_mutex_lock(&mesh_lock);
	struct mesh_rule *r = rule_find(gl_rules, (unsigned) n);
	pthread_mutex_unlock(&mesh_lock);
	if(!r) {
		r = malloc(sizeof(*r));
		r->x = malloc(n * sizeof(*r->x));
		r->w = malloc(n * sizeof(*r->w));
		sphr_gl_rule(n, r->x, r->w);
		r->l_max = (unsigned) n;
		r = rule_insert(&gl_rules, r);
	}
	*x = r->x;
	*w = r->w;
}

void sphr_mesh_rule(unsigned l_max, const double **cos_theta, const double **weights)
{
	pthread_mutex_lock(&mesh_lock);
	struct mesh_rule *r = rule_find(mesh_rules, l_max);
	pthread_mutex_unlock(&mesh_lock);
	if(!r) {
		const int n = 2 * ((int) l_max + 1);
		double *xa = malloc(n * sizeof(*xa)), *wa = malloc(n * sizeof(*wa));
		r = malloc(sizeof(*r));
		r->x = malloc(n * sizeof(*r->x));
		r->w = malloc(n * sizeof(*r->w));
		sphr_gl_rule(n, xa, wa);
		for(int i = 0; i < n; i++) {
			r->x[i] = xa[n - 1 - i];
			r->w[i] = wa[n - 1 - i];
		}
		free(xa);
		free(wa);
		r->l_max = l_max;
		r = rule_insert(&mesh_rules, r);
	}
	*cos_theta = r->x;
	if(weights)
		*weights = r->w;
}

/* fill both rule caches for a set of tap-mesh orders and product-rule
 * node counts, the missing rules computed in parallel */
void sphr_rules_prefetch(const int *mesh_orders, int n_mesh, const int *gl_counts, int n_gl)
{
	#pragma omp parallel for schedule(dynamic, 1)
	for(int i = 0; i < n_mesh + n_gl; i++) {
		const double *x, *w;
		if(i < n_mesh)
			sphr_mesh_rule((unsigned) mesh_orders[i], &x, &w);
		else
			sphr_gl_rule_cached(gl_counts[i - n_mesh], &x, &w);
	}
}

/*
 * Normalised associated Legendre functions
 *   Pbar_lm(x) = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m(x)
 * with the Condon-Shortley phase in P_l^m.  Sectoral start
 *   Pbar_mm = (-1)^m sqrt((2m+1)!!/(4 pi (2m)!!)) (1-x^2)^{m/2}
 * accumulated as a product of ratios, then the l recurrence
 *   Pbar_lm = alpha_lm (x Pbar_{l-1,m} - Pbar_{l-2,m} / alpha_{l-1,m}),
 *   alpha_lm = sqrt((4l^2 - 1)/(l^2 - m^2)).
 */
void sphr_plm(unsigned l_max, int m, double x, double *out)
{
	if((unsigned) m > l_max)
		return;
	const double sin_t = sqrt((1.0 - x) * (1.0 + x));
	double p = 0.28209479177387814347;	/* 1 / sqrt(4 pi) */
	for(int k = 1; k <= m; k++)
		p *= -sqrt(1.0 + 0.5 / k) * sin_t;
	out[0] = p;
	if((unsigned) m == l_max)
		return;
	double prev = p;
	double alpha_prev = sqrt(2.0 * m + 3.0);
	double cur = alpha_prev * x * p;
	out[1] = cur;
	for(unsigned l = m + 2; l <= l_max; l++) {
		const double dl = l, dm = m;
		const double alpha = sqrt((4.0 * dl * dl - 1.0) / ((dl - dm) * (dl + dm)));
		const double next = alpha * (x * cur - prev / alpha_prev);
		out[l - m] = next;
		prev = cur;
		cur = next;
		alpha_prev = alpha;
	}
}

/*
 * Wigner 3j symbols (j1 j2 j; m1 m2 m), m = -m1 - m2, for every j at once
 * by the three-term recurrence of Schulten & Gordon (J. Math. Phys. 16,
 * 1961, 1975), as organised by Luscombe & Luban (Phys. Rev. E 57, 7274,
 * 1998):
 *   j A(j+1) f(j+1) + B(j) f(j) + (j+1) A(j) f(j-1) = 0,
 *   A(j) = sqrt((j^2 - (j1-j2)^2) ((j1+j2+1)^2 - j^2) (j^2 - m^2)),
 *   B(j) = -(2j+1) (j1(j1+1) m - j2(j2+1) m - j(j+1)(m2 - m1)).
 * Forward from jmin while |f| grows (the non-classical region), backward
 * from jmax to that point, match, normalise sum (2j+1) f^2 = 1 and fix the
 * sign by f(jmax) having sign (-1)^(j1-j2-m).
 */
static double sg_A(int j, int j1, int j2, int m)
{
	const double a = (double) j * j - (double) (j1 - j2) * (j1 - j2);
	const double b = (double) (j1 + j2 + 1) * (j1 + j2 + 1) - (double) j * j;
	const double c = (double) j * j - (double) m * m;
	const double v = a * b * c;
	return v > 0 ? sqrt(v) : 0.0;
}

static double sg_B(int j, int j1, int j2, int m1, int m2)
{
	const int m = -m1 - m2;
	return -(2.0 * j + 1.0) * ((double) j1 * (j1 + 1) * m - (double) j2 * (j2 + 1) * m - (double) j * (j + 1) * (m2 - m1));
}

int sphr_3j_range(int j1, int j2, int m1, int m2, double *out)
{
	const int m = -m1 - m2;
	if(abs(m1) > j1 || abs(m2) > j2)
		return -1;
	const int jmin = abs(j1 - j2) > abs(m) ? abs(j1 - j2) : abs(m);
	const int jmax = j1 + j2;
	if(jmin > jmax)
		return -1;
	const int n = jmax - jmin + 1;
	double *f = out;
	if(n == 1) {
		f[0] = 1.0;
	} else {
		/* forward from jmin while the magnitude grows */
		int top = 0;	/* last index computed by the forward pass */
		if(jmin > 0) {
			f[0] = 1.0;
			f[1] = -sg_B(jmin, j1, j2, m1, m2) / ((double) jmin * sg_A(jmin + 1, j1, j2, m));
			top = 1;
			while(jmin + top < jmax && fabs(f[top]) > fabs(f[top - 1])) {
				const int j = jmin + top;
				f[top + 1] = -(sg_B(j, j1, j2, m1, m2) * f[top] + (j + 1) * sg_A(j, j1, j2, m) * f[top - 1]) / ((double) j * sg_A(j + 1, j1, j2, m));
				top++;
			}
		}
		/* backward from jmax down to one below the forward top */
		const int low = top > 0 ? top - 1 : 0;
		double *g = malloc(n * sizeof(*g));
		g[n - 1] = 1.0;
		if(n - 2 >= low)
			g[n - 2] = -sg_B(jmax, j1, j2, m1, m2) / ((jmax + 1.0) * sg_A(jmax, j1, j2, m));
		for(int i = n - 2; i - 1 >= low; i--) {
			const int j = jmin + i;
			g[i - 1] = -(sg_B(j, j1, j2, m1, m2) * g[i] + (double) j * sg_A(j + 1, j1, j2, m) * g[i + 1]) / ((j + 1.0) * sg_A(j, j1, j2, m));
		}
		if(top == 0) {
			memcpy(f, g, n * sizeof(*f));
		} else {
			/* match where the backward values are larger */
			const int im = fabs(g[top]) >= fabs(g[top - 1]) ? top : top - 1;
			const double s = f[im] / g[im];
			for(int i = im; i < n; i++)
				f[i] = s * g[i];
		}
		free(g);
	}
	double norm = 0.0;
	for(int j = jmin; j <= jmax; j++)
		norm += (2.0 * j + 1.0) * f[j - jmin] * f[j - jmin];
	norm = 1.0 / sqrt(norm);
	const int want_neg = (j1 - j2 - m) & 1;
	if((f[n - 1] < 0) != want_neg)
		norm = -norm;
	for(int j = 0; j < n; j++)
		f[j] *= norm;
	return jmin;
}

/*
 * Host DFT: iterative radix-2 for powers of two, Bluestein (chirp-z, with
 * k^2 reduced mod 2n) for every other length.  Twiddles are computed from
 * exactly reduced integer angles.
 */
static void fft_pow2(double *d, int n, int sign)
{
	for(int i = 1, j = 0; i < n; i++) {
		int bit = n >> 1;
		for(; j & bit; bit >>= 1)
			j ^= bit;
		j ^= bit;
		if(i < j) {
			double t0 = d[2 * i], t1 = d[2 * i + 1];
			d[2 * i] = d[2 * j];
			d[2 * i + 1] = d[2 * j + 1];
			d[2 * j] = t0;
			d[2 * j + 1] = t1;
		}
	}
	double *tw = malloc(n * sizeof(*tw));	/* n/2 complex twiddles */
	for(int k = 0; k < n / 2; k++) {
		const double ang = 2.0 * M_PI * k / n;
		tw[2 * k] = cos(ang);
		tw[2 * k + 1] = sign * sin(ang);
	}
	for(int len = 2; len <= n; len <<= 1) {
		const int half = len >> 1, step = n / len;
		for(int i = 0; i < n; i += len)
			for(int k = 0; k < half; k++) {
				const double wr = tw[2 * k * step], wi = tw[2 * k * step + 1];
				double *u = d + 2 * (i + k), *v = d + 2 * (i + k + half);
				const double vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
				v[0] = u[0] - vr;
				v[1] = u[1] - vi;
				u[0] += vr;
				u[1] += vi;
			}
	}
	free(tw);
}

void sphr_host_dft(double *d, int n, int sign)
{
	if(n <= 1)
		return;
	if(!(n & (n - 1))) {
		fft_pow2(d, n, sign);
		return;
	}
	int m = 1;
	while(m < 2 * n - 1)
		m <<= 1;
	double *ch = malloc(2 * n * sizeof(*ch));
	double *a = calloc(2 * m, sizeof(*a)), *b = calloc(2 * m, sizeof(*b));
	for(long k = 0; k < n; k++) {
		const double ang = M_PI * (double) ((k * k) % (2L * n)) / n;
		ch[2 * k] = cos(ang);
		ch[2 * k + 1] = sign * sin(ang);
	}
	for(int k = 0; k < n; k++) {
		a[2 * k] = d[2 * k] * ch[2 * k] - d[2 * k + 1] * ch[2 * k + 1];
		a[2 * k + 1] = d[2 * k] * ch[2 * k + 1] + d[2 * k + 1] * ch[2 * k];
		b[2 * k] = ch[2 * k];
		b[2 * k + 1] = -ch[2 * k + 1];
		if(k) {
			b[2 * (m - k)] = ch[2 * k];
			b[2 * (m - k) + 1] = -ch[2 * k + 1];
		}
	}
	fft_pow2(a, m, -1);
	fft_pow2(b, m, -1);
	for(int k = 0; k < m; k++) {
		const double ar = a[2 * k], ai = a[2 * k + 1];
		a[2 * k] = ar * b[2 * k] - ai * b[2 * k + 1];
		a[2 * k + 1] = ar * b[2 * k + 1] + ai * b[2 * k];
	}
	fft_pow2(a, m, +1);
	for(int k = 0; k < n; k++) {
		const double ar = a[2 * k] / m, ai = a[2 * k + 1] / m;
		d[2 * k] = ar * ch[2 * k] - ai * ch[2 * k + 1];
		d[2 * k + 1] = ar * ch[2 * k + 1] + ai * ch[2 * k];
	}
	free(a);
	free(b);
	free(ch);
}
