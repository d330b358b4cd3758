/*
 * The precalc tree of the localization plans (inspiral_sky_map.c:1113-1811
 * of the reference): the FD network plans of the beta = +1 and -1
 * polarisations after the projection multiply, their baselines, the log
 * prior, the time-series length and the autocorrelator projections, saved
 * so that a later run skips plan generation.
 *
 * Layout under <path>; every .dat file is the native-endian bytes of the
 * listed fields written one after another (no headers, no padding), as
 * the reference's fwrite sequences:
 *
 *   time_series_length.dat                      unsigned int
 *   sh_series/logprior.fits                     HEALPix a_lm table
 *   correlator_network_plan_fd/correlator_network_baselines/
 *     n_baselines.dat                           int
 *     baselines/<k>/correlator_baseline.dat     int index_a, int index_b,
 *                                               double theta, double phi
 *     baselines/<k>/d.dat                       double d[3]
 *   correlator_network_plan_fd/correlator_plan_fd/<k>/
 *     correlator_plan_fd.dat                    double delta_t, int transient
 *     rotation_plan/l_max.dat                   unsigned int
 *     rotation_plan/D.dat                       for l = 1..l_max the
 *                                               (2l+1)^2 complex doubles of
 *                                               D^l, element (m, m') at
 *                                               (2l+1)(m'+l) + (m+l)
 *     delay_product_p/sh_series_array.dat       unsigned l_max, int polar,
 *     delay_product_n/sh_series_array.dat       int n, int stride, then
 *                                               n * stride complex doubles
 *   autocorrelator_network_plan_fd/length.dat   int
 *   autocorrelator_network_plan_fd/psh_series/<i>/sh_series_{p,n}/
 *     projections<k>.fits                       HEALPix a_lm table
 *
 * The two polarisations share baselines, delta_t, transient and rotation
 * plan; only the delay products differ, so one rotation plan is stored.
 * Readers check every read (the reference checks only the baseline's), and
 * fail with a message for sphr_last_error().
 */
#include <complex.h>
#include <errno.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <sphradiometer/correlator.h>
#include <sphradiometer/inspiral_sky_map.h>
#include <sphradiometer/instrument.h>
#include <sphradiometer/sh_series.h>
#include "../sphr_internal.h"

#define PLAN_DIR "correlator_network_plan_fd"
#define BASELINES_DIR PLAN_DIR "/correlator_network_baselines"
#define AUTO_DIR "autocorrelator_network_plan_fd"

/* <parent>/<formatted rest> in a new buffer */
static char *path_of(const char *parent, const char *fmt, va_list ap)
{
	va_list aq;
	va_copy(aq, ap);
	const int n = vsnprintf(NULL, 0, fmt, aq);
	va_end(aq);
	const size_t lp = strlen(parent);
	char *s = malloc(lp + 1 + (size_t) n + 1);
	if(!s)
		return NULL;
	memcpy(s, parent, lp);
	s[lp] = '/';
	vsnprintf(s + lp + 1, (size_t) n + 1, fmt, ap);
	return s;
}

static char *path_fmt(const char *parent, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	char *s = path_of(parent, fmt, ap);
	va_end(ap);
	return s;
}

/* fopen of <parent>/<fmt ...>; NULL with an error message on failure */
static FILE *open_at(const char *mode, const char *parent, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	char *s = path_of(parent, fmt, ap);
	va_end(ap);
	if(!s) {
		sphr_set_error("precalc: out of memory");
		return NULL;
	}
	FILE *fp = fopen(s, mode);
	if(!fp)
		sphr_set_error("%s: %s", s, strerror(errno));
	free(s);
	return fp;
}

/* fwrite / fread of n items, recording the failure */
static int put(FILE *fp, const void *p, size_t size, size_t n, int *ok)
{
	if(*ok && fwrite(p, size, n, fp) != n) {
		sphr_set_error("precalc: write failed");
		*ok = 0;
	}
	return *ok;
}

static int get(FILE *fp, void *p, size_t size, size_t n, int *ok)
{
	if(*ok && fread(p, size, n, fp) != n) {
		sphr_set_error("precalc: short or unreadable file");
		*ok = 0;
	}
	return *ok;
}

/* fclose folding a failed close (unflushed write) into ok; 0 or -1 */
static int finish(FILE *fp, int ok)
{
	if(fclose(fp) && ok) {
		sphr_set_error("precalc: write failed");
		ok = 0;
	}
	return ok ? 0 : -1;
}

/* ------------------------------------------------------------ directories */

static int make_dir(const char *parent, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	char *s = fmt ? path_of(parent, fmt, ap) : strdup(parent);
	va_end(ap);
	if(!s) {
		sphr_set_error("precalc: out of memory");
		return -1;
	}
	const int r = mkdir(s, S_IRWXU | S_IRWXG | S_IRWXO);
	if(r)
		sphr_set_error("error making %s: %s", s, strerror(errno));
	free(s);
	return r ? -1 : 0;
}

/*
 * inspiral_sky_map.c:1458-1553.  The reference makes the per-baseline
 * folders for k < n_ifo, which covers the n_ifo (n_ifo - 1) / 2 baselines
 * only up to three detectors; here they are made for k up to the larger
 * of the two counts, so the tree holds every baseline of larger networks.
 * The folder must not exist yet (mkdir fails on an existing one, as in the
 * reference).
 */
int make_precalc_directories(char *precalc_path, int n_ifo)
{
	if(!precalc_path || n_ifo < 0) {
		sphr_set_error("make_precalc_directories: bad arguments");
		return -1;
	}
	const int n_bl = n_ifo * (n_ifo - 1) / 2;
	const int n_k = n_bl > n_ifo ? n_bl : n_ifo;
	if(make_dir(precalc_path, NULL) || make_dir(precalc_path, PLAN_DIR) || make_dir(precalc_path, BASELINES_DIR) || make_dir(precalc_path, BASELINES_DIR "/baselines"))
		return -1;
	for(int k = 0; k < n_k; k++)
		if(make_dir(precalc_path, BASELINES_DIR "/baselines/%d", k))
			return -1;
	if(make_dir(precalc_path, PLAN_DIR "/correlator_plan_fd"))
		return -1;
	for(int k = 0; k < n_k; k++)
		if(make_dir(precalc_path, PLAN_DIR "/correlator_plan_fd/%d", k) || make_dir(precalc_path, PLAN_DIR "/correlator_plan_fd/%d/delay_product_n", k) || make_dir(precalc_path, PLAN_DIR "/correlator_plan_fd/%d/delay_product_p", k) || make_dir(precalc_path, PLAN_DIR "/correlator_plan_fd/%d/rotation_plan", k))
			return -1;
	if(make_dir(precalc_path, AUTO_DIR) || make_dir(precalc_path, AUTO_DIR "/psh_series"))
		return -1;
	for(int k = 0; k < n_ifo; k++)
		if(make_dir(precalc_path, AUTO_DIR "/psh_series/%d", k) || make_dir(precalc_path, AUTO_DIR "/psh_series/%d/sh_series_n", k) || make_dir(precalc_path, AUTO_DIR "/psh_series/%d/sh_series_p", k))
			return -1;
	return make_dir(precalc_path, "sh_series");
}

/* --------------------------------------------------- time series, prior */

/* inspiral_sky_map.c:1795-1811 */
int write_precalc_time_series_length(unsigned int length, char *parent_dir)
{
	FILE *fp = open_at("wb", parent_dir, "time_series_length.dat");
	if(!fp)
		return -1;
	int ok = 1;
	put(fp, &length, sizeof(length), 1, &ok);
	return finish(fp, ok);
}

/* inspiral_sky_map.c:1439-1456 */
int read_precalc_time_series_length(unsigned int *length, char *parent_dir)
{
	FILE *fp = open_at("rb", parent_dir, "time_series_length.dat");
	if(!fp)
		return -1;
	int ok = 1;
	get(fp, length, sizeof(*length), 1, &ok);
	return finish(fp, ok);
}

/* inspiral_sky_map.c:1787-1792 */
int write_precalc_logprior(const struct sh_series *series, char *parent_dir)
{
	char *s = path_fmt(parent_dir, "sh_series/logprior.fits");
	const int r = s ? sh_series_write_healpix_alm(series, s) : -1;
	free(s);
	return r;
}

/* inspiral_sky_map.c:1429-1434 */
struct sh_series *read_precalc_logprior(char *parent_dir)
{
	char *s = path_fmt(parent_dir, "sh_series/logprior.fits");
	struct sh_series *r = s ? sh_series_read_healpix_alm(s) : NULL;
	free(s);
	return r;
}

/* -------------------------------------------------------------- baselines */

/* inspiral_sky_map.c:1556-1588 */
static int write_baseline(const struct correlator_baseline *b, int k, const char *parent)
{
	FILE *fp = open_at("wb", parent, BASELINES_DIR "/baselines/%d/correlator_baseline.dat", k);
	if(!fp)
		return -1;
	int ok = 1;
	put(fp, &b->index_a, sizeof(b->index_a), 1, &ok);
	put(fp, &b->index_b, sizeof(b->index_b), 1, &ok);
	put(fp, &b->theta, sizeof(b->theta), 1, &ok);
	put(fp, &b->phi, sizeof(b->phi), 1, &ok);
	if(finish(fp, ok))
		return -1;
	/* gsl_vector_fwrite: the elements in index order */
	fp = open_at("wb", parent, BASELINES_DIR "/baselines/%d/d.dat", k);
	if(!fp)
		return -1;
	for(size_t i = 0; i < b->d->size; i++) {
		const double v = gsl_vector_get(b->d, i);
		put(fp, &v, sizeof(v), 1, &ok);
	}
	return finish(fp, ok);
}

/* inspiral_sky_map.c:1591-1615 */
static int write_network_baselines(const struct correlator_network_baselines *bl, const char *parent)
{
	FILE *fp = open_at("wb", parent, BASELINES_DIR "/n_baselines.dat");
	if(!fp)
		return -1;
	int ok = 1;
	put(fp, &bl->n_baselines, sizeof(bl->n_baselines), 1, &ok);
	if(finish(fp, ok))
		return -1;
	for(int k = 0; k < bl->n_baselines; k++)
		if(write_baseline(bl->baselines[k], k, parent))
			return -1;
	return 0;
}

/*
 * inspiral_sky_map.c:1113-1167: the baseline is rebuilt from the
 * instruments and must reproduce the stored direction exactly; d.dat is
 * not read (the reference's reader of it is compiled out).
 */
static struct correlator_baseline *read_baseline(const struct instrument_array *instruments, int k, const char *parent)
{
	FILE *fp = open_at("rb", parent, BASELINES_DIR "/baselines/%d/correlator_baseline.dat", k);
	if(!fp)
		return NULL;
	int index_a, index_b, ok = 1;
	double theta, phi;
	get(fp, &index_a, sizeof(index_a), 1, &ok);
	get(fp, &index_b, sizeof(index_b), 1, &ok);
	get(fp, &theta, sizeof(theta), 1, &ok);
	get(fp, &phi, sizeof(phi), 1, &ok);
	fclose(fp);
	if(!ok)
		return NULL;
	const int n = instrument_array_len(instruments);
	if(index_a < 0 || index_a >= n || index_b < 0 || index_b >= n) {
		sphr_set_error("precalc baseline %d: instrument index (%d, %d) outside the %d instruments", k, index_a, index_b, n);
		return NULL;
	}
	struct correlator_baseline *b = correlator_baseline_new(instruments, index_a, index_b);
	if(!b) {
		sphr_set_error("precalc baseline %d: correlator_baseline_new failed", k);
		return NULL;
	}
	if(b->theta != theta || b->phi != phi) {
		sphr_set_error("inconsistent data in correlator_baseline.dat of baseline %d", k);
		correlator_baseline_free(b);
		return NULL;
	}
	return b;
}

/* inspiral_sky_map.c:1170-1190; NULL if any baseline fails (the
 * reference keeps a NULL entry) */
struct correlator_network_baselines *read_precalc_correlator_network_baselines(const struct instrument_array *instruments, char *parent_dir)
{
	FILE *fp = open_at("rb", parent_dir, BASELINES_DIR "/n_baselines.dat");
	if(!fp)
		return NULL;
	int n = 0, ok = 1;
	get(fp, &n, sizeof(n), 1, &ok);
	fclose(fp);
	if(!ok)
		return NULL;
	if(n < 0) {
		sphr_set_error("precalc: negative baseline count %d", n);
		return NULL;
	}
	struct correlator_network_baselines *net = malloc(sizeof(*net));
	struct correlator_baseline **b = calloc(n > 0 ? n : 1, sizeof(*b));
	if(!net || !b) {
		free(net);
		free(b);
		sphr_set_error("precalc: out of memory");
		return NULL;
	}
	net->baselines = b;
	net->n_baselines = 0;
	for(int k = 0; k < n; k++, net->n_baselines++)
		if(!(b[k] = read_baseline(instruments, k, parent_dir))) {
			correlator_network_baselines_free(net);
			return NULL;
		}
	return net;
}

/* ----------------------------------------------------------- plan pieces */

/* inspiral_sky_map.c:1619-1649 */
static int write_rotation_plan(const struct sh_series_rotation_plan *p, int k, const char *parent)
{
	FILE *fp = open_at("wb", parent, PLAN_DIR "/correlator_plan_fd/%d/rotation_plan/l_max.dat", k);
	if(!fp)
		return -1;
	int ok = 1;
	put(fp, &p->l_max, sizeof(p->l_max), 1, &ok);
	if(finish(fp, ok))
		return -1;
	fp = open_at("wb", parent, PLAN_DIR "/correlator_plan_fd/%d/rotation_plan/D.dat", k);
	if(!fp)
		return -1;
	for(unsigned l = 1; l <= p->l_max; l++) {
		const size_t w = 2 * l + 1;
		put(fp, p->D[l] - (w * l + l), sizeof(double complex), w * w, &ok);
	}
	return finish(fp, ok);
}

/* inspiral_sky_map.c:1193-1218 */
static struct sh_series_rotation_plan *read_rotation_plan(int k, const char *parent)
{
	FILE *fp = open_at("rb", parent, PLAN_DIR "/correlator_plan_fd/%d/rotation_plan/l_max.dat", k);
	if(!fp)
		return NULL;
	unsigned l_max = 0;
	int ok = 1;
	get(fp, &l_max, sizeof(l_max), 1, &ok);
	fclose(fp);
	if(!ok)
		return NULL;
	if(l_max > 65535) {
		sphr_set_error("precalc rotation plan %d: implausible l_max %u", k, l_max);
		return NULL;
	}
	fp = open_at("rb", parent, PLAN_DIR "/correlator_plan_fd/%d/rotation_plan/D.dat", k);
	if(!fp)
		return NULL;
	struct sh_series_rotation_plan *p = sphr_rotation_plan_alloc(l_max);
	if(!p) {
		fclose(fp);
		sphr_set_error("precalc: out of memory");
		return NULL;
	}
	for(unsigned l = 1; l <= l_max; l++) {
		const size_t w = 2 * l + 1;
		get(fp, p->D[l] - (w * l + l), sizeof(double complex), w * w, &ok);
	}
	fclose(fp);
	if(!ok) {
		sh_series_rotation_plan_free(p);
		return NULL;
	}
	return p;
}

/* inspiral_sky_map.c:1676-1707 (one polarisation) */
static int write_delay_product(const struct sh_series_array *a, int k, const char *sign, const char *parent)
{
	FILE *fp = open_at("wb", parent, PLAN_DIR "/correlator_plan_fd/%d/delay_product_%s/sh_series_array.dat", k, sign);
	if(!fp)
		return -1;
	int ok = 1;
	put(fp, &a->l_max, sizeof(a->l_max), 1, &ok);
	put(fp, &a->polar, sizeof(a->polar), 1, &ok);
	put(fp, &a->n, sizeof(a->n), 1, &ok);
	put(fp, &a->stride, sizeof(a->stride), 1, &ok);
	put(fp, a->coeff, sizeof(*a->coeff), (size_t) a->n * a->stride, &ok);
	return finish(fp, ok);
}

/* inspiral_sky_map.c:1239-1292 (one polarisation).  The array is built
 * at its packed stride; a stored stride beyond the series length (padding)
 * is dropped row by row. */
static struct sh_series_array *read_delay_product(int k, const char *sign, const char *parent)
{
	FILE *fp = open_at("rb", parent, PLAN_DIR "/correlator_plan_fd/%d/delay_product_%s/sh_series_array.dat", k, sign);
	if(!fp)
		return NULL;
	unsigned l_max = 0;
	int polar = 0, n = 0, stride = 0, ok = 1;
	get(fp, &l_max, sizeof(l_max), 1, &ok);
	get(fp, &polar, sizeof(polar), 1, &ok);
	get(fp, &n, sizeof(n), 1, &ok);
	get(fp, &stride, sizeof(stride), 1, &ok);
	struct sh_series_array *a = NULL;
	if(ok && (l_max > 65535 || n < 0 || stride < 0 || (size_t) stride < sh_series_length(l_max, polar))) {
		sphr_set_error("precalc delay product %d%s: bad header (l_max %u, polar %d, n %d, stride %d)", k, sign, l_max, polar, n, stride);
		ok = 0;
	}
	if(ok && !(a = sh_series_array_new(n, l_max, polar))) {
		sphr_set_error("precalc: out of memory");
		ok = 0;
	}
	if(ok) {
		const size_t len = (size_t) a->stride;
		if((size_t) stride == len) {
			get(fp, a->coeff, sizeof(*a->coeff), (size_t) n * len, &ok);
		} else {
			double complex *row = malloc((size_t) stride * sizeof(*row));
			if(!row) {
				sphr_set_error("precalc: out of memory");
				ok = 0;
			}
			for(int j = 0; ok && j < n; j++)
				if(get(fp, row, sizeof(*row), (size_t) stride, &ok))
					memcpy(a->coeff + (size_t) j * len, row, len * sizeof(*row));
			free(row);
		}
	}
	fclose(fp);
	if(!ok) {
		sh_series_array_free(a);
		return NULL;
	}
	return a;
}

/* --------------------------------------------------------- network plans */

/* inspiral_sky_map.c:1652-1722 */
static int write_plan(const struct correlator_plan_fd *pp, const struct correlator_plan_fd *pn, int k, const char *parent)
{
	FILE *fp = open_at("wb", parent, PLAN_DIR "/correlator_plan_fd/%d/correlator_plan_fd.dat", k);
	if(!fp)
		return -1;
	int ok = 1;
	put(fp, &pp->delta_t, sizeof(pp->delta_t), 1, &ok);
	put(fp, &pp->transient, sizeof(pp->transient), 1, &ok);
	if(finish(fp, ok))
		return -1;
	if(write_rotation_plan(pp->rotation_plan, k, parent) || write_delay_product(pp->delay_product, k, "p", parent) || write_delay_product(pn->delay_product, k, "n", parent))
		return -1;
	return 0;
}

/* inspiral_sky_map.c:1725-1742 */
int write_precalc_correlator_network_plan_fd(const struct correlator_network_plan_fd *fdplansp, const struct correlator_network_plan_fd *fdplansn, char *parent_dir)
{
	if(!fdplansp || !fdplansn || !fdplansp->baselines || fdplansn->baselines->n_baselines != fdplansp->baselines->n_baselines) {
		sphr_set_error("write_precalc_correlator_network_plan_fd: bad arguments");
		return -1;
	}
	if(write_network_baselines(fdplansp->baselines, parent_dir))
		return -1;
	for(int k = 0; k < fdplansp->baselines->n_baselines; k++)
		if(write_plan(fdplansp->plans[k], fdplansn->plans[k], k, parent_dir))
			return -1;
	return 0;
}

/*
 * inspiral_sky_map.c:1221-1309: the +1 plan takes the stored rotation
 * plan, the -1 plan a copy; power_1d is a zero series at the rotation
 * plan's order and the delay products' polarity; fseries_product has room
 * for the larger of tseries_length and the delay products' length.
 */
static struct correlator_plan_fd *read_plan(const struct correlator_baseline *b, int tseries_length, int k, const char *sign, const char *parent, struct sh_series_rotation_plan *rot, double delta_t, int transient)
{
	struct correlator_plan_fd *p = calloc(1, sizeof(*p));
	if(!p) {
		sphr_set_error("precalc: out of memory");
		sh_series_rotation_plan_free(rot);
		return NULL;
	}
	p->baseline = b;
	p->delta_t = delta_t;
	p->transient = transient;
	p->rotation_plan = rot;
	if(!rot || !(p->delay_product = read_delay_product(k, sign, parent)))
		goto fail;
	const int nf = tseries_length > p->delay_product->n ? tseries_length : p->delay_product->n;
	p->fseries_product = malloc((size_t) (nf > 0 ? nf : 1) * sizeof(*p->fseries_product));
	p->power_1d = sh_series_new_zero(rot->l_max, p->delay_product->polar);
	if(!p->fseries_product || !p->power_1d) {
		sphr_set_error("precalc: out of memory");
		goto fail;
	}
	return p;
fail:
	correlator_plan_fd_free(p);
	return NULL;
}

/* inspiral_sky_map.c:1312-1337.  fdplansp and fdplansn are caller-owned
 * structures whose baselines and plans are filled in; on failure every
 * plan read so far is released and plans is NULL, so the structures are
 * then released with free(), not correlator_network_plan_fd_free. */
int read_precalc_correlator_network_plan_fd(struct correlator_network_plan_fd *fdplansp, struct correlator_network_plan_fd *fdplansn, const struct correlator_network_baselines *baselines, int tseries_length, char *parent_dir)
{
	if(!fdplansp || !fdplansn || !baselines) {
		sphr_set_error("read_precalc_correlator_network_plan_fd: bad arguments");
		return -1;
	}
	const int nb = baselines->n_baselines;
	fdplansp->baselines = fdplansn->baselines = (struct correlator_network_baselines *) baselines;
	fdplansp->plans = calloc(nb > 0 ? nb : 1, sizeof(*fdplansp->plans));
	fdplansn->plans = calloc(nb > 0 ? nb : 1, sizeof(*fdplansn->plans));
	if(!fdplansp->plans || !fdplansn->plans) {
		sphr_set_error("precalc: out of memory");
		goto fail;
	}
	for(int k = 0; k < nb; k++) {
		FILE *fp = open_at("rb", parent_dir, PLAN_DIR "/correlator_plan_fd/%d/correlator_plan_fd.dat", k);
		if(!fp)
			goto fail;
		double delta_t = 0;
		int transient = 0, ok = 1;
		get(fp, &delta_t, sizeof(delta_t), 1, &ok);
		get(fp, &transient, sizeof(transient), 1, &ok);
		fclose(fp);
		if(!ok)
			goto fail;
		struct sh_series_rotation_plan *rot = read_rotation_plan(k, parent_dir);
		if(!rot)
			goto fail;
		struct sh_series_rotation_plan *rot_n = sh_series_rotation_plan_copy(rot);
		fdplansp->plans[k] = read_plan(baselines->baselines[k], tseries_length, k, "p", parent_dir, rot, delta_t, transient);
		fdplansn->plans[k] = read_plan(baselines->baselines[k], tseries_length, k, "n", parent_dir, rot_n, delta_t, transient);
		if(!fdplansp->plans[k] || !fdplansn->plans[k])
			goto fail;
	}
	return 0;
fail:
	for(int k = 0; k < nb; k++) {
		if(fdplansp->plans)
			correlator_plan_fd_free(fdplansp->plans[k]);
		if(fdplansn->plans)
			correlator_plan_fd_free(fdplansn->plans[k]);
	}
	free(fdplansp->plans);
	free(fdplansn->plans);
	fdplansp->plans = fdplansn->plans = NULL;
	return -1;
}

/* ------------------------------------------------------ autocorrelator */

/* inspiral_sky_map.c:1745-1784 */
int write_precalc_autocorrelator_network_plan_fd(const struct autocorrelator_network_plan_fd *fdplansp, const struct autocorrelator_network_plan_fd *fdplansn, char *parent_dir)
{
	FILE *fp = open_at("wb", parent_dir, AUTO_DIR "/length.dat");
	if(!fp)
		return -1;
	int ok = 1;
	put(fp, &fdplansp->length, sizeof(fdplansp->length), 1, &ok);
	if(finish(fp, ok))
		return -1;
	const struct autocorrelator_network_plan_fd *pl[2] = { fdplansp, fdplansn };
	const char *sign[2] = { "p", "n" };
	for(int s = 0; s < 2; s++)
		for(int i = 0; i < instrument_array_len(pl[s]->instruments); i++)
			for(int k = 0; k < pl[s]->length; k++) {
				char *f = path_fmt(parent_dir, AUTO_DIR "/psh_series/%d/sh_series_%s/projections%d.fits", i, sign[s], k);
				const int r = f ? sh_series_write_healpix_alm(pl[s]->projections[i][k], f) : -1;
				free(f);
				if(r)
					return -1;
			}
	return 0;
}

/* the projections of one polarisation; NULL on failure */
static struct sh_series ***read_projections(const struct instrument_array *instruments, int length, const char *sign, const char *parent)
{
	const int n = instrument_array_len(instruments);
	struct sh_series ***pr = calloc(n > 0 ? n : 1, sizeof(*pr));
	int ok = pr != NULL;
	for(int i = 0; ok && i < n; i++) {
		ok = (pr[i] = calloc(length > 0 ? length : 1, sizeof(**pr))) != NULL;
		for(int k = 0; ok && k < length; k++) {
			char *f = path_fmt(parent, AUTO_DIR "/psh_series/%d/sh_series_%s/projections%d.fits", i, sign, k);
			ok = f && (pr[i][k] = sh_series_read_healpix_alm(f)) != NULL;
			free(f);
		}
	}
	if(ok)
		return pr;
	for(int i = 0; pr && i < n; i++) {
		for(int k = 0; pr[i] && k < length; k++)
			sh_series_free(pr[i][k]);
		free(pr[i]);
	}
	free(pr);
	return NULL;
}

/* inspiral_sky_map.c:1362-1426.  fdplansp and fdplansn are caller-owned
 * structures whose instruments, length and projections are filled in
 * (projections NULL on failure). */
int read_precalc_autocorrelator_network_plan_fd(struct autocorrelator_network_plan_fd *fdplansp, struct autocorrelator_network_plan_fd *fdplansn, const struct instrument_array *instruments, int tseries_length, char *parent_dir)
{
	(void) tseries_length;
	if(!fdplansp || !fdplansn || !instruments) {
		sphr_set_error("read_precalc_autocorrelator_network_plan_fd: bad arguments");
		return -1;
	}
	fdplansp->projections = fdplansn->projections = NULL;
	fdplansp->instruments = fdplansn->instruments = instruments;
	FILE *fp = open_at("rb", parent_dir, AUTO_DIR "/length.dat");
	if(!fp)
		return -1;
	int length = 0, ok = 1;
	get(fp, &length, sizeof(length), 1, &ok);
	fclose(fp);
	if(!ok)
		return -1;
	if(length < 0) {
		sphr_set_error("precalc: negative autocorrelator length %d", length);
		return -1;
	}
	fdplansp->length = fdplansn->length = length;
	if(!(fdplansp->projections = read_projections(instruments, length, "p", parent_dir)))
		return -1;
	if(!(fdplansn->projections = read_projections(instruments, length, "n", parent_dir))) {
		struct sh_series ***pr = fdplansp->projections;
		for(int i = 0; i < instrument_array_len(instruments); i++) {
			for(int k = 0; k < length; k++)
				sh_series_free(pr[i][k]);
			free(pr[i]);
		}
		free(pr);
		fdplansp->projections = NULL;
		return -1;
	}
	return 0;
}
