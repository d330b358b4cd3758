/*
 * HEALPix FITS files of sh_series (sh_series_healpix.cc of the
 * reference, which goes through healpix_cxx): the a_lm table and the
 * RING-ordered map.  Written and read here directly in the FITS binary
 * table layout healpix uses:
 *
 *   a_lm:  BINTABLE, columns index (1J) = l^2 + l + m + 1, real (1D),
 *          imag (1D), one row per (l, 0 <= m <= m_max), keywords MAX-LPOL,
 *          MAX-MPOL; the coefficients are those of the real part of the
 *          series (sh_series_real), m < 0 being implied by the symmetry
 *          of a real function;
 *   map:   BINTABLE, one column signal (1D), one row per pixel, keywords
 *          PIXTYPE = 'HEALPIX', ORDERING = 'RING', NSIDE, FIRSTPIX,
 *          LASTPIX, INDXSCHM = 'IMPLICIT', OBJECT = 'FULLSKY'.
 *
 * FITS: 2880-byte records of 80-character header cards ending in END,
 * big-endian binary data padded with zeros to whole records.
 */
#include <complex.h>
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <unistd.h>
#include <sphradiometer/sh_series.h>
#include "../sphr_internal.h"

#define FITS_REC 2880
#define FITS_CARD 80

/* ------------------------------------------------------------- writing */

struct fits_hdr {
	char buf[64 * FITS_CARD];
	int n;
};

static void card_raw(struct fits_hdr *h, const char *text)
{
	char c[FITS_CARD + 1];
	snprintf(c, sizeof(c), "%-80s", text);
	memcpy(h->buf + (size_t) h->n * FITS_CARD, c, FITS_CARD);
	h->n++;
}

static void card_int(struct fits_hdr *h, const char *key, long v)
{
	char c[FITS_CARD + 1];
	snprintf(c, sizeof(c), "%-8.8s= %20ld", key, v);
	card_raw(h, c);
}

static void card_bool(struct fits_hdr *h, const char *key, int v)
{
	char c[FITS_CARD + 1];
	snprintf(c, sizeof(c), "%-8.8s= %20s", key, v ? "T" : "F");
	card_raw(h, c);
}

static void card_str(struct fits_hdr *h, const char *key, const char *v)
{
	char c[FITS_CARD + 1];
	snprintf(c, sizeof(c), "%-8.8s= '%-8s'", key, v);
	card_raw(h, c);
}

static int write_hdr(FILE *f, struct fits_hdr *h)
{
	card_raw(h, "END");
	const size_t bytes = (size_t) h->n * FITS_CARD, pad = (FITS_REC - bytes % FITS_REC) % FITS_REC;
	char sp[FITS_REC];
	memset(sp, ' ', sizeof(sp));
	return fwrite(h->buf, 1, bytes, f) != bytes || fwrite(sp, 1, pad, f) != pad ? -1 : 0;
}

static int write_pad(FILE *f, size_t bytes)
{
	const size_t pad = (FITS_REC - bytes % FITS_REC) % FITS_REC;
	char z[FITS_REC];
	memset(z, 0, sizeof(z));
	return fwrite(z, 1, pad, f) != pad ? -1 : 0;
}

static void be64(unsigned char *p, double v)
{
	uint64_t u;
	memcpy(&u, &v, 8);
	for(int i = 0; i < 8; i++)
		p[i] = (unsigned char) (u >> (56 - 8 * i));
}

static void be32(unsigned char *p, int32_t v)
{
	const uint32_t u = (uint32_t) v;
	for(int i = 0; i < 4; i++)
		p[i] = (unsigned char) (u >> (24 - 8 * i));
}

static FILE *create(const char *filename)
{
	/* the reference removes an existing file first (healpix refuses to
	 * overwrite) and reports the write error otherwise */
	const int errsv = errno;
	unlink(filename);
	errno = errsv;
	FILE *f = fopen(filename, "wb");
	if(!f)
		sphr_set_error("%s: %s", filename, strerror(errno));
	return f;
}

static int write_primary(FILE *f)
{
	struct fits_hdr h = { .n = 0 };
	card_bool(&h, "SIMPLE", 1);
	card_int(&h, "BITPIX", 8);
	card_int(&h, "NAXIS", 0);
	card_bool(&h, "EXTEND", 1);
	return write_hdr(f, &h);
}

int sphr_fits_write_healpix_map(const char *filename, const double *map, int nside)
{
	const long npix = 12L * nside * nside;
	FILE *f = create(filename);
	if(!f)
		return -1;
	struct fits_hdr h = { .n = 0 };
	card_str(&h, "XTENSION", "BINTABLE");
	card_int(&h, "BITPIX", 8);
	card_int(&h, "NAXIS", 2);
	card_int(&h, "NAXIS1", 8);
	card_int(&h, "NAXIS2", npix);
	card_int(&h, "PCOUNT", 0);
	card_int(&h, "GCOUNT", 1);
	card_int(&h, "TFIELDS", 1);
	card_str(&h, "TTYPE1", "signal");
	card_str(&h, "TFORM1", "1D");
	card_str(&h, "PIXTYPE", "HEALPIX");
	card_str(&h, "ORDERING", "RING");
	card_int(&h, "NSIDE", nside);
	card_int(&h, "FIRSTPIX", 0);
	card_int(&h, "LASTPIX", npix - 1);
	card_str(&h, "INDXSCHM", "IMPLICIT");
	card_str(&h, "OBJECT", "FULLSKY");
	int rc = write_primary(f) || write_hdr(f, &h);
	unsigned char b[8];
	for(long p = 0; p < npix && !rc; p++) {
		be64(b, map[p]);
		rc = fwrite(b, 1, 8, f) != 8;
	}
	rc = rc || write_pad(f, (size_t) npix * 8);
	if(fclose(f))
		rc = 1;
	if(rc)
		sphr_set_error("%s: write failed", filename);
	return rc ? -1 : 0;
}

/* sh_series_write_healpix_alm (sh_series_healpix.cc:131-163) */
int sh_series_write_healpix_alm(const struct sh_series *series, const char *filename)
{
	if(!series || !filename) {
		sphr_set_error("sh_series_write_healpix_alm: bad arguments");
		return -1;
	}
	struct sh_series *real = sh_series_copy(series);
	if(!real)
		return -1;
	real = sh_series_real(real);
	const int L = (int) series->l_max, mmax = series->polar ? 0 : L;
	long nrow = 0;
	for(int m = 0; m <= mmax; m++)
		nrow += L - m + 1;
	FILE *f = create(filename);
	if(!f) {
		sh_series_free(real);
		return -1;
	}
	struct fits_hdr h = { .n = 0 };
	card_str(&h, "XTENSION", "BINTABLE");
	card_int(&h, "BITPIX", 8);
	card_int(&h, "NAXIS", 2);
	card_int(&h, "NAXIS1", 20);
	card_int(&h, "NAXIS2", nrow);
	card_int(&h, "PCOUNT", 0);
	card_int(&h, "GCOUNT", 1);
	card_int(&h, "TFIELDS", 3);
	card_str(&h, "TTYPE1", "index");
	card_str(&h, "TFORM1", "1J");
	card_str(&h, "TTYPE2", "real");
	card_str(&h, "TFORM2", "1D");
	card_str(&h, "TTYPE3", "imag");
	card_str(&h, "TFORM3", "1D");
	card_int(&h, "MAX-LPOL", L);
	card_int(&h, "MAX-MPOL", mmax);
	int rc = write_primary(f) || write_hdr(f, &h);
	unsigned char row[20];
	for(int m = 0; m <= mmax && !rc; m++)
		for(int l = m; l <= L && !rc; l++) {
			const double complex a = sh_series_get(real, l, m);
			be32(row, l * l + l + m + 1);
			be64(row + 4, creal(a));
			be64(row + 12, cimag(a));
			rc = fwrite(row, 1, 20, f) != 20;
		}
	rc = rc || write_pad(f, (size_t) nrow * 20);
	if(fclose(f))
		rc = 1;
	sh_series_free(real);
	if(rc)
		sphr_set_error("%s: write failed", filename);
	return rc ? -1 : 0;
}

/* ------------------------------------------------------------- reading */

/* value of keyword key in a header of ncard cards, or NULL */
static const char *kw(const char *hdr, int ncard, const char *key)
{
	const size_t kl = strlen(key);
	for(int i = 0; i < ncard; i++) {
		const char *c = hdr + (size_t) i * FITS_CARD;
		if(!strncmp(c, key, kl) && (kl == 8 || c[kl] == ' ') && c[8] == '=')
			return c + 10;
	}
	return NULL;
}

static long kw_int(const char *hdr, int ncard, const char *key, long dflt)
{
	const char *v = kw(hdr, ncard, key);
	return v ? strtol(v, NULL, 10) : dflt;
}

/* quoted string value, trailing blanks removed */
static void kw_str(const char *hdr, int ncard, const char *key, char *out, size_t n)
{
	const char *v = kw(hdr, ncard, key);
	out[0] = 0;
	if(!v)
		return;
	/* the value field is the card's last 70 characters; a string ends
	 * at its closing quote */
	const char *end = v + 70;
	while(v < end && *v == ' ')
		v++;
	if(v >= end || *v != '\'')
		return;
	v++;
	size_t k = 0;
	while(k + 1 < n && v < end && *v != '\'')
		out[k++] = *v++;
	while(k && out[k - 1] == ' ')
		k--;
	out[k] = 0;
}

/* read one header (whole records) into *hdr; returns its card count */
static int read_hdr(FILE *f, char **hdr)
{
	size_t cap = 0, used = 0;
	char *b = NULL;
	for(;;) {
		if(used + FITS_REC > cap) {
			cap = cap ? 2 * cap : 4 * FITS_REC;
			char *nb = realloc(b, cap);
			if(!nb) {
				free(b);
				return -1;
			}
			b = nb;
		}
		if(fread(b + used, 1, FITS_REC, f) != FITS_REC) {
			free(b);
			return -1;
		}
		used += FITS_REC;
		for(size_t c = used - FITS_REC; c < used; c += FITS_CARD)
			if(!strncmp(b + c, "END", 3) && (b[c + 3] == ' ' || c + 3 == used)) {
				*hdr = b;
				return (int) (c / FITS_CARD) + 1;
			}
	}
}

static double rd_be(const unsigned char *p, char type)
{
	uint64_t u = 0;
	const int n = type == 'D' || type == 'K' ? 8 : type == 'E' || type == 'J' ? 4 : type == 'I' ? 2 : 1;
	for(int i = 0; i < n; i++)
		u = (u << 8) | p[i];
	switch(type) {
	case 'D': {
		double d;
		memcpy(&d, &u, 8);
		return d;
	}
	case 'E': {
		const uint32_t w = (uint32_t) u;
		float x;
		memcpy(&x, &w, 4);
		return x;
	}
	case 'K':
		return (double) (int64_t) u;
	case 'J':
		return (double) (int32_t) (uint32_t) u;
	case 'I':
		return (double) (int16_t) (uint16_t) u;
	default:
		return (double) (uint8_t) u;
	}
}

static int tform_bytes(const char *tf, int *rep, char *type)
{
	char *end;
	long r = strtol(tf, &end, 10);
	if(end == tf)
		r = 1;
	*rep = (int) r;
	*type = *end;
	switch(*end) {
	case 'L':
	case 'B':
		return (int) r;
	case 'I':
		return 2 * (int) r;
	case 'J':
	case 'E':
		return 4 * (int) r;
	case 'K':
	case 'D':
		return 8 * (int) r;
	case 'C':
		return 8 * (int) r;
	case 'M':
		return 16 * (int) r;
	default:
		return -1;
	}
}

/* sh_series_read_healpix_alm (sh_series_healpix.cc:166-187): the table
 * of the second HDU, columns index, real, imag (index = l^2 + l + m + 1);
 * m_max must be 0 (polar series) or l_max; m < 0 filled in for a real
 * function (sh_series_from_healpix_Alm, :80-118) */
struct sh_series *sh_series_read_healpix_alm(const char *filename)
{
	FILE *f = filename ? fopen(filename, "rb") : NULL;
	char *h0 = NULL, *h1 = NULL;
	unsigned char *data = NULL;
	struct sh_series *series = NULL;
	if(!f) {
		sphr_set_error("%s: %s", filename ? filename : "(null)", strerror(errno));
		return NULL;
	}
	const int n0 = read_hdr(f, &h0);
	if(n0 < 0)
		goto bad;
	{
		/* skip the primary data unit (NAXIS 0 in healpix files) */
		long pbytes = 0;
		const long naxis = kw_int(h0, n0, "NAXIS", 0);
		if(naxis > 0) {
			pbytes = labs(kw_int(h0, n0, "BITPIX", 8)) / 8;
			for(int a = 1; a <= naxis; a++) {
				char k[16];
				snprintf(k, sizeof(k), "NAXIS%d", a);
				pbytes *= kw_int(h0, n0, k, 0);
			}
		}
		if(fseek(f, (pbytes + FITS_REC - 1) / FITS_REC * FITS_REC, SEEK_CUR))
			goto bad;
	}
	const int n1 = read_hdr(f, &h1);
	if(n1 < 0)
		goto bad;
	{
		char xt[32];
		kw_str(h1, n1, "XTENSION", xt, sizeof(xt));
		const long rowb = kw_int(h1, n1, "NAXIS1", -1), nrow = kw_int(h1, n1, "NAXIS2", -1), nf = kw_int(h1, n1, "TFIELDS", 0);
		if(strcmp(xt, "BINTABLE") || rowb <= 0 || nrow < 1 || nf < 3 || nf > 999)
			goto bad;
		int off[3] = { -1, -1, -1 };
		char typ[3] = { 0, 0, 0 };
		int o = 0;
		for(int i = 1; i <= nf; i++) {
			char k[16], name[72], tf[72];
			snprintf(k, sizeof(k), "TTYPE%d", i);
			kw_str(h1, n1, k, name, sizeof(name));
			snprintf(k, sizeof(k), "TFORM%d", i);
			kw_str(h1, n1, k, tf, sizeof(tf));
			int rep;
			char t;
			const int nb = tform_bytes(tf, &rep, &t);
			if(nb < 0)
				goto bad;
			const int c = !strcasecmp(name, "index") ? 0 : !strcasecmp(name, "real") ? 1 : !strcasecmp(name, "imag") ? 2 : -1;
			if(c >= 0 && rep >= 1) {
				off[c] = o;
				typ[c] = t;
			}
			o += nb;
		}
		if(off[0] < 0 || off[1] < 0 || off[2] < 0 || o > rowb)
			goto bad;
		data = malloc((size_t) rowb * nrow);
		if(!data || fread(data, (size_t) rowb, (size_t) nrow, f) != (size_t) nrow)
			goto bad;
		long lmax = -1, mmax = -1;
		for(long r = 0; r < nrow; r++) {
			const long idx = (long) rd_be(data + r * rowb + off[0], typ[0]);
			if(idx < 1)
				goto bad;
			long l = (long) sqrt((double) (idx - 1));
			while(l * l > idx - 1)
				l--;
			while((l + 1) * (l + 1) <= idx - 1)
				l++;
			const long m = idx - 1 - l * l - l;
			if(m < 0 || m > l)
				goto bad;
			lmax = l > lmax ? l : lmax;
			mmax = m > mmax ? m : mmax;
		}
		if(lmax < 0 || (mmax != 0 && mmax != lmax)) {
			sphr_set_error("%s: unsupported (l_max, m_max) = (%ld, %ld)", filename, lmax, mmax);
			goto out;
		}
		series = sh_series_new((unsigned) lmax, mmax == 0);
		if(!series)
			goto out;
		sh_series_zero(series);
		for(long r = 0; r < nrow; r++) {
			const unsigned char *p = data + r * rowb;
			const long idx = (long) rd_be(p + off[0], typ[0]);
			long l = (long) sqrt((double) (idx - 1));
			while(l * l > idx - 1)
				l--;
			while((l + 1) * (l + 1) <= idx - 1)
				l++;
			const int m = (int) (idx - 1 - l * l - l);
			const double complex x = rd_be(p + off[1], typ[1]) + I * rd_be(p + off[2], typ[2]);
			sh_series_set(series, (unsigned) l, m, x);
			if(m)
				sh_series_set(series, (unsigned) l, -m, (m & 1) ? -conj(x) : conj(x));
		}
	}
	goto out;
bad:
	sphr_set_error("%s: not a HEALPix a_lm FITS table", filename);
out:
	free(h0);
	free(h1);
	free(data);
	fclose(f);
	return series;
}
