/*
 * Comparison helpers the reference's tests use (sphradiometer
 * src/lib/diagnostics.c:97-175).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <sphradiometer/sh_series.h>
#include <sphradiometer/diagnostics.h>

/* diagnostics.c:97-137: 0 when both describe the same function */
int sh_series_cmp(const struct sh_series *a, const struct sh_series *b)
{
	const int la = a->l_max, lb = b->l_max;
	const int ma = a->polar ? 0 : la, mb = b->polar ? 0 : lb;
	const int lt = la > lb ? la : lb;
	for(int l = 0; l <= lt; l++) {
		const int mt = l < (ma > mb ? ma : mb) ? l : (ma > mb ? ma : mb);
		for(int m = -mt; m <= mt; m++) {
			const double complex va = (l <= la && abs(m) <= ma) ? sh_series_get(a, l, m) : 0.0;
			const double complex vb = (l <= lb && abs(m) <= mb) ? sh_series_get(b, l, m) : 0.0;
			if(va != vb)
				return 1;
		}
	}
	return 0;
}

/* diagnostics.c:145-175: sqrt(sum |a - b|^2 / 4 pi) at the larger order */
double diagnostics_rms_error(const struct sh_series *measured, const struct sh_series *exact)
{
	const unsigned l = measured->l_max > exact->l_max ? measured->l_max : exact->l_max;
	struct sh_series *m = sh_series_copy(measured), *e = sh_series_copy(exact);
	double r = NAN;
	if(!m || !e || !sh_series_resize(m, l) || !sh_series_resize(e, l))
		goto done;
	if(!(measured->polar && exact->polar) && (!sh_series_set_polar(m, 0) || !sh_series_set_polar(e, 0)))
		goto done;
	sh_series_add(e, -1., m);
	r = sqrt(creal(sh_series_dot(e, e)) / (4. * M_PI));
done:
	sh_series_free(m);
	sh_series_free(e);
	return r;
}
