/*
 * diagnostics.h -- the comparison helpers of sphradiometer's
 * src/include/sphradiometer/diagnostics.h that its tests rely on.
 */
#ifndef __RADIOMETER_DIAGNOSTICS_H__
#define __RADIOMETER_DIAGNOSTICS_H__

#include <sphradiometer/sh_series.h>

#ifdef __cplusplus
extern "C" {
#endif

int sh_series_cmp(const struct sh_series *, const struct sh_series *);
double diagnostics_rms_error(const struct sh_series *, const struct sh_series *);

#ifdef __cplusplus
}
#endif

#endif
