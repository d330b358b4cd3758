/*
 * sphradiometer_batch.h -- batched, device-resident form of the
 * frequency-domain network integrator plus map synthesis.  No counterpart
 * in the reference: it runs, for many segments in one set of launches,
 * what a caller of the reference does per segment,
 *
 *   for each detector: correlator_tseries_to_fseries   (correlator.c:224)
 *   correlator_network_integrate_power_fd              (correlator.c:993)
 *   sh_series_rotate_z(sky, sky, gmst)                 (caller's step)
 *   sh_series_to_mesh(sky)                             (sh_series_trans.c:796)
 *
 * with the plan built by correlator_network_plan_fd_new / _set_l.  Inputs
 * and outputs are device buffers; nothing is copied to or from the host.
 *
 * Layouts (row-major):
 *   tseries  [n_det][n_segments][N]        float64 (or float32)
 *   window   [N] float64, applied before the FFT, may be NULL
 *   gmst     [n_segments] float64 radians, may be NULL (no rotate_z)
 *   alm      [n_segments][(L+1)^2]         complex (re, im) in the
 *                                          reference coefficient order
 *   map      [n_segments][2(L+1)][2(L+1)]  complex, rows theta ascending
 *                                          (Gauss-Legendre), columns phi
 * Complex values are float64 pairs for SPHR_FP64 and float32 pairs for
 * SPHR_MIXED.
 */
#ifndef SPHRADIOMETER_BATCH_H
#define SPHRADIOMETER_BATCH_H

#include <sphradiometer/correlator.h>

#ifdef __cplusplus
extern "C" {
#endif

enum sphr_precision {
	/* every stage in float64: the drop-in arithmetic */
	SPHR_FP64 = 0,
	/* FFT, projection, rotation and synthesis in float32 with float32
	 * accumulation; plan tables rounded from the float64 plan */
	SPHR_MIXED = 1
};

struct sphr_batch;

/* Device tables from a network FD plan whose baselines are all polar (a
 * real-data plan, before any mult_by_projection), segment length N a
 * power of two, every baseline's l_max <= sky_l_max.  max_segments sizes
 * the scratch; NULL on failure (sphr_last_error() says why). */
struct sphr_batch *sphr_batch_new(const struct correlator_network_plan_fd *plan, unsigned sky_l_max, int precision, int max_segments);
void sphr_batch_free(struct sphr_batch *);

/* One pass over n_segments <= max_segments segments, asynchronous on
 * stream (hipStream_t; NULL = the library stream).  alm may be NULL only
 * when map is NULL too.  Returns 0 when the launches were queued. */
int sphr_batch_run(struct sphr_batch *, const void *tseries, int tseries_is_f32, int n_segments, const double *window, const double *gmst, void *alm, void *map, void *stream);

/* Per-stage timing: while enabled, every run records HIP events on its
 * stream around its stages (FFT, projection, [unused], sky: coefficients
 * and synthesis); _read synchronises and returns the summed milliseconds
 * per stage and the number of runs since enabling. */
int sphr_batch_profile(struct sphr_batch *, int enable);
int sphr_batch_profile_read(struct sphr_batch *, double ms_per_stage[4], int *runs);

/* Shape information: [0] detectors, [1] baselines, [2] N, [3] sky l_max,
 * [4] projection columns, [5] K splits, [6] precision */
void sphr_batch_info(const struct sphr_batch *, int info[8]);

#ifdef __cplusplus
}
#endif

#endif
