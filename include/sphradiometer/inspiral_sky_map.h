/*
 * inspiral_sky_map.h -- the localization log-likelihood sky of
 * sphradiometer's src/include/sphradiometer/inspiral_sky_map.h:44-69, the
 * subset on the cross-power path, with plain arrays where the reference
 * takes LAL time-series types.
 *
 * Detectors: correlator_network_plan_mult_by_projection reads each
 * instrument's data pointer as a struct sphr_detector, whose fields are
 * the leading fields of LAL's LALDetector (location, response), so a
 * LALDetector * stored there by a reference caller works unchanged.
 */
#ifndef __RADIOMETER_INSPIRAL_SKY_MAP_H__
#define __RADIOMETER_INSPIRAL_SKY_MAP_H__

#include <complex.h>
#include <sphradiometer/correlator.h>
#include <sphradiometer/sh_series.h>

#ifdef __cplusplus
extern "C" {
#define complex _Complex
#endif

/* LALDetector's leading fields: phase centre in metres (unused here) and
 * the response tensor D (F+ = X.D.X - Y.D.Y, Fx = X.D.Y + Y.D.X) */
struct sphr_detector {
	double location[3];
	float response[3][3];
};

/* inspiral_sky_map.c:125-128: length + round(2 R_earth / c / deltaT) */
long precalculated_TimeSeries_length(long, double);

/* inspiral_sky_map.c:840-856: SH series of log(sin(theta) / 4 pi) */
struct sh_series *sh_series_log_uniformsky_prior(int);

/* inspiral_sky_map.c:423-549: multiply every baseline's delay product by
 * the CBC projection operator for polarisation (beta, psi), with psd[f]
 * the n_instruments PSD values of frequency bin f; the plan's power_1d
 * becomes non-polar with a new rotation plan.  0 on success, -1 on
 * failure (the plan is then undefined, as in the reference). */
int correlator_network_plan_mult_by_projection(struct correlator_network_plan_fd *, double, double, double **);

/* inspiral_sky_map.c:192-219 for n_series series of one length sharing
 * one epoch: Tukey(0.1) window, zero-padding to
 * precalculated_TimeSeries_length(length, delta_t) centred as
 * time_series_pad does, and the sqrt(padded / length) rescale.  out[k]
 * must hold the padded length.  0 on success. */
int preprocess_SNR_arrays(double complex *const *series, int n_series, int length, double delta_t, double complex **out);

/* inspiral_sky_map.c:942-1101 (generate_alm_skys) for preprocessed
 * series of one length: c2c FFT per detector, then for each of the
 * beta = +1 / -1 plans the network integrate, x2, x(n_baselines 4),
 * + logprior and rotate_z(gmst).  *skyp and *skyn are new series at
 * logprior's l_max.  0 on success, -1 on failure. */
int generate_alm_skys_arrays(struct sh_series **skyp, struct sh_series **skyn, struct correlator_network_plan_fd *plansp, struct correlator_network_plan_fd *plansn, double complex *const *series, int length, double gmst, const struct sh_series *logprior);

/*
 * The precalc tree (inspiral_sky_map.c:1113-1811, declared at
 * inspiral_sky_map.h:59-69 of the reference): the same folder layout and
 * the same raw field order in each .dat file, FITS a_lm tables for the
 * log prior and the autocorrelator projections.  0 / non-NULL on success;
 * on failure -1 / NULL with a message for sphr_last_error().
 */

/* inspiral_sky_map.c:1458-1553; per-baseline folders for every baseline
 * of n_ifo detectors (the reference makes n_ifo of them) */
int make_precalc_directories(char *precalc_path, int n_ifo);
int write_precalc_time_series_length(unsigned int length, char *parent_dir);
int read_precalc_time_series_length(unsigned int *length, char *parent_dir);
int write_precalc_logprior(const struct sh_series *series, char *parent_dir);
struct sh_series *read_precalc_logprior(char *parent_dir);
/* the two polarisations' plans share baselines and rotation plan */
int write_precalc_correlator_network_plan_fd(const struct correlator_network_plan_fd *fdplansp, const struct correlator_network_plan_fd *fdplansn, char *parent_dir);
/* the baselines are rebuilt from instruments and must reproduce the stored
 * directions exactly */
struct correlator_network_baselines *read_precalc_correlator_network_baselines(const struct instrument_array *instruments, char *parent_dir);
/* fills caller-owned structures; on failure their plans are NULL */
int read_precalc_correlator_network_plan_fd(struct correlator_network_plan_fd *fdplansp, struct correlator_network_plan_fd *fdplansn, const struct correlator_network_baselines *baselines, int tseries_length, char *parent_dir);
int write_precalc_autocorrelator_network_plan_fd(const struct autocorrelator_network_plan_fd *fdplansp, const struct autocorrelator_network_plan_fd *fdplansn, char *parent_dir);
int read_precalc_autocorrelator_network_plan_fd(struct autocorrelator_network_plan_fd *fdplansp, struct autocorrelator_network_plan_fd *fdplansn, const struct instrument_array *instruments, int tseries_length, char *parent_dir);

/* inspiral_sky_map.c:561-660: projections[i][f] = the (i, i) element of
 * the projection operator of polarisation beta for PSD values psd[f]
 * (n_instruments each), expanded at order l_max; the detectors are the
 * instruments' data pointers (struct sphr_detector) */
struct autocorrelator_network_plan_fd *autocorrelator_network_plan_fd_new(const struct instrument_array *instruments, double beta, double psi, double **psd, int length, int l_max);
/* inspiral_sky_map.c:663-675: every projection resized to l_max */
struct autocorrelator_network_plan_fd *autocorrelator_network_plan_fd_set_l(struct autocorrelator_network_plan_fd *plan, unsigned l_max);
/* inspiral_sky_map.c:678-691 */
void autocorrelator_network_plan_fd_free(struct autocorrelator_network_plan_fd *plan);

#ifdef __cplusplus
#undef complex
}
#endif

#endif
