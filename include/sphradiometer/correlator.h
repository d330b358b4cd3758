/*
 * correlator.h -- baselines, frequency-domain correlation plans and the
 * network cross-power integrator.  Drop-in for sphradiometer's
 * src/include/sphradiometer/correlator.h (structs 59-164, prototypes
 * 190-272).
 *
 * correlator_baseline_integrate_power_fd, correlator_network_integrate_
 * power_fd, the time-domain integrators and the FFT helpers run on the GPU
 * and are synchronous.  Plan construction runs in host code.
 */
#ifndef __RADIOMETER_CORRELATOR_H__
#define __RADIOMETER_CORRELATOR_H__

#include <complex.h>
#include <sphradiometer/sphr_compat.h>
#include <sphradiometer/instrument.h>
#include <sphradiometer/sh_series.h>

#ifdef __cplusplus
extern "C" {
#define complex _Complex
#endif

struct correlator_baseline {
	const struct instrument_array *instruments;
	int index_a;
	int index_b;
	gsl_vector *d;
	double theta;
	double phi;
};

struct correlator_plan_td {
	const struct correlator_baseline *baseline;
	double delta_t;
	int transient;
	struct sh_series_array *proj_a;
	struct sh_series_array *proj_b;
	struct sh_series *sample_a;
	struct sh_series *sample_b;
	struct sh_series *product;
	struct sh_series *power_1d;
	struct sh_series_product_plan *product_plan;
	struct sh_series_rotation_plan *rotation_plan;
};

struct correlator_plan_fd {
	const struct correlator_baseline *baseline;
	double delta_t;
	int transient;
	struct sh_series_array *delay_product;
	double complex *fseries_product;
	struct sh_series *power_1d;
	struct sh_series_rotation_plan *rotation_plan;
};

struct correlator_network_baselines {
	struct correlator_baseline **baselines;
	int n_baselines;
};

struct correlator_network_plan_td {
	struct correlator_network_baselines *baselines;
	struct correlator_plan_td **plans;
};

struct correlator_network_plan_fd {
	struct correlator_network_baselines *baselines;
	struct correlator_plan_fd **plans;
};

struct autocorrelator_network_plan_fd {
	const struct instrument_array *instruments;
	struct sh_series ***projections;
	int length;
};

unsigned int correlator_baseline_power_l_max_naive(double, double);
unsigned int correlator_baseline_power_l_max(const struct correlator_baseline *, double);
double correlator_dump_interval(unsigned int, unsigned int);
int correlator_transient(const struct sh_series_array *, const struct sh_series_array *);

double *correlator_tukey_window_new(int, int, int, double);
double *correlator_square_window_new(int, int, double);

fftw_plan correlator_tseries_to_fseries_plan(double *, double complex *, int);
void correlator_tseries_to_fseries(double *, double complex *, int, fftw_plan);
fftw_plan correlator_ctseries_to_fseries_plan(double complex *, double complex *, int);
void correlator_ctseries_to_fseries(fftw_plan);
/* releases a plan of the two helpers above (no counterpart in the
 * reference, whose plans are FFTW's and freed with fftw_destroy_plan) */
void correlator_fft_plan_free(fftw_plan);

struct correlator_baseline *correlator_baseline_new(const struct instrument_array *, int, int);
void correlator_baseline_free(struct correlator_baseline *);
struct correlator_baseline *correlator_baseline_copy(const struct correlator_baseline *);

struct correlator_network_baselines *correlator_network_baselines_new(const struct instrument_array *);
void correlator_network_baselines_free(struct correlator_network_baselines *);
struct correlator_network_baselines *correlator_network_baselines_copy(const struct correlator_network_baselines *);
unsigned int correlator_network_l_max(struct correlator_network_baselines *, double);

/*
 * Time-domain correlator (correlator.c:325-400, 611-673, 820-858, 967-984):
 * the integration runs on the GPU and is synchronous; the plan's struct is
 * the reference's, allocated inside a larger one that also holds the
 * device tables, so it is created and freed only by these functions.
 */
struct correlator_plan_td *correlator_plan_td_new(const struct correlator_baseline *, double);
void correlator_plan_td_free(struct correlator_plan_td *);
struct sh_series *correlator_baseline_integrate_power_td(const double *, const double *, const double *, int, struct correlator_plan_td *);
struct correlator_network_plan_td *correlator_network_plan_td_new(struct correlator_network_baselines *, double);
void correlator_network_plan_td_free(struct correlator_network_plan_td *);
struct sh_series *correlator_network_integrate_power_td(struct sh_series *, double **, int, double **, struct correlator_network_plan_td *);

struct correlator_plan_fd *correlator_plan_fd_new(const struct correlator_baseline *, int, double);
void correlator_plan_fd_free(struct correlator_plan_fd *);
struct correlator_plan_fd *correlator_plan_fd_copy(const struct correlator_plan_fd *);
struct correlator_plan_fd *correlator_plan_fd_set_l(struct correlator_plan_fd *, unsigned);

struct correlator_network_plan_fd *correlator_network_plan_fd_new(struct correlator_network_baselines *, int, double);
void correlator_network_plan_fd_free(struct correlator_network_plan_fd *);
struct correlator_network_plan_fd *correlator_network_plan_fd_copy(const struct correlator_network_plan_fd *);
struct correlator_network_plan_fd *correlator_network_plan_fd_set_l(struct correlator_network_plan_fd *, unsigned);

struct sh_series *correlator_baseline_integrate_power_fd(const double complex *, const double complex *, struct correlator_plan_fd *);
struct sh_series *correlator_network_integrate_power_fd(struct sh_series *, double complex **, struct correlator_network_plan_fd *);
/* No counterpart in the reference.  The FD integrators keep a device copy
 * of each plan's delay product and rotation table from one call to the
 * next; the plan functions above (free, set_l, mult_by_projection) drop
 * it when they change the plan.  A caller that writes into a plan's
 * delay_product or rotation_plan in place calls this before the next
 * integrate. */
void correlator_plan_fd_device_invalidate(struct correlator_plan_fd *);

#ifdef __cplusplus
#undef complex
}
#endif

#endif
