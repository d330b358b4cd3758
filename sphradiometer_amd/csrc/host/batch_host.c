/*
 * Host half of the batched API: flattens a network FD plan (built by
 * correlator_network_plan_fd_new / _set_l) for the device pipeline.
 */
#include <sphradiometer/correlator.h>
#include <sphradiometer/sphradiometer_batch.h>
#include "../sphr_internal.h"

struct sphr_batch *sphr_batch_new(const struct correlator_network_plan_fd *plan, unsigned sky_l_max, int precision, int max_segments)
{
	struct sphr_net_desc d;
	const int nb = plan->baselines->n_baselines;
	if(nb < 1 || nb > SPHR_DESC_MAX) {
		sphr_set_error("sphr_batch_new: %d baselines (1..%d supported)", nb, SPHR_DESC_MAX);
		return NULL;
	}
	d.n_det = instrument_array_len(plan->baselines->baselines[0]->instruments);
	d.nb = nb;
	d.N = plan->plans[0]->delay_product->n;
	/* baseline k is the pair (i, j), i = 1..n-1, j < i, as the network
	 * integrator walks them (correlator.c:998-1001) */
	int k = 0;
	for(int i = 1; i < d.n_det && k < nb; i++)
		for(int j = 0; j < i && k < nb; j++, k++) {
			const struct correlator_plan_fd *p = plan->plans[k];
			if(p->delay_product->n != d.N || p->delay_product->l_max != p->power_1d->l_max || p->rotation_plan->l_max < p->power_1d->l_max) {
				sphr_set_error("sphr_batch_new: baseline %d plan is inconsistent", k);
				return NULL;
			}
			d.ia[k] = i;
			d.ib[k] = j;
			d.Lb[k] = (int) p->power_1d->l_max;
			d.polar[k] = p->power_1d->polar;
			d.transient[k] = p->transient;
			d.dp[k] = (const double *) p->delay_product->coeff;
			d.blocks[k] = sphr_rotation_blocks(p->rotation_plan);
		}
	if(k != nb) {
		sphr_set_error("sphr_batch_new: the instrument array has too few detectors for %d baselines", nb);
		return NULL;
	}
	return sphr_batch_create(&d, sky_l_max, precision, max_segments);
}
