/*
 * Device tables of the mesh synthesis at one order L: the positive-half
 * Gauss-Legendre nodes, the sectoral Pbar_mm values at those nodes, the
 * coefficients of the l recurrence and the phi-DFT twiddles.
 */
#ifndef SPHR_SYNTH_TABLES_H
#define SPHR_SYNTH_TABLES_H

#include "kernels.h"

struct synth_dev {
	struct synth_geom g;
	struct synth_tables tb;
	double *x, *smm, *alpha, *twp;
	double *cm;	/* L+1: Pbar_mm(cos theta) = cm[m] sin^m theta */
};

/* the l_seed low-pass mesh of the extremizer's seed search */
#define SPHR_L_SEED 6

int synth_tables_build(int L, struct synth_dev *t);
void synth_tables_free(struct synth_dev *t);

#endif
