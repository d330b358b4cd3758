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
	/* phi DFT as a real-expanded matrix for the GEMM form of the
	 * synthesis: rows 2 i, 2 i + 1 for column i of F (m = i for i <= L,
	 * m = i - (2L+1) above), columns (re, im) of point k:
	 * E[i][k] = exp(2 pi i k m / n), expanded as (Er, Ei) / (-Ei, Er) */
	double *E64;
	float *E32;
	struct sky_fused_dev fz;	/* fused sky kernel tables (orders it is compiled for) */
	void *blob;	/* the one device allocation holding x ... E32 */
};

/* the l_seed low-pass mesh of the extremizer's seed search */
#define SPHR_L_SEED 6

/* gemm_tables = 0: without the phi-DFT matrices of the GEMM-form
 * synthesis and the fused sky kernel's tables (callers that synthesise
 * with launch_synth only) */
int synth_tables_build(int L, struct synth_dev *t, int gemm_tables = 1);
void synth_tables_free(struct synth_dev *t);

#endif
