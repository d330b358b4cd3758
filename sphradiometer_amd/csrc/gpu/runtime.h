/*
 * Runtime internals shared by the HIP sources.
 */
#ifndef SPHR_GPU_RUNTIME_H
#define SPHR_GPU_RUNTIME_H

#include <hip/hip_runtime.h>
#include <pthread.h>

#define SPHR_NSLOTS 16

/* serialises the synchronous entry points (they share scratch) */
extern pthread_mutex_t sphr_dropin_lock;

hipStream_t sphr_rt_stream(void);
size_t sphr_rt_lds_max(void);
int sphr_rt_cus(void);
void *sphr_rt_scratch(int slot, size_t bytes);
void *sphr_rt_pinned(int slot, size_t bytes);
int sphr_rt_check(hipError_t e, const char *what);

/* twiddle tables exp(-2 pi i k / n) (sign -1) or exp(+2 pi i k / n)
 * (sign +1), interleaved, computed in long double from exact integer
 * angles */
void sphr_twiddles(int n, int sign, double *out);
/* factor n into the radices of the phi DFT; returns the count */
int sphr_factor(int n, int *fac, int max);

#endif
