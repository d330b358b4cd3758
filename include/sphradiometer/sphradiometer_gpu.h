/*
 * sphradiometer_gpu.h -- runtime controls of the MI355X library: device
 * status, the last error message, device memory, the library stream and
 * HIP event timers.  No counterpart in the reference (which has no GPU);
 * these exist so callers of the batched API (sphradiometer_batch.h) and
 * language bindings can manage device buffers without a HIP toolchain.
 * Every function is a plain C entry point; streams and events are passed
 * as opaque pointers (hipStream_t / hipEvent_t).
 */
#ifndef SPHRADIOMETER_GPU_H
#define SPHRADIOMETER_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 1 when a GPU is usable (initialises the runtime on first call) */
int sphr_gpu_available(void);
/* message of the last failure on this thread ("" if none) */
const char *sphr_last_error(void);
/* device name, compute units, LDS bytes per workgroup; 0 on success */
int sphr_device_info(char *name, int name_len, int *compute_units, size_t *lds_per_block);

void *sphr_device_alloc(size_t bytes);
void sphr_device_free(void *p);
/* synchronous copies and fill, ordered after the work queued on the
 * library's stream (sphr_stream) */
int sphr_memcpy_htod(void *dst, const void *src, size_t bytes);
int sphr_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int sphr_memcpy_dtod(void *dst, const void *src, size_t bytes);
int sphr_memset_device(void *dst, int value, size_t bytes);

/* the stream the synchronous entry points use; any stream may be passed
 * to the batched API */
void *sphr_stream(void);
int sphr_stream_synchronize(void *stream);
int sphr_device_synchronize(void);

void *sphr_event_new(void);
void sphr_event_free(void *event);
int sphr_event_record(void *event, void *stream);
int sphr_event_synchronize(void *event);
/* milliseconds between two recorded events, or a negative value */
float sphr_event_elapsed_ms(void *start, void *stop);

#ifdef __cplusplus
}
#endif

#endif
