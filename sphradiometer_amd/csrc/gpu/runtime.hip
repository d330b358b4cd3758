/*
 * Device runtime of the library: one device (the current HIP device),
 * one non-blocking stream at the default priority for the synchronous
 * entry points, grow-only scratch buffers, pinned staging, error text,
 * and the public helpers of sphradiometer_gpu.h.
 */
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sphradiometer/sphradiometer_gpu.h>
#include "../sphr_internal.h"
#include "runtime.h"

static thread_local char err_msg[512];

extern "C" void sphr_set_error(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(err_msg, sizeof(err_msg), fmt, ap);
	va_end(ap);
}

extern "C" const char *sphr_last_error(void)
{
	return err_msg;
}

static struct {
	int state;	/* 0 = untried, 1 = ready, -1 = unusable */
	int device;
	hipStream_t stream;
	size_t lds_max;
	int cus;
	char name[256];
} ctx;
static pthread_mutex_t init_lock = PTHREAD_MUTEX_INITIALIZER;
pthread_mutex_t sphr_dropin_lock = PTHREAD_MUTEX_INITIALIZER;

extern "C" int sphr_gpu_ensure(void)
{
	if(ctx.state == 1)
		return 0;
	pthread_mutex_lock(&init_lock);
	if(ctx.state == 0) {
		int n = 0;
		hipError_t e = hipGetDeviceCount(&n);
		if(e != hipSuccess || n <= 0) {
			ctx.state = -1;
			snprintf(ctx.name, sizeof(ctx.name), "no HIP device (%s)", e != hipSuccess ? hipGetErrorString(e) : "count 0");
		} else {
			hipDeviceProp_t prop;
			hipGetDevice(&ctx.device);
			if(hipGetDeviceProperties(&prop, ctx.device) == hipSuccess && hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking) == hipSuccess) {
				ctx.lds_max = prop.sharedMemPerBlock;
				ctx.cus = prop.multiProcessorCount;
				snprintf(ctx.name, sizeof(ctx.name), "%s", prop.gcnArchName);
				ctx.state = 1;
			} else {
				ctx.state = -1;
				snprintf(ctx.name, sizeof(ctx.name), "HIP device initialisation failed");
			}
		}
	}
	pthread_mutex_unlock(&init_lock);
	if(ctx.state != 1) {
		sphr_set_error("sphradiometer_amd: GPU unavailable: %s", ctx.name);
		return -1;
	}
	return 0;
}

hipStream_t sphr_rt_stream(void)
{
	return ctx.stream;
}

/* LDS a kernel may request without raising its dynamic-LDS attribute */
size_t sphr_rt_lds_max(void)
{
	return ctx.lds_max < 65536 ? ctx.lds_max : 65536;
}

int sphr_rt_cus(void)
{
	return ctx.cus;
}

/* grow-only device scratch by slot (callers hold sphr_dropin_lock) */
static struct {
	void *p;
	size_t n;
} dev_slots[SPHR_NSLOTS], host_slots[SPHR_NSLOTS];

void *sphr_rt_scratch(int slot, size_t bytes)
{
	if(bytes == 0)
		bytes = 1;
	if(dev_slots[slot].n < bytes) {
		if(dev_slots[slot].p)
			hipFree(dev_slots[slot].p);
		dev_slots[slot].p = NULL;
		dev_slots[slot].n = 0;
		size_t want = bytes + bytes / 4;
		if(hipMalloc(&dev_slots[slot].p, want) != hipSuccess) {
			dev_slots[slot].p = NULL;
			sphr_set_error("hipMalloc of %zu bytes failed", want);
			return NULL;
		}
		dev_slots[slot].n = want;
	}
	return dev_slots[slot].p;
}

void *sphr_rt_pinned(int slot, size_t bytes)
{
	if(bytes == 0)
		bytes = 1;
	if(host_slots[slot].n < bytes) {
		if(host_slots[slot].p)
			hipHostFree(host_slots[slot].p);
		host_slots[slot].p = NULL;
		host_slots[slot].n = 0;
		size_t want = bytes + bytes / 4;
		if(hipHostMalloc(&host_slots[slot].p, want, hipHostMallocDefault) != hipSuccess) {
			host_slots[slot].p = NULL;
			sphr_set_error("hipHostMalloc of %zu bytes failed", want);
			return NULL;
		}
		host_slots[slot].n = want;
	}
	return host_slots[slot].p;
}

int sphr_rt_check(hipError_t e, const char *what)
{
	if(e != hipSuccess) {
		sphr_set_error("%s: %s", what, hipGetErrorString(e));
		return -1;
	}
	return 0;
}

/* ---- public helpers ---- */

extern "C" int sphr_gpu_available(void)
{
	return sphr_gpu_ensure() == 0;
}

extern "C" int sphr_device_info(char *name, int name_len, int *cus, size_t *lds)
{
	if(sphr_gpu_ensure())
		return -1;
	if(name && name_len > 0)
		snprintf(name, name_len, "%s", ctx.name);
	if(cus)
		*cus = ctx.cus;
	if(lds)
		*lds = ctx.lds_max;
	return 0;
}

extern "C" void *sphr_device_alloc(size_t bytes)
{
	void *p = NULL;
	if(sphr_gpu_ensure())
		return NULL;
	if(hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
		sphr_set_error("hipMalloc of %zu bytes failed", bytes);
		return NULL;
	}
	return p;
}

extern "C" void sphr_device_free(void *p)
{
	if(p)
		hipFree(p);
}

/* The copies and the fill are ordered after the work already queued on
 * the library's stream (a non-blocking stream, which the null-stream
 * copies would not wait for): a host read sees the results of that work,
 * and a host write cannot overtake a kernel still using the buffer. */
extern "C" int sphr_memcpy_htod(void *dst, const void *src, size_t bytes)
{
	if(sphr_gpu_ensure())
		return -1;
	return sphr_rt_check(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize") || sphr_rt_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy HtoD") ? -1 : 0;
}

extern "C" int sphr_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
	if(sphr_gpu_ensure())
		return -1;
	return sphr_rt_check(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize") || sphr_rt_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy DtoH") ? -1 : 0;
}

extern "C" int sphr_memcpy_dtod(void *dst, const void *src, size_t bytes)
{
	if(sphr_gpu_ensure())
		return -1;
	return sphr_rt_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx.stream), "hipMemcpy DtoD") || sphr_rt_check(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize") ? -1 : 0;
}

extern "C" int sphr_memset_device(void *dst, int value, size_t bytes)
{
	if(sphr_gpu_ensure())
		return -1;
	return sphr_rt_check(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize") || sphr_rt_check(hipMemset(dst, value, bytes), "hipMemset")
		       || sphr_rt_check(hipDeviceSynchronize(), "hipMemset") ? -1 : 0;
}

extern "C" void *sphr_stream(void)
{
	if(sphr_gpu_ensure())
		return NULL;
	return (void *) ctx.stream;
}

extern "C" int sphr_stream_synchronize(void *stream)
{
	if(sphr_gpu_ensure())
		return -1;
	return sphr_rt_check(hipStreamSynchronize((hipStream_t) stream), "hipStreamSynchronize");
}

extern "C" int sphr_device_synchronize(void)
{
	if(sphr_gpu_ensure())
		return -1;
	return sphr_rt_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

extern "C" void *sphr_event_new(void)
{
	hipEvent_t e;
	if(sphr_gpu_ensure() || hipEventCreate(&e) != hipSuccess)
		return NULL;
	return (void *) e;
}

extern "C" void sphr_event_free(void *e)
{
	if(e)
		hipEventDestroy((hipEvent_t) e);
}

extern "C" int sphr_event_record(void *e, void *stream)
{
	return sphr_rt_check(hipEventRecord((hipEvent_t) e, (hipStream_t) stream), "hipEventRecord");
}

extern "C" int sphr_event_synchronize(void *e)
{
	return sphr_rt_check(hipEventSynchronize((hipEvent_t) e), "hipEventSynchronize");
}

extern "C" float sphr_event_elapsed_ms(void *a, void *b)
{
	float ms = -1.f;
	if(hipEventElapsedTime(&ms, (hipEvent_t) a, (hipEvent_t) b) != hipSuccess)
		return -1.f;
	return ms;
}
