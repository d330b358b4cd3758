/*
 * Synthetic detector data on the device, for training and evaluating the
 * learned peak proposal at scale: white Gaussian noise from a
 * counter-based generator plus one sine-Gaussian plane wave per segment
 * (the burst model of the three-site workload), written straight into
 * the batched pipeline's input layout [n_det][S][n] float32.
 *
 * Noise: Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC 2011) keyed by the seed, one counter
 * per group of four samples, Box-Muller to normals.  Burst:
 *   h_d(t) = amp sin(2 pi f0 u) exp(-(u / tau)^2),  u = t - t_mid + delay_d,
 * with per-segment (amp, f0, tau, delay_0 .. delay_{D-1}) given.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include "../sphr_internal.h"
#include "runtime.h"

__device__ inline void philox_round(uint32_t c[4], const uint32_t k[2])
{
	const uint64_t p0 = (uint64_t) 0xD2511F53u * c[0];
	const uint64_t p1 = (uint64_t) 0xCD9E8D57u * c[2];
	const uint32_t hi0 = (uint32_t) (p0 >> 32), lo0 = (uint32_t) p0;
	const uint32_t hi1 = (uint32_t) (p1 >> 32), lo1 = (uint32_t) p1;
	const uint32_t n0 = hi1 ^ c[1] ^ k[0], n2 = hi0 ^ c[3] ^ k[1];
	c[0] = n0;
	c[1] = lo1;
	c[2] = n2;
	c[3] = lo0;
}

__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
	uint32_t k[2] = { k0, k1 };
	for(int r = 0; r < 10; r++) {
		philox_round(c, k);
		k[0] += 0x9E3779B9u;
		k[1] += 0xBB67AE85u;
	}
}

__global__ void k_inject_bursts(const float *__restrict__ par, int D, int S, int n, float delta_t, float noise, uint32_t k0, uint32_t k1, float *__restrict__ out)
{
	const long g = (long) blockIdx.x * blockDim.x + threadIdx.x;
	const long total = (long) D * S * n;
	if(4 * g >= total)
		return;
	uint32_t c[4] = { (uint32_t) g, (uint32_t) (g >> 32), 0u, 0u };
	philox4x32_10(c, k0, k1);
	float z[4];
	for(int q = 0; q < 2; q++) {
		/* 24-bit uniforms in (0, 1), exact in float32 */
		const float u1 = ((float) (c[2 * q] >> 8) + 0.5f) * 5.9604644775390625e-8f;
		const float u2 = ((float) (c[2 * q + 1] >> 8) + 0.5f) * 5.9604644775390625e-8f;
		const float r = sqrtf(-2.f * logf(u1));
		float sn, cs;
		sincosf(6.283185307179586f * u2, &sn, &cs);
		z[2 * q] = r * cs;
		z[2 * q + 1] = r * sn;
	}
	const int np = D + 3;
	for(int q = 0; q < 4; q++) {
		const long i = 4 * g + q;
		if(i >= total)
			break;
		const int k = (int) (i % n);
		const long ds = i / n;
		const int s = (int) (ds % S), d = (int) (ds / S);
		const float *p = par + (size_t) s * np;
		const float u = (k - 0.5f * n) * delta_t + p[3 + d];
		const float e = u / p[2];
		out[i] = noise * z[q] + p[0] * sinf(6.283185307179586f * p[1] * u) * expf(-e * e);
	}
}

extern "C" int sphr_inject_bursts(const float *params, int n_det, int n_segments, int n, double delta_t, double noise, unsigned long long seed, float *out, void *stream)
{
	if(sphr_gpu_ensure())
		return -1;
	if(!params || !out || n_det < 1 || n_segments < 1 || n < 1) {
		sphr_set_error("sphr_inject_bursts: bad arguments");
		return -1;
	}
	hipStream_t st = stream ? (hipStream_t) stream : sphr_rt_stream();
	const long groups = ((long) n_det * n_segments * n + 3) / 4;
	hipLaunchKernelGGL(k_inject_bursts, dim3((unsigned) ((groups + 255) / 256)), dim3(256), 0, st, params, n_det, n_segments, n, (float) delta_t, (float) noise, (uint32_t) seed, (uint32_t) (seed >> 32), out);
	return sphr_rt_check(hipGetLastError(), "inject launch");
}
