/*
 * Plan-time dense products on the GPU: complex float64 C = A B for host
 * operands, on the MFMA GEMM (the complex B expanded to its real
 * [2K][2N] form on the device).  Used where plan construction reduces to
 * large products (the localization plan's projection multiply).
 */
#include <hip/hip_runtime.h>
#include "../sphr_internal.h"
#include "kernels.h"
#include "runtime.h"

/* B complex [K][N] -> E real [2K][2N]: E[2k][2n] = Br, E[2k][2n+1] = Bi,
 * E[2k+1][2n] = -Bi, E[2k+1][2n+1] = Br */
__global__ void k_cexpand(const cplx<double> *__restrict__ B, long K, long N, double *__restrict__ E)
{
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= K * N)
		return;
	const long k = i / N, n = i - k * N;
	const cplx<double> b = B[i];
	double *r0 = E + (2 * k) * (2 * N) + 2 * n;
	double *r1 = r0 + 2 * N;
	r0[0] = b.re;
	r0[1] = b.im;
	r1[0] = -b.im;
	r1[1] = b.re;
}

extern "C" int sphr_gpu_zgemm(int M, int N, int K, const double *A, const double *B, double *C)
{
	if(sphr_gpu_ensure())
		return -1;
	if(M <= 0 || N <= 0 || K <= 0)
		return 0;
	hipStream_t st = sphr_rt_stream();
	void *dA = NULL, *dB = NULL, *dE = NULL, *dC = NULL;
	const size_t szA = (size_t) M * K * 16, szB = (size_t) K * N * 16, szE = (size_t) 4 * K * N * 8, szC = (size_t) M * N * 16;
	int rc = -1;
	if(sphr_rt_check(hipMalloc(&dA, szA), "hipMalloc") || sphr_rt_check(hipMalloc(&dB, szB), "hipMalloc")
	   || sphr_rt_check(hipMalloc(&dE, szE), "hipMalloc") || sphr_rt_check(hipMalloc(&dC, szC), "hipMalloc"))
		goto done;
	if(sphr_rt_check(hipMemcpyAsync(dA, A, szA, hipMemcpyHostToDevice, st), "hipMemcpy")
	   || sphr_rt_check(hipMemcpyAsync(dB, B, szB, hipMemcpyHostToDevice, st), "hipMemcpy"))
		goto done;
	{
		const long kn = (long) K * N;
		hipLaunchKernelGGL(k_cexpand, dim3((unsigned) ((kn + 255) / 256)), dim3(256), 0, st, (const cplx<double> *) dB, (long) K, (long) N, (double *) dE);
		if(sphr_rt_check(hipGetLastError(), "expand"))
			goto done;
	}
	if(launch_gemm<double, double>(M, 2 * N, 2 * K, (const double *) dA, 2L * K, (const double *) dE, 2L * N, (double *) dC, 2L * N, st)) {
		sphr_set_error("gemm launch failed");
		goto done;
	}
	if(sphr_rt_check(hipMemcpyAsync(C, dC, szC, hipMemcpyDeviceToHost, st), "hipMemcpy") || sphr_rt_check(hipStreamSynchronize(st), "synchronize"))
		goto done;
	rc = 0;
done:
	hipFree(dA);
	hipFree(dB);
	hipFree(dE);
	hipFree(dC);
	return rc;
}
