/*
 * Real GEMM on MFMA for gfx950: C[M][N] = A[M][K] B[K][N], all row-major
 * with leading dimensions, float64 (v_mfma_f64_16x16x4f64) or float32
 * (v_mfma_f32_16x16x4f32) with accumulation in the same type.  A may be
 * stored in float32 and widened on the way into LDS.
 *
 * A complex product with interleaved (re, im) storage is the same kernel:
 * read the complex A[M][K] as real [M][2K] and expand the complex B[K][N]
 * to real [2K][2N] with rows 2k = (Br, Bi) pairs and 2k+1 = (-Bi, Br)
 * pairs; the real [M][2N] result is the complex C.  Eight real flops per
 * complex multiply-add, as in the complex product itself.
 *
 * Tiling: workgroup 256 threads = 4 waves in 2 x 2, block tile 128 x 128,
 * wave tile 64 x 64 = 4 x 4 MFMA tiles of 16 x 16; K chunks of KC (8 in
 * float64, 32 in float32) staged global -> registers -> LDS through two
 * LDS buffers: the next chunk's global loads are in flight during the
 * current chunk's MFMAs and are stored to the other buffer after them, one
 * barrier per chunk.  Both tiles stay in their global orientation (A as
 * [row][k], B as [k][col]) with padded rows: the A rows are KC + 2 (float64)
 * or KC + 4 (float32) elements long, the B rows 128 + 8 or 128 + 16, so
 * the stores of a wave and the MFMA fragment reads (16 rows x 4 k of A,
 * 4 k x 16 columns of B) each hit every LDS bank equally (a transposed A
 * tile put the 16 k of a row on one bank).  Workgroup ids are
 * remapped so the 8 XCDs each work on a contiguous range of tiles (each
 * XCD has its own L2), ordered in groups of 8 row panels: the workgroups
 * in flight on an XCD cover 8 row panels x 8 column panels and share them
 * in its L2.
 *
 * MFMA 16x16x4 operand layout (64 lanes): A[i][k] in lane i + 16 k, B[k][j]
 * in lane j + 16 k; accumulator element e of lane j + 16 q holds D[i][j]
 * with i = 4 q + e (f32) or i = q + 4 e (f64).
 */
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <rocblas/rocblas.h>
#include "kernels.h"

namespace {

template <typename T> struct mf;
template <> struct mf<float> {
	typedef float v4 __attribute__((ext_vector_type(4)));
	static __device__ inline v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
	static __device__ inline int row(int q, int e) { return 4 * q + e; }
	enum { KC = 32 };
};
template <> struct mf<double> {
	typedef double v4 __attribute__((ext_vector_type(4)));
	static __device__ inline v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
	static __device__ inline int row(int q, int e) { return q + 4 * e; }
	/* 8: the staging registers of a 16-deep chunk would push the kernel
	 * past 256 registers per lane (one workgroup per CU) */
	enum { KC = 8 };
};

#define GM_BM 128
#define GM_BN 128
#define GM_NT 256
#define GM_GM 8		/* row panels per tile group */

template <typename T> struct gm_pad;
template <> struct gm_pad<float> {
	enum { A = 4, B = 16 };
};
template <> struct gm_pad<double> {
	enum { A = 2, B = 8 };
};

template <typename T, typename TA>
__global__ void __launch_bounds__(GM_NT) k_gemm(int M, int N, int K, const TA *__restrict__ A, long lda, const T *__restrict__ B, long ldb, T *__restrict__ C, long ldc, int tiles_m, int tiles_n)
{
	typedef typename mf<T>::v4 v4;
	constexpr int KC = mf<T>::KC;
	constexpr int RA = GM_BM * KC / GM_NT, RB = KC * GM_BN / GM_NT;	/* staged elements per thread */
	constexpr int SA = KC + gm_pad<T>::A, SB = GM_BN + gm_pad<T>::B;
	__shared__ T sA[2][GM_BM * SA];	/* [row][k] */
	__shared__ T sB[2][KC * SB];	/* [k][col] */
	/* XCD-aware remap: hardware places workgroup id on XCD id % 8; give
	 * each XCD a contiguous range of tiles (column panel major) */
	const int nwg = tiles_m * tiles_n;
	const int id = blockIdx.x;
	int t = id;
	if(nwg % 8 == 0)
		t = (id % 8) * (nwg / 8) + id / 8;
	/* groups of GM_GM row panels side by side, all column tiles of a
	 * group before the next: the workgroups in flight on an XCD share
	 * both operands' panels in its L2 */
	const int gid = t / (GM_GM * tiles_n), first = gid * GM_GM;
	const int gs = tiles_m - first < GM_GM ? tiles_m - first : GM_GM;
	const int w = t - gid * GM_GM * tiles_n;
	const int tm = first + w % gs, tn = w / gs;
	const int m0 = tm * GM_BM, n0 = tn * GM_BN;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
	const int li = lane & 15, lk = lane >> 4;
	/* global -> register staging; A coalesced along k, B along n */
	T ra[RA], rb[RB];
	auto load = [&](int k0) {
#pragma unroll
		for(int v = 0; v < RA; v++) {
			const int i = v * GM_NT + threadIdx.x;
			const int r = i / KC, c = i - r * KC;
			const int gm = m0 + r, gk = k0 + c;
			ra[v] = (gm < M && gk < K) ? (T) A[(size_t) gm * lda + gk] : (T) 0;
		}
#pragma unroll
		for(int v = 0; v < RB; v++) {
			const int i = v * GM_NT + threadIdx.x;
			const int r = i / GM_BN, c = i - r * GM_BN;
			const int gk = k0 + r, gn = n0 + c;
			rb[v] = (gk < K && gn < N) ? B[(size_t) gk * ldb + gn] : (T) 0;
		}
	};
	auto store = [&](int buf) {
#pragma unroll
		for(int v = 0; v < RA; v++) {
			const int i = v * GM_NT + threadIdx.x;
			const int r = i / KC, c = i - r * KC;
			sA[buf][r * SA + c] = ra[v];
		}
#pragma unroll
		for(int v = 0; v < RB; v++) {
			const int i = v * GM_NT + threadIdx.x;
			const int r = i / GM_BN, c = i - r * GM_BN;
			sB[buf][r * SB + c] = rb[v];
		}
	};
	v4 acc[4][4];
#pragma unroll
	for(int a = 0; a < 4; a++)
#pragma unroll
		for(int b = 0; b < 4; b++)
			acc[a][b] = (v4) 0;
	load(0);
	store(0);
	__syncthreads();
	int cur = 0;
	for(int k0 = 0; k0 < K; k0 += KC) {
		const bool next = k0 + KC < K;
		if(next)
			load(k0 + KC);
		const T *pa = sA[cur] + (wm + li) * SA + lk;
		const T *pb = sB[cur] + lk * SB + wn + li;
#pragma unroll
		for(int ks = 0; ks < KC; ks += 4) {
			T af[4], bf[4];
#pragma unroll
			for(int a = 0; a < 4; a++)
				af[a] = pa[16 * a * SA + ks];
#pragma unroll
			for(int b = 0; b < 4; b++)
				bf[b] = pb[ks * SB + 16 * b];
#pragma unroll
			for(int a = 0; a < 4; a++)
#pragma unroll
				for(int b = 0; b < 4; b++)
					acc[a][b] = mf<T>::mma(af[a], bf[b], acc[a][b]);
		}
		if(next)
			store(cur ^ 1);
		__syncthreads();
		cur ^= 1;
	}
	for(int a = 0; a < 4; a++)
		for(int b = 0; b < 4; b++) {
			const int gn = n0 + wn + 16 * b + li;
			if(gn >= N)
				continue;
			for(int e = 0; e < 4; e++) {
				const int gm = m0 + wm + 16 * a + mf<T>::row(lk, e);
				if(gm < M)
					C[(size_t) gm * ldc + gn] = acc[a][b][e];
			}
		}
}

}	/* namespace */

template <typename T, typename TA>
int launch_gemm(int M, int N, int K, const TA *A, long lda, const T *B, long ldb, T *C, long ldc, hipStream_t st)
{
	if(M <= 0 || N <= 0 || K <= 0)
		return 0;
	const int tm = (M + GM_BM - 1) / GM_BM, tn = (N + GM_BN - 1) / GM_BN;
	hipLaunchKernelGGL((k_gemm<T, TA>), dim3(tm * tn), dim3(GM_NT), 0, st, M, N, K, A, lda, B, ldb, C, ldc, tm, tn);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_gemm<double, double>(int, int, int, const double *, long, const double *, long, double *, long, hipStream_t);
template int launch_gemm<double, float>(int, int, int, const float *, long, const double *, long, double *, long, hipStream_t);
template int launch_gemm<float, float>(int, int, int, const float *, long, const float *, long, float *, long, hipStream_t);
template int launch_gemm<float, double>(int, int, int, const double *, long, const float *, long, float *, long, hipStream_t);

/*
 * The same row-major product through rocBLAS (column-major): C^T = B^T A^T,
 * i.e. gemm(N, M, K) with B as the first operand.  For the plain
 * float32 / float64 products whose shape the library kernel above does
 * not fill well.  One handle per device, created on first use; the
 * stream is set under a lock, so calls from several threads each run on
 * their own stream.
 */
static pthread_mutex_t blas_lock = PTHREAD_MUTEX_INITIALIZER;
static rocblas_handle blas_handles[64];

static rocblas_handle blas_handle(void)
{
	int dev = 0;
	if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
		return NULL;
	if(!blas_handles[dev] && rocblas_create_handle(&blas_handles[dev]) != rocblas_status_success)
		blas_handles[dev] = NULL;
	return blas_handles[dev];
}

static rocblas_status blas_call(int M, int N, int K, const float *A, long lda, const float *B, long ldb, float *C, long ldc, rocblas_handle h)
{
	const float one = 1.0f, zero = 0.0f;
	return rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, N, M, K, &one, B, (rocblas_int) ldb, A, (rocblas_int) lda, &zero, C, (rocblas_int) ldc);
}

static rocblas_status blas_call(int M, int N, int K, const double *A, long lda, const double *B, long ldb, double *C, long ldc, rocblas_handle h)
{
	const double one = 1.0, zero = 0.0;
	return rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, N, M, K, &one, B, (rocblas_int) ldb, A, (rocblas_int) lda, &zero, C, (rocblas_int) ldc);
}

template <typename T>
int launch_gemm_blas(int M, int N, int K, const T *A, long lda, const T *B, long ldb, T *C, long ldc, hipStream_t st)
{
	if(M <= 0 || N <= 0 || K <= 0)
		return 0;
	pthread_mutex_lock(&blas_lock);
	rocblas_handle h = blas_handle();
	rocblas_status r = h ? rocblas_set_stream(h, st) : rocblas_status_invalid_handle;
	if(r == rocblas_status_success)
		r = blas_call(M, N, K, A, lda, B, ldb, C, ldc, h);
	pthread_mutex_unlock(&blas_lock);
	return r == rocblas_status_success ? 0 : -1;
}

template int launch_gemm_blas<float>(int, int, int, const float *, long, const float *, long, float *, long, hipStream_t);
template int launch_gemm_blas<double>(int, int, int, const double *, long, const double *, long, double *, long, hipStream_t);

/*
 * Test and timing hook: C [M][N] = A [M][K] B [K][N] (row-major device
 * arrays, float64 when f64 else float32) on the library kernel (impl 0)
 * or rocBLAS (impl 1), queued on stream.
 */
extern "C" int sphr_gemm_run(int impl, int f64, int M, int N, int K, const void *A, long lda, const void *B, long ldb, void *C, long ldc, void *stream)
{
	hipStream_t st = (hipStream_t) stream;
	if(f64)
		return impl ? launch_gemm_blas<double>(M, N, K, (const double *) A, lda, (const double *) B, ldb, (double *) C, ldc, st)
			    : launch_gemm<double, double>(M, N, K, (const double *) A, lda, (const double *) B, ldb, (double *) C, ldc, st);
	return impl ? launch_gemm_blas<float>(M, N, K, (const float *) A, lda, (const float *) B, ldb, (float *) C, ldc, st)
		    : launch_gemm<float, float>(M, N, K, (const float *) A, lda, (const float *) B, ldb, (float *) C, ldc, st);
}
