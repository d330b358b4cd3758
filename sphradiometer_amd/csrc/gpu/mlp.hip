/*
 * The learned peak proposal: a multilayer perceptron from the real-form SH
 * coefficients of a sky to K candidate peak directions, trained and run
 * with the kernels of this file.
 *
 *   h_0 = x (n_in), z_i = h_{i-1} W_i + b_i, h_i = GELU(z_i) (i = 1..depth),
 *   y = h_depth W_out + b_out (3 K outputs), candidate k = y[3k..3k+2] / |.|
 *
 * Weights are kept in float32 (master copy, Adam state) and in the
 * operand type of the products: bfloat16 (MFMA 16x16x16 bf16, float32
 * accumulation) or float32 (MFMA 16x16x4 f32, used to check gradients).
 *
 * Loss per sample, with t the unit target and d_k = 1 - c_k . t:
 *   min_k d_k + (lambda / K) sum_k d_k,
 * so the best candidate is pulled onto the target and the others stay
 * trained.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <sphradiometer/sphradiometer_batch.h>
#include "../sphr_internal.h"
#include "kernels.h"
#include "runtime.h"

typedef uint16_t bf16_t;

extern "C" int sphr_mlp_set_params(struct sphr_mlp *, const float *);

__host__ __device__ inline bf16_t f2bf(float f)
{
	/* round to nearest even; NaN stays NaN */
	union {
		float f;
		uint32_t u;
	} v;
	v.f = f;
	if((v.u & 0x7fffffffu) > 0x7f800000u)
		return (bf16_t) ((v.u >> 16) | 0x40);
	const uint32_t r = 0x7fffu + ((v.u >> 16) & 1u);
	return (bf16_t) ((v.u + r) >> 16);
}

__host__ __device__ inline float bf2f(bf16_t h)
{
	union {
		float f;
		uint32_t u;
	} v;
	v.u = (uint32_t) h << 16;
	return v.f;
}

template <typename TI> __device__ inline TI to_op(float f);
template <> __device__ inline float to_op<float>(float f) { return f; }
template <> __device__ inline bf16_t to_op<bf16_t>(float f) { return f2bf(f); }

/* ------------------------------------------------------------------ */
/* GEMM C[M][N] (f32) = op(A)[M][K] op(B)[K][N], operands of type TI   */
/* ------------------------------------------------------------------ */

#define MG_BM 64
#define MG_BN 64
#define MG_KC 32
#define MG_PAD 4

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename TI> struct mg_mma;
template <> struct mg_mma<float> {
	enum { KS = 4 };
	static __device__ inline f4v step(const float *a, const float *b, int lq, f4v c)
	{
		return __builtin_amdgcn_mfma_f32_16x16x4f32(a[lq], b[lq], c, 0, 0, 0);
	}
};
template <> struct mg_mma<bf16_t> {
	enum { KS = 16 };
	/* lane i + 16 q holds k = 4 q .. 4 q + 3 of its row / column */
	static __device__ inline f4v step(const bf16_t *a, const bf16_t *b, int lq, f4v c)
	{
		const s4v av = *reinterpret_cast<const s4v *>(a + 4 * lq);
		const s4v bv = *reinterpret_cast<const s4v *>(b + 4 * lq);
		return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, c, 0, 0, 0);
	}
};

template <typename TI, int TA, int TB>
__global__ void __launch_bounds__(256) k_mlp_gemm(int M, int N, int K, const TI *__restrict__ A, int lda, const TI *__restrict__ B, int ldb, float *__restrict__ C, int ldc)
{
	__shared__ __align__(16) TI sA[MG_BM][MG_KC + MG_PAD];	/* [row][k] */
	__shared__ __align__(16) TI sB[MG_BN][MG_KC + MG_PAD];	/* [col][k] */
	const int m0 = blockIdx.y * MG_BM, n0 = blockIdx.x * MG_BN;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
	const int li = lane & 15, lq = lane >> 4;
	f4v acc[2][2];
	for(int a = 0; a < 2; a++)
		for(int b = 0; b < 2; b++)
			acc[a][b] = (f4v) 0;
	const TI zero = to_op<TI>(0.f);
	for(int k0 = 0; k0 < K; k0 += MG_KC) {
		for(int idx = threadIdx.x; idx < MG_BM * MG_KC; idx += 256) {
			int i, k;
			if(TA) {
				k = idx / MG_BM;
				i = idx - k * MG_BM;
			} else {
				i = idx / MG_KC;
				k = idx - i * MG_KC;
			}
			const int gi = m0 + i, gk = k0 + k;
			TI v = zero;
			if(gi < M && gk < K)
				v = TA ? A[(size_t) gk * lda + gi] : A[(size_t) gi * lda + gk];
			sA[i][k] = v;
		}
		for(int idx = threadIdx.x; idx < MG_BN * MG_KC; idx += 256) {
			int j, k;
			if(TB) {
				j = idx / MG_KC;
				k = idx - j * MG_KC;
			} else {
				k = idx / MG_BN;
				j = idx - k * MG_BN;
			}
			const int gj = n0 + j, gk = k0 + k;
			TI v = zero;
			if(gj < N && gk < K)
				v = TB ? B[(size_t) gj * ldb + gk] : B[(size_t) gk * ldb + gj];
			sB[j][k] = v;
		}
		__syncthreads();
		for(int ks = 0; ks < MG_KC; ks += mg_mma<TI>::KS)
			for(int a = 0; a < 2; a++)
				for(int b = 0; b < 2; b++)
					acc[a][b] = mg_mma<TI>::step(&sA[wm + 16 * a + li][ks], &sB[wn + 16 * b + li][ks], lq, acc[a][b]);
		__syncthreads();
	}
	for(int a = 0; a < 2; a++)
		for(int b = 0; b < 2; b++) {
			const int gj = n0 + wn + 16 * b + li;
			if(gj >= N)
				continue;
			for(int e = 0; e < 4; e++) {
				const int gi = m0 + wm + 16 * a + 4 * lq + e;
				if(gi < M)
					C[(size_t) gi * ldc + gj] = acc[a][b][e];
			}
		}
}

template <typename TI>
static int mlp_gemm(int ta, int tb, int M, int N, int K, const TI *A, int lda, const TI *B, int ldb, float *C, int ldc, hipStream_t st)
{
	dim3 grid((N + MG_BN - 1) / MG_BN, (M + MG_BM - 1) / MG_BM);
	if(!ta && !tb)
		hipLaunchKernelGGL((k_mlp_gemm<TI, 0, 0>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc);
	else if(ta && !tb)
		hipLaunchKernelGGL((k_mlp_gemm<TI, 1, 0>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc);
	else if(!ta && tb)
		hipLaunchKernelGGL((k_mlp_gemm<TI, 0, 1>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc);
	else
		return -1;
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* ------------------------------------------------------------------ */
/* elementwise and reductions                                          */
/* ------------------------------------------------------------------ */

__device__ inline float gelu(float z)
{
	return 0.5f * z * (1.f + erff(z * 0.70710678118654752f));
}

__device__ inline float gelu_d(float z)
{
	return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

/* z += b; h = GELU(z) as the operand type (hidden layers) */
template <typename TI>
__global__ void k_bias_gelu(float *__restrict__ Z, const float *__restrict__ b, int rows, int n, TI *__restrict__ H)
{
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= (long) rows * n)
		return;
	const float z = Z[i] + b[i % n];
	Z[i] = z;
	H[i] = to_op<TI>(gelu(z));
}

__global__ void k_bias(float *__restrict__ Z, const float *__restrict__ b, int rows, int n)
{
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= (long) rows * n)
		return;
	Z[i] += b[i % n];
}

/* dZ = dH * GELU'(Z), also as the operand type */
template <typename TI>
__global__ void k_gelu_back(const float *__restrict__ dH, const float *__restrict__ Z, long count, float *__restrict__ dZ, TI *__restrict__ dZo)
{
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= count)
		return;
	const float d = dH[i] * gelu_d(Z[i]);
	dZ[i] = d;
	dZo[i] = to_op<TI>(d);
}

template <typename TI>
__global__ void k_to_op(const float *__restrict__ x, long count, TI *__restrict__ y)
{
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if(i < count)
		y[i] = to_op<TI>(x[i]);
}

/* db[j] = sum_r dZ[r][j] */
__global__ void k_colsum(const float *__restrict__ dZ, int rows, int n, float *__restrict__ db)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= n)
		return;
	float s = 0.f;
	for(int r = 0; r < rows; r++)
		s += dZ[(size_t) r * n + j];
	db[j] = s;
}

/* candidate loss and its gradient; one thread per sample */
__global__ void k_loss(const float *__restrict__ Y, const float *__restrict__ T, int rows, int K, float lambda, float *__restrict__ loss, float *__restrict__ dY)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if(r >= rows)
		return;
	const float *y = Y + (size_t) r * 3 * K;
	const float t0 = T[3 * r], t1 = T[3 * r + 1], t2 = T[3 * r + 2];
	float d[16], nrm[16], dot[16];
	int best = 0;
	for(int k = 0; k < K; k++) {
		const float n = sqrtf(y[3 * k] * y[3 * k] + y[3 * k + 1] * y[3 * k + 1] + y[3 * k + 2] * y[3 * k + 2]) + 1e-12f;
		const float c = (y[3 * k] * t0 + y[3 * k + 1] * t1 + y[3 * k + 2] * t2) / n;
		nrm[k] = n;
		dot[k] = c;
		d[k] = 1.f - c;
		if(d[k] < d[best])
			best = k;
	}
	float l = d[best], sum = 0.f;
	for(int k = 0; k < K; k++)
		sum += d[k];
	l += lambda / K * sum;
	if(loss)
		loss[r] = l;
	if(dY) {
		const float inv_rows = 1.f / rows;
		for(int k = 0; k < K; k++) {
			const float w = ((k == best) ? 1.f : 0.f) + lambda / K;
			const float n = nrm[k];
			const float c0 = y[3 * k] / n, c1 = y[3 * k + 1] / n, c2 = y[3 * k + 2] / n;
			/* d(1 - c.t)/dy = -(t - (c.t) c) / |y| */
			float *g = dY + (size_t) r * 3 * K + 3 * k;
			g[0] = -w * (t0 - dot[k] * c0) / n * inv_rows;
			g[1] = -w * (t1 - dot[k] * c1) / n * inv_rows;
			g[2] = -w * (t2 - dot[k] * c2) / n * inv_rows;
		}
	}
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2)
{
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n)
		return;
	const float gi = g[i];
	const float mi = b1 * m[i] + (1.f - b1) * gi;
	const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

/* real-form SH coefficients of a real sky, scaled to unit RMS:
 * Re a_l0 (l = 0..L), then for m = 1..L, l = m..L: Re a_lm, Im a_lm */
template <typename TA>
__global__ void __launch_bounds__(256) k_features(const cplx<TA> *__restrict__ alm, int L, float *__restrict__ out)
{
	const int s = blockIdx.x;
	const int ncoef = (L + 1) * (L + 1);
	const cplx<TA> *a = alm + (size_t) s * ncoef;
	float *o = out + (size_t) s * ncoef;
	__shared__ float red[256];
	float ss = 0.f;
	for(int i = threadIdx.x; i < ncoef; i += 256) {
		float v;
		if(i <= L)
			v = (float) a[i].re;
		else {
			const int c = (L + 1) + (i - (L + 1)) / 2;	/* the m > 0 coefficient */
			v = ((i - (L + 1)) & 1) ? (float) a[c].im : (float) a[c].re;
		}
		o[i] = v;
		ss += v * v;
	}
	red[threadIdx.x] = ss;
	__syncthreads();
	for(int w = 128; w > 0; w >>= 1) {
		if(threadIdx.x < w)
			red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	const float scale = red[0] > 0.f ? rsqrtf(red[0] / ncoef) : 0.f;
	for(int i = threadIdx.x; i < ncoef; i += 256)
		o[i] *= scale;
}

/* candidate directions as (theta, phi) seeds, [rows][K][2] */
__global__ void k_seeds(const float *__restrict__ Y, int rows, int K, double *__restrict__ seeds)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= rows * K)
		return;
	const float *y = Y + 3 * (size_t) i;
	const double x = y[0], yy = y[1], z = y[2];
	double ph = atan2(yy, x);
	if(ph < 0)
		ph += 2 * M_PI;
	seeds[2 * (size_t) i] = atan2(sqrt(x * x + yy * yy), z);
	seeds[2 * (size_t) i + 1] = ph;
}

/* ------------------------------------------------------------------ */
/* the network object                                                  */
/* ------------------------------------------------------------------ */

#define MLP_MAX_LAYERS 8

struct sphr_mlp {
	int n_in, width, depth, K, max_batch, bf16;
	int nl;				/* layers = depth + 1 */
	int din[MLP_MAX_LAYERS], dout[MLP_MAX_LAYERS];
	size_t ow[MLP_MAX_LAYERS], ob[MLP_MAX_LAYERS];
	size_t nparam;
	float *p, *g, *m, *v;		/* params, grads, Adam moments (float32) */
	void *wop;			/* weights in the operand type (same offsets as p) */
	void *xin;			/* [max_batch][n_in] operand type */
	float *Z[MLP_MAX_LAYERS];	/* pre-activations [max_batch][dout] */
	void *H[MLP_MAX_LAYERS];	/* activations (operand type) */
	float *dA, *dZ;			/* [max_batch][max width] */
	void *dZo;			/* operand type */
	float *loss;			/* [max_batch] */
};

static size_t op_size(const struct sphr_mlp *m)
{
	return m->bf16 ? sizeof(bf16_t) : sizeof(float);
}

extern "C" void sphr_mlp_free(struct sphr_mlp *m)
{
	if(!m)
		return;
	void *ptrs[] = { m->p, m->g, m->m, m->v, m->wop, m->xin, m->dA, m->dZ, m->dZo, m->loss };
	for(unsigned i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++)
		if(ptrs[i])
			hipFree(ptrs[i]);
	for(int l = 0; l < m->nl; l++) {
		if(m->Z[l])
			hipFree(m->Z[l]);
		if(m->H[l])
			hipFree(m->H[l]);
	}
	free(m);
}

extern "C" struct sphr_mlp *sphr_mlp_new(int n_in, int width, int depth, int n_candidates, int max_batch, int bf16, unsigned long long seed)
{
	if(sphr_gpu_ensure())
		return NULL;
	if(n_in < 1 || width < 1 || depth < 1 || depth + 1 > MLP_MAX_LAYERS || n_candidates < 1 || n_candidates > 16 || max_batch < 1) {
		sphr_set_error("sphr_mlp_new: bad shape");
		return NULL;
	}
	struct sphr_mlp *m = (struct sphr_mlp *) calloc(1, sizeof(*m));
	if(!m)
		return NULL;
	m->n_in = n_in;
	m->width = width;
	m->depth = depth;
	m->K = n_candidates;
	m->max_batch = max_batch;
	m->bf16 = bf16 ? 1 : 0;
	m->nl = depth + 1;
	size_t off = 0;
	for(int l = 0; l < m->nl; l++) {
		m->din[l] = l ? width : n_in;
		m->dout[l] = l < depth ? width : 3 * n_candidates;
		m->ow[l] = off;
		off += (size_t) m->din[l] * m->dout[l];
		m->ob[l] = off;
		off += m->dout[l];
	}
	m->nparam = off;
	const size_t wmax = width > 3 * n_candidates ? width : 3 * n_candidates;
	const size_t B = max_batch;
	int bad = sphr_rt_check(hipMalloc(&m->p, off * 4), "hipMalloc") || sphr_rt_check(hipMalloc(&m->g, off * 4), "hipMalloc")
		  || sphr_rt_check(hipMalloc(&m->m, off * 4), "hipMalloc") || sphr_rt_check(hipMalloc(&m->v, off * 4), "hipMalloc")
		  || sphr_rt_check(hipMalloc(&m->wop, off * op_size(m)), "hipMalloc")
		  || sphr_rt_check(hipMalloc(&m->xin, B * n_in * op_size(m)), "hipMalloc")
		  || sphr_rt_check(hipMalloc(&m->dA, B * wmax * 4), "hipMalloc") || sphr_rt_check(hipMalloc(&m->dZ, B * wmax * 4), "hipMalloc")
		  || sphr_rt_check(hipMalloc(&m->dZo, B * wmax * op_size(m)), "hipMalloc") || sphr_rt_check(hipMalloc(&m->loss, B * 4), "hipMalloc");
	for(int l = 0; l < m->nl && !bad; l++)
		bad = sphr_rt_check(hipMalloc(&m->Z[l], B * m->dout[l] * 4), "hipMalloc") || sphr_rt_check(hipMalloc(&m->H[l], B * m->dout[l] * op_size(m)), "hipMalloc");
	if(bad) {
		sphr_mlp_free(m);
		return NULL;
	}
	/* He-style initialisation from a seeded splitmix64 stream */
	float *h = (float *) malloc(off * sizeof(float));
	if(!h) {
		sphr_mlp_free(m);
		return NULL;
	}
	uint64_t st = seed * 0x9e3779b97f4a7c15ull + 1;
	for(int l = 0; l < m->nl; l++) {
		const double sc = sqrt(2.0 / m->din[l]) * (l == m->nl - 1 ? 0.5 : 1.0);
		for(size_t i = 0; i < (size_t) m->din[l] * m->dout[l]; i++) {
			double u[2];
			for(int q = 0; q < 2; q++) {
				uint64_t z = (st += 0x9e3779b97f4a7c15ull);
				z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
				z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
				z ^= z >> 31;
				u[q] = ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
			}
			h[m->ow[l] + i] = (float) (sc * sqrt(-2.0 * log(u[0])) * cos(2 * M_PI * u[1]));
		}
		for(int j = 0; j < m->dout[l]; j++)
			h[m->ob[l] + j] = 0.f;
	}
	const int rc = sphr_mlp_set_params(m, h);
	free(h);
	if(rc || sphr_rt_check(hipMemset(m->m, 0, off * 4), "hipMemset") || sphr_rt_check(hipMemset(m->v, 0, off * 4), "hipMemset")) {
		sphr_mlp_free(m);
		return NULL;
	}
	return m;
}

static int refresh_op_weights(struct sphr_mlp *m, hipStream_t st)
{
	const long n = (long) m->nparam;
	if(m->bf16)
		hipLaunchKernelGGL((k_to_op<bf16_t>), dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, m->p, n, (bf16_t *) m->wop);
	else
		hipLaunchKernelGGL((k_to_op<float>), dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, m->p, n, (float *) m->wop);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int sphr_mlp_set_params(struct sphr_mlp *m, const float *host)
{
	hipStream_t st = sphr_rt_stream();
	if(sphr_rt_check(hipMemcpyAsync(m->p, host, m->nparam * 4, hipMemcpyHostToDevice, st), "hipMemcpy") || refresh_op_weights(m, st))
		return -1;
	return sphr_rt_check(hipStreamSynchronize(st), "synchronize");
}

extern "C" int sphr_mlp_get_params(const struct sphr_mlp *m, float *host)
{
	return sphr_rt_check(hipMemcpy(host, m->p, m->nparam * 4, hipMemcpyDeviceToHost), "hipMemcpy");
}

extern "C" int sphr_mlp_get_grads(const struct sphr_mlp *m, float *host)
{
	return sphr_rt_check(hipMemcpy(host, m->g, m->nparam * 4, hipMemcpyDeviceToHost), "hipMemcpy");
}

extern "C" long sphr_mlp_n_params(const struct sphr_mlp *m)
{
	return (long) m->nparam;
}

template <typename TI>
static int forward_t(struct sphr_mlp *m, const float *x, int B, hipStream_t st)
{
	const long nx = (long) B * m->n_in;
	hipLaunchKernelGGL((k_to_op<TI>), dim3((unsigned) ((nx + 255) / 256)), dim3(256), 0, st, x, nx, (TI *) m->xin);
	const TI *in = (const TI *) m->xin;
	const TI *W = (const TI *) m->wop;
	for(int l = 0; l < m->nl; l++) {
		if(mlp_gemm<TI>(0, 0, B, m->dout[l], m->din[l], in, m->din[l], W + m->ow[l], m->dout[l], m->Z[l], m->dout[l], st))
			return -1;
		const long n = (long) B * m->dout[l];
		if(l < m->depth)
			hipLaunchKernelGGL((k_bias_gelu<TI>), dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, m->Z[l], m->p + m->ob[l], B, m->dout[l], (TI *) m->H[l]);
		else
			hipLaunchKernelGGL(k_bias, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, m->Z[l], m->p + m->ob[l], B, m->dout[l]);
		in = (const TI *) m->H[l];
	}
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename TI>
static int backward_t(struct sphr_mlp *m, int B, hipStream_t st)
{
	const TI *W = (const TI *) m->wop;
	/* dY is in dZ (float) on entry; its operand copy in dZo */
	for(int l = m->nl - 1; l >= 0; l--) {
		const int di = m->din[l], dn = m->dout[l];
		const TI *in = l ? (const TI *) m->H[l - 1] : (const TI *) m->xin;
		/* dW = in^T dZ, db = colsum dZ */
		if(mlp_gemm<TI>(1, 0, di, dn, B, in, di, (const TI *) m->dZo, dn, m->g + m->ow[l], dn, st))
			return -1;
		hipLaunchKernelGGL(k_colsum, dim3((dn + 255) / 256), dim3(256), 0, st, m->dZ, B, dn, m->g + m->ob[l]);
		if(l == 0)
			break;
		/* dA = dZ W^T, then through GELU of layer l-1 */
		if(mlp_gemm<TI>(0, 1, B, di, dn, (const TI *) m->dZo, dn, W + m->ow[l], dn, m->dA, di, st))
			return -1;
		const long n = (long) B * di;
		hipLaunchKernelGGL((k_gelu_back<TI>), dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, m->dA, m->Z[l - 1], n, m->dZ, (TI *) m->dZo);
	}
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

static hipStream_t pick(void *stream)
{
	return stream ? (hipStream_t) stream : sphr_rt_stream();
}

extern "C" int sphr_mlp_forward(struct sphr_mlp *m, const float *x, int B, float *y, void *stream)
{
	if(!m || !x || B < 1 || B > m->max_batch)
		return -1;
	hipStream_t st = pick(stream);
	const int r = m->bf16 ? forward_t<bf16_t>(m, x, B, st) : forward_t<float>(m, x, B, st);
	if(r)
		return -1;
	if(y)
		return sphr_rt_check(hipMemcpyAsync(y, m->Z[m->nl - 1], (size_t) B * 3 * m->K * 4, hipMemcpyDeviceToDevice, st), "hipMemcpy");
	return 0;
}

/* forward, loss and gradients (into the grad buffer) for a batch; per
 * sample losses to loss (device, may be NULL) */
extern "C" int sphr_mlp_loss_grad(struct sphr_mlp *m, const float *x, const float *target, int B, float lambda, float *loss, void *stream)
{
	if(!m || !x || !target || B < 1 || B > m->max_batch)
		return -1;
	hipStream_t st = pick(stream);
	if(sphr_mlp_forward(m, x, B, NULL, st))
		return -1;
	const int L = m->nl - 1;
	hipLaunchKernelGGL(k_loss, dim3((B + 255) / 256), dim3(256), 0, st, m->Z[L], target, B, m->K, lambda, loss ? loss : m->loss, m->dZ);
	const long n = (long) B * 3 * m->K;
	if(m->bf16)
		hipLaunchKernelGGL((k_to_op<bf16_t>), dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, m->dZ, n, (bf16_t *) m->dZo);
	else
		hipLaunchKernelGGL((k_to_op<float>), dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, m->dZ, n, (float *) m->dZo);
	if(hipGetLastError() != hipSuccess)
		return -1;
	return m->bf16 ? backward_t<bf16_t>(m, B, st) : backward_t<float>(m, B, st);
}

extern "C" int sphr_mlp_adam(struct sphr_mlp *m, float lr, float beta1, float beta2, float eps, int step, void *stream)
{
	hipStream_t st = pick(stream);
	const long n = (long) m->nparam;
	const float c1 = 1.f - powf(beta1, (float) step), c2 = 1.f - powf(beta2, (float) step);
	hipLaunchKernelGGL(k_adam, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, m->p, m->g, m->m, m->v, n, lr, beta1, beta2, eps, c1, c2);
	if(hipGetLastError() != hipSuccess)
		return -1;
	return refresh_op_weights(m, st);
}

extern "C" int sphr_mlp_train_step(struct sphr_mlp *m, const float *x, const float *target, int B, float lambda, float lr, int step, float *loss, void *stream)
{
	if(sphr_mlp_loss_grad(m, x, target, B, lambda, loss, stream))
		return -1;
	return sphr_mlp_adam(m, lr, 0.9f, 0.999f, 1e-8f, step, stream);
}

/* candidate seeds (theta, phi) for rows x K from the last forward */
extern "C" int sphr_mlp_seeds(struct sphr_mlp *m, int B, double *seeds, void *stream)
{
	hipStream_t st = pick(stream);
	const int n = B * m->K;
	hipLaunchKernelGGL(k_seeds, dim3((n + 255) / 256), dim3(256), 0, st, m->Z[m->nl - 1], B, m->K, seeds);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int sphr_features(const void *alm, int alm_f32, int S, int L, float *out, void *stream)
{
	hipStream_t st = pick(stream);
	if(alm_f32)
		hipLaunchKernelGGL((k_features<float>), dim3(S), dim3(256), 0, st, (const cplx<float> *) alm, L, out);
	else
		hipLaunchKernelGGL((k_features<double>), dim3(S), dim3(256), 0, st, (const cplx<double> *) alm, L, out);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* test hook: one GEMM of the network's operand type on device float
 * inputs (converted on the way), C = op(A) op(B) */
extern "C" int sphr_mlp_gemm_check(int bf16, int ta, int tb, int M, int N, int K, const float *A, const float *B, float *C)
{
	hipStream_t st = sphr_rt_stream();
	const size_t na = (size_t) M * K, nb = (size_t) K * N, es = bf16 ? 2 : 4;
	void *dA = NULL, *dB = NULL;
	int rc = -1;
	if(sphr_rt_check(hipMalloc(&dA, na * es), "hipMalloc") || sphr_rt_check(hipMalloc(&dB, nb * es), "hipMalloc"))
		goto done;
	if(bf16) {
		hipLaunchKernelGGL((k_to_op<bf16_t>), dim3((unsigned) ((na + 255) / 256)), dim3(256), 0, st, A, (long) na, (bf16_t *) dA);
		hipLaunchKernelGGL((k_to_op<bf16_t>), dim3((unsigned) ((nb + 255) / 256)), dim3(256), 0, st, B, (long) nb, (bf16_t *) dB);
		rc = mlp_gemm<bf16_t>(ta, tb, M, N, K, (const bf16_t *) dA, ta ? M : K, (const bf16_t *) dB, tb ? K : N, C, N, st);
	} else {
		hipLaunchKernelGGL((k_to_op<float>), dim3((unsigned) ((na + 255) / 256)), dim3(256), 0, st, A, (long) na, (float *) dA);
		hipLaunchKernelGGL((k_to_op<float>), dim3((unsigned) ((nb + 255) / 256)), dim3(256), 0, st, B, (long) nb, (float *) dB);
		rc = mlp_gemm<float>(ta, tb, M, N, K, (const float *) dA, ta ? M : K, (const float *) dB, tb ? K : N, C, N, st);
	}
	if(!rc)
		rc = sphr_rt_check(hipStreamSynchronize(st), "synchronize");
done:
	hipFree(dA);
	hipFree(dB);
	return rc;
}
