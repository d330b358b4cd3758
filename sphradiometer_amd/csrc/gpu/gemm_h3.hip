/*
 * Split-FP16 GEMM on gfx950 MFMA (v_mfma_f32_16x16x32_f16), the mixed
 * precision form of the localization products:
 *
 *   C[i][j] = ira[i] icb[j] sum_k (Ah[i][k] Bh[j][k] + Ah[i][k] Bl[j][k] + Al[i][k] Bh[j][k])
 *
 * where a row of A scaled by ra[i] = 1/ira[i] (a power of two) is split into
 * fp16 parts a s = ah + al (ah = fp16(a s), al = fp16(a s - ah)) and a column
 * of B likewise with its own power of two.  The three products keep ~22
 * significant bits of each operand (the dropped al bl term is 2^-22 of the
 * product); the scales put every row's and column's largest element at
 * 2^13..2^14, far from fp16's overflow, and keep the low parts normal
 * down to 2^-17 of it.  Accumulation is float32 in the MFMA.
 *
 * Operands are K-contiguous: A [M][lda], B [N][ldb] (B transposed), K a
 * multiple of 64 (producers zero-pad), lda and ldb multiples of 8.  A
 * second form reads A as float32 [M][lda] (lda even, K columns, no
 * padding) with the inverse row scales given, and splits it while staging
 * it into LDS: no split copy of A goes through HBM.
 *
 * Tiling: 256 threads = 4 waves in 2 x 2, block tile 128 x 128, wave tile
 * 64 x 64 = 4 x 4 MFMA tiles; K chunks of 64 staged global -> registers
 * (the next chunk's loads in flight during this chunk's MFMAs) -> LDS
 * (rows padded to 72 halves: the 16-byte fragment reads of 16 rows land
 * on disjoint banks).  Workgroup ids are remapped so each of the 8 XCDs
 * works on a contiguous range of tiles (its own L2), in groups of 8 row
 * panels: the 64 workgroups in flight on an XCD cover 8 x 8 tiles, so
 * each operand panel is read from HBM once per 8 tiles of the other
 * operand (column-panel order re-read the split A, 115 MB for the
 * localization sky product, once per column panel: 82 times).
 *
 * MFMA 16x16x32 f16 operand layout (64 lanes): lane i + 16 q holds
 * A[i][8q .. 8q+7] and B[8q .. 8q+7][i]; accumulator element e of lane
 * j + 16 q holds D[4q + e][j].
 */
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

#define H3_BM 128
#define H3_BN 128
#define H3_BK 64
#define H3_LD 72	/* LDS row stride in halves */
#define H3_NT 256
#define H3_GM 8		/* row panels per tile group */

/* AF32: A is float32 (Af, K columns, split while staged, inverse row
 * scales ira); else the fp16 pair (Ah, Al).  NFAST: the tiles of one row
 * panel run next to each other on one XCD (A read once from HBM when N
 * spans few tiles); else those of one column panel (B reused) */
template <typename TO, bool AF32, bool NFAST>
__global__ void __launch_bounds__(H3_NT) k_gemm_h3(int M, int N, int K, const _Float16 *__restrict__ Ah, const _Float16 *__restrict__ Al, const float *__restrict__ Af, long lda, const float *__restrict__ ira,
						  const _Float16 *__restrict__ Bh, const _Float16 *__restrict__ Bl, long ldb, const float *__restrict__ icb, TO *__restrict__ C, long ldc, int tiles_m, int tiles_n)
{
	__shared__ __align__(16) _Float16 sA[2][H3_BM * H3_LD];	/* [hi/lo][row][k] */
	__shared__ __align__(16) _Float16 sB[2][H3_BN * H3_LD];	/* [hi/lo][col][k] */
	const int nwg = tiles_m * tiles_n;
	const int id = blockIdx.x;
	int t = id;
	if(nwg % 8 == 0)
		t = (id % 8) * (nwg / 8) + id / 8;
	/* NFAST: the tiles of a row panel side by side; else groups of
	 * H3_GM row panels side by side, all column tiles of a group before
	 * the next group, so the workgroups in flight on an XCD share both
	 * operands' panels in its L2 (A is re-read once per H3_GM x 8
	 * tiles, not once per column panel) */
	int tm, tn;
	if(NFAST) {
		tm = t / tiles_n;
		tn = t % tiles_n;
	} else {
		const int gid = t / (H3_GM * tiles_n), first = gid * H3_GM;
		const int gs = tiles_m - first < H3_GM ? tiles_m - first : H3_GM;
		const int w = t - gid * H3_GM * tiles_n;
		tm = first + w % gs;
		tn = w / gs;
	}
	const int m0 = tm * H3_BM, n0 = tn * H3_BN;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
	const int li = lane & 15, lq = lane >> 4;
	/* global -> register staging: each thread moves 4 rows x 16 bytes of
	 * each of the four operand tiles per chunk (128 rows x 8 vectors); a
	 * float32 A as 4 rows x 8 values, split when stored */
	h8 ra[2][4], rb[2][4];
	float2 rf[AF32 ? 4 : 1][4];
	float rs[AF32 ? 4 : 1];
	const int vr = threadIdx.x >> 3, vc = (threadIdx.x & 7) * 8;	/* row 0..31 (+32 per step), k offset */
	if constexpr(AF32)
#pragma unroll
		for(int s = 0; s < 4; s++) {
			const int gm = m0 + vr + 32 * s;
			rs[s] = gm < M ? 1.0f / ira[gm] : 0.f;
		}
	auto load = [&](int k0) {
#pragma unroll
		for(int s = 0; s < 4; s++) {
			const int r = vr + 32 * s;
			const int gm = m0 + r, gn = n0 + r;
			if constexpr(AF32) {
				const float *a = Af + (size_t) gm * lda + k0 + vc;
#pragma unroll
				for(int u = 0; u < 4; u++)
					rf[s][u] = gm < M && k0 + vc + 2 * u < K ? *reinterpret_cast<const float2 *>(a + 2 * u) : make_float2(0.f, 0.f);
			} else if(gm < M) {
				ra[0][s] = *reinterpret_cast<const h8 *>(Ah + (size_t) gm * lda + k0 + vc);
				ra[1][s] = *reinterpret_cast<const h8 *>(Al + (size_t) gm * lda + k0 + vc);
			} else {
				ra[0][s] = (h8) 0;
				ra[1][s] = (h8) 0;
			}
			if(gn < N) {
				rb[0][s] = *reinterpret_cast<const h8 *>(Bh + (size_t) gn * ldb + k0 + vc);
				rb[1][s] = *reinterpret_cast<const h8 *>(Bl + (size_t) gn * ldb + k0 + vc);
			} else {
				rb[0][s] = (h8) 0;
				rb[1][s] = (h8) 0;
			}
		}
	};
	auto store = [&]() {
		if constexpr(AF32)
#pragma unroll
			for(int s = 0; s < 4; s++)
#pragma unroll
				for(int u = 0; u < 4; u++) {
					const float v0 = rf[s][u].x * rs[s], v1 = rf[s][u].y * rs[s];
					const _Float16 h0 = (_Float16) v0, h1 = (_Float16) v1;
					ra[0][s][2 * u] = h0;
					ra[0][s][2 * u + 1] = h1;
					ra[1][s][2 * u] = (_Float16) (v0 - (float) h0);
					ra[1][s][2 * u + 1] = (_Float16) (v1 - (float) h1);
				}
#pragma unroll
		for(int p = 0; p < 2; p++)
#pragma unroll
			for(int s = 0; s < 4; s++) {
				const int r = vr + 32 * s;
				*reinterpret_cast<h8 *>(&sA[p][r * H3_LD + vc]) = ra[p][s];
				*reinterpret_cast<h8 *>(&sB[p][r * H3_LD + vc]) = rb[p][s];
			}
	};
	f4 acc[4][4];
#pragma unroll
	for(int a = 0; a < 4; a++)
#pragma unroll
		for(int b = 0; b < 4; b++)
			acc[a][b] = (f4) 0;
	load(0);
	for(int k0 = 0; k0 < K; k0 += H3_BK) {
		store();
		__syncthreads();
		if(k0 + H3_BK < K)
			load(k0 + H3_BK);
#pragma unroll
		for(int kk = 0; kk < H3_BK; kk += 32) {
			h8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
			for(int a = 0; a < 4; a++) {
				const int r = (wm + 16 * a + li) * H3_LD + kk + 8 * lq;
				ah[a] = *reinterpret_cast<const h8 *>(&sA[0][r]);
				al[a] = *reinterpret_cast<const h8 *>(&sA[1][r]);
			}
#pragma unroll
			for(int b = 0; b < 4; b++) {
				const int r = (wn + 16 * b + li) * H3_LD + kk + 8 * lq;
				bh[b] = *reinterpret_cast<const h8 *>(&sB[0][r]);
				bl[b] = *reinterpret_cast<const h8 *>(&sB[1][r]);
			}
#pragma unroll
			for(int a = 0; a < 4; a++)
#pragma unroll
				for(int b = 0; b < 4; b++) {
					acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
					acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
					acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh[b], acc[a][b], 0, 0, 0);
				}
		}
		__syncthreads();
	}
#pragma unroll
	for(int b = 0; b < 4; b++) {
		const int gn = n0 + wn + 16 * b + li;
		if(gn >= N)
			continue;
		const float cs = icb[gn];
#pragma unroll
		for(int a = 0; a < 4; a++)
#pragma unroll
			for(int e = 0; e < 4; e++) {
				const int gm = m0 + wm + 16 * a + 4 * lq + e;
				if(gm < M)
					C[(size_t) gm * ldc + gn] = (TO) (acc[a][b][e] * (ira[gm] * cs));
			}
	}
}

/*
 * The phi DFT of the meshes with the half-turn split (float32 A, split
 * while staged, as the AF32 form above).  A mesh has n points
 * phi_k = 2 pi k / n, so exp(i m phi_{k + n/2}) = (-1)^m exp(i m phi_k):
 * with Ev and Od the sums over the even and the odd orders,
 *     map[r][k] = Ev[r][k] + Od[r][k],  map[r][k + n/2] = Ev[r][k] - Od[r][k],
 * k < n/2.  A row of A holds the ring sums of orders m = 0..L, -L..-1 as
 * (re, im) pairs; the loader gathers one parity class at a time: class
 * index i < n0 at column c0 + 2 i, the rest at c1 + 2 (i - n0) (the
 * orders m >= 0 of the class ascending, then m < 0).  Be and Bo [NH][ldb]
 * are the two half-turn DFT matrices (NH = n real columns: n/2 points x
 * re/im) over the classes in that order, transposed and split per column.
 *
 * One workgroup per 64-row panel computes both sums over all NH (<= 128)
 * columns and writes both halves of each map row: half the MFMA work of
 * the full [2 nm] x [2 n] product.  4 waves in 2 x 2, wave tile 32 x 64,
 * so the two classes' accumulators take 64 registers and two workgroups
 * share a CU (one's map stores overlap the other's loads and MFMAs).  The
 * K chunks alternate between the classes (even 0, odd 0, even 1, ..):
 * chunk c of either class reads the same stretch of each A row, so the
 * second read finds it in cache.
 */
#define PH_BM 64
#define PH_NT 256
/* LDS: the operand tiles, or the map rows of a panel at NH = 128 */
#define PH_SMEM (PH_BM * 2 * 128 * 4 > 2 * (PH_BM + H3_BN) * H3_LD * 2 ? PH_BM * 2 * 128 * 4 : 2 * (PH_BM + H3_BN) * H3_LD * 2)

struct phi_cls {
	int K, c0, n0, c1;	/* real K of the class; its column map */
};

__global__ void __launch_bounds__(PH_NT, 2) k_phi_h3(int M, int NH, struct phi_cls ce, struct phi_cls co, const float *__restrict__ Af, long lda, const float *__restrict__ ira,
						     const _Float16 *__restrict__ Beh, const _Float16 *__restrict__ Bel, const float *__restrict__ icbe,
						     const _Float16 *__restrict__ Boh, const _Float16 *__restrict__ Bol, const float *__restrict__ icbo, long ldb,
						     float *__restrict__ C, int tiles_m)
{
	/* the operand tiles, then the panel's map rows [PH_BM][2 NH] */
	__shared__ __align__(16) char smem[PH_SMEM];
	_Float16 (*sA)[PH_BM * H3_LD] = reinterpret_cast<_Float16 (*)[PH_BM * H3_LD]>(smem);	/* [hi/lo][row][k] */
	_Float16 (*sB)[H3_BN * H3_LD] = reinterpret_cast<_Float16 (*)[H3_BN * H3_LD]>(smem + 2 * PH_BM * H3_LD * sizeof(_Float16));	/* [hi/lo][col][k] */
	float *sC = reinterpret_cast<float *>(smem);
	const int id = blockIdx.x;
	int t = id;
	if(tiles_m % 8 == 0)
		t = (id % 8) * (tiles_m / 8) + id / 8;
	const int m0 = t * PH_BM;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int wm = (wave & 1) * 32, wn = (wave >> 1) * 64;
	const int li = lane & 15, lq = lane >> 4;
	h8 rb[2][4];
	float2 rfe[2][4], rfo[2][4];	/* the next chunks of A of either class */
	float rs[2];
	const int vr = threadIdx.x >> 3, vc = (threadIdx.x & 7) * 8;	/* row 0..31 (+32 per step), k offset */
#pragma unroll
	for(int q = 0; q < 2; q++) {
		const int gm = m0 + vr + 32 * q;
		rs[q] = gm < M ? 1.0f / ira[gm] : 0.f;
	}
	/* A at k0 of a parity class: its complex index i = (k0 + vc) / 2 + u
	 * from column c0 + 2 i or c1 + 2 (i - n0).  Every load is issued,
	 * from a clamped address, and the values out of range are dropped
	 * when the chunk is stored (bit u of km: column u in range; rs: zero
	 * for rows past M): with no branch around a load, the wait for one
	 * chunk leaves the loads issued after it in flight */
	auto load_a = [&](float2 (&rf)[2][4], int &km, const struct phi_cls &c, int k0) {
		int col[4];
		km = 0;
#pragma unroll
		for(int u = 0; u < 4; u++) {
			const int i = (k0 + vc) / 2 + u;
			const bool ok = 2 * i < c.K;
			km |= (int) ok << u;
			const int cp = c.c0 + 2 * i, cn = c.c1 + 2 * (i - c.n0);
			col[u] = ok ? (i < c.n0 ? cp : cn) : 0;
		}
#pragma unroll
		for(int q = 0; q < 2; q++) {
			const int gm = m0 + vr + 32 * q;
			const float *a = Af + (size_t) (gm < M ? gm : M - 1) * lda;
#pragma unroll
			for(int u = 0; u < 4; u++)
				rf[q][u] = *reinterpret_cast<const float2 *>(a + 2 * col[u]);
		}
	};
	/* B rows past NH (clamped to row 0 here) are stored as zeros */
	auto load_b = [&](int k0, const _Float16 *__restrict__ Bh, const _Float16 *__restrict__ Bl) {
#pragma unroll
		for(int q = 0; q < 4; q++) {
			const int r = vr + 32 * q;
			const size_t o = (size_t) (r < NH ? r : 0) * ldb + k0 + vc;
			rb[0][q] = *reinterpret_cast<const h8 *>(Bh + o);
			rb[1][q] = *reinterpret_cast<const h8 *>(Bl + o);
		}
	};
	auto store = [&](const float2 (&rf)[2][4], int km) {
#pragma unroll
		for(int q = 0; q < 2; q++) {
			h8 hi, lo;
#pragma unroll
			for(int u = 0; u < 4; u++) {
				const float sc = (km >> u) & 1 ? rs[q] : 0.f;
				const float v0 = rf[q][u].x * sc, v1 = rf[q][u].y * sc;
				const _Float16 h0 = (_Float16) v0, h1 = (_Float16) v1;
				hi[2 * u] = h0;
				hi[2 * u + 1] = h1;
				lo[2 * u] = (_Float16) (v0 - (float) h0);
				lo[2 * u + 1] = (_Float16) (v1 - (float) h1);
			}
			const int r = vr + 32 * q;
			*reinterpret_cast<h8 *>(&sA[0][r * H3_LD + vc]) = hi;
			*reinterpret_cast<h8 *>(&sA[1][r * H3_LD + vc]) = lo;
		}
#pragma unroll
		for(int p = 0; p < 2; p++)
#pragma unroll
			for(int q = 0; q < 4; q++) {
				const int r = vr + 32 * q;
				*reinterpret_cast<h8 *>(&sB[p][r * H3_LD + vc]) = r < NH ? rb[p][q] : (h8) 0;
			}
	};
	auto mma = [&](f4 (&acc)[2][4]) {
#pragma unroll
		for(int kk = 0; kk < H3_BK; kk += 32) {
			h8 ah[2], al[2], bh[4], bl[4];
#pragma unroll
			for(int a = 0; a < 2; a++) {
				const int r = (wm + 16 * a + li) * H3_LD + kk + 8 * lq;
				ah[a] = *reinterpret_cast<const h8 *>(&sA[0][r]);
				al[a] = *reinterpret_cast<const h8 *>(&sA[1][r]);
			}
#pragma unroll
			for(int b = 0; b < 4; b++) {
				const int r = (wn + 16 * b + li) * H3_LD + kk + 8 * lq;
				bh[b] = *reinterpret_cast<const h8 *>(&sB[0][r]);
				bl[b] = *reinterpret_cast<const h8 *>(&sB[1][r]);
			}
#pragma unroll
			for(int a = 0; a < 2; a++)
#pragma unroll
				for(int b = 0; b < 4; b++) {
					acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
					acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
					acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh[b], acc[a][b], 0, 0, 0);
				}
		}
	};
	f4 ae[2][4], ao[2][4];
#pragma unroll
	for(int a = 0; a < 2; a++)
#pragma unroll
		for(int b = 0; b < 4; b++) {
			ae[a][b] = (f4) 0;
			ao[a][b] = (f4) 0;
		}
	/* chunks even 0, odd 0, even 1, ..; nc chunks of each class (the
	 * odd class's last one may be all padding).  A is loaded two chunks
	 * ahead (one chunk of each class in flight), B one chunk ahead from
	 * L2; each chunk's B loads go out before the next A chunk's, so the
	 * wait for a chunk leaves the later A loads in flight.  The last pass
	 * loads its own chunk again instead of branching around the loads */
	const int ne = (ce.K + H3_BK - 1) / H3_BK, no = (co.K + H3_BK - 1) / H3_BK;
	const int nc = ne > no ? ne : no;
	int kme, kmo;
	load_b(0, Beh, Bel);
	load_a(rfe, kme, ce, 0);
	load_a(rfo, kmo, co, 0);
	for(int c = 0; c < nc; c++) {
		const int kn = (c + 1 < nc ? c + 1 : c) * H3_BK;
		store(rfe, kme);
		__syncthreads();
		load_b(c * H3_BK, Boh, Bol);
		load_a(rfe, kme, ce, kn);
		mma(ae);
		__syncthreads();
		store(rfo, kmo);
		__syncthreads();
		load_b(kn, Beh, Bel);
		load_a(rfo, kmo, co, kn);
		mma(ao);
		__syncthreads();
	}
	/* the map rows into LDS (the loop's last barrier has freed the
	 * operand tiles), then out as one contiguous block: C rows are 2 NH
	 * floats with no gap, so the panel's rows m0 .. m0 + 63 are
	 * 64 x 8 NH bytes in a row, written whole 16 bytes per lane (every
	 * cache line of the block written once, by one workgroup) */
	const int W = 2 * NH;
#pragma unroll
	for(int b = 0; b < 4; b++) {
		const int gn = wn + 16 * b + li;
		if(gn >= NH)
			continue;
		const float se = icbe[gn], so = icbo[gn];
#pragma unroll
		for(int a = 0; a < 2; a++)
#pragma unroll
			for(int e = 0; e < 4; e++) {
				const int r = wm + 16 * a + 4 * lq + e;
				const float ir = ira[m0 + r < M ? m0 + r : M - 1];
				const float ev = ae[a][b][e] * (ir * se), od = ao[a][b][e] * (ir * so);
				sC[r * W + gn] = ev + od;
				sC[r * W + NH + gn] = ev - od;
			}
	}
	__syncthreads();
	const int rv = M - m0 < PH_BM ? M - m0 : PH_BM;
	const int n4 = rv * W / 4;
	float4 *dst = reinterpret_cast<float4 *>(C + (size_t) m0 * W);
	const float4 *src = reinterpret_cast<const float4 *>(sC);
	for(int j = threadIdx.x; j < n4; j += PH_NT)
		dst[j] = src[j];
}

/*
 * Row split of a real matrix X [rows][K] (float or double, leading dimension
 * ldx) into the fp16 pair and the inverse row scale, one workgroup per row;
 * columns K .. Kp-1 of the outputs are zero.
 */
template <typename TI>
__global__ void __launch_bounds__(256) k_split_rows(int rows, int K, int Kp, const TI *__restrict__ X, long ldx, _Float16 *__restrict__ Xh, _Float16 *__restrict__ Xl, long ldh, float *__restrict__ irs)
{
	__shared__ float red[4];
	const int r = blockIdx.x;
	if(r >= rows)
		return;
	const TI *x = X + (size_t) r * ldx;
	float mx = 0.f;
	for(int k = threadIdx.x; k < K; k += 256)
		mx = fmaxf(mx, fabsf((float) x[k]));
	for(int o = 32; o > 0; o >>= 1)
		mx = fmaxf(mx, __shfl_xor(mx, o));
	if((threadIdx.x & 63) == 0)
		red[threadIdx.x >> 6] = mx;
	__syncthreads();
	mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
	int ex = 0;
	if(mx > 0.f) {
		frexpf(mx, &ex);	/* mx = f 2^ex, f in [0.5, 1) */
		ex = 14 - ex;		/* mx 2^ex in [2^13, 2^14) */
	}
	const float s = ldexpf(1.0f, ex);
	for(int k = threadIdx.x; k < Kp; k += 256) {
		const float v = k < K ? (float) x[k] * s : 0.f;
		const _Float16 h = (_Float16) v;
		Xh[(size_t) r * ldh + k] = h;
		Xl[(size_t) r * ldh + k] = (_Float16) (v - (float) h);
	}
	if(threadIdx.x == 0)
		irs[r] = ldexpf(1.0f, -ex);
}

}	/* namespace */

template <typename TO>
int launch_gemm_h3(int M, int N, int K, const _Float16 *Ah, const _Float16 *Al, long lda, const float *ira, const _Float16 *Bh, const _Float16 *Bl, long ldb, const float *icb, TO *C, long ldc, hipStream_t st)
{
	if(M <= 0 || N <= 0 || K <= 0)
		return 0;
	if(K % H3_BK || lda % 8 || ldb % 8 || lda < K || ldb < K)
		return -1;
	const int tm = (M + H3_BM - 1) / H3_BM, tn = (N + H3_BN - 1) / H3_BN;
	hipLaunchKernelGGL((k_gemm_h3<TO, false, false>), dim3(tm * tn), dim3(H3_NT), 0, st, M, N, K, Ah, Al, (const float *) NULL, lda, ira, Bh, Bl, ldb, icb, C, ldc, tm, tn);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename TO>
int launch_gemm_h3a(int M, int N, int K, const float *A, long lda, const float *ira, const _Float16 *Bh, const _Float16 *Bl, long ldb, const float *icb, TO *C, long ldc, hipStream_t st)
{
	if(M <= 0 || N <= 0 || K <= 0)
		return 0;
	/* the chunks run to K rounded up to 64, where B's padding is zero; a
	 * float2 of A never straddles K */
	const int Kp = (K + H3_BK - 1) / H3_BK * H3_BK;
	if(K % 2 || lda % 2 || ldb % 8 || lda < K || ldb < Kp)
		return -1;
	const int tm = (M + H3_BM - 1) / H3_BM, tn = (N + H3_BN - 1) / H3_BN;
	hipLaunchKernelGGL((k_gemm_h3<TO, true, true>), dim3(tm * tn), dim3(H3_NT), 0, st, M, N, K, (const _Float16 *) NULL, (const _Float16 *) NULL, A, lda, ira, Bh, Bl, ldb, icb, C, ldc, tm, tn);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_phi_h3(int M, int L, const float *A, long lda, const float *ira, const _Float16 *Beh, const _Float16 *Bel, const float *icbe, const _Float16 *Boh, const _Float16 *Bol, const float *icbo, long ldb, float *C, hipStream_t st)
{
	if(M <= 0)
		return 0;
	const int nm = 2 * L + 1, NH = 2 * (L + 1), ev = L / 2 * 2, om = (L & 1) ? L : L - 1;
	/* even orders: 0, 2, .., ev at columns 0, 2, ..; -ev, .., -2 at
	 * nm - ev, ..; odd orders: 1, 3, .., om at 1, 3, ..; -om, .., -1 at
	 * nm - om, .. */
	const struct phi_cls ce = { 2 * (ev + 1), 0, ev / 2 + 1, nm - ev };
	const struct phi_cls co = { 2 * (nm - ev - 1), 1, (om + 1) / 2, nm - om };
	/* both classes run max(ne, no) chunks: B rows hold that many */
	const int ne = (ce.K + H3_BK - 1) / H3_BK, no = (co.K + H3_BK - 1) / H3_BK;
	if(L < 1 || NH > H3_BN || lda < 2 * nm || ldb % 8 || ldb < (ne > no ? ne : no) * H3_BK)
		return -1;
	const int tm = (M + PH_BM - 1) / PH_BM;
	hipLaunchKernelGGL(k_phi_h3, dim3(tm), dim3(PH_NT), 0, st, M, NH, ce, co, A, lda, ira, Beh, Bel, icbe, Boh, Bol, icbo, ldb, C, tm);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename TI>
int launch_split_rows(int rows, int K, int Kp, const TI *X, long ldx, _Float16 *Xh, _Float16 *Xl, long ldh, float *irs, hipStream_t st)
{
	if(rows <= 0)
		return 0;
	if(Kp < K || ldh < Kp)
		return -1;
	hipLaunchKernelGGL((k_split_rows<TI>), dim3(rows), dim3(256), 0, st, rows, K, Kp, X, ldx, Xh, Xl, ldh, irs);
	return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_gemm_h3<float>(int, int, int, const _Float16 *, const _Float16 *, long, const float *, const _Float16 *, const _Float16 *, long, const float *, float *, long, hipStream_t);
template int launch_gemm_h3<double>(int, int, int, const _Float16 *, const _Float16 *, long, const float *, const _Float16 *, const _Float16 *, long, const float *, double *, long, hipStream_t);
template int launch_gemm_h3a<float>(int, int, int, const float *, long, const float *, const _Float16 *, const _Float16 *, long, const float *, float *, long, hipStream_t);
template int launch_split_rows<float>(int, int, int, const float *, long, _Float16 *, _Float16 *, long, float *, hipStream_t);
template int launch_split_rows<double>(int, int, int, const double *, long, _Float16 *, _Float16 *, long, float *, hipStream_t);

/*
 * Test hook: C [M][N] (float32) = A [M][K] B^T with A and B [N][K] given
 * as float32 device arrays, through the split-FP16 path (rows of A and of
 * B split with their own scales, K zero-padded to a multiple of 64); with
 * a_f32, A goes in as float32 (copied to an even row length) and is split
 * while staged, its row scales from the row split.  Allocates its
 * buffers; synchronous.
 */
extern "C" int sphr_gemm_h3_check(int a_f32, int M, int N, int K, const float *A, const float *B, float *C)
{
	if(M <= 0 || N <= 0 || K <= 0 || !A || !B || !C)
		return -1;
	const int Kp = (K + H3_BK - 1) / H3_BK * H3_BK, Ke = (K + 1) / 2 * 2;
	_Float16 *ah = NULL, *al = NULL, *bh = NULL, *bl = NULL;
	float *ra = NULL, *cb = NULL, *ae = NULL;
	int rc = -1;
	if(hipMalloc((void **) &ah, (size_t) M * Kp * 2) != hipSuccess || hipMalloc((void **) &al, (size_t) M * Kp * 2) != hipSuccess
	   || hipMalloc((void **) &bh, (size_t) N * Kp * 2) != hipSuccess || hipMalloc((void **) &bl, (size_t) N * Kp * 2) != hipSuccess
	   || hipMalloc((void **) &ra, (size_t) M * 4) != hipSuccess || hipMalloc((void **) &cb, (size_t) N * 4) != hipSuccess
	   || hipMalloc((void **) &ae, (size_t) M * Ke * 4) != hipSuccess)
		goto done;
	if(launch_split_rows<float>(M, K, Kp, A, K, ah, al, Kp, ra, 0) || launch_split_rows<float>(N, K, Kp, B, K, bh, bl, Kp, cb, 0))
		goto done;
	if(a_f32) {
		if(hipMemset(ae, 0, (size_t) M * Ke * 4) != hipSuccess
		   || hipMemcpy2D(ae, (size_t) Ke * 4, A, (size_t) K * 4, (size_t) K * 4, M, hipMemcpyDeviceToDevice) != hipSuccess
		   || launch_gemm_h3a<float>(M, N, Ke, ae, Ke, ra, bh, bl, Kp, cb, C, N, 0))
			goto done;
	} else if(launch_gemm_h3<float>(M, N, Kp, ah, al, Kp, ra, bh, bl, Kp, cb, C, N, 0))
		goto done;
	if(hipDeviceSynchronize() != hipSuccess)
		goto done;
	rc = 0;
done:
	hipFree(ah);
	hipFree(al);
	hipFree(bh);
	hipFree(bl);
	hipFree(ra);
	hipFree(cb);
	hipFree(ae);
	return rc;
}
