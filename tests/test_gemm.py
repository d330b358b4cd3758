"""The split-FP16 product (gemm_h3.hip) against a float64 numpy product of
the same float32 operands.  Rows of A and of B are split into fp16 pairs
with their own power-of-two scales, so the error bound is relative to each
row's and column's own size: |C - C64|[i][j] <= 3e-6 |A_i| |B_j|.  Rows and
columns are given magnitudes spread over 12 decades, the shape has edges
on every side of the 128 x 128 tiles, K is not a multiple of the 64-wide
chunks, and one row and one column are zero."""
import numpy as np
import pytest

import sphradiometer_amd as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    S.lib()
    if not S.gpu_available():
        pytest.fail("no GPU: " + S.last_error())


@pytest.mark.parametrize("a_f32", [0, 1])
@pytest.mark.parametrize("M,N,K", [(200, 150, 300), (128, 128, 64), (1, 17, 5), (333, 260, 1000)])
def test_gemm_h3_matches_float64(M, N, K, a_f32):
    # a_f32: A staged as float32 and split in the loader (the localization
    # meshes' form), with the row scales of the separate split
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K)) * 10.0 ** rng.uniform(-6, 6, (M, 1))
    B = rng.standard_normal((N, K)) * 10.0 ** rng.uniform(-6, 6, (N, 1))
    A[M // 2] = 0.0
    B[N // 3] = 0.0
    A32, B32 = A.astype(np.float32), B.astype(np.float32)
    dA, dB, dC = S.DeviceBuffer.from_numpy(A32), S.DeviceBuffer.from_numpy(B32), S.DeviceBuffer(M * N * 4)
    print("gemm_h3", a_f32, M, N, K, flush=True)
    assert S.lib().sphr_gemm_h3_check(a_f32, M, N, K, dA.ptr, dB.ptr, dC.ptr) == 0, S.last_error()
    C = dC.to_numpy(np.float32, (M, N)).astype(np.float64)
    ref = A32.astype(np.float64) @ B32.astype(np.float64).T
    bound = np.linalg.norm(A32.astype(np.float64), axis=1)[:, None] * np.linalg.norm(B32.astype(np.float64), axis=1)[None, :]
    err = np.abs(C - ref)
    assert np.all(err <= 3e-6 * bound), float(np.max(err / np.where(bound > 0, bound, 1.0)))
    assert np.all(C[M // 2] == 0.0) and np.all(C[:, N // 3] == 0.0)


@pytest.mark.parametrize("f64", [0, 1])
@pytest.mark.parametrize("M,N,K", [(200, 150, 300), (129, 257, 33), (1, 5, 1)])
def test_library_gemm_matches_numpy(f64, M, N, K):
    """the library's MFMA GEMM (gemm.hip), the one that replaces the
    vendor library in the float64 products and the phi transform"""
    rng = np.random.default_rng(7 * M + N)
    dt = np.float64 if f64 else np.float32
    A = rng.standard_normal((M, K)).astype(dt)
    B = rng.standard_normal((K, N)).astype(dt)
    dA, dB, dC = S.DeviceBuffer.from_numpy(A), S.DeviceBuffer.from_numpy(B), S.DeviceBuffer(M * N * A.itemsize)
    print("gemm f64=%d" % f64, M, N, K, flush=True)
    assert S.lib().sphr_gemm_run(0, f64, M, N, K, dA.ptr, K, dB.ptr, N, dC.ptr, N, None) == 0
    S.synchronize()
    C = dC.to_numpy(dt, (M, N)).astype(np.float64)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    tol = 1e-13 if f64 else 1e-5
    assert np.linalg.norm(C - ref) <= tol * np.linalg.norm(ref)
