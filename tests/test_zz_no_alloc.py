"""sphr_batch_run_static, the allocation-free run: its scratch lives in the
library's static device array instead of the batch object's buffers.
Through the C ABI, against sphr_batch_run on the same inputs (bit for bit:
the same kernels on the same data) and against the oracle, on two calls in
a row, the second with fewer segments, in the fused sky form (l_max 32)
and the two-kernel form whose ring sums also go to the static array
(l_max 20); a call too large for the array is refused before any launch.
"""
import ctypes

import numpy as np
import pytest

import sphradiometer_amd as S
from sphradiometer_amd import inject

pytestmark = pytest.mark.gpu

DT = 1.0 / 4096


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module", autouse=True)
def gpu():
    S.lib()
    if not S.gpu_available():
        pytest.fail("no GPU: " + S.last_error())


def _run(entry, b, d_ts, nseg, d_gmst, L, cdt):
    n = 2 * (L + 1)
    esz = np.dtype(cdt).itemsize
    d_alm = S.DeviceBuffer(nseg * (L + 1) ** 2 * esz)
    d_map = S.DeviceBuffer(nseg * n * n * esz)
    r = entry(ctypes.c_void_p(b.h), ctypes.c_void_p(d_ts.ptr), 0, nseg, None, ctypes.c_void_p(d_gmst.ptr),
              ctypes.c_void_p(d_alm.ptr), ctypes.c_void_p(d_map.ptr), None)
    assert r == 0, S.last_error()
    S.synchronize()
    return d_alm.to_numpy(cdt, (nseg, (L + 1) ** 2)), d_map.to_numpy(cdt, (nseg, n, n))


@pytest.mark.parametrize("lag", [True, False])
@pytest.mark.parametrize("L", [32, 20])
@pytest.mark.parametrize("prec", [S.FP64, S.MIXED])
def test_static_run_matches_regular_run(oracle, L, prec, lag):
    # both projection routes: the lag-domain kernel's one partial-sum row
    # per segment, and the spectra + K splits of the frequency-domain one
    lib = S.lib()
    cdt = np.complex128 if prec == S.FP64 else np.complex64
    pos = inject.positions("ABC")
    net = S.Network(pos, 1024, DT, L)
    b = S.Batch(net, L, prec, max_segments=64)
    b.set_lag(lag)
    ts, _ = inject.three_site(80, 1024, seed=5)
    rng = np.random.default_rng(L)
    gmst = rng.uniform(0, 2 * np.pi, 80)
    onet = oracle.Network(pos, 1024, DT, L)
    for first, nseg in ((0, 64), (64, 16)):
        print("static run L=%d prec=%d segments %d..%d" % (L, prec, first, first + nseg), flush=True)
        d_ts = S.DeviceBuffer.from_numpy(np.ascontiguousarray(ts[:, first:first + nseg]))
        d_g = S.DeviceBuffer.from_numpy(np.ascontiguousarray(gmst[first:first + nseg]))
        a_ref, m_ref = _run(lib.sphr_batch_run, b, d_ts, nseg, d_g, L, cdt)
        a_st, m_st = _run(lib.sphr_batch_run_static, b, d_ts, nseg, d_g, L, cdt)
        assert np.array_equal(a_st, a_ref) and np.array_equal(m_st, m_ref)
        s = nseg - 1
        want, mesh = onet.segment(ts[:, first + s], L, gmst=gmst[first + s])
        tol = 1e-12 if prec == S.FP64 else 1e-5
        assert rel(a_st[s], want) < tol
        assert rel(m_st[s], mesh) < tol


def test_static_run_refuses_more_than_its_array():
    lib = S.lib()
    net = S.Network(inject.positions("ABC"), 1024, DT, 32)
    b = S.Batch(net, 32, S.FP64, max_segments=16)
    # the frequency-domain route keeps spectra and K-split partial sums
    # (about 37 KB per segment here); the lag-domain route only its one
    # partial-sum row per segment (1.8 KB): each refused past 64 MiB
    for lag, nseg in ((False, 2048), (True, 40000)):
        b.set_lag(lag)
        d_ts = S.DeviceBuffer(3 * nseg * 1024 * 8)
        d_alm = S.DeviceBuffer(nseg * 33 * 33 * 16)
        r = lib.sphr_batch_run_static(ctypes.c_void_p(b.h), ctypes.c_void_p(d_ts.ptr), 0, nseg, None, None,
                                      ctypes.c_void_p(d_alm.ptr), None, None)
        assert r == -1
        assert "scratch" in S.last_error()
        del d_ts, d_alm
