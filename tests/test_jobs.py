"""The uneven batch (sphr_jobs_*): every job has its own network, length and
order, and its plan is generated on the GPU in the lag-domain form.

CPU pin: the lag-domain identity itself.  From the oracle's tap projections
(projection.c:99-131) and its frequency-domain delay product
(correlator.c:424-519), power_1d computed as (1/N) sum_f dp[f] X[f] equals
sum_u c[u] M[u] with c the circular cross-correlation and M built from the
taps on a product Gauss-Legendre rule; a transposed lag orientation fails.

GPU: each job's a_lm and map against the oracle's own plan + segment
pipeline (1e-12 in float64, 1e-5 relative in float32), graph replays and
the tile autotuner.
"""
import math

import numpy as np
import pytest

import sphradiometer_amd as S
from sphradiometer_amd import inject

DT = 1.0 / 4096


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_lag_domain_identity(oracle):
    N, L = 512, 16
    net = oracle.Network(inject.positions("AB"), N, DT, L)
    g = net.geometry(0)
    D = float(np.linalg.norm(g[:3]))
    pr = oracle.params(DT, D)
    a, n_p = pr["a"], pr["n_p"]
    Lt = min(L, pr["P"])
    ta = oracle.projection_matrix_delay(n_p, a, [0, 0, +D / 2], DT).real
    nq = (2 * a + Lt + 2) // 2 + 1
    x, w = oracle.gl_nodes(nq)
    Pl = np.array([oracle.sphPlm_array(max(a, Lt), 0, xx) for xx in x])
    T = ta @ Pl[:, :a + 1].T                      # [tap][q]
    Tm = T[:, ::-1]                               # the other phase centre: T(-x)
    lags = np.arange(-(n_p - 1), n_p)
    K = np.array([[np.dot(T[max(0, -u):n_p - max(0, u), q], Tm[max(0, u):n_p - max(0, -u), q]) for q in range(nq)] for u in lags])
    M = (K * (2 * math.pi * w)[None, :]) @ Pl[:, :Lt + 1] / (N * N)
    rng = np.random.default_rng(0)
    xa, xb = rng.standard_normal(N), rng.standard_normal(N)
    X = np.fft.fft(xb) * np.conj(np.fft.fft(xa))
    p_fd = (net.delay_product(0) * X[:, None]).sum(axis=0) / N
    c = N * np.fft.ifft(X)
    assert rel(c[lags % N].real @ M, p_fd) < 1e-12
    assert rel(c[(-lags) % N].real @ M, p_fd) > 0.1


# ------------------------------------------------------------------- GPU


def small_jobs():
    rng = np.random.default_rng(12)
    r = inject.EARTH_RADIUS_SPHERE / inject.C_SI
    jobs = []
    for j, (D, N, L) in enumerate([(2, 512, 16), (3, 1024, 24), (2, 1024, 20), (3, 512, 18), (4, 512, 16)]):
        v = rng.standard_normal((D, 3))
        jobs.append({"positions": r * v / np.linalg.norm(v, axis=1, keepdims=True), "N": N, "l_max": L})
    data = [rng.standard_normal((j["positions"].shape[0], j["N"])).astype(np.float32) for j in jobs]
    return jobs, data


@pytest.mark.gpu
def test_gpu_jobs_match_oracle(oracle):
    jobs, data = small_jobs()
    packed = np.concatenate([d.ravel() for d in data])
    d_in = S.DeviceBuffer.from_numpy(packed)
    want = []
    for j, d in zip(jobs, data):
        onet = oracle.Network(j["positions"], j["N"], DT, j["l_max"])
        want.append(onet.segment(d.astype(np.float64), j["l_max"]))
    for prec, cdt, tol in ((S.FP64, np.complex128, 1e-12), (S.MIXED, np.complex64, 1e-5)):
        J = S.Jobs(jobs, DT, precision=prec)
        info = J.info()
        print("jobs", prec, info, flush=True)
        d_alm = S.DeviceBuffer(info["alm_elems"] * np.dtype(cdt).itemsize)
        d_map = S.DeviceBuffer(info["map_elems"] * np.dtype(cdt).itemsize)
        J.run(d_in.ptr, d_alm.ptr, d_map.ptr)
        S.synchronize()
        alm = d_alm.to_numpy(cdt, (info["alm_elems"],))
        mp = d_map.to_numpy(cdt, (info["map_elems"],))
        for k, (j, (sky, mesh)) in enumerate(zip(jobs, want)):
            L = j["l_max"]
            a = alm[J.off_alm[k]:J.off_alm[k] + (L + 1) ** 2]
            m = mp[J.off_map[k]:J.off_map[k] + (2 * L + 2) ** 2].reshape(2 * L + 2, 2 * L + 2)
            assert rel(a, sky) < tol, (prec, k, rel(a, sky))
            assert rel(m, mesh) < tol, (prec, k)
        # graph replay: a second run gives the same numbers
        J.run(d_in.ptr, d_alm.ptr, d_map.ptr)
        S.synchronize()
        assert np.array_equal(d_alm.to_numpy(cdt, (info["alm_elems"],)), alm)


@pytest.mark.gpu
def test_gpu_jobs_tune_and_uneven_workload():
    jobs, packed = inject.uneven_jobs(64, seed=3)
    J = S.Jobs(jobs, DT, precision=S.MIXED)
    info = J.info()
    d_in = S.DeviceBuffer.from_numpy(packed)
    d_alm = S.DeviceBuffer(info["alm_elems"] * 8)
    d_map = S.DeviceBuffer(info["map_elems"] * 8)
    J.run(d_in.ptr, d_alm.ptr, d_map.ptr)
    S.synchronize()
    ref = d_alm.to_numpy(np.complex64, (info["alm_elems"],))
    k, ms = J.tune(d_in.ptr, d_alm.ptr)
    print("tuned candidate", k, "%.3f ms" % ms, J.info(), flush=True)
    J.run(d_in.ptr, d_alm.ptr, d_map.ptr)
    S.synchronize()
    got = d_alm.to_numpy(np.complex64, (info["alm_elems"],))
    assert np.all(np.isfinite(got))
    assert rel(got, ref) < 1e-5


def top_jobs():
    """Jobs at the top of the uneven workload's range: five sites, 8192
    samples, l_max 64 (ten baselines, natural orders up to ~548); an
    antipodal pair (the longest baseline, |d| = 42.5 ms); a close pair
    whose natural order is below the target l_max, so its power stays at
    its own order."""
    rng = np.random.default_rng(31)
    r = inject.EARTH_RADIUS_SPHERE / inject.C_SI
    v5 = rng.standard_normal((5, 3))
    a = rng.standard_normal(3)
    anti = np.stack([a, -a])
    b = rng.standard_normal(3)
    b /= np.linalg.norm(b)
    t = np.cross(b, [0.0, 0.0, 1.0])
    close = np.stack([b, b + 0.02 * t / np.linalg.norm(t)])
    jobs = []
    for v, N, L in ((v5, 8192, 64), (anti, 8192, 64), (close, 512, 64)):
        jobs.append({"positions": r * v / np.linalg.norm(v, axis=1, keepdims=True), "N": N, "l_max": L})
    data = [rng.standard_normal((j["positions"].shape[0], j["N"])).astype(np.float32) for j in jobs]
    return jobs, data


@pytest.fixture(scope="module")
def top_case(oracle):
    jobs, data = top_jobs()
    want = []
    for j, d in zip(jobs, data):
        onet = oracle.Network(j["positions"], j["N"], DT, j["l_max"])
        want.append(onet.segment(d.astype(np.float64), j["l_max"]))
    # the close pair's natural order is below its target
    onet = oracle.Network(jobs[2]["positions"], jobs[2]["N"], DT)
    assert onet.info(0)["l_max"] < jobs[2]["l_max"]
    return jobs, data, want


def _check_jobs(J, jobs, want, d_in, cdt, tol):
    info = J.info()
    d_alm = S.DeviceBuffer(info["alm_elems"] * np.dtype(cdt).itemsize)
    d_map = S.DeviceBuffer(info["map_elems"] * np.dtype(cdt).itemsize)
    J.run(d_in.ptr, d_alm.ptr, d_map.ptr)
    S.synchronize()
    alm = d_alm.to_numpy(cdt, (info["alm_elems"],))
    mp = d_map.to_numpy(cdt, (info["map_elems"],))
    errs = []
    for k, (j, (sky, mesh)) in enumerate(zip(jobs, want)):
        L = j["l_max"]
        a = alm[J.off_alm[k]:J.off_alm[k] + (L + 1) ** 2]
        m = mp[J.off_map[k]:J.off_map[k] + (2 * L + 2) ** 2].reshape(2 * L + 2, 2 * L + 2)
        errs.append((rel(a, sky), rel(m, mesh)))
    print("errors", errs, flush=True)
    for ea, em in errs:
        assert ea < tol and em < tol, errs


# float64 agreement at the top of the range.  The GPU builds each plan in
# the lag domain (three float64 quadratures over ~550 Gauss-Legendre nodes
# and ~590 lags at |d| = 42.5 ms, natural order ~548), the oracle in the
# frequency domain (a pixel-domain product on a 1098-node mesh per bin,
# 8192 bins).  The two round-off paths agree to 2e-15 on the close pair
# (natural order below 64) and to 1.2e-12 (five sites) and 1.8e-12
# (antipodal pair) on the long baselines; the bar here is 4e-12.
TOL64_TOP = 4e-12


@pytest.mark.gpu
def test_gpu_jobs_top_of_range_match_oracle(top_case):
    jobs, data, want = top_case
    d_in = S.DeviceBuffer.from_numpy(np.concatenate([d.ravel() for d in data]))
    for prec, cdt, tol in ((S.FP64, np.complex128, TOL64_TOP), (S.MIXED, np.complex64, 1e-5)):
        print("top-of-range jobs", prec, flush=True)
        J = S.Jobs(jobs, DT, precision=prec)
        _check_jobs(J, jobs, want, d_in, cdt, tol)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [S.FP64, S.MIXED])
def test_gpu_jobs_tuned_tiles_match_oracle(top_case, prec):
    """every lag tile the tuner can keep gives the oracle's answer: the
    float64 lag kernel splits samples over idle lanes differently per
    tile"""
    jobs, data, want = top_case
    cdt, tol = (np.complex128, TOL64_TOP) if prec == S.FP64 else (np.complex64, 1e-5)
    d_in = S.DeviceBuffer.from_numpy(np.concatenate([d.ravel() for d in data]))
    J = S.Jobs(jobs, DT, precision=prec)
    info = J.info()
    d_alm = S.DeviceBuffer(info["alm_elems"] * np.dtype(cdt).itemsize)
    k, ms = J.tune(d_in.ptr, d_alm.ptr)
    print("tuned", prec, k, "%.3f ms" % ms, J.info(), flush=True)
    _check_jobs(J, jobs, want, d_in, cdt, tol)
    for sc, nt in ((512, 128), (2048, 128), (1024, 256)):
        J.set_lag_tile(sc, nt)
        print("lag tile", sc, nt, flush=True)
        _check_jobs(J, jobs, want, d_in, cdt, tol)
