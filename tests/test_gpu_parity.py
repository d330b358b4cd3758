"""GPU parity: the HIP path of the C library against the FP64 oracle.

Every call goes through the C ABI of libsphradiometer_amd.so.  Bars:
float64 paths match the oracle to 1e-12 relative L2 (the drop-in
tolerance); the SPHR_MIXED batched path matches the float64 path to
1e-5 relative L2 on a_lm.
"""
import math
import os

import numpy as np
import pytest

import sphradiometer_amd as S
from sphradiometer_amd import inject

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "oracle_cases.npz")
DT = 1.0 / 4096
TOL64 = 1e-12
TOL_MIXED = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module", autouse=True)
def gpu():
    S.lib()
    if not S.gpu_available():
        pytest.fail("no GPU: " + S.last_error())
    print("device:", S.last_error() or "ok", flush=True)


@pytest.fixture(scope="module")
def two_site():
    net = S.Network(inject.positions("AB"), 4096, DT, 25)
    return net


@pytest.fixture(scope="module")
def three_site():
    return S.Network(inject.positions("ABC"), 1024, DT, 32)


def test_fft_helpers_match_oracle(oracle):
    rng = np.random.default_rng(1)
    for n in (64, 1024, 4096, 584, 1000):
        x = rng.standard_normal(n)
        assert rel(S.tseries_to_fseries(x), oracle.tseries_to_fseries(x)) < TOL64
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert rel(S.ctseries_to_fseries(z), oracle.ctseries_to_fseries(z)) < TOL64


@pytest.mark.parametrize("L,polar", [(0, 0), (1, 0), (7, 1), (25, 0), (32, 0), (50, 0), (64, 0), (130, 1)])
def test_to_mesh_matches_oracle(oracle, L, polar):
    rng = np.random.default_rng(L + 100 * polar)
    n = L + 1 if polar else (L + 1) ** 2
    c = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    got = S.to_mesh(c, L, polar)
    ref = oracle.to_mesh(c, L, polar)
    assert got.shape == ref.shape
    assert rel(got, ref) < TOL64


def test_dotc_matches_numpy():
    L = S.lib()
    rng = np.random.default_rng(2)
    for n, lmax, polar in [(1, 3, 1), (1000, 20, 1), (584, 12, 0)]:
        arr = L.sh_series_array_new(n, lmax, polar)
        A = S.array_view(arr)
        A[:] = rng.standard_normal(A.shape) + 1j * rng.standard_normal(A.shape)
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        res = L.sh_series_new(lmax, polar)
        assert L.sh_series_array_dotc(res, arr, S._ptr(v))
        ref = (A * v[:, None]).sum(axis=0)
        assert rel(S.coeff_view(res), ref) < TOL64
        L.sh_series_free(res)
        L.sh_series_array_free(arr)


def test_rotate_and_rotate_z_match_oracle(oracle):
    L = S.lib()
    rng = np.random.default_rng(3)
    for lmax, polar in [(1, 0), (13, 0), (32, 1), (50, 0)]:
        n = lmax + 1 if polar else (lmax + 1) ** 2
        c = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        th, ph = rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi)
        s = S.series_from_numpy(c, lmax, polar)
        R = L.sh_series_rot_matrix(th, ph)
        plan = L.sh_series_rotation_plan_new(s, R)
        out = L.sh_series_new(lmax, 0)
        assert L.sh_series_rotate(out, s, plan)
        ref = oracle.rotate(c, lmax, polar, np.ctypeslib.as_array(R, shape=(9,)))
        assert rel(S.coeff_view(out), ref) < TOL64
        if not polar:
            assert L.sh_series_rotate_z(out, s, 0.37)
            assert rel(S.coeff_view(out), oracle.rotate_z(c, lmax, 0.37)) < TOL64
        for p in (out, s):
            L.sh_series_free(p)
        L.sh_series_rotation_plan_free(plan)


def test_rotated_impulse_on_gpu():
    # rotation_validate.c:396-437 through the GPU rotation
    L = S.lib()
    rng = np.random.default_rng(4)
    for lmax in (1, 13):
        for _ in range(5):
            th, ph = rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi)
            dth, dph = rng.uniform(0, math.pi) - th, rng.uniform(0, 2 * math.pi) - ph
            s = L.sh_series_impulse(lmax, th, ph)
            c = L.sh_series_impulse(lmax, th + dth, ph + dph)
            R = L.euler_rotation_matrix(-ph, dth, ph + dph)
            plan = L.sh_series_rotation_plan_new(s, R)
            out = L.sh_series_new(lmax, 0)
            assert L.sh_series_rotate(out, s, plan)
            assert L.diagnostics_rms_error(out, c) < 1e-10
            for p in (out, s, c):
                L.sh_series_free(p)
            L.sh_series_rotation_plan_free(plan)


def _segment_oracle(oracle, pos, n, L, ts, window=None, gmst=0.0):
    net = oracle.Network(pos, n, DT, L)
    return net.segment(ts, L, window=window, gmst=gmst)


def test_network_integrate_two_site(oracle, two_site):
    ts = inject.two_site(16, 4096, seed=0)[:, 0, :]
    w = inject.hann(4096)
    fs = np.array([oracle.tseries_to_fseries(ts[d] * w) for d in range(2)])
    sky = two_site.integrate(fs, 25)
    onet = oracle.Network(inject.positions("AB"), 4096, DT, 25)
    ref = onet.integrate(fs, 25)
    assert rel(sky, ref) < TOL64
    # the plan's power_1d is left holding the baseline power (correlator.c:703)
    assert rel(two_site.power_1d(0), onet.power_1d(0)) < TOL64
    gold = np.load(GOLD)
    assert rel(sky, gold["two_site_alm_0"]) < TOL64


def test_network_integrate_three_site_and_golden(oracle, three_site):
    ts, _ = inject.three_site(2, 1024, seed=0)
    gold = np.load(GOLD)
    for s in (0, 1):
        fs = np.array([S.tseries_to_fseries(ts[d, s]) for d in range(3)])
        sky = three_site.integrate(fs, 32)
        assert rel(sky, gold["three_site_alm_%d" % s]) < TOL64
    mesh = S.to_mesh(three_site.integrate(np.array([S.tseries_to_fseries(ts[d, 0]) for d in range(3)]), 32), 32)
    assert rel(mesh, gold["three_site_map_0"]) < TOL64


def _run_batch(net, L, ts, precision, window=None, gmst=None, fused=True, occ=1, nsub=0):
    D, Sg, N = ts.shape
    b = S.Batch(net, L, precision, max_segments=Sg)
    b.set_fused(fused)
    b.set_sky_occupancy(occ)
    b.set_sky_segments(nsub)
    cdt = np.complex128 if precision == S.FP64 else np.complex64
    d_ts = S.DeviceBuffer.from_numpy(ts.astype(np.float64))
    d_alm = S.DeviceBuffer(Sg * (L + 1) ** 2 * np.dtype(cdt).itemsize)
    d_map = S.DeviceBuffer(Sg * (2 * L + 2) ** 2 * np.dtype(cdt).itemsize)
    d_w = S.DeviceBuffer.from_numpy(window) if window is not None else None
    d_g = S.DeviceBuffer.from_numpy(np.asarray(gmst, dtype=np.float64)) if gmst is not None else None
    b.run(d_ts.ptr, Sg, d_alm.ptr, d_map.ptr, d_window=d_w.ptr if d_w else None, d_gmst=d_g.ptr if d_g else None)
    S.synchronize()
    alm = d_alm.to_numpy(cdt, (Sg, (L + 1) ** 2))
    mp = d_map.to_numpy(cdt, (Sg, 2 * L + 2, 2 * L + 2))
    return alm, mp, b.info()


def test_batch_fp64_two_site_matches_oracle_and_golden(oracle, two_site):
    ts = inject.two_site(16, 4096, seed=0)
    w = inject.hann(4096)
    alm, mp, info = _run_batch(two_site, 25, ts, S.FP64, window=w)
    print("batch info", info, flush=True)
    gold = np.load(GOLD)
    assert rel(alm[0], gold["two_site_alm_0"]) < TOL64
    assert rel(alm[1], gold["two_site_alm_1"]) < TOL64
    assert rel(mp[0], gold["two_site_map_0"]) < TOL64
    onet = oracle.Network(inject.positions("AB"), 4096, DT, 25)
    for s in (5, 15):
        sky, mesh = onet.segment(ts[:, s], 25, window=w)
        assert rel(alm[s], sky) < TOL64
        assert rel(mp[s], mesh) < TOL64


@pytest.mark.parametrize("nseg", [1, 3, 37])
def test_batch_fp64_three_site_ragged_and_gmst(oracle, three_site, nseg):
    ts, _ = inject.three_site(nseg, 1024, seed=nseg)
    gm = np.linspace(0.0, 2.0, nseg)
    alm, mp, _ = _run_batch(three_site, 32, ts, S.FP64, gmst=gm)
    onet = oracle.Network(inject.positions("ABC"), 1024, DT, 32)
    for s in sorted({0, nseg // 2, nseg - 1}):
        sky, mesh = onet.segment(ts[:, s], 32, gmst=float(gm[s]))
        assert rel(alm[s], sky) < TOL64
        assert rel(mp[s], mesh) < TOL64


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_gmst_with_few_k_splits(oracle, three_site, precision):
    # the projection's fewest K splits (shape 0: about one workgroup per CU)
    # keep the partial sums at most 8 deep, where the fused sky kernel
    # prefetches each workgroup's first segment before its table prologue;
    # with a GMST per segment, in each sky-kernel form (the two-workgroup
    # form keeps the rotation rows in its F tiles); 1100 segments, so every
    # form's persistent workgroups also run a second segment (their first
    # is the prefetched one)
    prec, tol = (S.FP64, TOL64) if precision == "fp64" else (S.MIXED, 1e-5)
    nseg = 1100
    ts, _ = inject.three_site(nseg, 1024, seed=5)
    gm = np.linspace(0.1, 6.1, nseg)
    onet = oracle.Network(inject.positions("ABC"), 1024, DT, 32)
    check = list(range(0, nseg, 37)) + [1023, 1024, 1050, nseg - 1]
    want = {s: onet.segment(ts[:, s], 32, gmst=float(gm[s])) for s in check}
    cdt = np.complex128 if prec == S.FP64 else np.complex64
    b = S.Batch(three_site, 32, prec, max_segments=nseg)
    d_ts = S.DeviceBuffer.from_numpy(ts.astype(np.float64))
    d_g = S.DeviceBuffer.from_numpy(gm)
    d_alm = S.DeviceBuffer(nseg * 33 ** 2 * np.dtype(cdt).itemsize)
    d_map = S.DeviceBuffer(nseg * 66 ** 2 * np.dtype(cdt).itemsize)
    b.set_lag(False)
    b.set_shape(0)
    assert b.info()["k_splits"] <= 8
    forms = [(1, 0), (2, 0), (3, 1)] + ([(1, 4)] if prec == S.MIXED else [])
    for occ, nsub in forms:
        b.set_sky_occupancy(occ)
        b.set_sky_segments(nsub)
        print("gmst, few splits", precision, occ, nsub, flush=True)
        b.run(d_ts.ptr, nseg, d_alm.ptr, d_map.ptr, d_gmst=d_g.ptr)
        S.synchronize()
        alm = d_alm.to_numpy(cdt, (nseg, 33 ** 2))
        mp = d_map.to_numpy(cdt, (nseg, 66, 66))
        for s in check:
            assert rel(alm[s], want[s][0]) < tol, (occ, nsub, s)
            assert rel(mp[s], want[s][1]) < (tol if prec == S.FP64 else 5e-6), (occ, nsub, s)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_sky_stage_forms_agree(oracle, three_site, precision):
    # the fused sky kernel (default at l_max 25 and 32) and the two-kernel
    # form (ring sums, then the phi DFT as a GEMM) against the oracle
    prec, tol = (S.FP64, TOL64) if precision == "fp64" else (S.MIXED, 1e-5)
    ts, _ = inject.three_site(5, 1024, seed=11)
    gm = np.linspace(0.3, 1.7, 5)
    a1, m1, i1 = _run_batch(three_site, 32, ts, prec, gmst=gm, fused=True)
    a2, m2, i2 = _run_batch(three_site, 32, ts, prec, gmst=gm, fused=False)
    # two persistent workgroups per CU (rotation rows from global memory),
    # and three of one segment each
    a3, m3, i3 = _run_batch(three_site, 32, ts, prec, gmst=gm, fused=True, occ=2)
    a4, m4, i4 = _run_batch(three_site, 32, ts, prec, gmst=gm, fused=True, occ=3)
    assert i1["fused"] == 1 and i2["fused"] == 0 and i3["sky_workgroups_per_cu"] == 2 and i4["sky_workgroups_per_cu"] == 3
    forms = [(a1, m1), (a2, m2), (a3, m3), (a4, m4)]
    if prec == S.MIXED:
        # four segments per workgroup sharing the tables and rotation rows
        # (the second workgroup has one active segment of four)
        a5, m5, i5 = _run_batch(three_site, 32, ts, prec, gmst=gm, fused=True, occ=1, nsub=4)
        assert i5["sky_segments_per_workgroup"] == 4
        forms.append((a5, m5))
        assert np.array_equal(a5, a1) and np.array_equal(m5, m1)
    else:
        b = S.Batch(three_site, 32, prec, max_segments=5)
        with pytest.raises(RuntimeError):
            b.set_sky_segments(4)
    onet = oracle.Network(inject.positions("ABC"), 1024, DT, 32)
    for s in range(5):
        sky, mesh = onet.segment(ts[:, s], 32, gmst=float(gm[s]))
        for alm, mp in forms:
            assert rel(alm[s], sky) < tol
            assert rel(mp[s], mesh) < (tol if prec == S.FP64 else 5e-6)


def test_batch_every_tuner_shape_matches_oracle(oracle, three_site):
    # each projection shape and K split the autotuner can pick, in both
    # precisions, against the oracle on the same segments
    ts, _ = inject.three_site(3, 1024, seed=21)
    onet = oracle.Network(inject.positions("ABC"), 1024, DT, 32)
    want = [onet.segment(ts[:, s], 32) for s in range(3)]
    d_ts = S.DeviceBuffer.from_numpy(ts.astype(np.float64))
    for prec, cdt, tol in ((S.FP64, np.complex128, TOL64), (S.MIXED, np.complex64, 1e-5)):
        b = S.Batch(three_site, 32, prec, max_segments=3)
        d_alm = S.DeviceBuffer(3 * 33 ** 2 * np.dtype(cdt).itemsize)
        d_map = S.DeviceBuffer(3 * 66 ** 2 * np.dtype(cdt).itemsize)
        # the shapes belong to the frequency-domain route
        b.set_lag(False)
        for shape in range(28):
            b.set_shape(shape)
            b.run(d_ts.ptr, 3, d_alm.ptr, d_map.ptr)
            S.synchronize()
            alm = d_alm.to_numpy(cdt, (3, 33 ** 2))
            mp = d_map.to_numpy(cdt, (3, 66, 66))
            for s in range(3):
                assert rel(alm[s], want[s][0]) < tol, (prec, shape, s)
                assert rel(mp[s], want[s][1]) < (tol if prec == S.FP64 else 5e-6), (prec, shape, s)
        with pytest.raises(ValueError):
            b.set_shape(28)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nseg", [1, 2, 5, 37])
def test_batch_lag_route_matches_oracle(oracle, three_site, precision, nseg):
    # the lag-domain route (one kernel: FFTs, cross spectra, inverse
    # transforms, folded lags x the table) is the default at three sites
    # and 1024 samples; odd counts leave the last pair with one segment,
    # counts that are not multiples of four leave the wave-level form's
    # last pass with a missing pair; with a Hann window and a GMST per
    # segment, in every kernel form, against the oracle and against the
    # frequency-domain route
    prec, tol = (S.FP64, TOL64) if precision == "fp64" else (S.MIXED, 1e-5)
    ts, _ = inject.three_site(nseg, 1024, seed=40 + nseg)
    w = inject.hann(1024)
    gm = np.linspace(0.2, 5.9, nseg)
    cdt = np.complex128 if prec == S.FP64 else np.complex64
    b = S.Batch(three_site, 32, prec, max_segments=nseg)
    assert b.lag
    # form 1 is the default; the wave-level form (3) and the form with two
    # pairs per pass (4) are compiled for mixed precision
    forms = (1, 2, 3, 4) if prec == S.MIXED else (1, 2)
    assert b.info()["xcorr_form"] == forms[0]
    d_ts = S.DeviceBuffer.from_numpy(ts.astype(np.float64))
    d_w = S.DeviceBuffer.from_numpy(w)
    d_g = S.DeviceBuffer.from_numpy(gm)
    d_alm = S.DeviceBuffer(nseg * 33 ** 2 * np.dtype(cdt).itemsize)
    d_map = S.DeviceBuffer(nseg * 66 ** 2 * np.dtype(cdt).itemsize)

    def run():
        b.run(d_ts.ptr, nseg, d_alm.ptr, d_map.ptr, d_window=d_w.ptr, d_gmst=d_g.ptr)
        S.synchronize()
        return d_alm.to_numpy(cdt, (nseg, 33 ** 2)), d_map.to_numpy(cdt, (nseg, 66, 66))

    onet = oracle.Network(inject.positions("ABC"), 1024, DT, 32)
    for s in range(nseg):
        if nseg > 5 and s not in (0, 1, 17, nseg - 2, nseg - 1):
            continue
        sky, mesh = onet.segment(ts[:, s], 32, window=w, gmst=float(gm[s]))
        for form in forms:
            b.set_xcorr_form(form)
            alm, mp = run()
            print("lag route", precision, nseg, s, form, rel(alm[s], sky), flush=True)
            assert rel(alm[s], sky) < tol, (form, s)
            assert rel(mp[s], mesh) < (tol if prec == S.FP64 else 5e-6), (form, s)
    b.set_lag(False)
    alm_fd, mp_fd = run()
    b.set_lag(True)
    alm, mp = run()
    assert rel(alm, alm_fd) < (1e-13 if prec == S.FP64 else 2e-6)
    with pytest.raises(RuntimeError):
        b.set_xcorr_form(5 if prec == S.MIXED else 3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_lag_route_two_detectors_1024(oracle, precision):
    # two detectors at 1024 samples: the kernel's two-group shape (one
    # baseline; the second group idle in the inverse transforms), every
    # form, against the oracle
    prec, tol = (S.FP64, TOL64) if precision == "fp64" else (S.MIXED, 1e-5)
    pos = inject.positions("AB")
    net = S.Network(pos, 1024, DT, 20)
    ts = inject.two_site(5, 1024, seed=9)
    cdt = np.complex128 if prec == S.FP64 else np.complex64
    b = S.Batch(net, 20, prec, max_segments=5)
    assert b.lag
    d_ts = S.DeviceBuffer.from_numpy(ts)
    d_alm = S.DeviceBuffer(5 * 21 ** 2 * np.dtype(cdt).itemsize)
    d_map = S.DeviceBuffer(5 * 42 ** 2 * np.dtype(cdt).itemsize)
    onet = oracle.Network(pos, 1024, DT, 20)
    want = [onet.segment(ts[:, s], 20) for s in range(5)]
    for form in ((1, 2, 3, 4) if prec == S.MIXED else (1, 2)):
        b.set_xcorr_form(form)
        b.run(d_ts.ptr, 5, d_alm.ptr, d_map.ptr)
        S.synchronize()
        alm = d_alm.to_numpy(cdt, (5, 21 ** 2))
        mp = d_map.to_numpy(cdt, (5, 42, 42))
        for s in range(5):
            assert rel(alm[s], want[s][0]) < tol, (form, s)
            assert rel(mp[s], want[s][1]) < (tol if prec == S.FP64 else 5e-6), (form, s)


@pytest.mark.parametrize("lag", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_float32_series(oracle, three_site, precision, lag):
    # float32 input series (the layout the learned proposal's training
    # batches use), on both projection routes, against the oracle on the
    # same float32 values widened
    prec, tol = (S.FP64, TOL64) if precision == "fp64" else (S.MIXED, 1e-5)
    ts, _ = inject.three_site(6, 1024, seed=77)
    ts32 = ts.astype(np.float32)
    cdt = np.complex128 if prec == S.FP64 else np.complex64
    b = S.Batch(three_site, 32, prec, max_segments=6)
    b.set_lag(lag)
    d_ts = S.DeviceBuffer.from_numpy(ts32)
    d_alm = S.DeviceBuffer(6 * 33 ** 2 * np.dtype(cdt).itemsize)
    d_map = S.DeviceBuffer(6 * 66 ** 2 * np.dtype(cdt).itemsize)
    b.run(d_ts.ptr, 6, d_alm.ptr, d_map.ptr, f32_input=True)
    S.synchronize()
    alm = d_alm.to_numpy(cdt, (6, 33 ** 2))
    mp = d_map.to_numpy(cdt, (6, 66, 66))
    onet = oracle.Network(inject.positions("ABC"), 1024, DT, 32)
    for s in (0, 3, 5):
        sky, mesh = onet.segment(ts32[:, s].astype(np.float64), 32)
        assert rel(alm[s], sky) < tol, s
        assert rel(mp[s], mesh) < (tol if prec == S.FP64 else 5e-6), s


def test_batch_lag_route_refused_when_lags_wrap(oracle):
    # 256 samples: the delay products' lags (2 transient + 1 of them) do
    # not fit the segment, so the batch keeps the frequency-domain route,
    # with the same results as the oracle
    pos = inject.positions("AB")
    net = S.Network(pos, 256, DT, 12)
    onet = oracle.Network(pos, 256, DT, 12)
    info = onet.info(0)
    print("transient", info["transient"], flush=True)
    assert 4 * info["transient"] + 1 > 256
    ts = inject.two_site(3, 256, seed=4)
    alm, mp, binfo = _run_batch(net, 12, ts, S.FP64)
    assert binfo["lag"] == 0 and binfo["xcorr_form"] == 0
    b = S.Batch(net, 12, S.FP64, max_segments=3)
    with pytest.raises(ValueError):
        b.set_lag(True)
    for s in range(3):
        sky, mesh = onet.segment(ts[:, s], 12)
        assert rel(alm[s], sky) < TOL64
        assert rel(mp[s], mesh) < TOL64


def test_batch_lag_route_two_site_mixed(oracle, two_site):
    # two detectors at 4096 samples: the lag-domain route in mixed
    # precision (float64 keeps the frequency-domain route there)
    ts = inject.two_site(16, 4096, seed=0)
    w = inject.hann(4096)
    b64 = S.Batch(two_site, 25, S.FP64, max_segments=16)
    assert not b64.lag
    with pytest.raises(ValueError):
        b64.set_lag(True)
    alm, mp, _ = _run_batch(two_site, 25, ts, S.MIXED, window=w)
    onet = oracle.Network(inject.positions("AB"), 4096, DT, 25)
    for s in (0, 7, 15):
        sky, mesh = onet.segment(ts[:, s], 25, window=w)
        assert rel(alm[s], sky) < 1e-5
        assert rel(mp[s], mesh) < 1e-5
    bm = S.Batch(two_site, 25, S.MIXED, max_segments=16)
    assert bm.lag


def test_batch_tuner_routes_match_oracle(oracle, two_site):
    # the autotuner times both projection routes while the lag-domain
    # route is in use and keeps the faster (16 segments of 4096 samples:
    # eight segment pairs, eight workgroups of the lag kernel); whichever
    # it keeps matches the oracle, and a batch set to the
    # frequency-domain route stays there
    ts = inject.two_site(16, 4096, seed=3)
    w = inject.hann(4096)
    d_ts = S.DeviceBuffer.from_numpy(ts)
    d_w = S.DeviceBuffer.from_numpy(w)
    d_alm = S.DeviceBuffer(16 * 26 ** 2 * 8)
    d_map = S.DeviceBuffer(16 * 52 ** 2 * 8)
    onet = oracle.Network(inject.positions("AB"), 4096, DT, 25)
    want = [onet.segment(ts[:, s], 25, window=w) for s in (0, 9, 15)]
    for lag in (True, False):
        b = S.Batch(two_site, 25, S.MIXED, max_segments=16)
        b.set_lag(lag)
        k, ms = b.tune(d_ts.ptr, 16, d_alm.ptr, d_map.ptr, d_window=d_w.ptr)
        info = b.info()
        print("tuned two-site: lag %s -> %d, shape %d, %.4f ms, xcorr form %d" % (lag, info["lag"], k, ms, info["xcorr_form"]), flush=True)
        assert 0 <= k < 28 and ms > 0
        if not lag:
            assert info["lag"] == 0
        b.run(d_ts.ptr, 16, d_alm.ptr, d_map.ptr, d_window=d_w.ptr)
        S.synchronize()
        alm = d_alm.to_numpy(np.complex64, (16, 26 ** 2))
        mp = d_map.to_numpy(np.complex64, (16, 52, 52))
        for (sky, mesh), s in zip(want, (0, 9, 15)):
            assert rel(alm[s], sky) < 1e-5, (lag, s)
            assert rel(mp[s], mesh) < 1e-5, (lag, s)


def test_batch_full_size_three_site_properties(three_site):
    # the 1024-segment workload: golden segments, linearity (x -> 2x gives
    # 4x power), and a real sky for real data
    ts, _ = inject.three_site(1024, 1024, seed=0)
    alm, mp, _ = _run_batch(three_site, 32, ts, S.FP64)
    gold = np.load(GOLD)
    assert rel(alm[0], gold["three_site_alm_0"]) < TOL64
    assert rel(alm[1], gold["three_site_alm_1"]) < TOL64
    alm2, _, _ = _run_batch(three_site, 32, 2.0 * ts, S.FP64)
    assert rel(alm2, 4.0 * alm) < 1e-14
    assert np.max(np.abs(mp.imag)) < 1e-12 * np.max(np.abs(mp.real))
    # mixed precision: a_lm within 1e-5 of float64 per segment
    almm, mpm, _ = _run_batch(three_site, 32, ts, S.MIXED)
    err = np.linalg.norm(almm - alm, axis=1) / np.linalg.norm(alm, axis=1)
    print("mixed a_lm rel err: max %.3g median %.3g; map %.3g" % (err.max(), np.median(err), rel(mpm, mp)), flush=True)
    assert err.max() < TOL_MIXED
