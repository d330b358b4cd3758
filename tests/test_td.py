"""Time-domain correlator (correlator.c:325-400, 611-673, 820-858, 967-984
of the reference) and the Tukey-overlap stitching of long streams it
documents (correlator.c:611-646).

The oracle (oracle/or_td.c) restates the reference sample by sample:
sh_series_array_dot of the two delay projections, sh_series_product,
accumulation with the window, rotation.  It is pinned against the FD
oracle (itself pinned on the reference's golden vectors): on a series
extended periodically by 2 transient samples with a unit window, the TD
power equals N times the FD power (both are the circular cross-power of
the same taps; the routes share no code).  The stitching identity is exact
algebra: two overlapping segments integrated with their own windows sum
to the whole series integrated with the two windows added at their
offsets.  The library's GPU integrators are compared with the oracle at
1e-12 relative, with random windows (the correlation over every sample)
and with square and Tukey windows (the correlation over the samples where
the window changes, the rest recovered along the diagonals of C).
"""
import ctypes

import numpy as np
import pytest

import sphradiometer_amd as S

TOL64 = 1e-12
DT = 1.0 / 4096
# two short baselines: small tap counts keep the sample-by-sample oracle fast
POS = np.array([[-0.001, 0.0, 0.0015], [0.002, 0.001, -0.0005], [0.0005, -0.002, 0.0]])


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def test_oracle_td_equals_n_times_fd_on_periodic_data(oracle):
    N = 512
    rng = np.random.default_rng(1)
    for (i, j) in ((1, 0), (2, 1)):
        td = oracle.PlanTD(POS[i], POS[j], DT)
        tr = td.transient
        net = oracle.Network(np.array([POS[j], POS[i]]), N, DT)
        ts = rng.standard_normal((2, N))
        net.integrate(np.array([oracle.tseries_to_fseries(ts[k]) for k in range(2)]), td.l_max)
        fd = net.power_1d(0)
        ext = lambda x: np.concatenate([x, x[: 2 * tr]])
        _, p1 = td.integrate(ext(ts[1]), ext(ts[0]), np.ones(N))
        L = min(len(p1), len(fd))
        assert rel(p1[:L], N * fd[:L]) < 1e-12, (i, j, rel(p1[:L], N * fd[:L]))


def _stitch_windows(n, tr, s, n1, taper):
    """Tukey windows of two segments [0, n1) and [s, n) of an n-sample
    series and their sum at the full series' output samples"""
    L = S.lib()
    m1, m2 = n1 - 2 * tr, n - s - 2 * tr
    w1 = np.ctypeslib.as_array(L.correlator_tukey_window_new(m1, 0, taper, 1.0 / m1), (m1,)).copy()
    w2 = np.ctypeslib.as_array(L.correlator_tukey_window_new(m2, 0, taper, 1.0 / m2), (m2,)).copy()
    wf = np.zeros(n - 2 * tr)
    wf[:m1] += w1
    wf[s: s + m2] += w2
    return w1, w2, wf


def test_oracle_td_stitching_identity(oracle):
    td = oracle.PlanTD(POS[1], POS[0], DT)
    tr = td.transient
    rng = np.random.default_rng(2)
    n, n1, s, taper = 2 * tr + 400, 2 * tr + 250, 150, 60
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    w1, w2, wf = _stitch_windows(n, tr, s, n1, taper)
    p_1, _ = td.integrate(a[:n1], b[:n1], w1)
    p_2, _ = td.integrate(a[s:], b[s:], w2)
    p_f, _ = td.integrate(a, b, wf)
    assert rel(p_1 + p_2, p_f) < 1e-13


# ------------------------------------------------------------------- GPU


class _Net:
    """instruments and baselines of the library for POS"""

    def __init__(self, pos):
        L = S.lib()
        self.inst = L.instrument_array_new(0)
        for x, y, z in pos:
            L.instrument_array_append(self.inst, L.instrument_new(float(x), float(y), float(z), None))
        self.bl = L.correlator_network_baselines_new(self.inst)

    def baseline(self, k):
        return self.bl.contents.baselines[k]


def _integrate(plan, a, b, w):
    a, b, w = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, w))
    r = S.lib().correlator_baseline_integrate_power_td(a.ctypes.data, b.ctypes.data, w.ctypes.data, a.size, plan)
    assert r, S.last_error()
    out = S.coeff_view(r).copy()
    S.lib().sh_series_free(r)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [1, 600, 5000])
def test_gpu_td_baseline_matches_oracle(oracle, extra):
    print("td baseline, output samples", extra, flush=True)
    net = _Net(POS[:2])
    plan = S.lib().correlator_plan_td_new(net.baseline(0), DT)
    assert plan, S.last_error()
    td = oracle.PlanTD(POS[1], POS[0], DT)
    tr = plan.contents.transient
    assert tr == td.transient and plan.contents.power_1d.contents.l_max == td.l_max
    rng = np.random.default_rng(extra)
    n = 2 * tr + extra
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    w = rng.uniform(0.1, 1.0, extra) / extra
    got = _integrate(plan, a, b, w)
    want, p1 = td.integrate(a, b, w)
    assert rel(got, want) < TOL64, rel(got, want)
    # power_1d holds the axially symmetric power after the call
    assert rel(S.coeff_view(plan.contents.power_1d), p1) < TOL64
    # too short a series: NULL, as the reference
    short = np.zeros(2 * tr - 1)
    assert not S.lib().correlator_baseline_integrate_power_td(short.ctypes.data, short.ctypes.data, w.ctypes.data, short.size, plan)
    S.lib().correlator_plan_td_free(plan)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["square", "tukey"])
@pytest.mark.parametrize("extra", [600, 5000])
def test_gpu_td_flat_windows_match_oracle(oracle, kind, extra):
    """windows flat over most of the series (square, Tukey): the library
    correlates only where the window changes and recovers the rest along
    the diagonals of C"""
    print("td window", kind, "output samples", extra, flush=True)
    net = _Net(POS[:2])
    L = S.lib()
    plan = L.correlator_plan_td_new(net.baseline(0), DT)
    td = oracle.PlanTD(POS[1], POS[0], DT)
    tr = plan.contents.transient
    rng = np.random.default_rng(extra + len(kind))
    n = 2 * tr + extra
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    if kind == "square":
        w = np.full(extra, 1.0 / extra)
    else:
        w = np.ctypeslib.as_array(L.correlator_tukey_window_new(extra, 0, extra // 8, 1.0 / extra), (extra,)).copy()
    got = _integrate(plan, a, b, w)
    want, _ = td.integrate(a, b, w)
    assert rel(got, want) < TOL64, rel(got, want)
    L.correlator_plan_td_free(plan)


def _window(L, kind, m, rng):
    if kind == "square":
        return np.full(m, 1.0 / m)
    if kind == "tukey":
        return np.ctypeslib.as_array(L.correlator_tukey_window_new(m, 0, m // 8, 1.0 / m), (m,)).copy()
    return rng.uniform(0.5, 1.0, m) / m


@pytest.mark.gpu
@pytest.mark.parametrize("kinds", [("random",) * 3, ("square",) * 3, ("tukey",) * 3, ("square", "random", "tukey")])
def test_gpu_td_network_matches_oracle(oracle, kinds):
    """the network integrator takes, per baseline, the window-difference
    route for flat windows (square, Tukey) and the full correlation for
    the others; every mix against the oracle"""
    print("td network", kinds, flush=True)
    net = _Net(POS)
    L = S.lib()
    plan = L.correlator_network_plan_td_new(net.bl, DT)
    assert plan, S.last_error()
    onet = oracle.NetworkTD(POS, DT)
    trs = [plan.contents.plans[k].contents.transient for k in range(3)]
    rng = np.random.default_rng(7)
    n = 2 * max(trs) + 700
    ts = rng.standard_normal((3, n))
    wins = [_window(L, kind, n - 2 * t, rng) for kind, t in zip(kinds, trs)]
    stride = max(len(w) for w in wins)
    wpack = np.zeros((3, stride))
    for k, w in enumerate(wins):
        wpack[k, : len(w)] = w
    tp = (ctypes.c_void_p * 3)(*[ts[i].ctypes.data for i in range(3)])
    wp = (ctypes.c_void_p * 3)(*[wins[k].ctypes.data for k in range(3)])
    # the sky takes each baseline's power by sh_series_add, which needs an
    # order at least the power's (sh_series.c:366-391): a smaller sky
    # fails, as in the reference
    lmax = max(plan.contents.plans[k].contents.power_1d.contents.l_max for k in range(3))
    small = L.sh_series_new(lmax - 1, 0)
    assert not L.correlator_network_integrate_power_td(small, tp, n, wp, plan)
    L.sh_series_free(small)
    sky_l = lmax + 2
    sky = L.sh_series_new(sky_l, 0)
    assert L.correlator_network_integrate_power_td(sky, tp, n, wp, plan), S.last_error()
    want = onet.integrate(ts, wpack, sky_l)
    assert rel(S.coeff_view(sky), want) < TOL64
    L.sh_series_free(sky)
    L.correlator_network_plan_td_free(plan)


@pytest.mark.gpu
def test_gpu_td_stitching(oracle):
    print("td stitching", flush=True)
    net = _Net(POS[:2])
    plan = S.lib().correlator_plan_td_new(net.baseline(0), DT)
    tr = plan.contents.transient
    rng = np.random.default_rng(3)
    n, n1, s, taper = 2 * tr + 3000, 2 * tr + 1800, 1400, 300
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    w1, w2, wf = _stitch_windows(n, tr, s, n1, taper)
    p_1 = _integrate(plan, a[:n1], b[:n1], w1)
    p_2 = _integrate(plan, a[s:], b[s:], w2)
    p_f = _integrate(plan, a, b, wf)
    assert rel(p_1 + p_2, p_f) < 1e-12
    td = oracle.PlanTD(POS[1], POS[0], DT)
    assert rel(p_f, td.integrate(a, b, wf)[0]) < TOL64
    S.lib().correlator_plan_td_free(plan)


@pytest.mark.gpu
def test_gpu_td_long_stream_bounded_splits():
    """a stream long enough that the correlation's K splits hit their cap
    (eight workgroups per CU's worth): the capped call over 245760 output
    samples equals the sum of the two uncapped calls over its halves with
    the same window (the TD power is linear in the window and local in
    time, correlator.c:611-646); the different summation orders agree to
    1e-11"""
    from sphradiometer_amd import inject
    print("td long stream", flush=True)
    net = _Net(inject.positions("AB"))
    plan = S.lib().correlator_plan_td_new(net.baseline(0), DT)
    tr = plan.contents.transient
    nt, s = 240 * 1024, 100000
    rng = np.random.default_rng(17)
    a, b = rng.standard_normal(nt + 2 * tr), rng.standard_normal(nt + 2 * tr)
    w = rng.uniform(0.5, 1.0, nt) / nt
    p_full = _integrate(plan, a, b, w)
    p_1 = _integrate(plan, a[:s + 2 * tr], b[:s + 2 * tr], w[:s])
    p_2 = _integrate(plan, a[s:], b[s:], w[s:])
    assert rel(p_1 + p_2, p_full) < 1e-11, rel(p_1 + p_2, p_full)
    S.lib().correlator_plan_td_free(plan)
