"""The localization log-likelihood sky (inspiral_sky_map.c,
RapidLocalization.py:246-298): oracle pins on the CPU, then the library's
projection multiply, per-event skies and the batched GPU path against the
oracle.

Pins: the Tukey window's closed form and scipy.signal.windows.tukey (the
textbook definition, computed by code that shares nothing with the
oracle); the padded length at 4096 Hz (174 samples of light-travel guard
for 410 -> 584); the antenna response of a zenith source on a polar
detector (F+ = -cos 2psi, Fx = sin 2psi), the textbook pattern of an
L-shaped detector at the pole for sources anywhere on the sky, covariance
under a rotation of the Earth frame about z (detector and source turned
together) and the psi-invariance of F+^2 + Fx^2; the projection operator is a rank-one
projector minus the identity (trace 1 - n, Hermitian, (p + 1)^2 = p + 1);
the log prior is axisymmetric and even, with a_00 close to the analytic
sqrt(4 pi) (log 2 - 1 - log 4 pi); a sky from zero data is the rotated,
l-weighted prior alone, and the data part scales as the square of the data.

The LAL pieces (antenna response formula, Tukey definition) have no
reference test in the reference; they are pinned by the closed forms and
independent implementations above.
"""
import math

import numpy as np
import pytest

import sphradiometer_amd as S
from sphradiometer_amd import inject

DT_SMALL, N_SMALL = 1.0 / 256, 32


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


# ---------------------------------------------------------------- oracle pins


def test_oracle_tukey_closed_form(oracle):
    w = oracle.tukey(11, 0.1)
    assert w[0] == 0.0 and w[-1] == 0.0 and np.all(w[1:-1] == 1.0)
    # taper midpoint |y| = 1 - beta/2 sits at a sample for n = 401, beta = 0.1
    w = oracle.tukey(401, 0.1)
    assert abs(w[10] - 0.5) < 1e-15 and abs(w[390] - 0.5) < 1e-15
    assert np.all(w[20:381] == 1.0)


@pytest.mark.parametrize("n", [2, 3, 21, 410, 411, 584, 1000])
def test_oracle_tukey_matches_scipy(oracle, n):
    # scipy's Tukey window (sym=True) tapers n <= floor(beta (n-1)/2) with
    # 0.5 (1 + cos(pi (-1 + 2 n / (beta (n-1))))), the textbook form; the
    # oracle restates LAL's |y|-parametrised cos^2 taper.  A linear or
    # half-width taper fails this at every n > 21.
    from scipy.signal import windows
    for beta in (0.1, 0.5, 1.0):
        ref = windows.tukey(n, beta)
        assert np.max(np.abs(oracle.tukey(n, beta) - ref)) < 1e-14, (n, beta)


def test_oracle_precalc_length(oracle):
    assert oracle.precalc_length(410, 1.0 / 4096) == 584
    assert oracle.precalc_length(21, DT_SMALL) == N_SMALL


def test_oracle_preprocess_padding(oracle):
    x = np.ones((2, 410), dtype=np.complex128)
    y = oracle.preprocess_snr(x, 1.0 / 4096)
    assert y.shape == (2, 584)
    assert np.all(y[:, :87] == 0) and np.all(y[:, 497:] == 0)
    w = oracle.tukey(410, 0.1) * math.sqrt(584 / 410)
    assert np.max(np.abs(y[0, 87:497] - w)) < 1e-14


def test_oracle_antenna_response(oracle):
    D = np.diag([0.5, -0.5, 0.0])
    for psi in (0.0, 0.3, 1.1):
        fp, fx = oracle.det_response(D.ravel(), 0.0, math.pi / 2, psi, 0.0)
        assert abs(fp + math.cos(2 * psi)) < 1e-15 and abs(fx - math.sin(2 * psi)) < 1e-15
    rng = np.random.default_rng(2)
    Dr = inject.responses("ABC")[1]
    for _ in range(5):
        ra, dec = rng.uniform(0, 2 * math.pi), rng.uniform(-1.5, 1.5)
        a = oracle.det_response(Dr.ravel(), ra, dec, 0.0, 0.4)
        b = oracle.det_response(Dr.ravel(), ra, dec, 0.7, 0.4)
        assert abs(a[0] ** 2 + a[1] ** 2 - b[0] ** 2 - b[1] ** 2) < 1e-15


def test_oracle_antenna_response_textbook_pattern(oracle):
    # An L-shaped detector at the north pole with arms along x and y,
    # D = (xx - yy)/2, sees a source at polar angle theta = pi/2 - dec and
    # azimuth phi = ra - gmst with the textbook pattern (Thorne 1987,
    # eqs. 104a-b; Schutz and Tinto 1987):
    #   F+ = 1/2 (1 + cos^2 th) cos 2ph cos 2ps - cos th sin 2ph sin 2ps
    #   Fx = 1/2 (1 + cos^2 th) cos 2ph sin 2ps + cos th sin 2ph cos 2ps
    # whose azimuth is measured a quarter turn away from LAL's and whose
    # polarisation angle turns the other way: ph = phi + pi/2, ps = -psi.
    # One mapping has to fit both outputs at every sky position; a sign
    # error in any component of the wave frame, swapped X and Y or gmst
    # entering with the wrong sign fails it.  F+^2 + Fx^2 needs no mapping.
    D = np.diag([0.5, -0.5, 0.0])
    rng = np.random.default_rng(11)
    for _ in range(200):
        ra, dec = rng.uniform(0, 2 * math.pi), rng.uniform(-1.5, 1.5)
        psi, gmst = rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi)
        fp, fx = oracle.det_response(D.ravel(), ra, dec, psi, gmst)
        c = math.cos(math.pi / 2 - dec)
        ph, ps = ra - gmst + math.pi / 2, -psi
        tp = 0.5 * (1 + c * c) * math.cos(2 * ph) * math.cos(2 * ps) - c * math.sin(2 * ph) * math.sin(2 * ps)
        tx = 0.5 * (1 + c * c) * math.cos(2 * ph) * math.sin(2 * ps) + c * math.sin(2 * ph) * math.cos(2 * ps)
        assert abs(fp - tp) < 1e-15 and abs(fx - tx) < 1e-15
        mod = 0.25 * (1 + c * c) ** 2 * math.cos(2 * (ra - gmst)) ** 2 + c * c * math.sin(2 * (ra - gmst)) ** 2
        assert abs(fp * fp + fx * fx - mod) < 1e-15


def test_oracle_antenna_response_earth_rotation(oracle):
    # Turning the detector by gamma about the Earth's axis (D -> Rz D Rz^T)
    # and the source by the same angle (ra -> ra + gamma) leaves the
    # response unchanged: the polarisation frame is fixed to local north
    # and east, which turn with them.  Turning the source the other way
    # does not, so the check fixes the sign of the hour angle.
    rng = np.random.default_rng(12)
    for name in "ABC":
        Dr = inject.responses(name)[0]
        worst_wrong = 0.0
        for _ in range(30):
            g = rng.uniform(0, 2 * math.pi)
            Rz = np.array([[math.cos(g), -math.sin(g), 0.0], [math.sin(g), math.cos(g), 0.0], [0.0, 0.0, 1.0]])
            D2 = Rz @ Dr @ Rz.T
            ra, dec = rng.uniform(0, 2 * math.pi), rng.uniform(-1.5, 1.5)
            psi, gmst = rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi)
            a = oracle.det_response(Dr.ravel(), ra, dec, psi, gmst)
            b = oracle.det_response(D2.ravel(), ra + g, dec, psi, gmst)
            assert abs(a[0] - b[0]) < 1e-14 and abs(a[1] - b[1]) < 1e-14
            # gmst and ra enter only through the hour angle gmst - ra
            c = oracle.det_response(Dr.ravel(), ra + g, dec, psi, gmst + g)
            assert abs(a[0] - c[0]) < 1e-14 and abs(a[1] - c[1]) < 1e-14
            w = oracle.det_response(D2.ravel(), ra - g, dec, psi, gmst)
            worst_wrong = max(worst_wrong, abs(a[0] - w[0]), abs(a[1] - w[1]))
        assert worst_wrong > 0.1


def test_oracle_projection_is_projector_minus_identity(oracle):
    R = inject.responses("ABC").reshape(3, 9)
    rng = np.random.default_rng(4)
    for beta in (1.0, -1.0):
        th, ph = rng.uniform(0.1, 3.0), rng.uniform(0, 2 * math.pi)
        psd = rng.uniform(0.5, 2.0, 3)
        p = np.array([[oracle.cbc_projection(th, ph, i, j, R, beta, 0.0, psd) for j in range(3)] for i in range(3)])
        assert abs(np.trace(p) - (1 - 3)) < 1e-14
        assert np.max(np.abs(p - p.conj().T)) < 1e-15
        q = p + np.eye(3)
        assert np.max(np.abs(q @ q - q)) < 1e-14


def test_oracle_log_prior(oracle):
    L = 50
    c = oracle.log_uniformsky_prior(L)
    for l in range(L + 1):
        for m in range(-l, l + 1):
            v = c[oracle.lmoffset(L, l, m)]
            if m != 0 or l % 2:
                assert abs(v) < 1e-12
    a00 = math.sqrt(4 * math.pi) * (math.log(2) - 1 - math.log(4 * math.pi))
    assert abs(c[0].real - a00) < 2e-3 * abs(a00)


def det_responses(names):
    """The detectors' response tensors as the library holds them: single
    precision, as LALDetector's REAL4 response[3][3]."""
    return inject.responses(names).astype(np.float32).astype(np.float64)


def small_network(oracle, beta, target=None):
    net = oracle.Network(inject.positions("ABC"), N_SMALL, DT_SMALL)
    net.mult_by_projection(det_responses("ABC"), beta, 0.0, np.ones((N_SMALL, 3)))
    if target is not None:
        net.set_l(target)
    return net


def test_oracle_projection_product_against_quadrature(oracle):
    # dp'[f] after mult_by_projection equals the projection onto Y_lm of
    # dp[f](x) p(R x), with R = rot_matrix(baseline) taking the baseline
    # frame to the Earth frame and p the projection operator evaluated
    # directly, up to the truncation of p's expansion at the plan's order
    # (about 1e-5 here); the transposed rotation is off by order one
    pos, R = inject.positions("ABC"), det_responses("ABC").reshape(3, 9)
    net0 = oracle.Network(pos, N_SMALL, DT_SMALL)
    net = small_network(oracle, 1.0)
    k, f = 1, 3
    inf = net0.info(k)
    L, i, j = inf["l_max"], inf["index_a"], inf["index_b"]
    g = net0.geometry(k)
    rot = oracle.rot_matrix(g[3], g[4]).reshape(3, 3)
    dp, dpp = net0.delay_product(k)[f], net.delay_product(k)[f]
    nq = 2 * L + 10
    x, w = oracle.gl_nodes(nq)
    nph = 2 * nq
    phis = 2 * math.pi * np.arange(nph) / nph
    out = {}
    for name, M in (("rot", rot), ("transposed", rot.T)):
        V = np.zeros((nq, nph), dtype=np.complex128)
        for a in range(nq):
            th = math.acos(x[a])
            dpv = np.dot(dp, oracle.sphPlm_array(L, 0, x[a]))
            for b in range(nph):
                v = M @ np.array([math.sin(th) * math.cos(phis[b]), math.sin(th) * math.sin(phis[b]), math.cos(th)])
                te, pe = math.acos(max(-1.0, min(1.0, v[2]))), math.atan2(v[1], v[0]) % (2 * math.pi)
                V[a, b] = dpv * oracle.cbc_projection(te, pe, i, j, R, 1.0, 0.0, np.ones(3))
        Vm = np.fft.fft(V, axis=1) * (2 * math.pi / nph)   # sum_b V e^{-i m phi_b}
        I = np.zeros((L + 1) ** 2, dtype=np.complex128)
        for m in range(-L, L + 1):
            am = abs(m)
            s = -1.0 if (m < 0 and am % 2) else 1.0
            P = np.array([oracle.sphPlm_array(L, am, xa) for xa in x])   # [a][l - am]
            for l in range(am, L + 1):
                I[oracle.lmoffset(L, l, m)] = s * np.sum(w * Vm[:, m % nph] * P[:, l - am])
        out[name] = rel(I, dpp)
    assert out["rot"] < 1e-4, out
    assert out["transposed"] > 0.1, out


def test_oracle_localize_zero_data_and_scaling(oracle):
    net = small_network(oracle, 1.0)
    L = 22
    lp = oracle.log_uniformsky_prior(L)
    gm = 0.8
    zero = np.zeros((3, 21), dtype=np.complex128)
    sky0 = net.localize(zero, DT_SMALL, gm, lp, L)
    want = oracle.scale_power_l(oracle.rotate_z(lp, L, gm), L, 0.6, 2.7)
    assert rel(sky0, want) < 1e-14
    rng = np.random.default_rng(8)
    x = rng.standard_normal((3, 21)) + 1j * rng.standard_normal((3, 21))
    s1 = net.localize(x, DT_SMALL, gm, lp, L) - sky0
    s2 = net.localize(3.0 * x, DT_SMALL, gm, lp, L) - sky0
    assert rel(s2, 9.0 * s1) < 1e-12


# ------------------------------------------------------------------- GPU


def lib_network(beta, n=N_SMALL, dt=DT_SMALL, names="ABC"):
    net = S.Network(inject.positions(names), n, dt, responses=inject.responses(names))
    net.mult_by_projection(beta, 0.0, np.ones((n, len(names))))
    return net


@pytest.mark.gpu
@pytest.mark.parametrize("beta", [1.0, -1.0])
def test_gpu_mult_by_projection_matches_oracle(oracle, beta):
    net = lib_network(beta)
    onet = small_network(oracle, beta)
    for k in range(3):
        print(f"mult_by_projection beta={beta} baseline {k}", flush=True)
        assert rel(net.delay_product(k), onet.delay_product(k)) < 1e-12
        L = onet.info(k)["l_max"]
        assert net.power_1d(k).size == (L + 1) ** 2


@pytest.mark.gpu
def test_gpu_mult_by_projection_pixel_branch(oracle):
    # two sites at 1024 Hz: natural l = 33 (microcode) and a second
    # network at l 88 (pixel-domain products in the oracle)
    for names, n, dt in (("AB", 8, 1.0 / 1024), ("AC", 8, 1.0 / 1024)):
        net = lib_network(1.0, n, dt, names)
        onet = oracle.Network(inject.positions(names), n, dt)
        onet.mult_by_projection(det_responses(names), 1.0, 0.0, np.ones((n, 2)))
        print(f"mult_by_projection {names} l={onet.info(0)['l_max']}", flush=True)
        assert rel(net.delay_product(0), onet.delay_product(0)) < 1e-12


@pytest.mark.gpu
def test_gpu_generate_alm_skys_matches_oracle(oracle):
    L = 22
    netp, netn = lib_network(1.0), lib_network(-1.0)
    onp, onn = small_network(oracle, 1.0), small_network(oracle, -1.0)
    lp = oracle.log_uniformsky_prior(L)
    assert rel(S.log_uniformsky_prior(L), lp) < 1e-12
    rng = np.random.default_rng(9)
    x = rng.standard_normal((3, 21)) + 1j * rng.standard_normal((3, 21))
    pre = S.preprocess_snr(x, DT_SMALL)
    assert rel(pre, oracle.preprocess_snr(x, DT_SMALL)) < 1e-15
    sp, sn = S.generate_alm_skys(netp, netn, pre, 0.8, lp, L)
    # the oracle's localize applies scale_power_l; undo it on the library side
    want_p = onp.localize(x, DT_SMALL, 0.8, lp, L, power=0.0, overall=1.0)
    want_n = onn.localize(x, DT_SMALL, 0.8, lp, L, power=0.0, overall=1.0)
    assert rel(sp, want_p) < 1e-12
    assert rel(sn, want_n) < 1e-12


def _run_loc(loc, series, gmst, L, cdt):
    D, E, ln = series.shape
    d_x = S.DeviceBuffer.from_numpy(np.ascontiguousarray(series))
    d_g = S.DeviceBuffer.from_numpy(np.asarray(gmst, dtype=np.float64))
    n = 2 * (L + 1)
    esz = np.dtype(cdt).itemsize
    d_alm = S.DeviceBuffer(2 * E * (L + 1) ** 2 * esz)
    d_map = S.DeviceBuffer(2 * E * n * n * esz)
    loc.run(d_x.ptr, E, d_alm.ptr, d_map=d_map.ptr, d_gmst=d_g.ptr, series_f64=series.dtype == np.complex128)
    peaks = {}
    for seed in ("lowpass", "map"):
        d_t, d_p, d_v = (S.DeviceBuffer(2 * E * 8) for _ in range(3))
        loc.peaks(d_alm.ptr, E, d_t.ptr, d_p.ptr, d_v.ptr, d_map=d_map.ptr if seed == "map" else None)
        S.synchronize()
        peaks[seed] = [d.to_numpy(np.float64, (2, E)) for d in (d_t, d_p, d_v)]
    alm = d_alm.to_numpy(cdt, (2, E, (L + 1) ** 2))
    mp = d_map.to_numpy(cdt, (2, E, n, n))
    return alm, mp, peaks


@pytest.mark.gpu
@pytest.mark.parametrize("L", [22, 21])
def test_gpu_batched_localization_matches_oracle(oracle, local_minimum, L):
    # both parities of l_max: the mixed meshes' half-turn phi DFT gathers
    # the even and the odd orders from the ring sums by different maps
    netp, netn = lib_network(1.0), lib_network(-1.0)
    netp.set_l(L)
    netn.set_l(L)
    onp, onn = small_network(oracle, 1.0, L), small_network(oracle, -1.0, L)
    lp = oracle.log_uniformsky_prior(L)
    x, meta = inject.snr_events(6, length=21, delta_t=DT_SMALL, seed=1)
    for prec, cdt, tol in ((S.FP64, np.complex128, 1e-12), (S.MIXED, np.complex64, 1e-5)):
        loc = S.Localizer(netp, netn, lp, L, 21, DT_SMALL, precision=prec, max_events=6)
        print("localizer", loc.info(), flush=True)
        alm, mp, peaks = _run_loc(loc, x, meta["gmst"], L, cdt)
        # without meshes the a_lm come from the epilogue kernel, with them
        # from the ring-sum kernel as it reads the product: the same values
        d_x = S.DeviceBuffer.from_numpy(np.ascontiguousarray(x))
        d_g = S.DeviceBuffer.from_numpy(np.asarray(meta["gmst"], dtype=np.float64))
        d_a = S.DeviceBuffer(2 * 6 * (L + 1) ** 2 * np.dtype(cdt).itemsize)
        loc.run(d_x.ptr, 6, d_a.ptr, d_map=None, d_gmst=d_g.ptr, series_f64=x.dtype == np.complex128)
        S.synchronize()
        assert np.array_equal(d_a.to_numpy(cdt, (2, 6, (L + 1) ** 2)), alm)
        for e in range(6):
            want = [o.localize(x[:, e].astype(np.complex128), DT_SMALL, meta["gmst"][e], lp, L) for o in (onp, onn)]
            for bi in range(2):
                assert rel(alm[bi, e], want[bi]) < tol, (prec, e, bi)
                assert rel(mp[bi, e], oracle.to_mesh(want[bi], L)) < (1e-12 if prec == S.FP64 else 1e-5)
                if prec != S.FP64:
                    continue
                # reference seeding: a local maximum reached by ascent from
                # the reference's l = 6 seed
                t, p, v = (peaks["lowpass"][q][bi, e] for q in range(3))
                tol = 1e-9 * max(1.0, abs(v))
                local_minimum(-want[bi], L, t, p, -v, tol)
                # map seeding: a local maximum never below the map's maximum
                t, p, v = (peaks["map"][q][bi, e] for q in range(3))
                local_minimum(-want[bi], L, t, p, -v, tol, seed=False)
                assert v >= mp[bi, e].real.max() - tol


DT_1K, LEN_1K, L_CFG = 1.0 / 1024, 103, 50


@pytest.fixture(scope="module")
def plans_1k(oracle):
    """Plans at the localization order l_max = 50: three sites at 1024 Hz,
    +-50 ms series (103 samples, 147 bins after the light-travel padding),
    natural orders 33, 88 and 85, so two baselines are truncated to 50 by
    set_l after the projection multiply (pixel-domain products, as at
    4096 Hz) and one stays below it (microcode products)."""
    n = oracle.precalc_length(LEN_1K, DT_1K)
    out = {}
    for beta in (1.0, -1.0):
        onet = oracle.Network(inject.positions("ABC"), n, DT_1K)
        onet.mult_by_projection(det_responses("ABC"), beta, 0.0, np.ones((n, 3)))
        assert sorted(onet.info(k)["l_max"] for k in range(3)) == [33, 85, 88]
        onet.set_l(L_CFG)
        net = lib_network(beta, n, DT_1K)
        net.set_l(L_CFG)
        out[beta] = (net, onet)
    return n, out


@pytest.mark.gpu
def test_gpu_batched_localization_l50_matches_oracle(oracle, plans_1k):
    n, nets = plans_1k
    netp, onp = nets[1.0]
    netn, onn = nets[-1.0]
    for k in range(3):
        assert rel(netp.delay_product(k), onp.delay_product(k)) < 1e-12, k
        assert rel(netn.delay_product(k), onn.delay_product(k)) < 1e-12, k
    L = L_CFG
    lp = oracle.log_uniformsky_prior(L)
    E = 4
    x, meta = inject.snr_events(E, length=LEN_1K, delta_t=DT_1K, seed=21)
    for prec, cdt, tol in ((S.FP64, np.complex128, 1e-12), (S.MIXED, np.complex64, 1e-5)):
        loc = S.Localizer(netp, netn, lp, L, LEN_1K, DT_1K, precision=prec, max_events=E)
        print("localizer l=50", prec, loc.info(), flush=True)
        alm, mp, _ = _run_loc(loc, x, meta["gmst"], L, cdt)
        for e in range(E):
            want = [o.localize(x[:, e].astype(np.complex128), DT_1K, meta["gmst"][e], lp, L) for o in (onp, onn)]
            for bi in range(2):
                assert rel(alm[bi, e], want[bi]) < tol, (prec, e, bi, rel(alm[bi, e], want[bi]))
                assert rel(mp[bi, e], oracle.to_mesh(want[bi], L)) < tol, (prec, e, bi)
