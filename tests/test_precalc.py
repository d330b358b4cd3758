"""The precalc tree of the localization plans (inspiral_sky_map.c:1113-1811
of the reference; RapidLocalization.py:300-352 writes and reads it).

The reference holds no precalc tree among its files, so the format is
pinned three ways:
- the folder names against make_precalc_directories
  (inspiral_sky_map.c:1458-1553);
- the bytes of every .dat file against a reader written here with numpy
  from the reference's fwrite sequences (field order and types of
  inspiral_sky_map.c:1556-1811): baseline indices and direction, the
  baseline vector, delta_t and transient, the packed delay products, the
  length;
- the element order of the stored Wigner-D blocks against mathematics: the
  m = 0 row of D^l for the rotation to (theta, phi) is
  sqrt(4 pi / (2l+1)) conj(Y_lm'(theta, phi)), read from the file at
  (2l+1)(m'+l) + l, and every element against the public accessor
  sh_series_rotation_plan_wigner_D (sh_series_rotation.c:523-529).
Then a write -> read round trip reproduces every plan field bit for bit,
and on the GPU the plans read back give the same localization skies as the
plans they were written from.
"""
import ctypes
import math
import os

import numpy as np
import pytest

import sphradiometer_amd as S
import sphradiometer_amd.sphradiometer as sph
from sphradiometer_amd import inject
from sphradiometer_amd.healpix import _full_coefficients, _lm_offset, healpy_alm_length

DT, N = 1.0 / 256, 32
PLAN = "correlator_network_plan_fd"
BL = PLAN + "/correlator_network_baselines"


def _b(p):
    return str(p).encode()


def _fields(path, fmt):
    """the fields of a raw .dat file in struct order, and the rest"""
    raw = open(path, "rb").read()
    dt = np.dtype([("f%d" % i, c) for i, c in enumerate(fmt)])
    head = np.frombuffer(raw[: dt.itemsize], dtype=dt)[0]
    return tuple(head), raw[dt.itemsize:]


def _tree_dirs(root):
    out = set()
    for d, sub, _ in os.walk(root):
        for s in sub:
            out.add(os.path.relpath(os.path.join(d, s), root))
    return out


def _expected_dirs(n_ifo, n_k):
    e = {PLAN, BL, BL + "/baselines", PLAN + "/correlator_plan_fd", "autocorrelator_network_plan_fd",
         "autocorrelator_network_plan_fd/psh_series", "sh_series"}
    for k in range(n_k):
        e.add(BL + "/baselines/%d" % k)
        c = PLAN + "/correlator_plan_fd/%d" % k
        e |= {c, c + "/delay_product_n", c + "/delay_product_p", c + "/rotation_plan"}
    for k in range(n_ifo):
        a = "autocorrelator_network_plan_fd/psh_series/%d" % k
        e |= {a, a + "/sh_series_n", a + "/sh_series_p"}
    return e


def test_precalc_directories(tmp_path):
    L = S.lib()
    # three detectors: the reference's layout exactly (n_ifo folders of each
    # kind); four: six baseline folders, one per baseline
    for n_ifo, n_k in ((3, 3), (2, 2), (4, 6)):
        root = tmp_path / ("net%d" % n_ifo)
        assert L.make_precalc_directories(_b(root), n_ifo) == 0
        assert _tree_dirs(root) == _expected_dirs(n_ifo, n_k)
    # an existing folder is an error, as mkdir makes it in the reference
    assert L.make_precalc_directories(_b(tmp_path / "net3"), 3) == -1
    assert "error making" in S.last_error()


class _Net(S.Network):
    def scale_delay_products(self, f):
        for k in range(self.n_baselines):
            S.array_view(self.baseline_plan(k).contents.delay_product)[:] *= f


def _write_tree(root, netp, netn, logprior_l=8, length=N):
    L = S.lib()
    assert L.make_precalc_directories(_b(root), netp.n_instruments) == 0
    assert L.write_precalc_time_series_length(length, _b(root)) == 0
    lp = L.sh_series_log_uniformsky_prior(logprior_l)
    assert L.write_precalc_logprior(lp, _b(root)) == 0, S.last_error()
    assert L.write_precalc_correlator_network_plan_fd(netp.plan, netn.plan, _b(root)) == 0, S.last_error()
    return lp


def _read_tree(root, inst):
    """the reference's read sequence (RapidLocalization.py:320-352)"""
    L = S.lib()
    n = sph.new_uintp()
    assert L.read_precalc_time_series_length(n, _b(root)) == 0
    length = sph.uintp_value(n)
    sph.delete_uintp(n)
    bl = L.read_precalc_correlator_network_baselines(inst, _b(root))
    assert bl, S.last_error()
    fp, fn = sph.new_correlator_network_plan_fdp(), sph.new_correlator_network_plan_fdp()
    assert L.read_precalc_correlator_network_plan_fd(fp, fn, bl, length, _b(root)) == 0, S.last_error()
    return length, bl, fp, fn


def _free_read(bl, fp, fn):
    L = S.lib()
    L.correlator_network_plan_fd_free(fp)
    L.correlator_network_plan_fd_free(fn)
    L.correlator_network_baselines_free(bl)


def _pair(names="ABC"):
    pos = inject.positions(names)
    netp, netn = _Net(pos, N, DT), _Net(pos, N, DT)
    # the two polarisations' delay products must differ for the test to see
    # which file each one went to
    netn.scale_delay_products(-0.5)
    return netp, netn


def test_precalc_raw_layout(tmp_path, oracle):
    netp, netn = _pair()
    root = tmp_path / "tree"
    _write_tree(root, netp, netn, length=N)
    (length,), rest = _fields(root / "time_series_length.dat", ["<u4"])
    assert length == N and rest == b""
    (nb,), rest = _fields(root / BL / "n_baselines.dat", ["<i4"])
    assert nb == 3 and rest == b""
    for k in range(nb):
        p = netp.baseline_plan(k).contents
        b = p.baseline.contents
        (ia, ib, th, ph), rest = _fields(root / BL / "baselines" / str(k) / "correlator_baseline.dat", ["<i4", "<i4", "<f8", "<f8"])
        assert (ia, ib, th, ph) == (b.index_a, b.index_b, b.theta, b.phi) and rest == b""
        d = np.fromfile(root / BL / "baselines" / str(k) / "d.dat", dtype="<f8")
        want = inject.positions("ABC")[ia] - inject.positions("ABC")[ib]
        assert d.shape == (3,) and np.array_equal(d, want)
        c = root / PLAN / "correlator_plan_fd" / str(k)
        (dt, tr), rest = _fields(c / "correlator_plan_fd.dat", ["<f8", "<i4"])
        assert dt == DT and tr == p.transient and rest == b""
        (lr,), rest = _fields(c / "rotation_plan" / "l_max.dat", ["<u4"])
        assert lr == p.rotation_plan.contents.l_max and rest == b""
        D = np.fromfile(c / "rotation_plan" / "D.dat", dtype="<c16")
        assert D.size == sum((2 * l + 1) ** 2 for l in range(1, lr + 1))
        off = 0
        for l in range(1, lr + 1):
            w = 2 * l + 1
            blk = D[off: off + w * w]
            off += w * w
            for mp in range(-l, l + 1):
                # m = 0 row: the rotated polar impulse
                want = math.sqrt(4 * math.pi / w) * np.conj(oracle.Y(l, mp, b.theta, b.phi))
                assert abs(blk[w * (mp + l) + l] - want) < 1e-12
                for m in range(-l, l + 1):
                    v = S.lib().sh_series_rotation_plan_wigner_D(p.rotation_plan, l, m, mp)
                    assert blk[w * (mp + l) + m + l] == complex(v.re, v.im)
        for sign, net in (("p", netp), ("n", netn)):
            a = net.baseline_plan(k).contents.delay_product.contents
            (lm, pol, n, stride), rest = _fields(c / ("delay_product_" + sign) / "sh_series_array.dat", ["<u4", "<i4", "<i4", "<i4"])
            assert (lm, pol, n, stride) == (a.l_max, a.polar, a.n, a.stride)
            assert np.array_equal(np.frombuffer(rest, dtype="<c16").reshape(n, stride), net.delay_product(k))


def test_precalc_round_trip(tmp_path):
    L = S.lib()
    netp, netn = _pair()
    root = tmp_path / "tree"
    lp = _write_tree(root, netp, netn, length=N)
    length, bl, fp, fn = _read_tree(root, netp._inst)
    assert length == N
    assert bl.contents.n_baselines == netp.n_baselines
    for k in range(netp.n_baselines):
        b0, b1 = netp.baseline_plan(k).contents.baseline.contents, bl.contents.baselines[k].contents
        assert (b0.index_a, b0.index_b, b0.theta, b0.phi) == (b1.index_a, b1.index_b, b1.theta, b1.phi)
        for net, rd in ((netp, fp), (netn, fn)):
            p0, p1 = net.baseline_plan(k).contents, rd.contents.plans[k].contents
            assert (p0.delta_t, p0.transient) == (p1.delta_t, p1.transient)
            assert ctypes.addressof(p1.baseline.contents) == ctypes.addressof(b1)
            a0, a1 = p0.delay_product.contents, p1.delay_product.contents
            assert (a0.l_max, a0.polar, a0.n, a0.stride) == (a1.l_max, a1.polar, a1.n, a1.stride)
            assert np.array_equal(S.array_view(p0.delay_product), S.array_view(p1.delay_product))
            r0, r1 = S.rotation_blocks(p0.rotation_plan), S.rotation_blocks(p1.rotation_plan)
            assert sorted(r0) == sorted(r1) and all(np.array_equal(r0[l], r1[l]) for l in r0)
            pw = p1.power_1d.contents
            assert (pw.l_max, pw.polar) == (p1.rotation_plan.contents.l_max, a1.polar)
            assert not np.any(S.coeff_view(p1.power_1d))
        # the two polarisations own separate rotation plans
        assert ctypes.addressof(fp.contents.plans[k].contents.rotation_plan.contents) != ctypes.addressof(fn.contents.plans[k].contents.rotation_plan.contents)
    # the log prior goes through the a_lm FITS table: a real function, so
    # its coefficients come back to round-off
    rp = L.read_precalc_logprior(_b(root))
    assert rp, S.last_error()
    a, b = S.coeff_view(lp), S.coeff_view(rp)
    assert rp.contents.l_max == lp.contents.l_max
    n = min(a.size, b.size)
    assert np.max(np.abs(a[:n] - b[:n])) < 1e-15 * np.max(np.abs(a)) and not np.any(b[n:])
    L.sh_series_free(rp)
    L.sh_series_free(lp)
    _free_read(bl, fp, fn)
    # the plans read back serve set_l as freshly built ones do
    length, bl, fp, fn = _read_tree(root, netp._inst)
    lmin = min(fp.contents.plans[k].contents.power_1d.contents.l_max for k in range(3))
    assert L.correlator_network_plan_fd_set_l(fp, lmin - 1)
    netp.set_l(lmin - 1)
    for k in range(3):
        assert np.array_equal(S.array_view(fp.contents.plans[k].contents.delay_product), netp.delay_product(k))
    _free_read(bl, fp, fn)


def test_precalc_read_rejects(tmp_path):
    L = S.lib()
    netp, netn = _pair()
    root = tmp_path / "tree"
    lp = _write_tree(root, netp, netn)
    L.sh_series_free(lp)
    # a stored direction the instruments do not reproduce
    f = root / BL / "baselines" / "1" / "correlator_baseline.dat"
    good = f.read_bytes()
    bad = bytearray(good)
    bad[8:16] = np.float64(np.frombuffer(good[8:16], "<f8")[0] + 1e-9).tobytes()
    f.write_bytes(bytes(bad))
    assert not L.read_precalc_correlator_network_baselines(netp._inst, _b(root))
    assert "inconsistent" in S.last_error()
    # an instrument index outside the array
    bad = bytearray(good)
    bad[0:4] = np.int32(7).tobytes()
    f.write_bytes(bytes(bad))
    assert not L.read_precalc_correlator_network_baselines(netp._inst, _b(root))
    assert "outside" in S.last_error()
    f.write_bytes(good)
    bl = L.read_precalc_correlator_network_baselines(netp._inst, _b(root))
    assert bl
    # a truncated delay product
    g = root / PLAN / "correlator_plan_fd" / "2" / "delay_product_n" / "sh_series_array.dat"
    raw = g.read_bytes()
    g.write_bytes(raw[:-16])
    fp, fn = sph.new_correlator_network_plan_fdp(), sph.new_correlator_network_plan_fdp()
    assert L.read_precalc_correlator_network_plan_fd(fp, fn, bl, N, _b(root)) == -1
    assert "short" in S.last_error()
    assert not fp.contents.plans and not fn.contents.plans
    sph.delete_correlator_network_plan_fdp(fp)
    sph.delete_correlator_network_plan_fdp(fn)
    g.write_bytes(raw)
    L.correlator_network_baselines_free(bl)
    # missing files
    assert L.read_precalc_time_series_length(sph.new_uintp(), _b(tmp_path / "none")) == -1
    assert "No such file" in S.last_error()
    assert not L.read_precalc_logprior(_b(tmp_path / "none"))


def test_precalc_autocorrelator_round_trip(tmp_path):
    L = S.lib()
    rng = np.random.default_rng(4)
    pos = inject.positions("AB")
    net = S.Network(pos, N, DT)
    root = tmp_path / "tree"
    assert L.make_precalc_directories(_b(root), 2) == 0
    lmax, length = 6, 3
    keep = []

    def plan(sign):
        rows = (ctypes.POINTER(ctypes.POINTER(S.ShSeries)) * 2)()
        for i in range(2):
            r = (ctypes.POINTER(S.ShSeries) * length)()
            for k in range(length):
                # a real function: a_{l,-m} = (-1)^m conj(a_{l,m})
                alm = rng.standard_normal(healpy_alm_length(lmax, lmax)) * (1 + 0j)
                alm[lmax + 1:] += 1j * rng.standard_normal(alm.size - lmax - 1)
                r[k] = S.series_from_numpy(_full_coefficients(alm, lmax, lmax) * sign, lmax, 0)
            keep.append(r)
            rows[i] = ctypes.cast(r, ctypes.POINTER(ctypes.POINTER(S.ShSeries)))
        keep.append(rows)
        a = S.AutoPlanFD()
        a.instruments = net._inst
        a.projections = ctypes.cast(rows, ctypes.POINTER(ctypes.POINTER(ctypes.POINTER(S.ShSeries))))
        a.length = length
        keep.append(a)
        return a

    ap, an = plan(1.0), plan(-2.0)
    assert L.write_precalc_autocorrelator_network_plan_fd(ctypes.byref(ap), ctypes.byref(an), _b(root)) == 0, S.last_error()
    (ln,), _ = _fields(root / "autocorrelator_network_plan_fd" / "length.dat", ["<i4"])
    assert ln == length
    rp, rn = sph.new_autocorrelator_network_plan_fdp(), sph.new_autocorrelator_network_plan_fdp()
    assert L.read_precalc_autocorrelator_network_plan_fd(rp, rn, net._inst, N, _b(root)) == 0, S.last_error()
    for a, r in ((ap, rp), (an, rn)):
        assert r.contents.length == length
        for i in range(2):
            for k in range(length):
                x, y = S.coeff_view(a.projections[i][k]), S.coeff_view(r.contents.projections[i][k])
                assert np.max(np.abs(x - y)) < 1e-15 * np.max(np.abs(x))
    L.autocorrelator_network_plan_fd_free(rp)
    L.autocorrelator_network_plan_fd_free(rn)
    for a in (ap, an):
        for i in range(2):
            for k in range(length):
                L.sh_series_free(a.projections[i][k])
    # a missing projection file fails the read and leaves no projections
    os.remove(root / "autocorrelator_network_plan_fd/psh_series/1/sh_series_n/projections2.fits")
    rp, rn = sph.new_autocorrelator_network_plan_fdp(), sph.new_autocorrelator_network_plan_fdp()
    assert L.read_precalc_autocorrelator_network_plan_fd(rp, rn, net._inst, N, _b(root)) == -1
    assert not rp.contents.projections and not rn.contents.projections
    sph.delete_autocorrelator_network_plan_fdp(rp)
    sph.delete_autocorrelator_network_plan_fdp(rn)


def _auto_plans(beta, lmax, psd, names="ABC"):
    L = S.lib()
    net = S.Network(inject.positions(names), N, DT, responses=inject.responses(names))
    rows = (ctypes.POINTER(ctypes.c_double) * len(psd))(*[r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for r in psd])
    a = L.autocorrelator_network_plan_fd_new(net._inst, beta, 0.3, rows, len(psd), lmax)
    assert a, S.last_error()
    return net, a


def test_autocorrelator_plan_against_oracle(oracle):
    """projections[i][f] is the (i, i) element of the projection operator
    (inspiral_sky_map.c:561-660): the expansion converges to the oracle's
    pointwise operator (largest error over random points below 3e-3 at
    l 16, 3e-5 at 32, 1e-7 at 64), the diagonal sums to the trace 1 - n exactly (a constant: only
    a_00 = (1 - n) sqrt(4 pi)), and each element is a real function."""
    L = S.lib()
    rng = np.random.default_rng(5)
    psd = rng.uniform(0.5, 2.0, (2, 3))
    R = inject.responses("ABC").reshape(3, 9)
    for lmax, tol in ((16, 3e-3), (32, 3e-5), (64, 1e-7)):
        net, a = _auto_plans(-1.0, lmax, psd)
        assert a.contents.length == 2
        err = 0.0
        for _ in range(12):
            th, ph = rng.uniform(0.05, 3.1), rng.uniform(0, 2 * math.pi)
            for i in range(3):
                v = L.sh_series_eval(a.contents.projections[i][1], th, ph)
                err = max(err, abs(complex(v.re, v.im) - oracle.cbc_projection(th, ph, i, i, R, -1.0, 0.3, psd[1])))
        assert err < tol, (lmax, err)
        tot = sum(S.coeff_view(a.contents.projections[i][0]) for i in range(3))
        assert abs(tot[0] - (1 - 3) * math.sqrt(4 * math.pi)) < 1e-13 and np.max(np.abs(tot[1:])) < 1e-13
        c = S.coeff_view(a.contents.projections[2][0])
        full = _full_coefficients(c[: healpy_alm_length(lmax, lmax)], lmax, lmax)
        assert np.max(np.abs(c - full)) < 1e-14
        # set_l keeps the coefficients of l <= 9
        idx = [(l, m) for l in range(10) for m in range(-l, l + 1)]
        lead = [[S.coeff_view(a.contents.projections[i][f])[_lm_offset(lmax, l, m)] for l, m in idx] for i in range(3) for f in range(2)]
        assert L.autocorrelator_network_plan_fd_set_l(a, 9)
        got = [[S.coeff_view(a.contents.projections[i][f])[_lm_offset(9, l, m)] for l, m in idx] for i in range(3) for f in range(2)]
        assert lead == got and a.contents.projections[0][0].contents.l_max == 9
        L.autocorrelator_network_plan_fd_free(a)
        del net


def test_autocorrelator_plans_through_precalc(tmp_path):
    """the reference's RapidLocalization write / read of the
    autocorrelator plans (RapidLocalization.py:300-352) with plans built
    by autocorrelator_network_plan_fd_new"""
    L = S.lib()
    psd = np.random.default_rng(6).uniform(0.5, 2.0, (3, 3))
    net, ap = _auto_plans(1.0, 12, psd)
    _, an = _auto_plans(-1.0, 12, psd)
    root = tmp_path / "tree"
    assert L.make_precalc_directories(_b(root), 3) == 0
    assert L.write_precalc_autocorrelator_network_plan_fd(ap, an, _b(root)) == 0, S.last_error()
    rp, rn = sph.new_autocorrelator_network_plan_fdp(), sph.new_autocorrelator_network_plan_fdp()
    assert L.read_precalc_autocorrelator_network_plan_fd(rp, rn, net._inst, N, _b(root)) == 0, S.last_error()
    for a, r in ((ap, rp), (an, rn)):
        for i in range(3):
            for f in range(3):
                x, y = S.coeff_view(a.contents.projections[i][f]), S.coeff_view(r.contents.projections[i][f])
                assert np.max(np.abs(x - y)) < 1e-15 * np.max(np.abs(x))
    for a in (ap, an, rp, rn):
        L.autocorrelator_network_plan_fd_free(a)


# ------------------------------------------------------------------- GPU


class _Plan:
    def __init__(self, p):
        self.plan = p


@pytest.mark.gpu
def test_gpu_precalc_localization_plans(tmp_path):
    """plans after the projection multiply, saved and read back, give the
    skies of the plans they came from"""
    print("precalc localization plans", flush=True)
    names, L = "ABC", 22
    netp, netn = (S.Network(inject.positions(names), N, DT, responses=inject.responses(names)) for _ in range(2))
    netp.mult_by_projection(1.0, 0.0, np.ones((N, 3)))
    netn.mult_by_projection(-1.0, 0.0, np.ones((N, 3)))
    root = tmp_path / "tree"
    lp = _write_tree(root, netp, netn, logprior_l=L)
    S.lib().sh_series_free(lp)
    length, bl, fp, fn = _read_tree(root, netp._inst)
    assert length == N
    prior = S.lib().read_precalc_logprior(_b(root))
    lp_read = S.coeff_view(prior).copy()
    S.lib().sh_series_free(prior)
    lp_full = S.log_uniformsky_prior(L)
    assert lp_read.size == lp_full.size and np.max(np.abs(lp_read - lp_full)) < 1e-15 * np.max(np.abs(lp_full))
    rng = np.random.default_rng(9)
    x = rng.standard_normal((3, 21)) + 1j * rng.standard_normal((3, 21))
    pre = S.preprocess_snr(x, DT)
    want = S.generate_alm_skys(netp, netn, pre, 0.8, lp_read, L)
    got = S.generate_alm_skys(_Plan(fp), _Plan(fn), pre, 0.8, lp_read, L)
    for g, w in zip(got, want):
        assert np.array_equal(g, w) or np.linalg.norm(g - w) / np.linalg.norm(w) < 1e-15
    _free_read(bl, fp, fn)
