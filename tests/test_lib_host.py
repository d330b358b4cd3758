"""CPU tests of the C library: the ABI it exports, and the host-side plan
construction and object arithmetic against the oracle.  No GPU calls."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import sphradiometer_amd as S
from sphradiometer_amd import inject

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include", "sphradiometer")


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(INC)):
        txt = open(os.path.join(INC, fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", txt, flags=re.M):
            line = m.group(0)
            if line.startswith("static") or "typedef" in line or "#define" in line:
                continue
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    L = S.lib()
    names = declared_functions()
    assert len(names) > 100
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_offset_helpers_agree_with_oracle(oracle):
    # the header-inline helpers are ABI: check the same arithmetic through a
    # series object's layout
    L = S.lib()
    for lmax in (0, 1, 5, 32):
        s = L.sh_series_new(lmax, 0)
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                v = S.CDouble(l + 0.5, m)
                L.sh_series_set(s, l, m, v)
        c = S.coeff_view(s)
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                assert c[oracle.lmoffset(lmax, l, m)] == complex(l + 0.5, m)
        L.sh_series_free(s)


def test_wigner_3j_against_exact(oracle):
    pins = np.load(os.path.join(ROOT, "tests", "golden", "pins.npz"))
    L = S.lib()
    for a, v in zip(pins["threej_args"], pins["threej_vals"]):
        got = L.sh_series_wigner_3j(*[int(x) for x in a])
        assert abs(got - v) < 2e-14 * max(1.0, abs(v)) + 1e-15, (a, got, v)


def test_gauss_legendre_mesh(oracle):
    L = S.lib()
    for lmax in (3, 25, 32, 64):
        nt, nph = ctypes.c_int(), ctypes.c_int()
        ct = ctypes.POINTER(ctypes.c_double)()
        wt = ctypes.POINTER(ctypes.c_double)()
        m = L.sh_series_mesh_new if False else None
        fn = S.lib().sh_series_mesh_new
        fn.restype = ctypes.c_void_p
        fn.argtypes = [ctypes.c_uint, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                       ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.POINTER(ctypes.POINTER(ctypes.c_double))]
        mesh = fn(lmax, ctypes.byref(nt), ctypes.byref(nph), ctypes.byref(ct), ctypes.byref(wt))
        n = nt.value
        x = np.ctypeslib.as_array(ct, shape=(n,)).copy()
        w = np.ctypeslib.as_array(wt, shape=(n,)).copy()
        ox, ow = oracle.mesh_nodes(lmax)
        assert np.max(np.abs(x - ox)) < 2e-16 * 2
        assert np.max(np.abs(w - ow) / ow) < 1e-14
        libc = ctypes.CDLL(None)
        for p in (mesh, ctypes.cast(ct, ctypes.c_void_p).value, ctypes.cast(wt, ctypes.c_void_p).value):
            libc.free(ctypes.c_void_p(p))


@pytest.mark.parametrize("names,n,L", [("AB", 4096, 25), ("ABC", 1024, 32), ("ABC", 512, 9)])
def test_plan_matches_oracle(oracle, names, n, L):
    dt = 1.0 / 4096 if n != 512 else 1.0 / 512
    pos = inject.positions(names)
    net = S.Network(pos, n, dt, L)
    onet = oracle.Network(pos, n, dt, L)
    for k in range(net.n_baselines):
        p = net.baseline_plan(k).contents
        inf = onet.info(k)
        assert p.transient == inf["transient"]
        assert p.power_1d.contents.l_max == inf["l_max"]
        a, b = net.delay_product(k), onet.delay_product(k)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-13
        D = S.rotation_blocks(p.rotation_plan)
        flat = np.concatenate([D[l].ravel() for l in range(1, inf["rot_l_max"] + 1)]) if inf["rot_l_max"] else np.zeros(0)
        ob = onet.wigner_blocks(k)
        if ob.size:
            assert np.linalg.norm(flat - ob) / np.linalg.norm(ob) < 1e-13
        th, ph = p.baseline.contents.theta, p.baseline.contents.phi
        g = onet.geometry(k)
        assert abs(th - g[3]) < 1e-15 and abs(ph - g[4]) < 1e-15


def test_projection_matrix_delay_matches_oracle(oracle):
    L = S.lib()
    r = 0.0136
    dt = 1.0 / 4096
    n = L.projection_matrix_n_elements(r, dt)
    lm = L.projection_matrix_l_max(r, dt)
    v = (S.GslVector)()
    data = (ctypes.c_double * 3)(0.0, 0.0, r)
    v.size, v.stride, v.data, v.block, v.owner = 3, 1, ctypes.cast(data, ctypes.POINTER(ctypes.c_double)), None, 0
    arr = L.projection_matrix_delay(n, lm, ctypes.byref(v), dt)
    got = S.array_view(arr).copy()
    ref = oracle.projection_matrix_delay(n, lm, [0, 0, r], dt)
    assert np.max(np.abs(got - ref)) < 1e-14
    L.sh_series_array_free(arr)


def test_series_arithmetic_matches_oracle(oracle):
    L = S.lib()
    rng = np.random.default_rng(0)
    lmax = 9
    c = rng.standard_normal((lmax + 1) ** 2) + 1j * rng.standard_normal((lmax + 1) ** 2)
    for fn, ofn in [("sh_series_conj", lambda x: oracle.series_conj(x, lmax)),
                    ("sh_series_real", lambda x: oracle.series_real(x, lmax)),
                    ("sh_series_scale_power_l", lambda x: oracle.scale_power_l(x, lmax, 0.6, 2.7))]:
        s = S.series_from_numpy(c, lmax, 0)
        if fn == "sh_series_scale_power_l":
            L.sh_series_scale_power_l(s, 0.6, 2.7)
        else:
            getattr(L, fn)(s)
        assert np.max(np.abs(S.coeff_view(s) - ofn(c))) < 1e-15
        L.sh_series_free(s)
    # resize down and up preserves the function (sh_series.c:200-233)
    s = S.series_from_numpy(c, lmax, 0)
    L.sh_series_resize(s, 4)
    L.sh_series_resize(s, 12)
    got = S.coeff_view(s)
    for l in range(13):
        for m in range(-l, l + 1):
            want = c[oracle.lmoffset(lmax, l, m)] if l <= 4 else 0
            assert got[oracle.lmoffset(12, l, m)] == want
    L.sh_series_free(s)


def test_array_set_l_matches_oracle(oracle):
    L = S.lib()
    rng = np.random.default_rng(1)
    for polar in (0, 1):
        arr = L.sh_series_array_new(7, 11, polar)
        A = S.array_view(arr)
        A[:] = rng.standard_normal(A.shape) + 1j * rng.standard_normal(A.shape)
        before = A.copy()
        L.sh_series_array_set_l(arr, 5)
        assert np.array_equal(S.array_view(arr), oracle.array_set_l(before, 11, polar, 5))
        L.sh_series_array_free(arr)


@pytest.mark.parametrize("l1,l2,polar", [(5, 10, 0), (6, 9, 1), (40, 70, 1)])
def test_products_match_oracle(oracle, l1, l2, polar):
    # microcode below order 60, pixel domain above (sh_series_product.c:199);
    # the polar pixel product runs on the host, so this needs no GPU
    L = S.lib()
    rng = np.random.default_rng(l1 + l2)
    na = l1 + 1 if polar else (l1 + 1) ** 2
    nb_ = l2 + 1 if polar else (l2 + 1) ** 2
    a = rng.uniform(-1, 1, na) + 1j * rng.uniform(-1, 1, na)
    b = rng.uniform(-1, 1, nb_) + 1j * rng.uniform(-1, 1, nb_)
    sa, sb = S.series_from_numpy(a, l1, polar), S.series_from_numpy(b, l2, polar)
    d = L.sh_series_new(l1 + l2, polar)
    plan = L.sh_series_product_plan_new(d, sa, sb)
    assert L.sh_series_product(d, sa, sb, plan)
    ref = oracle.product(a, l1, polar, b, l2, polar, l1 + l2, polar)
    assert np.max(np.abs(S.coeff_view(d) - ref)) < 1e-13 * (l1 + l2)
    if l1 + l2 < 60:
        assert plan.contents.plan_length == oracle.lib().orx_product_plan_length(l1 + l2, polar, l1, polar, l2, polar)
    for p in (sa, sb, d):
        L.sh_series_free(p)
    L.sh_series_product_plan_free(plan)


def test_error_conventions():
    # correlator.c:571-591: raising l_max fails; correlator.c:68-76: bad
    # delta_t gives (unsigned) -1; sh_series_array.c:443-444: mismatched dotc
    L = S.lib()
    net = S.Network(inject.positions("AB"), 256, 1.0 / 512, 9)
    p = net.baseline_plan(0)
    assert not L.correlator_plan_fd_set_l(p, 20)
    bl = p.contents.baseline
    assert L.correlator_baseline_power_l_max(bl, -1.0) == 2 ** 32 - 1
    arr = L.sh_series_array_new(3, 4, 1)
    res = L.sh_series_new(5, 1)
    assert not L.sh_series_array_dotc(res, arr, None)
    L.sh_series_free(res)
    L.sh_series_array_free(arr)
    ia = L.instrument_array_new(1)
    assert not L.instrument_array_set(ia, 5, L.instrument_new(0, 0, 0, None))
    L.instrument_array_free(ia)


def test_windows():
    L = S.lib()
    w = L.correlator_tukey_window_new(32, 2, 5, 1.0)
    a = np.ctypeslib.as_array(w, shape=(32,)).copy()
    assert np.allclose(a, a[::-1]) and a[0] == 0 and a[16] == 1.0
    ctypes.CDLL(None).free(ctypes.cast(w, ctypes.c_void_p))
    assert not L.correlator_tukey_window_new(10, 3, 3, 1.0)
