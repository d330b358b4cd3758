"""HEALPix output (sh_series_healpix.cc, healpix.py, RapidLocalization.py:
498-530 of the reference).

healpix_cxx and healpy are not in the reference tree and not importable
here, so the oracle (oracle/or_healpix.c) is pinned on what the HEALPix
definition fixes: the nside rules of both reference files, the exact
centres of the nside = 1 map, the ring structure (4 nside - 1 rings, ring
sizes, z), the agreement of two independent constructions of the centres
(index inversion in the oracle, ring enumeration in the library), and
closed forms of Y_00, Y_10 and Y_11 (normalisation and Condon-Shortley
phase).  The FITS files are checked by an independent reader written
here.  GPU maps and posteriors are compared with the oracle: 1e-12
relative in float64, 1e-5 in float32.
"""
import math
import os

import numpy as np
import pytest

import sphradiometer_amd as S
from sphradiometer_amd import healpix as H

TOL64 = 1e-12
TOL_MIXED = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def rand_series(rng, L, polar=False):
    n = L + 1 if polar else (L + 1) ** 2
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


# ---------------------------------------------------------------- oracle pins


def test_nside_rules(oracle):
    # lmax2nside (sh_series_healpix.cc:220-238): smallest power of two
    # >= ceil(l_max / 2 + 0.75)
    assert [oracle.healpix_nside_c(l) for l in (0, 1, 2, 3, 16, 25, 32, 50, 64, 100)] == [1, 2, 2, 4, 16, 16, 32, 32, 64, 64]
    # healpy_alm_to_map (healpix.py:163): 2^ceil(log2(l_max / 2))
    assert [oracle.healpix_nside_py(l) for l in (2, 3, 16, 25, 32, 50, 64)] == [1, 2, 8, 16, 16, 32, 32]
    S.lib()
    for l in range(0, 400):
        assert S.lib().sphr_healpix_nside(l) == oracle.healpix_nside_c(l)
        if l >= 2:
            assert H.nside_for_map(l) == oracle.healpix_nside_py(l)


def test_pix2ang_pins(oracle):
    th, ph = oracle.healpix_pix2ang(1)
    z = np.cos(th)
    np.testing.assert_allclose(z, [2 / 3] * 4 + [0] * 4 + [-2 / 3] * 4, atol=1e-15)
    q = math.pi / 4
    np.testing.assert_allclose(ph, [q, 3 * q, 5 * q, 7 * q, 0, 2 * q, 4 * q, 6 * q, q, 3 * q, 5 * q, 7 * q], atol=1e-15)
    for N in (1, 2, 4, 8, 16):
        th, ph = oracle.healpix_pix2ang(N)
        assert th.size == 12 * N * N
        zs, counts = np.unique(np.round(np.cos(th), 12), return_counts=True)
        assert zs.size == 4 * N - 1
        expect = [4 * min(i, N, 4 * N - i) for i in range(1, 4 * N)]
        assert list(counts[::-1]) == expect
        assert abs(np.cos(th).sum()) < 1e-9
        assert (ph >= 0).all() and (ph < 2 * math.pi).all()
        # centres are ordered ring by ring from the north pole, phi
        # ascending within a ring
        assert (np.diff(th) >= -1e-15).all()
        # the library's ring enumeration gives the same centres
        lt, lp = H.pix2ang(N)
        assert np.abs(lt - th).max() < 1e-15 and np.abs(lp - ph).max() < 1e-13


def test_oracle_map_closed_forms(oracle):
    N = 4
    th, ph = oracle.healpix_pix2ang(N)
    L = 2
    c = np.zeros((L + 1) ** 2, complex)
    c[oracle.lmoffset(L, 0, 0)] = 1.0
    np.testing.assert_allclose(oracle.healpix_map(c, L, N), 1 / math.sqrt(4 * math.pi), rtol=1e-14)
    c[:] = 0
    c[oracle.lmoffset(L, 1, 0)] = 1.0
    np.testing.assert_allclose(oracle.healpix_map(c, L, N), math.sqrt(3 / (4 * math.pi)) * np.cos(th), atol=1e-14)
    # a complex function: the map is Re f, Y_11 = -sqrt(3 / 8 pi) sin(theta) e^{i phi}
    c[:] = 0
    c[oracle.lmoffset(L, 1, 1)] = 1.0 - 2.0j
    y = -math.sqrt(3 / (8 * math.pi)) * np.sin(th) * np.exp(1j * ph)
    np.testing.assert_allclose(oracle.healpix_map(c, L, N), ((1 - 2j) * y).real, atol=1e-14)


def test_oracle_posterior_properties(oracle):
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(192) * 5, rng.standard_normal(192) * 5
    p = oracle.healpix_posterior(a, b)
    assert abs(p.sum() - 1) < 1e-13 and (p > 0).all()
    np.testing.assert_allclose(oracle.healpix_posterior(a + 40.0, b - 7.0), p, rtol=1e-13)
    # each sky alone is a softmax: equal skies give that softmax
    q = oracle.healpix_posterior(a, a)
    e = np.exp(a - a.max())
    np.testing.assert_allclose(q, e / e.sum(), rtol=1e-13)


# ------------------------------------------------------------- FITS (host)


def _fits_hdus(path):
    """independent reader: list of (cards dict, data bytes) per HDU"""
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    hdus, pos = [], 0
    while pos < len(raw):
        cards = {}
        while True:
            block = raw[pos: pos + 2880]
            pos += 2880
            done = False
            for k in range(36):
                card = block[80 * k: 80 * k + 80].decode("ascii")
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= ":
                    v = card[10:].strip()
                    cards[key] = v.strip("'").strip() if v.startswith("'") else (v == "T" if v in ("T", "F") else int(v))
            if done:
                break
        nbytes = 0
        if cards.get("NAXIS", 0):
            nbytes = abs(cards["BITPIX"]) // 8
            for a in range(1, cards["NAXIS"] + 1):
                nbytes *= cards["NAXIS%d" % a]
        hdus.append((cards, raw[pos: pos + nbytes]))
        pos += (nbytes + 2879) // 2880 * 2880
    return hdus


def _series(coeff, L, polar=False):
    return S.series_from_numpy(coeff, L, polar)


def test_fits_alm_round_trip(tmp_path, oracle):
    rng = np.random.default_rng(5)
    for L, polar in ((7, False), (0, False), (5, True)):
        c = rand_series(rng, L, polar)
        s = _series(c, L, polar)
        path = str(tmp_path / ("alm_%d_%d.fits" % (L, polar)))
        assert S.lib().sh_series_write_healpix_alm(s, path.encode()) == 0
        hdus = _fits_hdus(path)
        assert len(hdus) == 2 and hdus[0][0]["SIMPLE"] is True
        hd, data = hdus[1]
        assert hd["XTENSION"] == "BINTABLE" and hd["TTYPE1"].lower() == "index" and hd["TTYPE2"].lower() == "real"
        mmax = 0 if polar else L
        assert hd["MAX-LPOL"] == L and hd["MAX-MPOL"] == mmax
        rows = np.frombuffer(data, dtype=[("i", ">i4"), ("re", ">f8"), ("im", ">f8")])
        # the coefficients of the real part, index l^2 + l + m + 1
        real = c.copy()
        if not polar:
            oracle.lib().orx_series_real(L, oracle._ptr(real))
        else:
            real = real.real + 0j
        got = {}
        for r in rows:
            l = int(math.isqrt(int(r["i"]) - 1))
            m = int(r["i"]) - 1 - l * l - l
            got[(l, m)] = r["re"] + 1j * r["im"]
        assert len(got) == sum(L - m + 1 for m in range(mmax + 1))
        for (l, m), v in got.items():
            assert v == real[oracle.lmoffset(L, l, m)]
        back = S.lib().sh_series_read_healpix_alm(path.encode())
        assert back
        b = S.coeff_view(back)
        # m_max = 0 reads back as a polar series (sh_series_healpix.cc:
        # 95-96), which includes l_max = 0
        assert back.contents.polar == int(polar or L == 0) and back.contents.l_max == L
        np.testing.assert_array_equal(b, real if not polar else real[: L + 1])
        S.lib().sh_series_free(back)
        S.lib().sh_series_free(s)


def test_read_healpix_alm_rejects(tmp_path):
    p = tmp_path / "junk.fits"
    p.write_bytes(b"x" * 100)
    assert not S.lib().sh_series_read_healpix_alm(str(p).encode())
    assert not S.lib().sh_series_read_healpix_alm(str(tmp_path / "missing.fits").encode())


def test_healpy_alm_conversions(oracle):
    rng = np.random.default_rng(9)
    L = 6
    c = rand_series(rng, L)
    s = _series(c, L)
    alm, l, m = H.sh_series_to_healpy_alm(s)
    assert (l, m) == (L, L) and alm.size == H.healpy_alm_length(L, L) == (L + 1) * (L + 2) // 2
    real = c.copy()
    oracle.lib().orx_series_real(L, oracle._ptr(real))
    np.testing.assert_array_equal(alm, real[: alm.size])
    back = H.healpy_alm_to_sh_series(alm, L, L)
    np.testing.assert_allclose(S.coeff_view(back), real, rtol=0, atol=1e-15)
    S.lib().sh_series_free(back)
    S.lib().sh_series_free(s)
    with pytest.raises(ValueError):
        H.healpy_alm_length(5, 3)


# ------------------------------------------------------------------- GPU


def _gpu_maps(series, L, nside, prec):
    n = len(series)
    cdt = np.complex128 if prec == S.FP64 else np.complex64
    rdt = np.float64 if prec == S.FP64 else np.float32
    hp = S.Healpix(L, nside, prec)
    d_alm = S.DeviceBuffer.from_numpy(np.ascontiguousarray(np.array(series).astype(cdt)))
    d_map = S.DeviceBuffer(n * hp.npix * np.dtype(rdt).itemsize)
    hp.maps(d_alm.ptr, n, d_map.ptr)
    S.synchronize()
    return d_map.to_numpy(rdt, (n, hp.npix))


@pytest.mark.gpu
@pytest.mark.parametrize("L,nside", [(0, 1), (1, 1), (7, 2), (7, 8), (25, 16), (32, 16), (32, 32), (50, 32), (100, 8), (130, 4)])
def test_gpu_healpix_maps_match_oracle(oracle, L, nside):
    # l_max 100 (float64) and 130 (both) exceed the LDS of the
    # one-workgroup-per-series form and run the per-ring form
    print("healpix maps L", L, "nside", nside, flush=True)
    rng = np.random.default_rng(L * 100 + nside)
    series = [rand_series(rng, L) for _ in range(5)]
    m64 = _gpu_maps(series, L, nside, S.FP64)
    m32 = _gpu_maps(series, L, nside, S.MIXED)
    for k in (0, 4):
        ref = oracle.healpix_map(series[k], L, nside)
        assert rel(m64[k], ref) < TOL64, (k, rel(m64[k], ref))
        ref32 = oracle.healpix_map(series[k].astype(np.complex64).astype(np.complex128), L, nside)
        assert rel(m32[k], ref32) < TOL_MIXED, (k, rel(m32[k], ref32))


@pytest.mark.gpu
def test_gpu_write_healpix_map_file(tmp_path, oracle):
    rng = np.random.default_rng(11)
    for L, polar in ((13, False), (9, True)):
        print("write_healpix_map L", L, "polar", polar, flush=True)
        c = rand_series(rng, L, polar)
        s = _series(c, L, polar)
        path = str(tmp_path / ("map_%d.fits" % L))
        assert S.lib().sh_series_write_healpix_map(s, path.encode()) == 0, S.last_error()
        hd, data = _fits_hdus(path)[1]
        N = oracle.healpix_nside_c(L)
        assert hd["NSIDE"] == N and hd["ORDERING"] == "RING" and hd["PIXTYPE"] == "HEALPIX" and hd["NAXIS2"] == 12 * N * N
        got = np.frombuffer(data, dtype=">f8")
        ref = oracle.healpix_map(c, L, N, polar=int(polar))
        assert rel(got, ref) < TOL64
        S.lib().sh_series_free(s)


@pytest.mark.gpu
@pytest.mark.parametrize("L", [12, 40, 100])
def test_gpu_posterior_matches_oracle(oracle, L):
    # nside 8, 32 and 64 (768 to 49152 pixels per map)
    rng = np.random.default_rng(13)
    E = 3
    N = oracle.healpix_nside_py(L)
    series = [rand_series(rng, L) * 3 for _ in range(2 * E)]
    for prec, tol in ((S.FP64, TOL64), (S.MIXED, TOL_MIXED)):
        print("posterior L", L, "precision", prec, flush=True)
        cdt = np.complex128 if prec == S.FP64 else np.complex64
        rdt = np.float64 if prec == S.FP64 else np.float32
        hp = S.Healpix(L, N, prec)
        d_alm = S.DeviceBuffer.from_numpy(np.array(series).astype(cdt))
        d_map = S.DeviceBuffer(2 * E * hp.npix * np.dtype(rdt).itemsize)
        d_prob = S.DeviceBuffer(E * hp.npix * np.dtype(rdt).itemsize)
        hp.maps(d_alm.ptr, 2 * E, d_map.ptr)
        hp.posterior(d_map.ptr, E, d_prob.ptr)
        S.synchronize()
        maps = d_map.to_numpy(rdt, (2, E, hp.npix)).astype(np.float64)
        prob = d_prob.to_numpy(rdt, (E, hp.npix))
        for e in range(E):
            ref = oracle.healpix_posterior(maps[0, e], maps[1, e])
            assert rel(prob[e], ref) < tol, (prec, e, rel(prob[e], ref))


@pytest.mark.gpu
def test_gpu_healpy_alm_to_map(oracle):
    rng = np.random.default_rng(17)
    L = 20
    c = rand_series(rng, L)
    real = c.copy()
    oracle.lib().orx_series_real(L, oracle._ptr(real))
    alm = real[: H.healpy_alm_length(L, L)]
    th, ph, m = H.healpy_alm_to_map(alm, L, L)
    N = oracle.healpix_nside_py(L)
    ot, op = oracle.healpix_pix2ang(N)
    assert np.abs(th - ot).max() < 1e-15 and np.abs(ph - op).max() < 1e-13
    # the map of the real part equals Re f of the original series
    assert rel(m, oracle.healpix_map(c, L, N)) < TOL64
