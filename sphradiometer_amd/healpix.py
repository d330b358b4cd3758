"""HEALPix conversions and pixel maps of sh_series, as the reference's
healpix.py (src/python/healpix.py:44-206), without healpy: the pixel maps
are computed on the GPU (sphr_healpix_maps), the a_lm files go through the
library's FITS reader and writer (sh_series_write_healpix_alm,
sh_series_read_healpix_alm).

healpy stores the a_lm of a real function for 0 <= m <= m_max only, in the
same (m-major, l ascending) order as the first coefficients of an
sh_series, so the conversions copy the leading coefficients of the real
part (healpix.py:91-114) and rebuild m < 0 by a_{l,-m} = (-1)^m conj
a_{l,m} (healpix.py:117-137).

Not reproduced: write_fits_healpy_map / read_fits_healpy_map
(healpix.py:183-206), which label RING data NESTED and reorder on reading;
they need the NESTED pixel scheme.
"""
import math

import numpy as np

from . import FP64, DeviceBuffer, Healpix, coeff_view, lib, last_error, length, series_from_numpy


def _check_mmax(l_max, m_max):
    if m_max not in (0, l_max):
        raise ValueError("unsupported (l_max, m_max) = (%d, %d)" % (l_max, m_max))


def healpy_alm_length(l_max, m_max):
    """Length of the healpy alm array (healpix.py:66-84)."""
    _check_mmax(l_max, m_max)
    return (m_max + 1) * (2 * l_max + 2 - m_max) // 2


def healpy_alm_l_max(alm, m_max):
    """l_max from the length of a healpy alm array (healpix.py:44-63)."""
    n, = np.shape(alm)
    # (m_max + 1)(2 l_max + 2 - m_max) / 2 = n
    l_max = (2 * n // (m_max + 1) + m_max - 2) // 2
    if healpy_alm_length(l_max, m_max) != n:
        raise ValueError("alm length %d matches no l_max for m_max %d" % (n, m_max))
    _check_mmax(l_max, m_max)
    return l_max


def sh_series_to_healpy_alm(series):
    """(alm, l_max, m_max) of an sh_series pointer (healpix.py:87-114): the
    leading coefficients of its real part."""
    s = series.contents
    l_max, m_max = s.l_max, 0 if s.polar else s.l_max
    real = lib().sh_series_real(lib().sh_series_copy(series))
    alm = coeff_view(real)[:healpy_alm_length(l_max, m_max)].copy()
    lib().sh_series_free(real)
    return alm, l_max, m_max


def healpy_alm_to_sh_series(alm, l_max, m_max):
    """sh_series pointer of a healpy alm array (healpix.py:117-137)."""
    if l_max != healpy_alm_l_max(alm, m_max):
        raise ValueError("invalid l_max=%d" % l_max)
    return series_from_numpy(_full_coefficients(alm, l_max, m_max)[: length(l_max, m_max == 0)], l_max, m_max == 0)


def _lm_offset(L, l, m):
    """offset of (l, m) in a series of order L (sh_series.h layout: m = 0,
    1, ..., L blocks, then m = -L, ..., -1, each l ascending)"""
    L1 = L + 1
    x = m * (m - 1)
    base = L1 * m - x // 2 if m >= 0 else L1 * (L1 + m) + x // 2
    return base + l - abs(m)


def _full_coefficients(alm, l_max, m_max):
    """the full (l_max + 1)^2 coefficient layout of the real function with
    healpy coefficients alm (m < 0 by conjugate symmetry)"""
    L1 = l_max + 1
    full = np.zeros(L1 * L1, dtype=np.complex128)
    a = np.asarray(alm, dtype=np.complex128)
    full[: a.size] = a
    for m in range(1, m_max + 1):
        blk = a[_lm_offset(l_max, m, m): _lm_offset(l_max, m, m) + L1 - m]
        o = _lm_offset(l_max, m, -m)
        full[o: o + L1 - m] = (-1.0 if m & 1 else 1.0) * np.conj(blk)
    return full


def pix2ang(nside):
    """(theta, phi) of the RING pixel centres of an nside map."""
    N = int(nside)
    th, ph = [], []
    for i in range(1, 4 * N):
        if i < N:
            z, nphi, phi0 = 1.0 - i * i / (3.0 * N * N), 4 * i, math.pi / (4 * i)
        elif i <= 3 * N:
            z, nphi = 4.0 / 3.0 - 2.0 * i / (3.0 * N), 4 * N
            phi0 = 0.0 if (i + N) & 1 else math.pi / (4 * N)
        else:
            k = 4 * N - i
            z, nphi, phi0 = k * k / (3.0 * N * N) - 1.0, 4 * k, math.pi / (4 * k)
        th.append(np.full(nphi, math.acos(z)))
        ph.append(phi0 + 2.0 * math.pi * np.arange(nphi) / nphi)
    return np.concatenate(th), np.concatenate(ph)


def nside_for_map(l_max):
    """nside of healpy_alm_to_map (healpix.py:163): 2^ceil(log2(l_max / 2))."""
    return int(round(2.0 ** math.ceil(math.log(l_max / 2.0, 2))))


def healpy_alm_to_map(alms, l_max, m_max):
    """(theta, phi, map) of the real function with healpy coefficients
    alms, RING order at nside 2^ceil(log2(l_max / 2)) (healpix.py:140-171),
    the map computed on the GPU in float64."""
    _check_mmax(l_max, m_max)
    nside = nside_for_map(l_max)
    hp = Healpix(l_max, nside, FP64)
    full = _full_coefficients(alms, l_max, m_max)
    d_alm = DeviceBuffer.from_numpy(full)
    d_map = DeviceBuffer(hp.npix * 8)
    hp.maps(d_alm.ptr, 1, d_map.ptr)
    m = d_map.to_numpy(np.float64, (hp.npix,))
    theta, phi = pix2ang(nside)
    return theta, phi, m


def read_fits_healpy_alm(filename):
    """(alms, l_max, m_max) of a HEALPix a_lm FITS file (healpix.py:
    192-206)."""
    s = lib().sh_series_read_healpix_alm(filename.encode())
    if not s:
        raise RuntimeError(last_error())
    alm, l_max, m_max = sh_series_to_healpy_alm(s)
    lib().sh_series_free(s)
    return alm, l_max, m_max


def write_fits_healpy_alm(filename, series):
    """the a_lm table of an sh_series pointer (sh_series_write_healpix_alm)."""
    if lib().sh_series_write_healpix_alm(series, filename.encode()):
        raise RuntimeError(last_error())
