"""ctypes view of the FP64 CPU oracle (liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg import this module, and only as the checker.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

c_int, c_uint, c_double, c_void_p = ctypes.c_int, ctypes.c_uint, ctypes.c_double, ctypes.c_void_p
_P = ctypes.POINTER


def _ptr(a):
    return a.ctypes.data_as(c_void_p)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.orx_network_new.restype = c_void_p
        L.orx_network_new.argtypes = [c_int, c_void_p, c_int, c_double, c_int]
        L.orx_network_free.argtypes = [c_void_p]
        L.orx_network_set_l.argtypes = [c_void_p, c_uint]
        L.orx_network_n_baselines.argtypes = [c_void_p]
        L.orx_plan_info.argtypes = [c_void_p, c_int, c_void_p]
        L.orx_plan_geometry.argtypes = [c_void_p, c_int, c_void_p]
        L.orx_plan_get_dp.argtypes = [c_void_p, c_int, c_void_p]
        L.orx_plan_get_power_1d.argtypes = [c_void_p, c_int, c_void_p]
        L.orx_plan_get_D.argtypes = [c_void_p, c_int, c_void_p]
        L.orx_network_integrate.argtypes = [c_void_p, c_void_p, c_uint, c_void_p]
        L.orx_segment_pipeline.argtypes = [c_void_p, c_void_p, c_void_p, c_uint, c_double, c_void_p, c_void_p]
        L.orx_segments_run.argtypes = [c_void_p, c_void_p, c_int, c_uint, c_int, c_int, c_void_p, c_void_p]
        L.orx_tseries_to_fseries.argtypes = [c_void_p, c_void_p, c_int]
        L.orx_ctseries_to_fseries.argtypes = [c_void_p, c_void_p, c_int]
        L.orx_dft.argtypes = [c_void_p, c_int, c_int]
        L.orx_gl_nodes.argtypes = [c_int, c_void_p, c_void_p]
        L.orx_sphPlm_array.argtypes = [c_uint, c_int, c_double, c_void_p]
        L.orx_wigner_3j.restype = c_double
        L.orx_wigner_3j.argtypes = [c_int] * 6
        L.orx_mesh_nodes.argtypes = [c_uint, c_void_p, c_void_p]
        L.orx_Y.argtypes = [c_uint, c_int, c_double, c_double, c_void_p]
        L.orx_eval.argtypes = [c_uint, c_int, c_void_p, c_double, c_double, c_void_p]
        L.orx_to_mesh.argtypes = [c_uint, c_int, c_void_p, c_void_p]
        L.orx_from_mesh.argtypes = [c_uint, c_int, c_void_p, c_void_p]
        L.orx_from_realmesh.argtypes = [c_uint, c_int, c_void_p, c_void_p]
        L.orx_impulse.argtypes = [c_uint, c_double, c_double, c_void_p]
        L.orx_euler_rotation_matrix.argtypes = [c_double, c_double, c_double, c_void_p]
        L.orx_rot_matrix.argtypes = [c_double, c_double, c_void_p]
        L.orx_invrot_matrix.argtypes = [c_double, c_double, c_void_p]
        L.orx_rotation_D.argtypes = [c_uint, c_void_p, c_void_p]
        L.orx_rotate.argtypes = [c_uint, c_int, c_void_p, c_void_p, c_void_p]
        L.orx_rotate_z.argtypes = [c_uint, c_void_p, c_double, c_void_p]
        L.orx_product.argtypes = [c_uint, c_int, c_uint, c_int, c_void_p, c_uint, c_int, c_void_p, c_int, c_void_p]
        L.orx_product_plan_length.argtypes = [c_uint, c_int, c_uint, c_int, c_uint, c_int]
        L.orx_projection_matrix_delay.argtypes = [c_uint, c_uint, c_void_p, c_double, c_void_p]
        L.orx_projection_element_general.argtypes = [c_uint, c_int, c_void_p, c_double, c_int, c_void_p]
        L.orx_params.argtypes = [c_double, c_double, c_void_p]
        L.orx_rms_error.restype = c_double
        L.orx_rms_error.argtypes = [c_uint, c_int, c_void_p, c_uint, c_int, c_void_p]
        L.orx_series_real.argtypes = [c_uint, c_void_p]
        L.orx_series_conj.argtypes = [c_uint, c_int, c_void_p]
        L.orx_scale_power_l.argtypes = [c_uint, c_void_p, c_double, c_double]
        L.orx_array_set_l.argtypes = [c_int, c_uint, c_int, c_void_p, c_uint, c_void_p]
        L.orx_real_minimum.argtypes = [c_uint, c_int, c_void_p, c_void_p]
        L.orx_find_near_minimum.argtypes = [c_uint, c_int, c_void_p, c_uint, c_void_p]
        L.orx_refine_minimum.argtypes = [c_uint, c_int, c_void_p, c_double, c_double, c_double, c_void_p]
        L.orx_sintheta_d_by_dtheta.argtypes = [c_uint, c_int, c_void_p, c_void_p]
        L.orx_det_response.argtypes = [c_void_p, c_double, c_double, c_double, c_double, c_void_p]
        L.orx_cbc_projection.argtypes = [c_double, c_double, c_int, c_int, c_void_p, c_int, c_double, c_double, c_void_p, c_void_p]
        L.orx_network_mult_by_projection.argtypes = [c_void_p, c_void_p, c_double, c_double, c_void_p]
        L.orx_log_uniformsky_prior.argtypes = [c_uint, c_void_p]
        L.orx_tukey.argtypes = [c_int, c_double, c_void_p]
        L.orx_precalc_length.restype = ctypes.c_long
        L.orx_precalc_length.argtypes = [ctypes.c_long, c_double]
        L.orx_preprocess_snr.argtypes = [c_void_p, c_int, c_int, c_double, c_void_p]
        L.orx_localize.argtypes = [c_void_p, c_void_p, c_int, c_double, c_double, c_void_p, c_uint, c_double, c_double, c_void_p]
        L.orx_healpix_nside_c.argtypes = [c_uint]
        L.orx_healpix_nside_py.argtypes = [c_uint]
        L.orx_healpix_pix2ang.argtypes = [c_int, ctypes.c_long, c_void_p, c_void_p]
        L.orx_healpix_map.argtypes = [c_uint, c_int, c_void_p, c_int, c_void_p]
        L.orx_healpix_posterior.argtypes = [c_void_p, c_void_p, ctypes.c_long, c_void_p]
        L.orx_plan_td_new.restype = c_void_p
        L.orx_plan_td_new.argtypes = [c_void_p, c_void_p, c_double]
        L.orx_plan_td_free.argtypes = [c_void_p]
        L.orx_plan_td_info.argtypes = [c_void_p, c_void_p]
        L.orx_baseline_integrate_power_td.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]
        L.orx_network_td_new.restype = c_void_p
        L.orx_network_td_new.argtypes = [c_int, c_void_p, c_double]
        L.orx_network_td_free.argtypes = [c_void_p]
        L.orx_network_integrate_power_td.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_int, c_uint, c_void_p]
        _LIB = L
    return _LIB


def length(l_max, polar):
    return l_max + 1 if polar else (l_max + 1) ** 2


def lmoffset(l_max, l, m):
    L1 = l_max + 1
    x = m * (m - 1)
    base = L1 * m - x // 2 if m >= 0 else L1 * (L1 + m) + x // 2
    return base + l - abs(m)


class Network:
    """A network FD plan (correlator_network_plan_fd) built by the oracle."""

    def __init__(self, positions, n, delta_t, target_l=-1):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        self.n_instruments = pos.shape[0]
        self.n = n
        self._h = lib().orx_network_new(self.n_instruments, _ptr(pos), n, delta_t, target_l)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orx_network_free(self._h)
            self._h = None

    @property
    def n_baselines(self):
        return lib().orx_network_n_baselines(self._h)

    def set_l(self, l):
        if lib().orx_network_set_l(self._h, l):
            raise RuntimeError("set_l failed")

    def info(self, k):
        a = np.zeros(8, dtype=np.int32)
        lib().orx_plan_info(self._h, k, _ptr(a))
        keys = ["l_max", "n", "stride", "transient", "polar", "rot_l_max", "index_a", "index_b"]
        return dict(zip(keys, [int(v) for v in a]))

    def geometry(self, k):
        g = np.zeros(5)
        lib().orx_plan_geometry(self._h, k, _ptr(g))
        return g

    def delay_product(self, k):
        inf = self.info(k)
        out = np.zeros((inf["n"], inf["stride"]), dtype=np.complex128)
        lib().orx_plan_get_dp(self._h, k, _ptr(out))
        return out

    def power_1d(self, k):
        inf = self.info(k)
        out = np.zeros(inf["l_max"] + 1, dtype=np.complex128)
        lib().orx_plan_get_power_1d(self._h, k, _ptr(out))
        return out

    def wigner_blocks(self, k):
        inf = self.info(k)
        L = inf["rot_l_max"]
        tot = sum((2 * l + 1) ** 2 for l in range(1, L + 1))
        out = np.zeros(tot, dtype=np.complex128)
        lib().orx_plan_get_D(self._h, k, _ptr(out))
        return out

    def integrate(self, fseries, sky_l):
        fs = np.ascontiguousarray(fseries, dtype=np.complex128)
        sky = np.zeros((sky_l + 1) ** 2, dtype=np.complex128)
        if lib().orx_network_integrate(self._h, _ptr(fs), sky_l, _ptr(sky)):
            raise RuntimeError("integrate failed")
        return sky

    def mult_by_projection(self, responses, beta, psi, psd):
        """correlator_network_plan_mult_by_projection (inspiral_sky_map.c:
        423-549); responses [n_inst][3][3], psd [n][n_inst]."""
        r = np.ascontiguousarray(responses, dtype=np.float64)
        p = np.ascontiguousarray(psd, dtype=np.float64)
        if lib().orx_network_mult_by_projection(self._h, _ptr(r), beta, psi, _ptr(p)):
            raise RuntimeError("mult_by_projection failed")

    def localize(self, series, delta_t, gmst, logprior, sky_l, power=0.6, overall=2.7):
        """generate_alm_sky + scale_power_l for one event: raw complex
        series [n_inst][len] (preprocessed here), logprior at sky_l."""
        x = np.ascontiguousarray(series, dtype=np.complex128)
        lp = np.ascontiguousarray(logprior, dtype=np.complex128)
        out = np.zeros((sky_l + 1) ** 2, dtype=np.complex128)
        if lib().orx_localize(self._h, _ptr(x), x.shape[1], delta_t, gmst, _ptr(lp), sky_l, power, overall, _ptr(out)):
            raise RuntimeError("localize failed")
        return out

    def segments_run(self, tseries, sky_l, seg_threads=1, mesh_threads=4, mesh=True):
        """The per-segment pipeline over tseries [n_inst][S][N] in C:
        seg_threads = 1 runs the segments in order with the ring loop of
        to_mesh on mesh_threads threads (the reference's split);
        seg_threads > 1 splits the segments over that many threads, each
        with its own plan scratch.  Returns (sky [S][(L+1)^2], mesh)."""
        ts = np.ascontiguousarray(tseries, dtype=np.float64)
        D, S, N = ts.shape
        sky = np.zeros((S, (sky_l + 1) ** 2), dtype=np.complex128)
        B2 = 2 * (sky_l + 1)
        m = np.zeros((S, B2, B2), dtype=np.complex128) if mesh else None
        if lib().orx_segments_run(self._h, _ptr(ts), S, sky_l, seg_threads, mesh_threads, _ptr(sky), None if m is None else _ptr(m)):
            raise RuntimeError("segments_run failed")
        return sky, m

    def segment(self, tseries, sky_l, window=None, gmst=0.0, mesh=True):
        ts = np.ascontiguousarray(tseries, dtype=np.float64)
        w = None if window is None else np.ascontiguousarray(window, dtype=np.float64)
        sky = np.zeros((sky_l + 1) ** 2, dtype=np.complex128)
        B2 = 2 * (sky_l + 1)
        m = np.zeros((B2, B2), dtype=np.complex128) if mesh else None
        rc = lib().orx_segment_pipeline(self._h, _ptr(ts), None if w is None else _ptr(w), sky_l, gmst,
                                        _ptr(sky), None if m is None else _ptr(m))
        if rc:
            raise RuntimeError("segment pipeline failed")
        return sky, m


def tseries_to_fseries(ts):
    ts = np.ascontiguousarray(ts, dtype=np.float64)
    out = np.zeros(ts.shape[0], dtype=np.complex128)
    lib().orx_tseries_to_fseries(_ptr(ts), _ptr(out), ts.shape[0])
    return out


def ctseries_to_fseries(ts):
    ts = np.ascontiguousarray(ts, dtype=np.complex128)
    out = np.zeros(ts.shape[0], dtype=np.complex128)
    lib().orx_ctseries_to_fseries(_ptr(ts), _ptr(out), ts.shape[0])
    return out


def dft(x, sign=-1):
    d = np.array(x, dtype=np.complex128)
    lib().orx_dft(_ptr(d), d.shape[0], sign)
    return d


def gl_nodes(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orx_gl_nodes(n, _ptr(x), _ptr(w))
    return x, w


def mesh_nodes(l):
    n = 2 * (l + 1)
    x, w = np.zeros(n), np.zeros(n)
    lib().orx_mesh_nodes(l, _ptr(x), _ptr(w))
    return x, w


def sphPlm_array(l_max, m, x):
    out = np.zeros(l_max - m + 1)
    lib().orx_sphPlm_array(l_max, m, x, _ptr(out))
    return out


def wigner_3j(*a):
    return lib().orx_wigner_3j(*[int(v) for v in a])


def Y(l, m, theta, phi):
    out = np.zeros(1, dtype=np.complex128)
    lib().orx_Y(l, m, theta, phi, _ptr(out))
    return out[0]


def eval_series(coeff, l, polar, theta, phi):
    c = np.ascontiguousarray(coeff, dtype=np.complex128)
    out = np.zeros(1, dtype=np.complex128)
    lib().orx_eval(l, polar, _ptr(c), theta, phi, _ptr(out))
    return out[0]


def to_mesh(coeff, l, polar=0):
    c = np.ascontiguousarray(coeff, dtype=np.complex128)
    n = 2 * (l + 1)
    out = np.zeros((n, n), dtype=np.complex128)
    lib().orx_to_mesh(l, polar, _ptr(c), _ptr(out))
    return out


def from_mesh(mesh, l, polar=0):
    m = np.ascontiguousarray(mesh, dtype=np.complex128)
    out = np.zeros(length(l, polar), dtype=np.complex128)
    lib().orx_from_mesh(l, polar, _ptr(m), _ptr(out))
    return out


def from_realmesh(mesh, l, polar=0):
    m = np.ascontiguousarray(mesh, dtype=np.float64)
    out = np.zeros(length(l, polar), dtype=np.complex128)
    lib().orx_from_realmesh(l, polar, _ptr(m), _ptr(out))
    return out


def impulse(l, theta, phi):
    out = np.zeros((l + 1) ** 2, dtype=np.complex128)
    lib().orx_impulse(l, theta, phi, _ptr(out))
    return out


def euler_rotation_matrix(gamma, beta, alpha):
    R = np.zeros(9)
    lib().orx_euler_rotation_matrix(gamma, beta, alpha, _ptr(R))
    return R.reshape(3, 3)


def rot_matrix(theta, phi):
    R = np.zeros(9)
    lib().orx_rot_matrix(theta, phi, _ptr(R))
    return R.reshape(3, 3)


def invrot_matrix(theta, phi):
    R = np.zeros(9)
    lib().orx_invrot_matrix(theta, phi, _ptr(R))
    return R.reshape(3, 3)


def rotation_D(l_max, R):
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(9)
    tot = sum((2 * l + 1) ** 2 for l in range(1, l_max + 1))
    out = np.zeros(tot, dtype=np.complex128)
    lib().orx_rotation_D(l_max, _ptr(R), _ptr(out))
    blocks, off = {}, 0
    for l in range(1, l_max + 1):
        n = 2 * l + 1
        # storage index (2l+1) * (m'+l) + (m+l): row m', column m
        blocks[l] = out[off:off + n * n].reshape(n, n)
        off += n * n
    return blocks


def rotate(coeff, l, polar, R):
    c = np.ascontiguousarray(coeff, dtype=np.complex128)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(9)
    out = np.zeros((l + 1) ** 2, dtype=np.complex128)
    if lib().orx_rotate(l, polar, _ptr(c), _ptr(R), _ptr(out)):
        raise RuntimeError("rotate failed")
    return out


def rotate_z(coeff, l, phi):
    c = np.ascontiguousarray(coeff, dtype=np.complex128)
    out = np.zeros_like(c)
    lib().orx_rotate_z(l, _ptr(c), phi, _ptr(out))
    return out


def product(a, al, ap, b, bl, bp, dl, dp, mode=-1):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    out = np.zeros(length(dl, dp), dtype=np.complex128)
    if lib().orx_product(dl, dp, al, ap, _ptr(a), bl, bp, _ptr(b), mode, _ptr(out)):
        raise RuntimeError("product failed")
    return out


def projection_matrix_delay(n, l_max, r, dt):
    r = np.ascontiguousarray(r, dtype=np.float64)
    polar = r[0] == 0.0 and r[1] == 0.0
    out = np.zeros((n, length(l_max, polar)), dtype=np.complex128)
    lib().orx_projection_matrix_delay(n, l_max, _ptr(r), dt, _ptr(out))
    return out


def projection_element_general(l_max, polar, r, dt, k):
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(length(l_max, polar), dtype=np.complex128)
    lib().orx_projection_element_general(l_max, polar, _ptr(r), dt, k, _ptr(out))
    return out


def params(dt, D):
    out = np.zeros(3, dtype=np.int32)
    lib().orx_params(dt, D, _ptr(out))
    return dict(P=int(out[0]), a=int(out[1]), n_p=int(out[2]))


def rms_error(m, lm, pm, e, le, pe):
    m = np.ascontiguousarray(m, dtype=np.complex128)
    e = np.ascontiguousarray(e, dtype=np.complex128)
    return lib().orx_rms_error(lm, pm, _ptr(m), le, pe, _ptr(e))


def series_real(coeff, l):
    c = np.array(coeff, dtype=np.complex128)
    lib().orx_series_real(l, _ptr(c))
    return c


def series_conj(coeff, l, polar=0):
    c = np.array(coeff, dtype=np.complex128)
    lib().orx_series_conj(l, polar, _ptr(c))
    return c


def scale_power_l(coeff, l, power, overall):
    c = np.array(coeff, dtype=np.complex128)
    lib().orx_scale_power_l(l, _ptr(c), power, overall)
    return c


def array_set_l(arr, l_old, polar, l_new):
    a = np.ascontiguousarray(arr, dtype=np.complex128)
    n = a.shape[0]
    out = np.zeros((n, length(l_new, polar)), dtype=np.complex128)
    lib().orx_array_set_l(n, l_old, polar, _ptr(a), l_new, _ptr(out))
    return out


def real_minimum(coeff, l, polar=0):
    """(theta, phi, value) of the minimum of Re f: reference seed search,
    Nelder-Mead refinement (sh_series_extremize.c:273-360)."""
    c = np.ascontiguousarray(coeff, dtype=np.complex128)
    out = np.zeros(3)
    lib().orx_real_minimum(l, polar, _ptr(c), _ptr(out))
    return tuple(out)


def refine_minimum(coeff, l, theta, phi, size=1e-3, polar=0):
    """(theta, phi, value): Nelder-Mead refinement of the minimum of Re f
    from a given point (simplex of `size` rad)."""
    c = np.ascontiguousarray(coeff, dtype=np.complex128)
    out = np.zeros(3)
    lib().orx_refine_minimum(l, polar, _ptr(c), theta, phi, size, _ptr(out))
    return tuple(out)


def find_near_minimum(coeff, l, polar=0, l_lowpass=6):
    """(theta, phi, value) of the minimum pixel of the low-passed series
    (sh_series_extremize.c:168-241)."""
    c = np.ascontiguousarray(coeff, dtype=np.complex128)
    out = np.zeros(3)
    lib().orx_find_near_minimum(l, polar, _ptr(c), l_lowpass, _ptr(out))
    return tuple(out)


def sintheta_d_by_dtheta(coeff, l, polar=0):
    """sin(theta) d/dtheta through the 3j form of sh_series_deriv.c:288-543;
    the result has l_max + 1."""
    c = np.ascontiguousarray(coeff, dtype=np.complex128)
    out = np.zeros(length(l + 1, polar), dtype=np.complex128)
    lib().orx_sintheta_d_by_dtheta(l, polar, _ptr(c), _ptr(out))
    return out


def det_response(resp, ra, dec, psi, gmst=0.0):
    r = np.ascontiguousarray(resp, dtype=np.float64)
    out = np.zeros(2)
    lib().orx_det_response(_ptr(r), ra, dec, psi, gmst, _ptr(out))
    return out[0], out[1]


def cbc_projection(theta, phi, i, j, responses, beta, psi, psds):
    r = np.ascontiguousarray(responses, dtype=np.float64)
    p = np.ascontiguousarray(psds, dtype=np.float64)
    out = np.zeros(1, dtype=np.complex128)
    lib().orx_cbc_projection(theta, phi, i, j, _ptr(r), r.shape[0], beta, psi, _ptr(p), _ptr(out))
    return complex(out[0])


def log_uniformsky_prior(l):
    out = np.zeros((l + 1) ** 2, dtype=np.complex128)
    lib().orx_log_uniformsky_prior(l, _ptr(out))
    return out


def tukey(n, beta):
    out = np.zeros(n)
    lib().orx_tukey(n, beta, _ptr(out))
    return out


def precalc_length(length, delta_t):
    return int(lib().orx_precalc_length(length, delta_t))


def preprocess_snr(series, delta_t):
    x = np.ascontiguousarray(series, dtype=np.complex128)
    n, ln = x.shape
    out = np.zeros((n, precalc_length(ln, delta_t)), dtype=np.complex128)
    lib().orx_preprocess_snr(_ptr(x), n, ln, delta_t, _ptr(out))
    return out


# HEALPix output (or_healpix.c; sh_series_healpix.cc, healpix.py,
# RapidLocalization.py:498-530)
def healpix_nside_c(l_max):
    return lib().orx_healpix_nside_c(l_max)


def healpix_nside_py(l_max):
    return lib().orx_healpix_nside_py(l_max)


def healpix_pix2ang(nside):
    n = 12 * nside * nside
    th, ph = np.zeros(n), np.zeros(n)
    lib().orx_healpix_pix2ang(nside, n, _ptr(th), _ptr(ph))
    return th, ph


def healpix_map(coeff, l, nside, polar=0):
    c = np.ascontiguousarray(coeff, dtype=np.complex128)
    out = np.zeros(12 * nside * nside)
    lib().orx_healpix_map(l, polar, _ptr(c), nside, _ptr(out))
    return out


def healpix_posterior(map_p, map_n):
    a = np.ascontiguousarray(map_p, dtype=np.float64)
    b = np.ascontiguousarray(map_n, dtype=np.float64)
    out = np.zeros(a.size)
    lib().orx_healpix_posterior(_ptr(a), _ptr(b), a.size, _ptr(out))
    return out


# time-domain correlator (or_td.c; correlator.c:325-400, 648-673, 820-858,
# 967-984)
class PlanTD:
    def __init__(self, ra, rb, delta_t):
        a = np.ascontiguousarray(ra, dtype=np.float64)
        b = np.ascontiguousarray(rb, dtype=np.float64)
        self.h = lib().orx_plan_td_new(_ptr(a), _ptr(b), delta_t)
        info = np.zeros(2, dtype=np.int32)
        lib().orx_plan_td_info(self.h, _ptr(info))
        self.transient, self.l_max = int(info[0]), int(info[1])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orx_plan_td_free(self.h)
            self.h = None

    def integrate(self, ta, tb, window):
        """(power_2d full layout, power_1d) of the baseline"""
        a = np.ascontiguousarray(ta, dtype=np.float64)
        b = np.ascontiguousarray(tb, dtype=np.float64)
        w = np.ascontiguousarray(window, dtype=np.float64)
        L = self.l_max
        p2 = np.zeros((L + 1) ** 2, dtype=np.complex128)
        p1 = np.zeros(L + 1, dtype=np.complex128)
        if lib().orx_baseline_integrate_power_td(_ptr(a), _ptr(b), _ptr(w), a.size, self.h, _ptr(p2), _ptr(p1)):
            raise RuntimeError("integrate_power_td failed")
        return p2, p1


class NetworkTD:
    def __init__(self, positions, delta_t):
        p = np.ascontiguousarray(positions, dtype=np.float64)
        self.n_instruments = p.shape[0]
        self.h = lib().orx_network_td_new(p.shape[0], _ptr(p), delta_t)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orx_network_td_free(self.h)
            self.h = None

    def integrate(self, ts, windows, sky_l):
        """sky of order sky_l; ts [n_instruments][n], windows [n_baselines][stride]"""
        t = np.ascontiguousarray(ts, dtype=np.float64)
        w = np.ascontiguousarray(windows, dtype=np.float64)
        sky = np.zeros((sky_l + 1) ** 2, dtype=np.complex128)
        if lib().orx_network_integrate_power_td(self.h, _ptr(t), t.shape[1], _ptr(w), w.shape[1], sky_l, _ptr(sky)):
            raise RuntimeError("network_integrate_power_td failed")
        return sky
