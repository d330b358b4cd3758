"""sphradiometer_amd: MI355X implementation of sphradiometer's time-resolved
full-sky cross-power path.

The product is the C-ABI library libsphradiometer_amd.so (headers under
include/sphradiometer/).  This package loads it with ctypes, declares the
public structs with the reference's layouts, and adds numpy conveniences.
``sphradiometer_amd.sphradiometer`` exposes the functions under the names of
the reference's SWIG module.

Importing works without a GPU (object management and plan construction are
host code); the entry points that run kernels return an error when no GPU
is present.  There is no CPU fallback for them.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphradiometer_amd.so")

c_int, c_uint, c_double, c_void_p, c_size_t, c_float = (ctypes.c_int, ctypes.c_uint, ctypes.c_double,
                                                        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float)
P = ctypes.POINTER


class CDouble(ctypes.Structure):
    """double complex by value (two doubles in SSE registers on x86-64,
    the same classification as a struct of two doubles)"""
    _fields_ = [("re", c_double), ("im", c_double)]

    @property
    def value(self):
        return complex(self.re, self.im)


class GslVector(ctypes.Structure):
    _fields_ = [("size", c_size_t), ("stride", c_size_t), ("data", P(c_double)), ("block", c_void_p), ("owner", c_int)]


class ShSeries(ctypes.Structure):
    _fields_ = [("l_max", c_uint), ("polar", c_int), ("coeff", c_void_p)]


class ProductOp(ctypes.Structure):
    _fields_ = [("factor", c_double), ("dest_offset", c_int), ("a_offset", c_int), ("b_offset", c_int)]


class ProductPlan(ctypes.Structure):
    _fields_ = [("a_l_max", c_uint), ("b_l_max", c_uint), ("dest_l_max", c_uint), ("a_polar", c_int),
                ("b_polar", c_int), ("dest_polar", c_int), ("plan_length", c_int), ("microcode", P(ProductOp))]


class RotationPlan(ctypes.Structure):
    _fields_ = [("l_max", c_uint), ("D", P(c_void_p))]


class ShSeriesArray(ctypes.Structure):
    _fields_ = [("l_max", c_uint), ("polar", c_int), ("n", c_int), ("series", P(ShSeries)),
                ("coeff", c_void_p), ("stride", c_int)]


class Instrument(ctypes.Structure):
    _fields_ = [("phase_centre", P(GslVector)), ("data", c_void_p)]


class InstrumentArray(ctypes.Structure):
    _fields_ = [("instruments", P(P(Instrument))), ("n", c_int)]


class Baseline(ctypes.Structure):
    _fields_ = [("instruments", P(InstrumentArray)), ("index_a", c_int), ("index_b", c_int),
                ("d", P(GslVector)), ("theta", c_double), ("phi", c_double)]


class PlanFD(ctypes.Structure):
    _fields_ = [("baseline", P(Baseline)), ("delta_t", c_double), ("transient", c_int),
                ("delay_product", P(ShSeriesArray)), ("fseries_product", c_void_p),
                ("power_1d", P(ShSeries)), ("rotation_plan", P(RotationPlan))]


class PlanTD(ctypes.Structure):
    _fields_ = [("baseline", P(Baseline)), ("delta_t", c_double), ("transient", c_int),
                ("proj_a", P(ShSeriesArray)), ("proj_b", P(ShSeriesArray)),
                ("sample_a", P(ShSeries)), ("sample_b", P(ShSeries)), ("product", P(ShSeries)),
                ("power_1d", P(ShSeries)), ("product_plan", c_void_p), ("rotation_plan", P(RotationPlan))]


class NetworkBaselines(ctypes.Structure):
    _fields_ = [("baselines", P(P(Baseline))), ("n_baselines", c_int)]


class NetworkPlanTD(ctypes.Structure):
    _fields_ = [("baselines", P(NetworkBaselines)), ("plans", P(P(PlanTD)))]


class NetworkPlanFD(ctypes.Structure):
    _fields_ = [("baselines", P(NetworkBaselines)), ("plans", P(P(PlanFD)))]


class AutoPlanFD(ctypes.Structure):
    _fields_ = [("instruments", P(InstrumentArray)), ("projections", P(P(P(ShSeries)))), ("length", c_int)]


class Detector(ctypes.Structure):
    """struct sphr_detector: LALDetector's leading fields."""
    _fields_ = [("location", c_double * 3), ("response", (ctypes.c_float * 3) * 3)]


# name: (restype, argtypes)
_PROTOS = {
    "sh_series_new": (P(ShSeries), [c_uint, c_int]),
    "sh_series_new_zero": (P(ShSeries), [c_uint, c_int]),
    "sh_series_copy": (P(ShSeries), [P(ShSeries)]),
    "sh_series_assign": (P(ShSeries), [P(ShSeries), P(ShSeries)]),
    "sh_series_resize": (P(ShSeries), [P(ShSeries), c_uint]),
    "sh_series_zero": (P(ShSeries), [P(ShSeries)]),
    "sh_series_set_polar": (P(ShSeries), [P(ShSeries), c_int]),
    "sh_series_free": (None, [P(ShSeries)]),
    "sh_series_eval": (CDouble, [P(ShSeries), c_double, c_double]),
    "sh_series_get": (CDouble, [P(ShSeries), c_uint, c_int]),
    "sh_series_set": (CDouble, [P(ShSeries), c_uint, c_int, CDouble]),
    "sh_series_Y": (CDouble, [c_uint, c_int, c_double, c_double]),
    "sh_series_dot": (CDouble, [P(ShSeries), P(ShSeries)]),
    "sh_series_rotation_plan_wigner_D": (CDouble, [P(RotationPlan), c_uint, c_int, c_int]),
    "sh_series_scale": (P(ShSeries), [P(ShSeries), CDouble]),
    "sh_series_to_mesh": (c_void_p, [P(ShSeries), P(c_int), P(c_int), P(P(c_double))]),
    "sh_series_from_mesh": (P(ShSeries), [P(ShSeries), c_void_p]),
    "sh_series_from_realmesh": (P(ShSeries), [P(ShSeries), c_void_p]),
    "sh_series_impulse": (P(ShSeries), [c_uint, c_double, c_double]),
    "sh_series_add": (P(ShSeries), [P(ShSeries), CDouble, P(ShSeries)]),
    "sh_series_scale_power_l": (P(ShSeries), [P(ShSeries), c_double, c_double]),
    "sh_series_conj": (P(ShSeries), [P(ShSeries)]),
    "sh_series_real": (P(ShSeries), [P(ShSeries)]),
    "sh_series_imag": (P(ShSeries), [P(ShSeries)]),
    "sh_series_product_plan_new": (P(ProductPlan), [P(ShSeries), P(ShSeries), P(ShSeries)]),
    "sh_series_product_plan_free": (None, [P(ProductPlan)]),
    "sh_series_product": (P(ShSeries), [P(ShSeries), P(ShSeries), P(ShSeries), P(ProductPlan)]),
    "sh_series_wigner_3j": (c_double, [c_int] * 6),
    "euler_rotation_matrix": (P(c_double), [c_double, c_double, c_double]),
    "sh_series_rot_matrix": (P(c_double), [c_double, c_double]),
    "sh_series_invrot_matrix": (P(c_double), [c_double, c_double]),
    "sh_series_rotation_plan_new": (P(RotationPlan), [P(ShSeries), P(c_double)]),
    "sh_series_rotation_plan_free": (None, [P(RotationPlan)]),
    "sh_series_rotation_plan_copy": (P(RotationPlan), [P(RotationPlan)]),
    "sh_series_rotation_plan_set_l": (P(RotationPlan), [P(RotationPlan), c_uint]),
    "sh_series_rotate": (P(ShSeries), [P(ShSeries), P(ShSeries), P(RotationPlan)]),
    "sh_series_rotate_z": (P(ShSeries), [P(ShSeries), P(ShSeries), c_double]),
    "sh_series_d_by_dphi": (P(ShSeries), [P(ShSeries)]),
    "sh_series_sintheta_d_by_dtheta": (P(ShSeries), [P(ShSeries)]),
    "sh_series_sintheta_grad": (P(ShSeries), [P(ShSeries), P(P(ShSeries)), P(P(ShSeries))]),
    "sh_series_real_minimum": (c_double, [P(ShSeries), P(c_double), P(c_double)]),
    "sh_series_array_new": (P(ShSeriesArray), [c_int, c_uint, c_int]),
    "sh_series_array_free": (None, [P(ShSeriesArray)]),
    "sh_series_array_set_l": (P(ShSeriesArray), [P(ShSeriesArray), c_uint]),
    "sh_series_array_set_len": (P(ShSeriesArray), [P(ShSeriesArray), c_int]),
    "sh_series_array_dotc": (P(ShSeries), [P(ShSeries), P(ShSeriesArray), c_void_p]),
    "sh_series_array_dot": (P(ShSeries), [P(ShSeries), P(ShSeriesArray), c_void_p]),
    "sh_series_array_forward_fft": (P(ShSeriesArray), [P(ShSeriesArray)]),
    "projection_matrix_n_elements": (c_int, [c_double, c_double]),
    "projection_matrix_l_max": (c_uint, [c_double, c_double]),
    "projection_matrix_delay": (P(ShSeriesArray), [c_uint, c_uint, P(GslVector), c_double]),
    "instrument_new": (P(Instrument), [c_double, c_double, c_double, c_void_p]),
    "instrument_new_from_r_theta_phi": (P(Instrument), [c_double, c_double, c_double, c_void_p]),
    "instrument_free": (None, [P(Instrument)]),
    "instrument_array_new": (P(InstrumentArray), [c_int]),
    "instrument_array_append": (P(Instrument), [P(InstrumentArray), P(Instrument)]),
    "instrument_array_set": (P(Instrument), [P(InstrumentArray), c_int, P(Instrument)]),
    "instrument_array_get": (P(Instrument), [P(InstrumentArray), c_int]),
    "instrument_array_len": (c_int, [P(InstrumentArray)]),
    "instrument_array_free": (None, [P(InstrumentArray)]),
    "vector_magnitude": (c_double, [P(GslVector)]),
    "correlator_baseline_power_l_max_naive": (c_uint, [c_double, c_double]),
    "correlator_baseline_power_l_max": (c_uint, [P(Baseline), c_double]),
    "correlator_dump_interval": (c_double, [c_uint, c_uint]),
    "correlator_network_baselines_new": (P(NetworkBaselines), [P(InstrumentArray)]),
    "correlator_network_baselines_free": (None, [P(NetworkBaselines)]),
    "correlator_network_l_max": (c_uint, [P(NetworkBaselines), c_double]),
    "correlator_plan_fd_new": (P(PlanFD), [P(Baseline), c_int, c_double]),
    "correlator_plan_fd_free": (None, [P(PlanFD)]),
    "correlator_plan_fd_set_l": (P(PlanFD), [P(PlanFD), c_uint]),
    "correlator_network_plan_fd_new": (P(NetworkPlanFD), [P(NetworkBaselines), c_int, c_double]),
    "correlator_network_plan_fd_free": (None, [P(NetworkPlanFD)]),
    "correlator_network_plan_fd_copy": (P(NetworkPlanFD), [P(NetworkPlanFD)]),
    "correlator_network_plan_fd_set_l": (P(NetworkPlanFD), [P(NetworkPlanFD), c_uint]),
    "correlator_baseline_integrate_power_fd": (P(ShSeries), [c_void_p, c_void_p, P(PlanFD)]),
    "correlator_plan_td_new": (P(PlanTD), [P(Baseline), c_double]),
    "correlator_plan_td_free": (None, [P(PlanTD)]),
    "correlator_baseline_integrate_power_td": (P(ShSeries), [c_void_p, c_void_p, c_void_p, c_int, P(PlanTD)]),
    "correlator_network_plan_td_new": (P(NetworkPlanTD), [P(NetworkBaselines), c_double]),
    "correlator_network_plan_td_free": (None, [P(NetworkPlanTD)]),
    "correlator_network_integrate_power_td": (P(ShSeries), [P(ShSeries), P(c_void_p), c_int, P(c_void_p), P(NetworkPlanTD)]),
    "correlator_network_integrate_power_fd": (P(ShSeries), [P(ShSeries), P(c_void_p), P(NetworkPlanFD)]),
    "correlator_plan_fd_device_invalidate": (None, [P(PlanFD)]),
    "sphr_gemm_h3_check": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "sphr_gemm_run": (c_int, [c_int, c_int, c_int, c_int, c_int, c_void_p, ctypes.c_long, c_void_p, ctypes.c_long, c_void_p, ctypes.c_long, c_void_p]),
    "correlator_tseries_to_fseries_plan": (c_void_p, [c_void_p, c_void_p, c_int]),
    "correlator_tseries_to_fseries": (None, [c_void_p, c_void_p, c_int, c_void_p]),
    "correlator_ctseries_to_fseries_plan": (c_void_p, [c_void_p, c_void_p, c_int]),
    "correlator_ctseries_to_fseries": (None, [c_void_p]),
    "correlator_fft_plan_free": (None, [c_void_p]),
    "correlator_square_window_new": (P(c_double), [c_int, c_int, c_double]),
    "correlator_tukey_window_new": (P(c_double), [c_int, c_int, c_int, c_double]),
    "diagnostics_rms_error": (c_double, [P(ShSeries), P(ShSeries)]),
    "sphr_gpu_available": (c_int, []),
    "sphr_last_error": (ctypes.c_char_p, []),
    "sphr_device_info": (c_int, [ctypes.c_char_p, c_int, P(c_int), P(c_size_t)]),
    "sphr_device_alloc": (c_void_p, [c_size_t]),
    "sphr_device_free": (None, [c_void_p]),
    "sphr_memcpy_htod": (c_int, [c_void_p, c_void_p, c_size_t]),
    "sphr_memcpy_dtoh": (c_int, [c_void_p, c_void_p, c_size_t]),
    "sphr_memcpy_dtod": (c_int, [c_void_p, c_void_p, c_size_t]),
    "sphr_memset_device": (c_int, [c_void_p, c_int, c_size_t]),
    "sphr_stream": (c_void_p, []),
    "sphr_stream_synchronize": (c_int, [c_void_p]),
    "sphr_device_synchronize": (c_int, []),
    "sphr_event_new": (c_void_p, []),
    "sphr_event_free": (None, [c_void_p]),
    "sphr_event_record": (c_int, [c_void_p, c_void_p]),
    "sphr_event_synchronize": (c_int, [c_void_p]),
    "sphr_event_elapsed_ms": (c_float, [c_void_p, c_void_p]),
    "sphr_batch_new": (c_void_p, [P(NetworkPlanFD), c_uint, c_int, c_int]),
    "sphr_batch_free": (None, [c_void_p]),
    "sphr_batch_run": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sphr_batch_run_static": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sphr_batch_profile": (c_int, [c_void_p, c_int]),
    "sphr_batch_profile_read": (c_int, [c_void_p, P(c_double), P(c_int)]),
    "sphr_batch_info": (None, [c_void_p, P(c_int)]),
    "sphr_batch_set_fused": (None, [c_void_p, c_int]),
    "sphr_batch_set_sky_occupancy": (c_int, [c_void_p, c_int]),
    "sphr_batch_set_sky_segments": (c_int, [c_void_p, c_int]),
    "sphr_batch_set_shape": (c_int, [c_void_p, c_int]),
    "sphr_batch_set_lag": (c_int, [c_void_p, c_int]),
    "sphr_batch_get_lag": (c_int, [c_void_p]),
    "sphr_batch_table_bytes": (c_size_t, [c_void_p]),
    "sphr_batch_set_xcorr_form": (c_int, [c_void_p, c_int]),
    "sphr_batch_get_xcorr_form": (c_int, [c_void_p]),
    "sphr_batch_peaks": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sphr_batch_tune": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, P(c_double)]),
    "sphr_jobs_new": (c_void_p, [c_int, P(c_int), P(c_double), P(c_int), P(c_int), c_double, c_int]),
    "sphr_jobs_free": (None, [c_void_p]),
    "sphr_jobs_run": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sphr_jobs_tune": (c_int, [c_void_p, c_void_p, c_void_p, P(c_double)]),
    "sphr_jobs_set_lag_tile": (c_int, [c_void_p, c_int, c_int]),
    "sphr_jobs_info": (None, [c_void_p, P(ctypes.c_long)]),
    "sphr_jobs_offsets": (None, [c_void_p, P(ctypes.c_long), P(ctypes.c_long), P(ctypes.c_long)]),
    "precalculated_TimeSeries_length": (ctypes.c_long, [ctypes.c_long, c_double]),
    "sh_series_log_uniformsky_prior": (P(ShSeries), [c_int]),
    "correlator_network_plan_mult_by_projection": (c_int, [P(NetworkPlanFD), c_double, c_double, P(P(c_double))]),
    "preprocess_SNR_arrays": (c_int, [c_void_p, c_int, c_int, c_double, c_void_p]),
    "generate_alm_skys_arrays": (c_int, [P(P(ShSeries)), P(P(ShSeries)), P(NetworkPlanFD), P(NetworkPlanFD), c_void_p, c_int, c_double, P(ShSeries)]),
    "sphr_loc_new": (c_void_p, [P(NetworkPlanFD), P(NetworkPlanFD), P(ShSeries), c_uint, c_double, c_double, c_int, c_double, c_int, c_int]),
    "sphr_loc_free": (None, [c_void_p]),
    "sphr_loc_run": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sphr_loc_peaks": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "make_precalc_directories": (c_int, [ctypes.c_char_p, c_int]),
    "write_precalc_time_series_length": (c_int, [c_uint, ctypes.c_char_p]),
    "read_precalc_time_series_length": (c_int, [P(c_uint), ctypes.c_char_p]),
    "write_precalc_logprior": (c_int, [P(ShSeries), ctypes.c_char_p]),
    "read_precalc_logprior": (P(ShSeries), [ctypes.c_char_p]),
    "write_precalc_correlator_network_plan_fd": (c_int, [P(NetworkPlanFD), P(NetworkPlanFD), ctypes.c_char_p]),
    "read_precalc_correlator_network_baselines": (P(NetworkBaselines), [P(InstrumentArray), ctypes.c_char_p]),
    "read_precalc_correlator_network_plan_fd": (c_int, [P(NetworkPlanFD), P(NetworkPlanFD), P(NetworkBaselines), c_int, ctypes.c_char_p]),
    "write_precalc_autocorrelator_network_plan_fd": (c_int, [P(AutoPlanFD), P(AutoPlanFD), ctypes.c_char_p]),
    "read_precalc_autocorrelator_network_plan_fd": (c_int, [P(AutoPlanFD), P(AutoPlanFD), P(InstrumentArray), c_int, ctypes.c_char_p]),
    "autocorrelator_network_plan_fd_new": (P(AutoPlanFD), [P(InstrumentArray), c_double, c_double, c_void_p, c_int, c_int]),
    "autocorrelator_network_plan_fd_set_l": (P(AutoPlanFD), [P(AutoPlanFD), c_uint]),
    "autocorrelator_network_plan_fd_free": (None, [P(AutoPlanFD)]),
    "sh_series_write_healpix_map": (c_int, [P(ShSeries), ctypes.c_char_p]),
    "sh_series_write_healpix_alm": (c_int, [P(ShSeries), ctypes.c_char_p]),
    "sh_series_read_healpix_alm": (P(ShSeries), [ctypes.c_char_p]),
    "sphr_healpix_nside": (c_int, [c_uint]),
    "sphr_healpix_new": (c_void_p, [c_uint, c_int, c_int]),
    "sphr_healpix_free": (None, [c_void_p]),
    "sphr_healpix_npix": (ctypes.c_long, [c_void_p]),
    "sphr_healpix_maps": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "sphr_healpix_posterior": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "sphr_loc_info": (None, [c_void_p, P(c_int)]),
}


_lib = None


def lib():
    """The loaded C library.  Raises if it is not built: nothing in this
    package substitutes for it."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsphradiometer_amd.so is not built; run __graft_entry__.build() "
                               "or make -C sphradiometer_amd")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def build(jobs=8):
    subprocess.check_call(["make", "-C", _HERE, "-j%d" % jobs])


def last_error():
    return lib().sphr_last_error().decode(errors="replace")


def gpu_available():
    return bool(lib().sphr_gpu_available())


def _ptr(a):
    return a.ctypes.data_as(c_void_p)


def length(l_max, polar):
    return l_max + 1 if polar else (l_max + 1) ** 2


def coeff_view(series_p):
    """numpy complex128 view of an sh_series' coefficients (no copy)."""
    s = series_p.contents
    n = length(s.l_max, s.polar)
    buf = (ctypes.c_double * (2 * n)).from_address(s.coeff)
    return np.frombuffer(buf, dtype=np.complex128, count=n)


def array_view(array_p):
    """numpy [n, stride] complex128 view of an sh_series_array."""
    a = array_p.contents
    buf = (ctypes.c_double * (2 * a.n * a.stride)).from_address(a.coeff)
    return np.frombuffer(buf, dtype=np.complex128).reshape(a.n, a.stride)


def series_from_numpy(coeff, l_max, polar):
    s = lib().sh_series_new(l_max, polar)
    coeff_view(s)[:] = np.asarray(coeff, dtype=np.complex128)
    return s


def rotation_blocks(plan_p):
    """The packed D blocks of a rotation plan as {l: (2l+1, 2l+1) array}
    indexed [m' + l, m + l]."""
    p = plan_p.contents
    out = {}
    for l in range(1, p.l_max + 1):
        n = 2 * l + 1
        centre = p.D[l]
        start = centre - 16 * (n * l + l)
        buf = (ctypes.c_double * (2 * n * n)).from_address(start)
        out[l] = np.frombuffer(buf, dtype=np.complex128).reshape(n, n).copy()
    return out


class Network:
    """Instruments + network baselines + FD network plan, built through the
    C API (correlator_network_plan_fd_new, then _set_l)."""

    def __init__(self, positions, n, delta_t, l_max=None, responses=None):
        L = lib()
        self._inst = L.instrument_array_new(0)
        # detector descriptors (struct sphr_detector) stay alive with the network
        self._det = []
        for k, (x, y, z) in enumerate(np.asarray(positions, dtype=np.float64)):
            data = None
            if responses is not None:
                d = Detector()
                for i in range(3):
                    for j in range(3):
                        d.response[i][j] = float(responses[k][i][j])
                self._det.append(d)
                data = ctypes.cast(ctypes.pointer(d), c_void_p)
            L.instrument_array_append(self._inst, L.instrument_new(float(x), float(y), float(z), data))
        self._bl = L.correlator_network_baselines_new(self._inst)
        self._plan = L.correlator_network_plan_fd_new(self._bl, n, delta_t)
        if not self._plan:
            raise RuntimeError("correlator_network_plan_fd_new failed")
        self.n = n
        self.n_instruments = len(positions)
        if l_max is not None:
            self.set_l(l_max)

    @property
    def plan(self):
        return self._plan

    @property
    def n_baselines(self):
        return self._bl.contents.n_baselines

    def set_l(self, l_max):
        if not lib().correlator_network_plan_fd_set_l(self._plan, l_max):
            raise RuntimeError("correlator_network_plan_fd_set_l failed")

    def baseline_plan(self, k):
        return self._plan.contents.plans[k]

    def delay_product(self, k):
        return array_view(self.baseline_plan(k).contents.delay_product).copy()

    def power_1d(self, k):
        return coeff_view(self.baseline_plan(k).contents.power_1d).copy()

    def mult_by_projection(self, beta, psi, psd):
        """correlator_network_plan_mult_by_projection; psd [n][n_inst]."""
        p = np.ascontiguousarray(psd, dtype=np.float64)
        rows = (P(c_double) * p.shape[0])(*[p[i].ctypes.data_as(P(c_double)) for i in range(p.shape[0])])
        if lib().correlator_network_plan_mult_by_projection(self._plan, beta, psi, rows):
            raise RuntimeError("mult_by_projection failed: " + last_error())

    def integrate(self, fseries, sky_l):
        """correlator_network_integrate_power_fd on full spectra [n_inst][n]."""
        L = lib()
        fs = np.ascontiguousarray(fseries, dtype=np.complex128)
        ptrs = (c_void_p * fs.shape[0])(*[fs[i].ctypes.data for i in range(fs.shape[0])])
        sky = L.sh_series_new_zero(sky_l, 0)
        r = L.correlator_network_integrate_power_fd(sky, ptrs, self._plan)
        if not r:
            msg = last_error()
            L.sh_series_free(sky)
            raise RuntimeError("correlator_network_integrate_power_fd failed: " + msg)
        out = coeff_view(sky).copy()
        L.sh_series_free(sky)
        return out

    def __del__(self):
        try:
            L = lib()
            if getattr(self, "_plan", None):
                L.correlator_network_plan_fd_free(self._plan)
            if getattr(self, "_bl", None):
                L.correlator_network_baselines_free(self._bl)
            if getattr(self, "_inst", None):
                L.instrument_array_free(self._inst)
        except Exception:
            pass
        self._plan = self._bl = self._inst = None


def tseries_to_fseries(ts):
    """correlator_tseries_to_fseries through a plan (GPU)."""
    L = lib()
    ts = np.ascontiguousarray(ts, dtype=np.float64)
    fs = np.zeros(ts.shape[0], dtype=np.complex128)
    plan = L.correlator_tseries_to_fseries_plan(_ptr(ts), _ptr(fs), ts.shape[0])
    if not plan:
        raise RuntimeError("FFT plan failed: " + last_error())
    L.correlator_tseries_to_fseries(_ptr(ts), _ptr(fs), ts.shape[0], plan)
    L.correlator_fft_plan_free(plan)
    return fs


def ctseries_to_fseries(ts):
    L = lib()
    ts = np.ascontiguousarray(ts, dtype=np.complex128)
    fs = np.zeros(ts.shape[0], dtype=np.complex128)
    plan = L.correlator_ctseries_to_fseries_plan(_ptr(ts), _ptr(fs), ts.shape[0])
    if not plan:
        raise RuntimeError("FFT plan failed: " + last_error())
    L.correlator_ctseries_to_fseries(plan)
    L.correlator_fft_plan_free(plan)
    return fs


def to_mesh(coeff, l_max, polar=0):
    """sh_series_to_mesh (GPU) -> [2(l+1), 2(l+1)] complex128."""
    L = lib()
    s = series_from_numpy(coeff, l_max, polar)
    nt, nph = c_int(), c_int()
    m = L.sh_series_to_mesh(s, ctypes.byref(nt), ctypes.byref(nph), None)
    L.sh_series_free(s)
    if not m:
        raise RuntimeError("sh_series_to_mesh failed: " + last_error())
    n = nt.value
    buf = (ctypes.c_double * (2 * n * n)).from_address(m)
    out = np.frombuffer(buf, dtype=np.complex128).reshape(n, n).copy()
    ctypes.CDLL(None).free(c_void_p(m))
    return out


class DeviceBuffer:
    """A device allocation made with sphr_device_alloc."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().sphr_device_alloc(self.nbytes)
        if not self.ptr:
            raise RuntimeError("device allocation failed: " + last_error())

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if lib().sphr_memcpy_htod(b.ptr, _ptr(a), a.nbytes):
            raise RuntimeError(last_error())
        return b

    def to_numpy(self, dtype, shape, offset=0):
        out = np.empty(shape, dtype=dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("read past the end of the device buffer")
        if lib().sphr_memcpy_dtoh(_ptr(out), self.ptr + offset, out.nbytes):
            raise RuntimeError(last_error())
        return out

    def upload(self, a, offset=0):
        """Copy a host array into the buffer at byte offset."""
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("write past the end of the device buffer")
        if lib().sphr_memcpy_htod(self.ptr + offset, _ptr(a), a.nbytes):
            raise RuntimeError(last_error())

    def free(self):
        if self.ptr:
            lib().sphr_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


FP64, MIXED = 0, 1


class Batch:
    """sphradiometer_batch.h: device-resident batched integrate + synthesis."""

    def __init__(self, network, sky_l, precision=FP64, max_segments=1024):
        self.h = lib().sphr_batch_new(network.plan, sky_l, precision, max_segments)
        if not self.h:
            raise RuntimeError("sphr_batch_new failed: " + last_error())
        self.sky_l = sky_l
        self.precision = precision
        self.network = network

    def info(self):
        a = (c_int * 13)()
        lib().sphr_batch_info(self.h, a)
        keys = ["detectors", "baselines", "N", "sky_l", "columns", "k_splits", "precision", "k_per_split", "fused",
                "tiles_per_wave", "segments_per_workgroup", "sky_workgroups_per_cu", "sky_segments_per_workgroup"]
        d = dict(zip(keys, list(a)))
        d["lag"] = int(lib().sphr_batch_get_lag(self.h))
        d["xcorr_form"] = int(lib().sphr_batch_get_xcorr_form(self.h))
        d["table_bytes"] = int(lib().sphr_batch_table_bytes(self.h))
        return d

    def set_shape(self, index):
        """One of the autotuner's projection shapes (0..27)."""
        if lib().sphr_batch_set_shape(self.h, int(index)):
            raise ValueError(last_error())

    def set_lag(self, on):
        """Projection route: the lag-domain kernel (the default where it
        applies) or the frequency-domain FFT + GEMM pair."""
        if lib().sphr_batch_set_lag(self.h, int(bool(on))):
            raise ValueError(last_error())

    @property
    def lag(self):
        """True while the lag-domain route is in use."""
        return bool(lib().sphr_batch_get_lag(self.h))

    def set_xcorr_form(self, form):
        """The lag-domain kernel's form: 1 (Stockham FFT, table rows in
        registers), 2 (rows from L2), 3 (one wavefront per transform;
        mixed precision at 1024 samples), 4 (two pairs at once, 128
        threads per transform; mixed precision at 1024 samples)."""
        if lib().sphr_batch_set_xcorr_form(self.h, int(form)):
            raise RuntimeError("sphr_batch_set_xcorr_form failed: " + last_error())

    def set_fused(self, on):
        """Sky stage as the one fused kernel (default, where compiled for
        the order) or as the two-kernel form."""
        lib().sphr_batch_set_fused(self.h, int(bool(on)))

    def set_sky_occupancy(self, workgroups_per_cu):
        if lib().sphr_batch_set_sky_occupancy(self.h, int(workgroups_per_cu)):
            raise RuntimeError("sphr_batch_set_sky_occupancy failed: " + last_error())

    def set_sky_segments(self, segments_per_workgroup):
        if lib().sphr_batch_set_sky_segments(self.h, int(segments_per_workgroup)):
            raise RuntimeError("sphr_batch_set_sky_segments failed: " + last_error())

    def run(self, d_tseries, n_segments, d_alm, d_map=None, d_window=None, d_gmst=None, f32_input=False, stream=None):
        r = lib().sphr_batch_run(self.h, d_tseries, int(f32_input), n_segments, d_window, d_gmst, d_alm, d_map, stream)
        if r:
            raise RuntimeError("sphr_batch_run failed: " + last_error())

    def run_static(self, d_tseries, n_segments, d_alm, d_map=None, d_window=None, d_gmst=None, f32_input=False, stream=None):
        """sphr_batch_run_static: the same run with the library's static
        device scratch (no allocation)."""
        r = lib().sphr_batch_run_static(self.h, d_tseries, int(f32_input), n_segments, d_window, d_gmst, d_alm, d_map, stream)
        if r:
            raise RuntimeError("sphr_batch_run_static failed: " + last_error())

    def peaks(self, d_alm, n_segments, d_theta, d_phi, d_value, d_map=None, find_max=True, stream=None):
        """Queue the per-segment peak search (sphr_batch_peaks); theta, phi
        and value are device float64 buffers of n_segments."""
        r = lib().sphr_batch_peaks(self.h, d_alm, d_map, n_segments, int(find_max), d_theta, d_phi, d_value, stream)
        if r:
            raise RuntimeError("sphr_batch_peaks failed: " + last_error())

    def tune(self, d_tseries, n_segments, d_alm, d_map=None, d_window=None, d_gmst=None, f32_input=False):
        """Autotune the projection tiles on these buffers (sphr_batch_tune);
        returns (candidate index, ms per run)."""
        ms = c_double()
        k = lib().sphr_batch_tune(self.h, d_tseries, int(f32_input), n_segments, d_window, d_gmst, d_alm, d_map, ctypes.byref(ms))
        if k < 0:
            raise RuntimeError("sphr_batch_tune failed: " + last_error())
        return k, ms.value

    def profile(self, enable):
        lib().sphr_batch_profile(self.h, int(enable))

    def profile_read(self):
        ms = (c_double * 4)()
        runs = c_int()
        if lib().sphr_batch_profile_read(self.h, ms, ctypes.byref(runs)):
            raise RuntimeError(last_error())
        return list(ms), runs.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().sphr_batch_free(self.h)
        except Exception:
            pass
        self.h = None


def log_uniformsky_prior(l_max):
    s = lib().sh_series_log_uniformsky_prior(l_max)
    if not s:
        raise RuntimeError("sh_series_log_uniformsky_prior failed")
    out = coeff_view(s).copy()
    lib().sh_series_free(s)
    return out


def precalculated_length(length, delta_t):
    return int(lib().precalculated_TimeSeries_length(length, delta_t))


def preprocess_snr(series, delta_t):
    """preprocess_SNR_arrays on [n][length] complex series."""
    x = np.ascontiguousarray(series, dtype=np.complex128)
    n, ln = x.shape
    out = np.zeros((n, precalculated_length(ln, delta_t)), dtype=np.complex128)
    ip = (c_void_p * n)(*[x[i].ctypes.data for i in range(n)])
    op = (c_void_p * n)(*[out[i].ctypes.data for i in range(n)])
    if lib().preprocess_SNR_arrays(ip, n, ln, delta_t, op):
        raise RuntimeError("preprocess_SNR_arrays failed")
    return out


def generate_alm_skys(net_p, net_n, series, gmst, logprior, l_max):
    """generate_alm_skys_arrays on preprocessed [n][N] series; returns
    (sky_p, sky_n) coefficient arrays."""
    L = lib()
    x = np.ascontiguousarray(series, dtype=np.complex128)
    n, ln = x.shape
    ip = (c_void_p * n)(*[x[i].ctypes.data for i in range(n)])
    lp = series_from_numpy(logprior, l_max, 0)
    sp, sn = P(ShSeries)(), P(ShSeries)()
    r = L.generate_alm_skys_arrays(ctypes.byref(sp), ctypes.byref(sn), net_p.plan, net_n.plan, ip, ln, gmst, lp)
    L.sh_series_free(lp)
    if r:
        raise RuntimeError("generate_alm_skys_arrays failed: " + last_error())
    out = coeff_view(sp).copy(), coeff_view(sn).copy()
    L.sh_series_free(sp)
    L.sh_series_free(sn)
    return out


class Localizer:
    """sphradiometer_batch.h sphr_loc_*: batched localization skies."""

    def __init__(self, net_p, net_n, logprior, l_max, length, delta_t, precision=FP64, max_events=1024, power=0.6, overall=2.7):
        lp = series_from_numpy(logprior, l_max, 0)
        self.h = lib().sphr_loc_new(net_p.plan, net_n.plan, lp, l_max, power, overall, length, delta_t, precision, max_events)
        lib().sh_series_free(lp)
        if not self.h:
            raise RuntimeError("sphr_loc_new failed: " + last_error())
        self.l_max = l_max
        self.precision = precision
        self._keep = (net_p, net_n)

    def info(self):
        a = (c_int * 8)()
        lib().sphr_loc_info(self.h, a)
        return dict(zip(["detectors", "baselines", "N", "length", "l_max", "precision", "max_events"], list(a)[:7]))

    def run(self, d_series, n_events, d_alm, d_map=None, d_gmst=None, series_f64=False, stream=None):
        if lib().sphr_loc_run(self.h, d_series, int(series_f64), n_events, d_gmst, d_alm, d_map, stream):
            raise RuntimeError("sphr_loc_run failed: " + last_error())

    def peaks(self, d_alm, n_events, d_theta, d_phi, d_value, d_map=None, stream=None):
        if lib().sphr_loc_peaks(self.h, d_alm, d_map, n_events, d_theta, d_phi, d_value, stream):
            raise RuntimeError("sphr_loc_peaks failed: " + last_error())

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().sphr_loc_free(self.h)
        except Exception:
            pass
        self.h = None


class Healpix:
    """sphradiometer_batch.h sphr_healpix_*: RING-ordered HEALPix maps of
    many series on the GPU (Re f at the pixel centres), and the posterior
    of the beta = +1 / -1 skies of localization events."""

    def __init__(self, l_max, nside=None, precision=FP64):
        self.l_max = int(l_max)
        self.nside = int(nside) if nside is not None else lib().sphr_healpix_nside(self.l_max)
        self.precision = precision
        self.h = lib().sphr_healpix_new(self.l_max, self.nside, precision)
        if not self.h:
            raise RuntimeError("sphr_healpix_new failed: " + last_error())
        self.npix = lib().sphr_healpix_npix(self.h)

    def maps(self, d_alm, n, d_map, stream=None):
        """d_alm [n][(l_max+1)^2] complex -> d_map [n][npix] real (device)."""
        if lib().sphr_healpix_maps(self.h, d_alm, n, d_map, stream):
            raise RuntimeError("sphr_healpix_maps failed: " + last_error())

    def posterior(self, d_maps, n_events, d_prob, stream=None):
        """d_maps [2][n_events][npix] -> d_prob [n_events][npix] (device)."""
        if lib().sphr_healpix_posterior(self.h, d_maps, n_events, d_prob, stream):
            raise RuntimeError("sphr_healpix_posterior failed: " + last_error())

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().sphr_healpix_free(self.h)
        except Exception:
            pass
        self.h = None


class Jobs:
    """sphradiometer_batch.h sphr_jobs_*: a batch of independent jobs with
    their own networks, lengths and orders (plans made on the GPU)."""

    def __init__(self, jobs, delta_t, precision=FP64):
        n = len(jobs)
        nd = (c_int * n)(*[len(j["positions"]) for j in jobs])
        pos = np.ascontiguousarray(np.concatenate([np.asarray(j["positions"], dtype=np.float64) for j in jobs]))
        Ns = (c_int * n)(*[int(j["N"]) for j in jobs])
        Ls = (c_int * n)(*[int(j["l_max"]) for j in jobs])
        self.h = lib().sphr_jobs_new(n, nd, pos.ctypes.data_as(P(c_double)), Ns, Ls, delta_t, precision)
        if not self.h:
            raise RuntimeError("sphr_jobs_new failed: " + last_error())
        self.n = n
        self.precision = precision
        self.l_max = [int(j["l_max"]) for j in jobs]
        a = (ctypes.c_long * n)()
        b = (ctypes.c_long * n)()
        c = (ctypes.c_long * n)()
        lib().sphr_jobs_offsets(self.h, a, b, c)
        self.off_in, self.off_alm, self.off_map = list(a), list(b), list(c)

    def info(self):
        a = (ctypes.c_long * 8)()
        lib().sphr_jobs_info(self.h, a)
        return dict(zip(["in_floats", "alm_elems", "map_elems", "baselines", "l_classes", "lag_samples", "lag_threads", "plan_bytes"], list(a)))

    def run(self, d_in, d_alm, d_map=None, stream=None):
        if lib().sphr_jobs_run(self.h, d_in, d_alm, d_map, stream):
            raise RuntimeError("sphr_jobs_run failed: " + last_error())

    def tune(self, d_in, d_alm):
        ms = c_double()
        k = lib().sphr_jobs_tune(self.h, d_in, d_alm, ctypes.byref(ms))
        if k < 0:
            raise RuntimeError("sphr_jobs_tune failed: " + last_error())
        return k, ms.value

    def set_lag_tile(self, samples, threads):
        if lib().sphr_jobs_set_lag_tile(self.h, int(samples), int(threads)):
            raise RuntimeError("sphr_jobs_set_lag_tile failed: " + last_error())

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().sphr_jobs_free(self.h)
        except Exception:
            pass
        self.h = None


def synchronize():
    if lib().sphr_device_synchronize():
        raise RuntimeError(last_error())
