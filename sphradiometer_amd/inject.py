"""Synthetic detector data for the cross-power path.

Restates the conventions of the reference's injection helpers
(src/lib/inject.c): white Gaussian noise (inject.c:48-56) and plane-wave
injection by a circular FFT phase shift, x_i(t) = h(t + r_i . s)
(inject_transform_point_source, inject.c:119-152; inject_into,
inject.c:187-203, with theta = pi/2 - dec and phi = ra - gmst).  Random
numbers come from numpy's PCG64 so a seed reproduces the same data on
every machine.

The workloads are the ones the throughput and parity runs use:

  two_site      2 sites (46.45N, 119.41W), (30.56N, 90.77W), 4096 Hz,
                16 x 4096-sample segments, Hann window, l_max 25.
  three_site    adds (43.63N, 10.50E); 1024 x 1024-sample segments,
                l_max 32, white noise plus one sine-Gaussian burst
                (f0 ~ U(100, 400) Hz, Q = 9, network SNR 20, isotropic
                direction) in every segment.
"""
import math

import numpy as np

C_SI = 299792458.0
EARTH_RADIUS_SPHERE = 6371e3

SITES = {
    "A": (46.45, -119.41),
    "B": (30.56, -90.77),
    "C": (43.63, 10.50),
}


def site_position(lat_deg, lon_deg, radius=EARTH_RADIUS_SPHERE):
    """Phase centre in light-seconds, as instrument_new_from_r_theta_phi
    (instrument.c:164-172) builds it from (r, theta = pi/2 - lat, phi = lon)."""
    r = radius / C_SI
    th = math.pi / 2 - math.radians(lat_deg)
    ph = math.radians(lon_deg)
    return np.array([r * math.sin(th) * math.cos(ph), r * math.sin(th) * math.sin(ph), r * math.cos(th)])


def positions(names):
    return np.array([site_position(*SITES[n]) for n in names])


def hann(n):
    """Symmetric Hann window w[k] = 0.5 - 0.5 cos(2 pi k / (n - 1))."""
    k = np.arange(n)
    return 0.5 - 0.5 * np.cos(2 * np.pi * k / (n - 1))


def delay_circular(x, delay_s, delta_t):
    """x(t + delay), circularly, by an FFT phase shift (inject.c:119-152)."""
    n = x.shape[-1]
    ft = np.fft.rfft(x)
    f = np.arange(ft.shape[-1])
    ft = ft * np.exp(2j * np.pi * delay_s * f / (delta_t * n))
    return np.fft.irfft(ft, n)


def r_dot_s(r, theta, phi):
    """instrument.c:77-80"""
    return (r[0] * math.cos(phi) + r[1] * math.sin(phi)) * math.sin(theta) + r[2] * math.cos(theta)


def sine_gaussian(n, delta_t, f0, q, t0):
    t = np.arange(n) * delta_t - t0
    tau = q / (math.sqrt(2.0) * math.pi * f0)
    return np.sin(2 * math.pi * f0 * t) * np.exp(-(t / tau) ** 2)


def isotropic_direction(rng):
    """(theta, phi) uniform on the sphere."""
    return math.acos(rng.uniform(-1.0, 1.0)), rng.uniform(0.0, 2 * math.pi)


def two_site(n_segments=16, n=4096, seed=0):
    """White N(0,1) float64 at 4096 Hz, one stream per detector (generator
    seeded with (seed, detector)), cut into non-overlapping segments, so the
    first k segments do not depend on n_segments.  Returns tseries
    [n_det][n_segments][n]."""
    out = np.empty((2, n_segments, n))
    for d in range(2):
        out[d] = np.random.default_rng([seed, d]).standard_normal(n_segments * n).reshape(n_segments, n)
    return out


def three_site(n_segments=1024, n=1024, seed=0, snr=20.0, delta_t=1.0 / 4096):
    """White N(0,1) float64 plus one sine-Gaussian plane wave per segment,
    delayed per site by r_i . s.  Returns (tseries [3][S][n], injections),
    injections = [(theta, phi, f0)] per segment.  Segment s draws from a
    generator seeded with (seed, s), so the first k segments do not depend
    on n_segments."""
    pos = positions("ABC")
    data = np.empty((3, n_segments, n))
    inj = []
    for s in range(n_segments):
        rng = np.random.default_rng([seed, s])
        data[:, s] = rng.standard_normal((3, n))
        f0 = rng.uniform(100.0, 400.0)
        theta, phi = isotropic_direction(rng)
        h = sine_gaussian(n, delta_t, f0, 9.0, 0.5 * n * delta_t)
        # network SNR^2 = sum over the 3 sites of sum_t h^2 (unit white noise)
        h *= snr / math.sqrt(3.0 * float(np.dot(h, h)))
        for d in range(3):
            data[d, s] += delay_circular(h, r_dot_s(pos[d], theta, phi), delta_t)
        inj.append((theta, phi, f0))
    return data, inj


# ---------------------------------------------------------------- localization

# x-arm azimuths (degrees east of north) of the L-shaped detectors placed at
# the three sites; the y arm is 90 degrees further east.  A fixed choice of
# this module, used only to give the projection operator a response tensor.
ARM_AZIMUTH = {"A": 36.0, "B": 108.0, "C": 19.0}


def detector_response(lat_deg, lon_deg, azimuth_deg):
    """Response tensor D = (u u^T - v v^T) / 2 of an L-shaped detector with
    arms u (azimuth) and v (azimuth + 90 deg) in the local horizontal plane,
    Earth-fixed Cartesian axes."""
    la, lo, az = math.radians(lat_deg), math.radians(lon_deg), math.radians(azimuth_deg)
    east = np.array([-math.sin(lo), math.cos(lo), 0.0])
    north = np.array([-math.sin(la) * math.cos(lo), -math.sin(la) * math.sin(lo), math.cos(la)])
    u = math.cos(az) * north + math.sin(az) * east
    v = math.cos(az + math.pi / 2) * north + math.sin(az + math.pi / 2) * east
    return 0.5 * (np.outer(u, u) - np.outer(v, v))


def responses(names):
    return np.array([detector_response(*SITES[n], ARM_AZIMUTH[n]) for n in names])


def snr_events(n_events, length=410, delta_t=1.0 / 4096, seed=0, snr_range=(8.0, 30.0), names="ABC"):
    """Matched-filter SNR series around a trigger, one event at a time:
    unit-variance complex Gaussian noise plus a chirp-like complex peak
    A_d exp(-(t - t_d)^2 / (2 s^2)) exp(2 pi i (f_c (t - t_d) + k (t - t_d)^2 / 2))
    with s = 2 ms, f_c = 150 Hz, k = 2000 Hz/s; t_d = -r_d . n for an
    isotropic direction n (Earth frame), A_d = rho a_d / |a| with
    a_d = (F+_d + i Fx_d) e^{i phi0}, F+_d and Fx_d uniform in [-1, 1] (random
    antenna factors), rho uniform in snr_range, phi0 uniform.  Samples at
    t_k = (k - length // 2) delta_t.  Event e draws from a generator seeded
    with (seed, e).  Returns (series [n_det][n_events][length] complex64,
    meta dict of per-event theta, phi, snr, gmst)."""
    pos = positions(names)
    D = len(names)
    out = np.empty((D, n_events, length), dtype=np.complex64)
    meta = {"theta": np.empty(n_events), "phi": np.empty(n_events), "snr": np.empty(n_events), "gmst": np.empty(n_events)}
    t = (np.arange(length) - length // 2) * delta_t
    s, fc, k = 2e-3, 150.0, 2000.0
    for e in range(n_events):
        rng = np.random.default_rng([seed, e])
        noise = (rng.standard_normal((D, length)) + 1j * rng.standard_normal((D, length))) / math.sqrt(2.0)
        theta, phi = isotropic_direction(rng)
        rho = rng.uniform(*snr_range)
        fpx = rng.uniform(-1.0, 1.0, (D, 2))
        phi0 = rng.uniform(0.0, 2 * math.pi)
        a = (fpx[:, 0] + 1j * fpx[:, 1]) * np.exp(1j * phi0)
        A = rho * a / np.linalg.norm(a)
        for d in range(D):
            td = -r_dot_s(pos[d], theta, phi)
            u = t - td
            noise[d] += A[d] * np.exp(-u * u / (2 * s * s)) * np.exp(2j * math.pi * (fc * u + 0.5 * k * u * u))
        out[:, e] = noise
        meta["theta"][e], meta["phi"][e], meta["snr"][e] = theta, phi, rho
        meta["gmst"][e] = rng.uniform(0.0, 2 * math.pi)
    return out, meta


def burst_batch(n_segments, n=1024, seed=0, snr_range=(8.0, 30.0), delta_t=1.0 / 4096, names="ABC"):
    """Training and evaluation skies for the learned peak proposal: white
    N(0,1) float32 plus one sine-Gaussian plane wave per segment (f0 ~
    U(100, 400) Hz, Q = 9, network SNR ~ U(snr_range), isotropic
    direction), delayed per site by r_d . s and evaluated in closed form
    (no circular shift; the burst sits mid-segment).  Vectorised over
    segments; one generator seeded with seed.  Returns (tseries
    [n_det][S][n] float32, dict of theta, phi, snr, f0 per segment)."""
    rng = np.random.default_rng(seed)
    pos = positions(names)
    D = len(names)
    out = rng.standard_normal((D, n_segments, n), dtype=np.float32)
    f0 = rng.uniform(100.0, 400.0, n_segments)
    cth = rng.uniform(-1.0, 1.0, n_segments)
    phi = rng.uniform(0.0, 2 * math.pi, n_segments)
    snr = rng.uniform(*snr_range, n_segments)
    theta = np.arccos(cth)
    tau = 9.0 / (math.sqrt(2.0) * math.pi * f0)
    # sum_t h^2 of a unit sine-Gaussian ~ tau sqrt(pi/2) / (2 delta_t)
    amp = snr / np.sqrt(D * tau * math.sqrt(math.pi / 2) / (2 * delta_t))
    st = np.sin(theta)
    sdir = np.stack([st * np.cos(phi), st * np.sin(phi), cth], axis=1)
    t = np.arange(n) * delta_t - 0.5 * n * delta_t
    for d in range(D):
        delay = sdir @ pos[d]
        u = t[None, :] + delay[:, None]
        h = amp[:, None] * np.sin(2 * math.pi * f0[:, None] * u) * np.exp(-(u / tau[:, None]) ** 2)
        out[d] += h.astype(np.float32)
    return out, {"theta": theta, "phi": phi, "snr": snr, "f0": f0}


def uneven_jobs(n_jobs, seed=0, delta_t=1.0 / 4096, l_range=(16, 64), det_range=(2, 5),
                durations=(0.125, 0.25, 0.5, 1.0, 2.0), snr=20.0):
    """The uneven batch: per job l_max ~ U{l_range}, detector count ~
    U{det_range} with sites uniform on the 6371 km sphere, segment length
    ~ U(durations) at 1 / delta_t, float32 white N(0,1) data and, in every
    other job (odd index), one sine-Gaussian plane wave (f0 ~ U(100, 400)
    Hz, Q = 9, network SNR snr, isotropic direction) delayed per site by
    r . s in closed form.  Job j draws from a generator seeded (seed, j).
    Returns (jobs, data): jobs a list of dicts (positions [D][3]
    light-seconds, N, l_max, burst), data the float32 samples packed job by
    job as [D][N]."""
    jobs, chunks = [], []
    r = EARTH_RADIUS_SPHERE / C_SI
    for j in range(n_jobs):
        rng = np.random.default_rng([seed, j])
        D = int(rng.integers(det_range[0], det_range[1] + 1))
        L = int(rng.integers(l_range[0], l_range[1] + 1))
        N = int(round(durations[int(rng.integers(0, len(durations)))] / delta_t))
        v = rng.standard_normal((D, 3))
        pos = r * v / np.linalg.norm(v, axis=1, keepdims=True)
        x = rng.standard_normal((D, N)).astype(np.float32)
        burst = bool(j & 1)
        if burst:
            f0 = rng.uniform(100.0, 400.0)
            th, ph = isotropic_direction(rng)
            sdir = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
            tau = 9.0 / (math.sqrt(2.0) * math.pi * f0)
            t = np.arange(N) * delta_t - 0.5 * N * delta_t
            amp = snr / math.sqrt(D * tau * math.sqrt(math.pi / 2) / (2 * delta_t))
            for d in range(D):
                u = t + float(pos[d] @ sdir)
                x[d] += (amp * np.sin(2 * math.pi * f0 * u) * np.exp(-(u / tau) ** 2)).astype(np.float32)
        jobs.append({"positions": pos, "N": N, "l_max": L, "burst": burst})
        chunks.append(x.ravel())
    return jobs, np.concatenate(chunks)


def proposal_jobs(n_jobs, seed=0, l_max=32, delta_t=1.0 / 4096, det_range=(2, 5),
                  durations=(0.125, 0.25, 0.5, 1.0, 2.0), snr_range=(8.0, 30.0)):
    """Skies of the uneven workload at one order, for the learned peak
    proposal: per job a detector count ~ U{det_range} with sites uniform on
    the 6371 km sphere, a segment length ~ U(durations) at 1 / delta_t, and
    one sine-Gaussian plane wave (f0 ~ U(100, 400) Hz, Q = 9, network SNR
    ~ U(snr_range), isotropic direction) delayed per site by r . s,
    mid-segment, in white N(0,1) float32 noise.  The jobs come out ordered
    by class of equal (detectors, length), so each class's packed data
    [job][detector][sample] is one contiguous block that the device
    injector (sphr_inject_bursts with one detector per series) fills.
    Returns (jobs, classes, injections): classes a list of (float offset,
    series = jobs x detectors, length, params float32 [series][4] = amp,
    f0, tau, delay); injections the theta, phi, snr per job in job order."""
    rng = np.random.default_rng(seed)
    r = EARTH_RADIUS_SPHERE / C_SI
    D = rng.integers(det_range[0], det_range[1] + 1, n_jobs)
    Nn = np.array([int(round(d / delta_t)) for d in durations])
    N = Nn[rng.integers(0, len(durations), n_jobs)]
    jobs, classes = [], []
    th_all, ph_all, snr_all = [], [], []
    off = 0
    for d in range(det_range[0], det_range[1] + 1):
        for n in Nn:
            S = int(np.sum((D == d) & (N == n)))
            if S == 0:
                continue
            v = rng.standard_normal((S, d, 3))
            pos = r * v / np.linalg.norm(v, axis=2, keepdims=True)
            f0 = rng.uniform(100.0, 400.0, S)
            cth = rng.uniform(-1.0, 1.0, S)
            phi = rng.uniform(0.0, 2 * math.pi, S)
            snr = rng.uniform(*snr_range, S)
            st = np.sqrt(1.0 - cth * cth)
            sdir = np.stack([st * np.cos(phi), st * np.sin(phi), cth], axis=1)
            tau = 9.0 / (math.sqrt(2.0) * math.pi * f0)
            amp = snr / np.sqrt(d * tau * math.sqrt(math.pi / 2) / (2 * delta_t))
            par = np.empty((S, d, 4), dtype=np.float32)
            par[:, :, 0] = amp[:, None]
            par[:, :, 1] = f0[:, None]
            par[:, :, 2] = tau[:, None]
            par[:, :, 3] = np.einsum("skj,sj->sk", pos, sdir)
            classes.append((off, S * d, int(n), par.reshape(S * d, 4)))
            off += S * d * int(n)
            for s in range(S):
                jobs.append({"positions": pos[s], "N": int(n), "l_max": l_max, "burst": True})
            th_all.append(np.arccos(cth))
            ph_all.append(phi)
            snr_all.append(snr)
    return jobs, classes, {"theta": np.concatenate(th_all), "phi": np.concatenate(ph_all), "snr": np.concatenate(snr_all)}
