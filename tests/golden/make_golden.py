"""Writes the small fixtures the tests pin against.

pins.npz (independent of every implementation in this repository):
  threej_args, threej_vals   exact Wigner 3j from sympy's rational
                             evaluation (the role gsl_sf_coupling_3j plays
                             in sh_series_product.c:80-83)
  gl_n, gl_x, gl_w           Gauss-Legendre nodes/weights in 50-digit
                             mpmath (pixels_from_l_max,
                             sh_series_trans.c:149-196), ascending
  ylm_args, ylm_vals         Y_lm(theta, phi) in 30-digit mpmath with the
                             Condon-Shortley phase (sh_series_trans.c:65-69)
  taps_*                     delay-kernel SH coefficients (projection.c:
                             99-131 through sh_series_from_realfunc) for a
                             small baseline, by the same 2(l+1)-node
                             quadrature evaluated in 40-digit mpmath

oracle_cases.npz (written by the oracle; the GPU parity tests compare
against these as well as against the oracle run live):
  two-site workload (l_max 25): a_lm of segments 0, 1 and the map of 0
  three-site workload (l_max 32): a_lm of segments 0, 1 and the map of 0

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def _mp_gl(n, mpmath):
    """n-point Gauss-Legendre nodes (ascending) and weights by Newton on
    the three-term recurrence in the current mpmath precision."""
    def pn(t):
        p0, p1 = mpmath.mpf(1), t
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * t * p1 - (k - 1) * p0) / k
        return p1, n * (t * p1 - p0) / (t * t - 1)
    xs, ws = [], []
    for i in range(n):
        t = -mpmath.cos(mpmath.pi * (i + mpmath.mpf(3) / 4) / (n + mpmath.mpf(1) / 2))
        for _ in range(200):
            p, dp = pn(t)
            dt = p / dp
            t -= dt
            if abs(dt) < mpmath.mpf(10) ** (-mpmath.mp.dps + 5):
                break
        p, dp = pn(t)
        xs.append(t)
        ws.append(2 / ((1 - t * t) * dp * dp))
    return xs, ws


def make_pins():
    import mpmath
    from sympy.physics.wigner import wigner_3j

    rng = np.random.default_rng(1234)
    args, vals = [], []
    fixed = [(2, 3, 4, 1, -1, 0), (10, 12, 20, 3, -5, 2), (45, 50, 59, -7, 3, 4),
             (59, 59, 59, -20, 30, -10), (60, 60, 118, 0, 0, 0), (30, 25, 41, 0, 0, 0),
             (1, 1, 2, 1, -1, 0), (0, 0, 0, 0, 0, 0), (7, 5, 3, 2, 1, -3)]
    for a in fixed:
        args.append(a)
    while len(args) < 80:
        j1 = int(rng.integers(0, 60))
        j2 = int(rng.integers(0, 60))
        j3 = int(rng.integers(abs(j1 - j2), j1 + j2 + 1))
        m1 = int(rng.integers(-j1, j1 + 1))
        m2 = int(rng.integers(-j2, j2 + 1))
        m3 = -m1 - m2
        if abs(m3) > j3:
            continue
        args.append((j1, j2, j3, m1, m2, m3))
    for a in args:
        vals.append(float(wigner_3j(*a)))

    mpmath.mp.dps = 50
    gl_n, gl_x, gl_w = [], [], []
    for n in (8, 52, 66):
        xs, ws = _mp_gl(n, mpmath)
        for r, w in zip(xs, ws):
            gl_n.append(n)
            gl_x.append(float(r))
            gl_w.append(float(w))

    mpmath.mp.dps = 30
    yargs, yvals = [], []
    for (l, m, th, ph) in [(0, 0, 0.3, 0.1), (1, 0, 1.1, 0.2), (2, 0, 2.0, 0.0), (7, 1, 0.4, 1.3),
                           (8, -7, 1.9, 0.7), (8, -6, 0.8, 2.1), (8, 6, 2.7, 5.5), (8, 7, 1.2, 3.3),
                           (10, -10, 1.4, 0.9), (10, 3, 0.6, 4.4), (10, 10, 2.2, 6.0),
                           (25, 13, 0.9, 1.7), (40, -21, 2.5, 0.3), (64, 33, 1.0, 2.0), (64, 0, 0.05, 0.0)]:
        v = mpmath.spherharm(l, m, th, ph)
        yargs.append((l, m, th, ph))
        yvals.append(complex(v))

    # delay kernel of a small baseline: D = 0.05 s, delta_t = 1/64 s
    mpmath.mp.dps = 40
    dt = 1.0 / 64
    r = 0.025
    l_max = int(np.pi * r / dt + 1)          # projection_matrix_l_max
    n_nodes = 2 * (l_max + 1)
    xs, ws = _mp_gl(n_nodes, mpmath)
    taps_k = [-3, -1, 0, 2, 5, 40]
    taps = np.zeros((len(taps_k), l_max + 1))
    for ik, k in enumerate(taps_k):
        for l in range(l_max + 1):
            acc = mpmath.mpf(0)
            for rt, wt in zip(xs, ws):
                px = mpmath.pi * (-k - r * rt / dt)
                f = 1 if px == 0 else mpmath.sin(px) / px
                # orthonormal Pbar_l0 = sqrt((2l+1)/(4pi)) P_l
                acc += 2 * mpmath.pi * wt * f * mpmath.sqrt((2 * l + 1) / (4 * mpmath.pi)) * mpmath.legendre(l, rt)
            taps[ik, l] = float(acc)

    np.savez(os.path.join(HERE, "pins.npz"),
             threej_args=np.array(args, dtype=np.int32), threej_vals=np.array(vals),
             gl_n=np.array(gl_n, dtype=np.int32), gl_x=np.array(gl_x), gl_w=np.array(gl_w),
             ylm_args=np.array(yargs), ylm_vals=np.array(yvals),
             taps_k=np.array(taps_k, dtype=np.int32), taps=taps,
             taps_params=np.array([dt, r, l_max]))


def make_oracle_cases():
    from oracle import oracle as O
    from sphradiometer_amd import inject

    out = {}
    dt = 1.0 / 4096
    ts = inject.two_site(16, 4096, seed=0)
    net = O.Network(inject.positions("AB"), 4096, dt, 25)
    w = inject.hann(4096)
    for s in (0, 1):
        sky, mesh = net.segment(ts[:, s, :], 25, window=w)
        out["two_site_alm_%d" % s] = sky
        if s == 0:
            out["two_site_map_0"] = mesh
    ts3, _ = inject.three_site(4, 1024, seed=0)
    net3 = O.Network(inject.positions("ABC"), 1024, dt, 32)
    for s in (0, 1):
        sky, mesh = net3.segment(ts3[:, s, :], 32)
        out["three_site_alm_%d" % s] = sky
        if s == 0:
            out["three_site_map_0"] = mesh
    np.savez(os.path.join(HERE, "oracle_cases.npz"), **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["pins", "oracle"]
    if "pins" in which:
        make_pins()
    if "oracle" in which:
        make_oracle_cases()
