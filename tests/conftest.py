import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def _sep(t1, p1, t2, p2):
    import math
    c = math.cos(t1) * math.cos(t2) + math.sin(t1) * math.sin(t2) * math.cos(p1 - p2)
    return math.acos(max(-1.0, min(1.0, c)))


def check_local_minimum(O, c, L, th, ph, v, tol, seed=True, polar=False):
    """(th, ph, v) is a local minimum of Re f for the full series c: v is
    f's value there; the oracle's Nelder-Mead started at (th, ph) finds
    nothing lower than v - tol and stays within 1e-5 rad; and v is not above
    f at the l = 6 seed the reference's search picks (sh_series_extremize.c:
    296-300), i.e. the point was reached by descent from that seed.

    Which basin a descent from the seed ends in depends on the method's
    steps (GSL BFGS2 with a pi/24 first step in the reference, a damped
    Newton in the library, Nelder-Mead in the oracle), so identity of the
    basin is not asserted.  For an axially symmetric (polar) surface only
    theta is compared: phi is free along the minimum's ring."""
    import math
    assert 0.0 <= th <= math.pi and 0.0 <= ph < 2 * math.pi, (th, ph)
    f = O.eval_series(c, L, 0, th, ph).real
    assert abs(f - v) <= tol, ("value", f, v)
    to, po, vo = O.refine_minimum(c, L, th, ph, size=1e-4)
    assert vo >= v - tol, ("lower point nearby", vo, v)
    moved = abs(th - to) if polar else _sep(th, ph, to, po)
    assert moved < 1e-5, ("moved", th, ph, to, po)
    if seed:
        ts, ps, _ = O.find_near_minimum(c, L)
        assert v <= O.eval_series(c, L, 0, ts, ps).real + tol, ("above the seed", v)


@pytest.fixture(scope="session")
def local_minimum(oracle):
    return lambda c, L, th, ph, v, tol, seed=True, polar=False: check_local_minimum(oracle, c, L, th, ph, v, tol, seed, polar)
