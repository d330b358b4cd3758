"""The ABI from C: tests/c/fd_validate.c includes only the public headers
and links libsphradiometer_amd.so directly, along the frequency-domain flow
of the reference's test/correlator_validate.c:221-406.

CPU: the program compiles warning-free against include/ with the system C
compiler (the headers stand alone, struct fields and prototypes as a C
caller sees them) and the built program loads the library.
GPU: the reference's accuracy test (a white-noise point source within
-21 dB RMS of the analytic cross power, correlator_validate.c:376) and
the three-site segment against the golden fixture and the oracle.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "c", "fd_validate.c")
PROG = os.path.join(HERE, "c", "fd_validate")
LIBDIR = os.path.join(ROOT, "sphradiometer_amd")
GOLD = os.path.join(HERE, "golden", "oracle_cases.npz")


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_c_caller_compiles_against_headers(tmp_path):
    cc = shutil.which("cc") or shutil.which("gcc")
    out = str(tmp_path / "fd_validate")
    r = subprocess.run([cc, "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-o", out, SRC, "-L" + LIBDIR, "-lsphradiometer_amd", "-Wl,-rpath," + LIBDIR,
                        "-Wl,-rpath-link,/opt/rocm/lib", "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # loads the library and rejects a bad command line without touching the GPU
    r = subprocess.run([out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def _prog():
    if not os.path.exists(PROG):
        pytest.fail("tests/c/fd_validate is not built (run build())")
    return PROG


@pytest.mark.gpu
def test_c_caller_white_noise_point_source():
    print("fd_validate analytic", flush=True)
    r = subprocess.run([_prog(), "analytic"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    dbs = [float(line.split("=")[-1].split()[0]) for line in r.stdout.splitlines() if "dB" in line]
    assert len(dbs) == 2 and max(dbs) <= -21.0, dbs


@pytest.mark.gpu
def test_c_caller_segment_matches_golden(oracle, tmp_path):
    from sphradiometer_amd import inject
    ts, _ = inject.three_site(1, 1024, seed=0)
    x = np.ascontiguousarray(ts[:, 0])
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    x.tofile(fin)
    print("fd_validate segment", flush=True)
    r = subprocess.run([_prog(), "segment", fin, fout], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(fout, dtype=np.complex128)
    L = 32
    alm, mesh = out[:(L + 1) ** 2], out[(L + 1) ** 2:].reshape(2 * L + 2, 2 * L + 2)
    gold = np.load(GOLD)
    assert rel(alm, gold["three_site_alm_0"]) < 1e-12
    assert rel(mesh, gold["three_site_map_0"]) < 1e-12
    want, wmesh = oracle.Network(inject.positions("ABC"), 1024, 1.0 / 4096, L).segment(x, L)
    assert rel(alm, want) < 1e-12 and rel(mesh, wmesh) < 1e-12
