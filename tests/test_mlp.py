"""The learned peak proposal: the network's GEMM, forward, loss and
gradients against the float64 numpy oracle (oracle/mlp.py), the feature
and data kernels, and a small end-to-end A/B run.

Oracle pins (CPU): central finite differences of the oracle's loss match
its analytic gradient; the real-form feature layout on a hand-written
L = 1 case.  The GPU checks use float32 operands for gradient parity
(1e-4) and bfloat16 operands against bf16-rounded inputs for the GEMM.
"""
import math

import numpy as np
import pytest

import sphradiometer_amd as S
from oracle import mlp as OM


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def small_problem(seed=0, n_in=13, width=9, depth=2, K=3, rows=7):
    rng = np.random.default_rng(seed)
    n = sum(i * o + o for i, o in OM.shapes(n_in, width, depth, K))
    p = rng.standard_normal(n) * 0.4
    x = rng.standard_normal((rows, n_in))
    t = rng.standard_normal((rows, 3))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    return p, x, t, (n_in, width, depth, K)


def test_oracle_gradient_finite_differences():
    p, x, t, shp = small_problem()
    lam = 0.05
    _, g, _ = OM.loss_grad(p, x, t, *shp, lam)
    rng = np.random.default_rng(1)
    h = 1e-6
    for i in rng.choice(p.size, 25, replace=False):
        e = np.zeros_like(p)
        e[i] = h
        num = (OM.loss_grad(p + e, x, t, *shp, lam)[0] - OM.loss_grad(p - e, x, t, *shp, lam)[0]) / (2 * h)
        assert abs(num - g[i]) < 1e-7 + 1e-5 * abs(g[i]), (i, num, g[i])


def test_oracle_feature_layout_l1():
    # L = 1: storage order a00, a10, a11, a1-1 -> [Re a00, Re a10, Re a11, Im a11]
    a = np.array([1 + 0j, 2 + 0j, 3 + 4j, -3 + 4j])
    f = OM.features(a, 1)
    raw = np.array([1.0, 2.0, 3.0, 4.0])
    assert np.allclose(f, raw / math.sqrt(np.mean(raw ** 2)), rtol=0, atol=1e-15)


# ------------------------------------------------------------------- GPU


def bf16_round(x):
    """float32 -> bfloat16 (round to nearest even) -> float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r << 16).astype(np.uint32).view(np.float32)


@pytest.fixture(scope="module")
def P():
    from sphradiometer_amd import proposal
    S.lib()
    if not S.gpu_available():
        pytest.fail("no GPU: " + S.last_error())
    proposal._protos()
    return proposal


@pytest.mark.gpu
@pytest.mark.parametrize("bf16", [0, 1])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1)])
def test_gpu_mlp_gemm(P, bf16, ta, tb):
    rng = np.random.default_rng(3)
    M, N, K = 77, 45, 70
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    dA, dB, dC = S.DeviceBuffer.from_numpy(A), S.DeviceBuffer.from_numpy(B), S.DeviceBuffer(M * N * 4)
    print(f"gemm bf16={bf16} ta={ta} tb={tb}", flush=True)
    assert S.lib().sphr_mlp_gemm_check(bf16, ta, tb, M, N, K, dA.ptr, dB.ptr, dC.ptr) == 0
    C = dC.to_numpy(np.float32, (M, N)).astype(np.float64)
    Ar = bf16_round(A) if bf16 else A
    Br = bf16_round(B) if bf16 else B
    Ao = Ar.astype(np.float64).T if ta else Ar.astype(np.float64)
    Bo = Br.astype(np.float64).T if tb else Br.astype(np.float64)
    assert rel(C, Ao @ Bo) < 1e-5


@pytest.mark.gpu
def test_gpu_mlp_loss_and_gradients_match_oracle(P):
    shp = (61, 40, 3, 4)
    rows = 50
    rng = np.random.default_rng(5)
    m = P.MLP(*shp, max_batch=64, bf16=False, seed=11)
    p = m.params().astype(np.float64)
    x = rng.standard_normal((rows, shp[0])).astype(np.float32)
    t = rng.standard_normal((rows, 3))
    t = (t / np.linalg.norm(t, axis=1, keepdims=True)).astype(np.float32)
    dx, dt, dl = S.DeviceBuffer.from_numpy(x), S.DeviceBuffer.from_numpy(t), S.DeviceBuffer(rows * 4)
    m.loss_grad(dx.ptr, dt.ptr, rows, 0.05, dl.ptr)
    S.synchronize()
    lo, go, lrow = OM.loss_grad(p, x.astype(np.float64), t.astype(np.float64), *shp, 0.05)
    lg = dl.to_numpy(np.float32, (rows,))
    assert rel(lg, lrow) < 1e-5
    assert rel(m.grads(), go) < 1e-4
    # bfloat16 operands: the same loss to bf16 accuracy
    mb = P.MLP(*shp, max_batch=64, bf16=True, seed=11)
    mb.loss_grad(dx.ptr, dt.ptr, rows, 0.05, dl.ptr)
    S.synchronize()
    assert rel(dl.to_numpy(np.float32, (rows,)), lrow) < 3e-2
    assert rel(mb.grads(), go) < 0.1


@pytest.mark.gpu
def test_gpu_mlp_full_size_gradients_match_oracle(P):
    """the proposal network at its real size: 1089 -> 3 x 512 (GELU) -> 4
    candidates, float32 operands on the MFMA path, against the float64
    oracle"""
    shp = (1089, 512, 3, 4)
    rows = 192
    rng = np.random.default_rng(15)
    m = P.MLP(*shp, max_batch=256, bf16=False, seed=3)
    p = m.params().astype(np.float64)
    x = rng.standard_normal((rows, shp[0])).astype(np.float32)
    t = rng.standard_normal((rows, 3))
    t = (t / np.linalg.norm(t, axis=1, keepdims=True)).astype(np.float32)
    dx, dt, dl = S.DeviceBuffer.from_numpy(x), S.DeviceBuffer.from_numpy(t), S.DeviceBuffer(rows * 4)
    m.loss_grad(dx.ptr, dt.ptr, rows, 0.05, dl.ptr)
    S.synchronize()
    lo, go, lrow = OM.loss_grad(p, x.astype(np.float64), t.astype(np.float64), *shp, 0.05)
    assert rel(dl.to_numpy(np.float32, (rows,)), lrow) < 1e-5
    g = m.grads()
    # per layer, so that a wrong block cannot hide behind a large one
    off = 0
    for din, dout in OM.shapes(*shp):
        for n in (din * dout, dout):
            e = rel(g[off:off + n], go[off:off + n])
            assert e < 1e-4, (din, dout, n, e)
            off += n
    assert off == g.size


@pytest.mark.gpu
def test_gpu_mlp_training_reduces_loss(P):
    # targets a fixed smooth function of the inputs: Adam must fit it
    shp = (32, 64, 3, 2)
    rng = np.random.default_rng(6)
    rows = 4096
    x = rng.standard_normal((rows, shp[0])).astype(np.float32)
    A = rng.standard_normal((shp[0], 3))
    t = np.tanh(x @ A)
    t = (t / np.linalg.norm(t, axis=1, keepdims=True)).astype(np.float32)
    m = P.MLP(*shp, max_batch=512, bf16=True, seed=2)
    dx, dt = S.DeviceBuffer.from_numpy(x), S.DeviceBuffer.from_numpy(t)
    l0 = P.mean_loss(m, dx, dt, rows)
    hist = P.train(m, dx, dt, rows, epochs=6, batch=512, lr=3e-3)
    l1 = P.mean_loss(m, dx, dt, rows)
    print("loss before %.4f after %.4f; per epoch %s" % (l0, l1, hist), flush=True)
    assert l1 < 0.7 * l0


@pytest.mark.gpu
def test_gpu_features_match_oracle(P):
    rng = np.random.default_rng(8)
    L, Sg = 32, 5
    a = rng.standard_normal((Sg, (L + 1) ** 2)) + 1j * rng.standard_normal((Sg, (L + 1) ** 2))
    for cdt, f32 in ((np.complex128, 0), (np.complex64, 1)):
        d_a = S.DeviceBuffer.from_numpy(a.astype(cdt))
        d_f = S.DeviceBuffer(Sg * (L + 1) ** 2 * 4)
        assert S.lib().sphr_features(d_a.ptr, f32, Sg, L, d_f.ptr, None) == 0
        S.synchronize()
        assert rel(d_f.to_numpy(np.float32, (Sg, (L + 1) ** 2)), OM.features(a.astype(cdt), L)) < 1e-6


@pytest.mark.gpu
def test_gpu_inject_bursts(P):
    Sg, n = 64, 1024
    par, _ = P.burst_params(Sg, seed=3)
    d_par = S.DeviceBuffer.from_numpy(par)
    d_out = S.DeviceBuffer(3 * Sg * n * 4)
    # burst alone against its closed form
    assert S.lib().sphr_inject_bursts(d_par.ptr, 3, Sg, n, 1.0 / 4096, 0.0, 1, d_out.ptr, None) == 0
    S.synchronize()
    h = d_out.to_numpy(np.float32, (3, Sg, n)).astype(np.float64)
    t = (np.arange(n) - 0.5 * n) / 4096
    for d in range(3):
        u = t[None, :] + par[:, 3 + d].astype(np.float64)[:, None]
        want = par[:, 0:1] * np.sin(2 * math.pi * par[:, 1:2] * u) * np.exp(-(u / par[:, 2:3]) ** 2)
        assert np.max(np.abs(h[d] - want)) < 2e-5 * np.max(np.abs(want))
    # noise alone: unit normal moments, different seeds differ
    par0 = par.copy()
    par0[:, 0] = 0.0
    d_par0 = S.DeviceBuffer.from_numpy(par0)
    assert S.lib().sphr_inject_bursts(d_par0.ptr, 3, Sg, n, 1.0 / 4096, 1.0, 42, d_out.ptr, None) == 0
    S.synchronize()
    z = d_out.to_numpy(np.float32, (3 * Sg * n,)).astype(np.float64)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.02
    assert abs(np.mean(z ** 4) - 3.0) < 0.1


@pytest.mark.gpu
def test_gpu_proposal_small_ab(P):
    # a short training run on method-generated skies, then the A/B against
    # the plain peak on held-out skies
    L = 32
    ncoef = (L + 1) ** 2
    fac = P.SkyFactory(chunk=4096)
    n_train = 16384
    d_x, d_t = S.DeviceBuffer(n_train * ncoef * 4), S.DeviceBuffer(n_train * 12)
    for off in range(0, n_train, 4096):
        fac.make(4096, 100 + off, d_x, d_t, feat_offset=off)
    m = P.MLP(max_batch=1024, bf16=True, seed=1)
    # the last 2048 skies are held out of training for the loss check
    n_fit = n_train - 2048
    d_xv = S.DeviceBuffer.from_numpy(d_x.to_numpy(np.float32, (n_train, ncoef))[n_fit:])
    d_tv = S.DeviceBuffer.from_numpy(d_t.to_numpy(np.float32, (n_train, 3))[n_fit:])
    l0 = P.mean_loss(m, d_xv, d_tv, 2048)
    hist = P.train(m, d_x, d_t, n_fit, epochs=4, batch=512, lr=1e-3)
    l1 = P.mean_loss(m, d_xv, d_tv, 2048)
    res = P.evaluate(m, fac, 2048, seed=999)
    summ = P.summary(res)
    print("held-out loss before %.4f after %.4f; A/B %s" % (l0, l1, summ), flush=True)
    assert l1 < l0
    # what the network gains: its refined proposals land on the plain peak
    # more often than the reference's l = 6 seeding does
    assert summ["proposal"]["within_0.10_deg"] > summ["reference_seed"]["within_0.10_deg"]
    # the fallback invariant, separately: the final peak is the plain one
    # or a maximum at least as high (a mirror image)
    assert summ["final"]["within_0.10_deg_or_not_lower"] >= 0.99


@pytest.mark.gpu
def test_gpu_lowpass_seed_matches_device_seed(P):
    """the host's l = 6 low-pass seed of the uneven-workload evaluation,
    refined, gives the peaks of the batch's own l = 6 seeding (the
    reference's find_near_minimum, sh_series_extremize.c:168-241)"""
    L, Sg = 32, 64
    fac = P.SkyFactory(chunk=Sg)
    d_f = S.DeviceBuffer(Sg * (L + 1) ** 2 * 4)
    plain = fac.make(Sg, 77, d_f, keep=True)
    d_pk = [S.DeviceBuffer(Sg * 8) for _ in range(3)]
    fac.batch.peaks(fac.d_alm.ptr, Sg, d_pk[0].ptr, d_pk[1].ptr, d_pk[2].ptr, d_map=None)
    seeds = P.LowPassSeeds(L)(plain["alm"])
    d_s = S.DeviceBuffer.from_numpy(np.ascontiguousarray(seeds))
    d_r = [S.DeviceBuffer(Sg * 8) for _ in range(3)]
    assert S.lib().sphr_refine_peaks(fac.d_alm.ptr, 1, Sg, 1, L, d_s.ptr, 1, d_r[0].ptr, d_r[1].ptr, d_r[2].ptr, None) == 0
    S.synchronize()
    dev = [d.to_numpy(np.float64, (Sg,)) for d in d_pk]
    host = [d.to_numpy(np.float64, (Sg,)) for d in d_r]
    sep = P.angle(dev[0], dev[1], host[0], host[1])
    assert np.mean(sep < 1e-6) >= 0.95, np.sort(sep)[-5:]


@pytest.mark.gpu
def test_gpu_jobs_sky_factory(P):
    """skies of the uneven workload at l_max 32: every job its own network;
    the plain peak is a local maximum at or above the grid's maximum"""
    L, Sg = 32, 256
    fac = P.JobsSkyFactory(chunk=Sg)
    d_f = S.DeviceBuffer(Sg * (L + 1) ** 2 * 4)
    d_t = S.DeviceBuffer(Sg * 12)
    out = fac.make(Sg, 5, d_f, d_t, keep=True)
    assert set(np.unique(out["detectors"])) <= {2, 3, 4, 5}
    assert np.all(np.isfinite(out["value"])) and np.all(np.isfinite(out["alm"]))
    assert np.all(out["value"] >= out["grid_max"] - 1e-5 * np.abs(out["grid_max"]))
    t = d_t.to_numpy(np.float32, (Sg, 3))
    assert np.max(np.abs(np.linalg.norm(t, axis=1) - 1.0)) < 1e-5
