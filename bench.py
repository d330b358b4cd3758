"""Throughput of the batched cross-power path on one MI355X.

Workload (the metric's configuration): three sites (46.45N 119.41W),
(30.56N 90.77W), (43.63N 10.50E) on a 6371 km sphere; 1024 segments of
1024 samples (0.25 s at 4096 Hz) of white noise plus one sine-Gaussian
burst per segment; sky l_max 32.  One step = one pass of the batched
pipeline over the 1024 segments (FFT, cross spectra x delay-product
projection, rotation and baseline average, mesh synthesis), inputs
resident in HBM.  value = sky maps/s of the mixed-precision (float32)
path; the float64 path's rate and the a_lm / map error of the mixed path
against it are reported beside it, with the CPU oracle's rate on a bounded
sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def stage_bytes(info, S, esz, in_bytes_per_sample=8):
    """Algorithmic HBM bytes per launch of each stage: FFT, projection,
    a_lm + Legendre ring sums, phi DFT (GEMM); with the fused sky kernel
    the third stage is partials + rotation rows in, a_lm + maps out, and
    the fourth is empty.  On the lag-domain route the first stage is the
    one kernel that reads the series and the lag table and writes one row
    of partial sums per segment, and the second is empty."""
    D, nb, N, L, CT = info["detectors"], info["baselines"], info["N"], info["sky_l"], info["columns"]
    ns = 1 if info.get("lag") else info["k_splits"]
    Nh = N // 2 + 1
    ncoef = (L + 1) ** 2
    n = 2 * (L + 1)
    nm = 2 * L + 1
    spec = D * S * Nh * 2 * esz
    fft = D * S * N * in_bytes_per_sample + spec
    xproj = spec + info["table_bytes"] + ns * S * CT * esz
    if info.get("lag"):
        fft, xproj = D * S * N * in_bytes_per_sample + info["table_bytes"] + S * CT * esz, 0
    ring = S * n * nm * 2 * esz
    lf = ns * S * CT * esz + nb * ncoef * 2 * esz + S * ncoef * 2 * esz + ring
    phi = ring + 2 * nm * 2 * n * esz + S * n * n * 2 * esz
    if info.get("fused"):
        return [fft, xproj, ns * S * CT * esz + nb * ncoef * 2 * esz + S * ncoef * 2 * esz + S * n * n * 2 * esz, 0]
    return [fft, xproj, lf, phi]


def run_path(S, net, L, d_ts, nseg, precision, steps, warmup, d_window=None):
    cdt = np.complex128 if precision == S.FP64 else np.complex64
    esz = np.dtype(cdt).itemsize // 2
    b = S.Batch(net, L, precision, max_segments=nseg)
    d_alm = S.DeviceBuffer(nseg * (L + 1) ** 2 * 2 * esz)
    d_map = S.DeviceBuffer(nseg * (2 * L + 2) ** 2 * 2 * esz)
    st = S.lib().sphr_stream()
    dw = d_window.ptr if d_window is not None else None
    tuned, tuned_ms = b.tune(d_ts.ptr, nseg, d_alm.ptr, d_map.ptr, d_window=dw)
    for _ in range(warmup):
        b.run(d_ts.ptr, nseg, d_alm.ptr, d_map.ptr, d_window=dw, stream=st)
    S.lib().sphr_stream_synchronize(st)
    e0, e1 = S.lib().sphr_event_new(), S.lib().sphr_event_new()
    S.lib().sphr_event_record(e0, st)
    for _ in range(steps):
        b.run(d_ts.ptr, nseg, d_alm.ptr, d_map.ptr, d_window=dw, stream=st)
    S.lib().sphr_event_record(e1, st)
    S.lib().sphr_event_synchronize(e1)
    ms = S.lib().sphr_event_elapsed_ms(e0, e1)
    # per-stage times on the same stream (separate pass, events around each stage)
    b.profile(True)
    for _ in range(steps):
        b.run(d_ts.ptr, nseg, d_alm.ptr, d_map.ptr, d_window=dw, stream=st)
    stage_ms, runs = b.profile_read()
    b.profile(False)
    alm = d_alm.to_numpy(cdt, (nseg, (L + 1) ** 2))
    mp = d_map.to_numpy(cdt, (nseg, 2 * L + 2, 2 * L + 2))
    # each sky's peak: map seed + extremize on the series (sphr_batch_peaks)
    d_pk = [S.DeviceBuffer(nseg * 8) for _ in range(3)]
    b.peaks(d_alm.ptr, nseg, d_pk[0].ptr, d_pk[1].ptr, d_pk[2].ptr, d_map=d_map.ptr)
    S.synchronize()
    peaks = [d.to_numpy(np.float64, (nseg,)) for d in d_pk]
    info = b.info()
    S.lib().sphr_event_free(e0)
    S.lib().sphr_event_free(e1)
    return dict(ms_per_step=ms / steps, stage_ms=[t / runs for t in stage_ms], alm=alm, map=mp, info=info, esz=esz, tuned=tuned,
                peaks=peaks)


def plain_fp64(S, net, ts, L, budget_s, max_segments, window=None):
    """The FP64 drop-in path as a reference caller drives it, one segment at
    a time through the C API: correlator_tseries_to_fseries per detector
    (plans made once), correlator_network_integrate_power_fd into a sky
    series, sh_series_to_mesh.  Each call is synchronous (host in, host
    out).  Returns maps/s over a bounded sample and the last segment's
    a_lm for a parity check against the batched path."""
    import ctypes
    lib = S.lib()
    D, nseg, N = ts.shape
    tsb = [np.zeros(N) for _ in range(D)]
    fsb = [np.zeros(N, dtype=np.complex128) for _ in range(D)]
    plans = [lib.correlator_tseries_to_fseries_plan(tsb[d].ctypes.data, fsb[d].ctypes.data, N) for d in range(D)]
    fptrs = (ctypes.c_void_p * D)(*[fsb[d].ctypes.data for d in range(D)])
    sky = lib.sh_series_new_zero(L, 0)
    nt, nph = ctypes.c_int(), ctypes.c_int()
    libc_free = ctypes.CDLL(None).free
    libc_free.argtypes = [ctypes.c_void_p]

    def one(s):
        for d in range(D):
            tsb[d][:] = ts[d, s] if window is None else ts[d, s] * window
            lib.correlator_tseries_to_fseries(tsb[d].ctypes.data, fsb[d].ctypes.data, N, plans[d])
        if not lib.correlator_network_integrate_power_fd(sky, fptrs, net.plan):
            raise RuntimeError("integrate failed: " + S.last_error())
        m = lib.sh_series_to_mesh(sky, ctypes.byref(nt), ctypes.byref(nph), None)
        if not m:
            raise RuntimeError("to_mesh failed: " + S.last_error())
        libc_free(m)

    one(0)
    done = 0
    t0 = time.time()
    while done < min(nseg, max_segments) and time.time() - t0 < budget_s:
        one(done)
        done += 1
    el = time.time() - t0
    alm = S.coeff_view(sky).copy()
    lib.sh_series_free(sky)
    for p in plans:
        lib.correlator_fft_plan_free(p)
    return dict(value=done / el, segments=done, last=done - 1, alm=alm)


def host_cpu():
    """The host's usable cores and CPU model, for the CPU baseline."""
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def _timed_segments(net, ts, L, budget_s, seg_threads, mesh_threads, chunk):
    """Runs chunks of segments through the oracle's pipeline until the
    budget is spent; returns (segments, seconds)."""
    done, el = 0, 0.0
    S = ts.shape[1]
    while done < S and el < budget_s:
        k = min(chunk, S - done)
        t0 = time.time()
        net.segments_run(ts[:, done:done + k], L, seg_threads=seg_threads, mesh_threads=mesh_threads)
        el += time.time() - t0
        done += k
    return done, el


def cpu_baseline(ts, L, budget_s, positions=None, delta_t=1.0 / 4096, label="three-site"):
    """The FP64 C oracle on this host in the two modes of the reference's
    own threading: reference-faithful (integrate on one thread, the ring
    loop of to_mesh on 4, sh_series_trans.c:829) and all host cores
    (segments split over threads with per-thread plan scratch, as
    correlator_network_plan_fd_copy callers do, correlator.c:908-935).
    value = the all-cores rate."""
    from oracle import oracle as O
    from sphradiometer_amd import inject
    pos = inject.positions("ABC") if positions is None else positions
    ncpu, model = host_cpu()
    threads = max(1, min(16, ncpu))
    t0 = time.time()
    net = O.Network(pos, ts.shape[2], delta_t, L)
    plan_s = time.time() - t0
    net.segments_run(ts[:, :2], L, seg_threads=1, mesh_threads=4)
    n1, e1 = _timed_segments(net, ts, L, budget_s / 2, 1, 4, 16)
    nall, eall = _timed_segments(net, ts, L, budget_s / 2, threads, 1, 16 * threads)
    return dict(value=nall / eall, unit="sky maps/s", cores=threads, kind="port",
                sample="%d of the %d %s segments (a_lm + map each) over %d threads, FP64 C oracle (gcc -O3, AVX2, OpenMP); "
                       "plan build %.2f s not included" % (nall, ts.shape[1], label, threads, plan_s),
                reference_faithful={"value": round(n1 / e1, 2), "cores": 1, "mesh_threads": 4,
                                    "sample": "%d segments in order, to_mesh's ring loop on 4 threads" % n1},
                host={"nproc": ncpu, "cpu_model": model})


FP64_MFMA_TFLOPS = 78.6
FP32_MFMA_TFLOPS = 157.3
FP16_MFMA_TFLOPS = 2500.0


def bench_healpix(S, args):
    """The HEALPix output of the localization workload (RapidLocalization.py:
    498-530): per event the beta = +1 and -1 skies at l_max 50 as RING
    maps at nside 32 (healpy_alm_to_map's rule, healpix.py:163) and their
    posterior.  One step = the maps of 2 x 8192 skies + the posterior of
    8192 events; value = events/s (float32 maps); HBM roofline of the map
    kernel (a_lm read + maps written)."""
    L, E = 50, args.events
    nside = 1 << max(0, math.ceil(math.log2(L / 2.0)))
    ncoef = (L + 1) ** 2
    rng = np.random.default_rng(0)
    lsc = np.array([1.0 / (1.0 + l) for l, m in _lm_list(L)])
    alm = (rng.standard_normal((2, E, ncoef)) + 1j * rng.standard_normal((2, E, ncoef))) * lsc * 8.0
    res = {}
    for name, prec, cdt, rdt in (("fp64", S.FP64, np.complex128, np.float64), ("mixed", S.MIXED, np.complex64, np.float32)):
        hp = S.Healpix(L, nside, prec)
        d_alm = S.DeviceBuffer.from_numpy(alm.astype(cdt))
        d_map = S.DeviceBuffer(2 * E * hp.npix * np.dtype(rdt).itemsize)
        d_prob = S.DeviceBuffer(E * hp.npix * np.dtype(rdt).itemsize)
        st = S.lib().sphr_stream()
        for _ in range(args.warmup):
            hp.maps(d_alm.ptr, 2 * E, d_map.ptr, stream=st)
            hp.posterior(d_map.ptr, E, d_prob.ptr, stream=st)
        S.lib().sphr_stream_synchronize(st)
        ev = [S.lib().sphr_event_new() for _ in range(3)]
        S.lib().sphr_event_record(ev[0], st)
        for _ in range(args.steps):
            hp.maps(d_alm.ptr, 2 * E, d_map.ptr, stream=st)
            hp.posterior(d_map.ptr, E, d_prob.ptr, stream=st)
        S.lib().sphr_event_record(ev[1], st)
        for _ in range(args.steps):
            hp.maps(d_alm.ptr, 2 * E, d_map.ptr, stream=st)
        S.lib().sphr_event_record(ev[2], st)
        S.lib().sphr_event_synchronize(ev[2])
        ms = S.lib().sphr_event_elapsed_ms(ev[0], ev[1]) / args.steps
        ms_maps = S.lib().sphr_event_elapsed_ms(ev[1], ev[2]) / args.steps
        res[name] = dict(ms=ms, ms_maps=ms_maps, prob=d_prob.to_numpy(rdt, (E, hp.npix)), npix=hp.npix,
                         esz=np.dtype(cdt).itemsize, rsz=np.dtype(rdt).itemsize)
        print("%s: %.3f ms per step (maps %.3f ms), %d events, nside %d" % (name, ms, ms_maps, E, nside), file=sys.stderr, flush=True)
    perr = np.linalg.norm(res["mixed"]["prob"] - res["fp64"]["prob"], axis=1) / np.linalg.norm(res["fp64"]["prob"], axis=1)
    mx = res["mixed"]
    map_bytes = 2 * E * ncoef * mx["esz"] + 2 * E * mx["npix"] * mx["rsz"]
    ach = map_bytes / (mx["ms_maps"] * 1e-3) / 1e9
    out = {
        "metric": "HEALPix posterior events/s, l_max=50 beta = +1 / -1 skies -> nside 32 RING maps + posterior",
        "value": round(E / (mx["ms"] * 1e-3), 1), "unit": "events/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(mx["ms"], 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: complex Gaussian a_lm scaled by 8 / (l + 1) (seed 0)",
        "config": {"workload": "%d events x 2 skies, l_max %d, nside %d (%d pixels)" % (E, L, nside, mx["npix"]),
                   "events": E, "l_max": L, "nside": nside},
        "fp64": {"value": round(E / (res["fp64"]["ms"] * 1e-3), 1), "ms_per_step": round(res["fp64"]["ms"], 4),
                 "maps_ms": round(res["fp64"]["ms_maps"], 4)},
        "mixed_vs_fp64": {"prob_rel_err_max": float(perr.max())},
        "roofline": {"bound": "hbm", "kernel": "k_healpix", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": map_bytes,
                     "kernel_ms": round(mx["ms_maps"], 4)},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        # the FP64 oracle on a bounded sample: the two skies of a few events
        t0, n = time.time(), 0
        while n < 64 and (n < 2 or time.time() - t0 < args.cpu_budget):
            a = alm[0, n].astype(np.complex128)
            mp = O.healpix_map(a, L, nside)
            mn = O.healpix_map(alm[1, n].astype(np.complex128), L, nside)
            O.healpix_posterior(mp, mn)
            n += 1
        dt = time.time() - t0
        out["cpu_baseline"] = {"value": round(n / dt, 3), "unit": "events/s", "cores": 1, "kind": "port",
                               "sample": "%d events (2 maps + posterior each), FP64 C oracle, 1 thread" % n}
    print(json.dumps(out), flush=True)


def bench_td(S, args):
    """The time-domain correlator (correlator.c:611-673, 967-984) over a
    long stream: three sites, 4096 Hz, a 16 s stream integrated with a
    square window in one network call (three baselines).  value = stream
    samples/s.  With a square window the correlation runs only where the
    window changes, so the roofline is the call's algorithmic bytes (series
    and windows in, sky out) against HBM.  The same call with a random
    window (every sample correlated, 2 n_p^2 float64 flops per output
    sample and baseline on MFMA) is reported beside it."""
    import ctypes
    from sphradiometer_amd import inject
    L = S.lib()
    dt, n = 1.0 / 4096, 16 * 4096
    pos = inject.positions("ABC")
    inst = L.instrument_array_new(0)
    for x, y, z in pos:
        L.instrument_array_append(inst, L.instrument_new(float(x), float(y), float(z), None))
    bl = L.correlator_network_baselines_new(inst)
    t0 = time.time()
    plan = L.correlator_network_plan_td_new(bl, dt)
    if not plan:
        raise RuntimeError(S.last_error())
    plan_s = time.time() - t0
    nb = bl.contents.n_baselines
    pl = [plan.contents.plans[k].contents for k in range(nb)]
    trs = [p.transient for p in pl]
    nps = [2 * t + 1 for t in trs]
    lmax = max(p.power_1d.contents.l_max for p in pl)
    print("td plans built in %.1f s: taps %s, power l_max %s" % (plan_s, nps, [p.power_1d.contents.l_max for p in pl]), file=sys.stderr, flush=True)
    rng = np.random.default_rng(0)
    ts = rng.standard_normal((3, n))
    wins = [np.full(n - 2 * t, 1.0 / (n - 2 * t)) for t in trs]
    tp = (ctypes.c_void_p * 3)(*[ts[i].ctypes.data for i in range(3)])
    wp = (ctypes.c_void_p * nb)(*[wins[k].ctypes.data for k in range(nb)])
    sky = L.sh_series_new(lmax, 0)
    for _ in range(args.warmup):
        if not L.correlator_network_integrate_power_td(sky, tp, n, wp, plan):
            raise RuntimeError(S.last_error())
    t0 = time.time()
    for _ in range(args.steps):
        L.correlator_network_integrate_power_td(sky, tp, n, wp, plan)
    sec = (time.time() - t0) / args.steps
    rwins = [rng.uniform(0.5, 1.5, n - 2 * t) / (n - 2 * t) for t in trs]
    rwp = (ctypes.c_void_p * nb)(*[rwins[k].ctypes.data for k in range(nb)])
    if not L.correlator_network_integrate_power_td(sky, tp, n, rwp, plan):
        raise RuntimeError(S.last_error())
    t0 = time.time()
    for _ in range(args.steps):
        L.correlator_network_integrate_power_td(sky, tp, n, rwp, plan)
    rsec = (time.time() - t0) / args.steps
    rflops = sum(2.0 * npk * npk * (n - 2 * t) for npk, t in zip(nps, trs))
    # algorithmic bytes of the call: the three series and the three windows
    # in, the sky out (float64); the square window makes the correlation
    # itself small (it runs only where the window changes), so the call is
    # bound by these transfers and their latency
    io_bytes = 8.0 * (3 * n + sum(n - 2 * t for t in trs)) + 16.0 * (lmax + 1) ** 2
    ach = io_bytes / sec / 1e9
    out = {
        "metric": "TD network cross-power, 3 sites at 4096 Hz: stream samples/s (16 s stream, one call)",
        "value": round(n / sec, 1), "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: white N(0,1) float64 (seed 0), square window",
        "config": {"workload": "three sites, %d samples at 4096 Hz, power l_max %d, sky l_max %d" % (n, lmax, lmax),
                   "samples": n, "taps": nps, "sky_l": lmax},
        "roofline": {"bound": "hbm", "kernel": "whole call (series and windows in, sky out)", "achieved": round(ach, 3),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None,
                     "algorithmic_bytes": int(io_bytes)},
        "random_window": {"value": round(n / rsec, 1), "ms_per_step": round(rsec * 1e3, 3),
                          "mfma_tflops_whole_call": round(rflops / rsec / 1e12, 3),
                          "frac": round(rflops / rsec / 1e12 / FP64_MFMA_TFLOPS, 4)},
        "plan_build_s": round(plan_s, 1),
    }
    if not args.no_cpu:
        from oracle import oracle as O
        # the sample-by-sample FP64 oracle on every baseline, a bounded
        # number of output samples each (after one call that builds the
        # product tables); network samples/s = 1 / sum of per-sample costs
        cost, m = 0.0, 256
        k = 0
        for i in range(1, 3):
            for j in range(i):
                td = O.PlanTD(pos[i], pos[j], dt)
                a = rng.standard_normal(2 * td.transient + m)
                td.integrate(a[: 2 * td.transient + 1], a[: 2 * td.transient + 1], np.ones(1))
                t0 = time.time()
                td.integrate(a, a, np.full(m, 1.0 / m))
                cost += (time.time() - t0) / m
                k += 1
        out["cpu_baseline"] = {"value": round(1.0 / cost, 3), "unit": "samples/s", "cores": 1, "kind": "port",
                               "sample": "%d output samples of each of the %d baselines (sh_series_array_dot x 2 + product per sample), FP64 C oracle, 1 thread" % (m, k)}
    L.sh_series_free(sky)
    L.correlator_network_plan_td_free(plan)
    print(json.dumps(out), flush=True)


def bench_localization(S, args):
    """The localization workload: three sites, complex SNR series of 410
    samples at 4096 Hz (+-50 ms, padded to 584), beta = +1 / -1 plans
    multiplied by the projection at their natural order then set_l(50),
    8192 events per step: spectra (DFT GEMM), cross spectra, the sky GEMM
    (K = 3 x 584, 2 x 2601 complex columns), prior, rotate_z, meshes.
    One JSON line, events/s; MFMA roofline of the sky GEMM."""
    from sphradiometer_amd import inject
    L, ln, dt, E = 50, 410, 1.0 / 4096, args.events
    N = S.precalculated_length(ln, dt)
    t0 = time.time()
    nets = []
    for beta in (1.0, -1.0):
        net = S.Network(inject.positions("ABC"), N, dt, responses=inject.responses("ABC"))
        nat = [int(net.baseline_plan(k).contents.delay_product.contents.l_max) for k in range(3)]
        net.mult_by_projection(beta, 0.0, np.ones((N, 3)))
        net.set_l(L)
        nets.append(net)
    lp_nat = S.log_uniformsky_prior(S.lib().correlator_network_l_max(nets[0]._bl, dt))
    lnat = int(round(math.sqrt(lp_nat.size))) - 1
    lp = np.array([lp_nat[_lmoff(lnat, l, m)] for l, m in _lm_list(L)])
    plan_s = time.time() - t0
    print("localization plans built in %.1f s (logprior natural l %d)" % (plan_s, lnat), file=sys.stderr, flush=True)
    x, meta = inject.snr_events(E, length=ln, delta_t=dt, seed=0)
    d_x = S.DeviceBuffer.from_numpy(x)
    d_g = S.DeviceBuffer.from_numpy(meta["gmst"])
    res = {}
    for name, prec, cdt in (("fp64", S.FP64, np.complex128), ("mixed", S.MIXED, np.complex64)):
        loc = S.Localizer(nets[0], nets[1], lp, L, ln, dt, precision=prec, max_events=E)
        esz = np.dtype(cdt).itemsize
        d_alm = S.DeviceBuffer(2 * E * (L + 1) ** 2 * esz)
        n = 2 * (L + 1)
        d_map = S.DeviceBuffer(2 * E * n * n * esz)
        st = S.lib().sphr_stream()
        for _ in range(args.warmup):
            loc.run(d_x.ptr, E, d_alm.ptr, d_map=d_map.ptr, d_gmst=d_g.ptr, stream=st)
        S.lib().sphr_stream_synchronize(st)
        e0, e1 = S.lib().sphr_event_new(), S.lib().sphr_event_new()
        S.lib().sphr_event_record(e0, st)
        for _ in range(args.steps):
            loc.run(d_x.ptr, E, d_alm.ptr, d_map=d_map.ptr, d_gmst=d_g.ptr, stream=st)
        S.lib().sphr_event_record(e1, st)
        S.lib().sphr_event_synchronize(e1)
        ms = S.lib().sphr_event_elapsed_ms(e0, e1) / args.steps
        res[name] = dict(ms=ms, alm=d_alm.to_numpy(cdt, (2, E, (L + 1) ** 2)))
        print("%s: %.3f ms per %d events" % (name, ms, E), file=sys.stderr, flush=True)
    err = np.linalg.norm(res["mixed"]["alm"] - res["fp64"]["alm"], axis=2) / np.linalg.norm(res["fp64"]["alm"], axis=2)
    # the drop-in path on a few events: generate_alm_skys_arrays + scale_power_l
    plain_err = []
    for e in range(3):
        pre = S.preprocess_snr(x[:, e].astype(np.complex128), dt)
        sp, sn = S.generate_alm_skys(nets[0], nets[1], pre, float(meta["gmst"][e]), lp, L)
        sl = np.array([2.7 * (l / L) ** 0.6 for l, m in _lm_list(L)])   # storage order
        for bi, s in enumerate((sp, sn)):
            want = s * sl
            plain_err.append(float(np.linalg.norm(res["fp64"]["alm"][bi, e] - want) / np.linalg.norm(want)))
    nb, ncoef = 3, (L + 1) ** 2
    flops = 2.0 * E * (2 * nb * N) * (4 * ncoef) + 2.0 * (3 * E) * (2 * ln) * (2 * N)
    value = E / (res["mixed"]["ms"] * 1e-3)
    # the mixed path runs both products as three fp16 MFMA products each
    # (split-FP16, float32 accumulation): its matrix work is 3 x flops
    ach = 3.0 * flops / (res["mixed"]["ms"] * 1e-3) / 1e12
    ach64 = flops / (res["fp64"]["ms"] * 1e-3) / 1e12
    out = {
        "metric": "localization sky pairs/s, 3-det l_max=50 (beta = +1 and -1 skies per event)",
        "value": round(value, 1), "unit": "events/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(res["mixed"]["ms"], 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f16x3 products, f32 accumulation",
        "data": "synthetic: unit complex Gaussian SNR series + chirp-like peak, random antenna factors (seed 0)",
        "config": {"workload": "three sites, %d events x 410 complex samples at 4096 Hz padded to %d, l_max %d, a_lm + 102x102 map per sky" % (E, N, L),
                   "events": E, "length": ln, "N": N, "l_max": L, "natural_l": nat},
        "fp64": {"value": round(E / (res["fp64"]["ms"] * 1e-3), 1), "ms_per_step": round(res["fp64"]["ms"], 4),
                 "mfma_tflops": round(ach64, 2), "frac": round(ach64 / FP64_MFMA_TFLOPS, 4),
                 "alm_rel_err_vs_dropin_max": max(plain_err)},
        "mixed_vs_fp64": {"alm_rel_err_max": float(err.max()), "alm_rel_err_median": float(np.median(err))},
        "roofline": {"bound": "mfma", "kernel": "whole step (split-FP16 products dominate)", "achieved": round(ach, 2),
                     "peak": FP16_MFMA_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / FP16_MFMA_TFLOPS, 4), "traffic": None,
                     "basis": "3 x (sky + DFT product flops) per step over the step time"},
        "plan_build_s": round(plan_s, 1),
    }
    print(json.dumps(out))


def bench_jobs(S, args):
    """The uneven workload: 4096 independent jobs, each with its own
    network (2-5 sites uniform on the sphere), length (0.125-2 s at 4096
    Hz) and order (l_max 16-64), float32 data, a burst in every other job.
    Every job needs its own plan, so one step = plan generation on the GPU
    (sphr_jobs_new) + one captured-graph run; value = jobs/s over both."""
    from sphradiometer_amd import inject
    nj = args.jobs
    jobs, packed = inject.uneven_jobs(nj, seed=0)
    d_in = S.DeviceBuffer.from_numpy(packed)
    res = {}
    for name, prec, esz in (("fp64", S.FP64, 16), ("mixed", S.MIXED, 8)):
        S.synchronize()
        t0 = time.time()
        J = S.Jobs(jobs, 1.0 / 4096, precision=prec)
        S.synchronize()
        plan_s = time.time() - t0
        info = J.info()
        d_alm = S.DeviceBuffer(info["alm_elems"] * esz)
        d_map = S.DeviceBuffer(info["map_elems"] * esz)
        k, tuned_ms = J.tune(d_in.ptr, d_alm.ptr)
        st = S.lib().sphr_stream()
        for _ in range(args.warmup):
            J.run(d_in.ptr, d_alm.ptr, d_map.ptr, stream=st)
        S.lib().sphr_stream_synchronize(st)
        e0, e1 = S.lib().sphr_event_new(), S.lib().sphr_event_new()
        S.lib().sphr_event_record(e0, st)
        for _ in range(args.steps):
            J.run(d_in.ptr, d_alm.ptr, d_map.ptr, stream=st)
        S.lib().sphr_event_record(e1, st)
        S.lib().sphr_event_synchronize(e1)
        run_ms = S.lib().sphr_event_elapsed_ms(e0, e1) / args.steps
        res[name] = dict(plan_s=plan_s, run_ms=run_ms, info=info, tuned=k,
                         alm=d_alm.to_numpy(np.complex128 if prec == S.FP64 else np.complex64, (info["alm_elems"],)))
        print("%s: plan %.3f s, run %.3f ms, %s" % (name, plan_s, run_ms, info), file=sys.stderr, flush=True)
        del J
    err = float(np.linalg.norm(res["mixed"]["alm"] - res["fp64"]["alm"]) / np.linalg.norm(res["fp64"]["alm"]))
    total_s = res["mixed"]["plan_s"] + res["mixed"]["run_ms"] * 1e-3
    value = nj / total_s
    # run-phase bytes: the input samples and the lag plans, read once
    run_bytes = packed.nbytes + res["mixed"]["info"]["plan_bytes"]
    ach = run_bytes / (res["mixed"]["run_ms"] * 1e-3) / 1e9
    out = {
        "metric": "uneven sky-map jobs/s (own plan per job: GPU plan generation + run)",
        "value": round(value, 1), "unit": "jobs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total_s * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: float32 white noise, burst in every other job (seed 0)",
        "config": {"workload": "%d jobs: 2-5 sites uniform on the sphere, 0.125-2 s at 4096 Hz, l_max 16-64, a_lm + map each" % nj,
                   "jobs": nj, "baselines": res["mixed"]["info"]["baselines"], "l_classes": res["mixed"]["info"]["l_classes"]},
        "mixed": {"plan_s": round(res["mixed"]["plan_s"], 3), "run_ms": round(res["mixed"]["run_ms"], 4),
                  "run_jobs_per_s": round(nj / (res["mixed"]["run_ms"] * 1e-3), 1), "tuned_tile": res["mixed"]["tuned"]},
        "fp64": {"plan_s": round(res["fp64"]["plan_s"], 3), "run_ms": round(res["fp64"]["run_ms"], 4),
                 "jobs_per_s": round(nj / (res["fp64"]["plan_s"] + res["fp64"]["run_ms"] * 1e-3), 1)},
        "mixed_vs_fp64": {"alm_rel_err": err},
        "roofline": {"bound": "hbm", "kernel": "run (lag correlation + plan read + synthesis)", "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None},
    }
    print(json.dumps(out))


def _lm_list(L):
    """(l, m) in the coefficient storage order."""
    out = [(l, 0) for l in range(L + 1)]
    for m in range(1, L + 1):
        out += [(l, m) for l in range(m, L + 1)]
    for m in range(-L, 0):
        out += [(l, m) for l in range(-m, L + 1)]
    return out


def _lmoff(L, l, m):
    L1 = L + 1
    x = m * (m - 1)
    base = L1 * m - x // 2 if m >= 0 else L1 * (L1 + m) + x // 2
    return base + l - abs(m)


def bench_two_site(S, args):
    """The two-site workload: sites (46.45N 119.41W) and (30.56N 90.77W) on
    a 6371 km sphere, white N(0,1) float64 at 4096 Hz from seed 0, 16
    non-overlapping 1 s segments (4096 samples, symmetric Hann window),
    sky l_max 25 (natural order 129 truncated by set_l).  One step = the
    batched pipeline over the 16 segments (a_lm + 52 x 52 map each);
    value = maps/s of the mixed-precision path, the float64 batch, the
    float64 drop-in per segment and the CPU oracle beside it."""
    from sphradiometer_amd import inject
    L, N, nseg = 25, 4096, 16
    pos = inject.positions("AB")
    t0 = time.time()
    net = S.Network(pos, N, 1.0 / 4096)
    net.set_l(L)
    plan_s = time.time() - t0
    ts = inject.two_site(nseg, N, seed=0)
    w = inject.hann(N)
    d_ts = S.DeviceBuffer.from_numpy(ts)
    d_w = S.DeviceBuffer.from_numpy(w)
    r64 = run_path(S, net, L, d_ts, nseg, S.FP64, args.steps, args.warmup, d_window=d_w)
    rmx = run_path(S, net, L, d_ts, nseg, S.MIXED, args.steps, args.warmup, d_window=d_w)
    print("fp64 %.4f ms, mixed %.4f ms per step of %d segments" % (r64["ms_per_step"], rmx["ms_per_step"], nseg), file=sys.stderr, flush=True)
    pl = plain_fp64(S, net, ts, L, args.plain_budget, nseg, window=w)
    plain_err = float(np.linalg.norm(pl["alm"] - r64["alm"][pl["last"]]) / np.linalg.norm(r64["alm"][pl["last"]]))
    alm_err = np.linalg.norm(rmx["alm"] - r64["alm"], axis=1) / np.linalg.norm(r64["alm"], axis=1)
    map_err = float(np.linalg.norm(rmx["map"] - r64["map"]) / np.linalg.norm(r64["map"]))
    value = nseg / (rmx["ms_per_step"] * 1e-3)
    ncoef, nmesh = (L + 1) ** 2, (2 * L + 2) ** 2
    b_map = 2 * N * 8 + (ncoef + nmesh) * 8 + rmx["info"]["table_bytes"] / nseg
    gbs = b_map * nseg / (rmx["ms_per_step"] * 1e-3) / 1e9
    out = {
        "metric": "sky maps/s, 2-det l_max=25, 16 x 4096-sample Hann segments (batched mixed precision)",
        "value": round(value, 1), "unit": "sky maps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(rmx["ms_per_step"], 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: white N(0,1) float64 per detector (seed 0), Hann-windowed",
        "config": {"workload": "two sites, 16 segments x 4096 samples at 4096 Hz, Hann window, sky l_max 25, a_lm + 52x52 map per segment",
                   "segments": nseg, "samples": N, "l_max": L, "detectors": 2, "baselines": 1},
        "fp64": {"value": round(nseg / (r64["ms_per_step"] * 1e-3), 1), "ms_per_step": round(r64["ms_per_step"], 4)},
        "fp64_plain": {"value": round(pl["value"], 1), "segments": pl["segments"], "alm_rel_err_vs_batched_fp64": plain_err},
        "mixed_vs_fp64": {"alm_rel_err_max": float(alm_err.max()), "map_rel_err": map_err},
        "stage_ms": [round(t, 5) for t in rmx["stage_ms"]],
        "roofline": {"bound": "hbm", "kernel": "whole step (16 segments: launch-latency bound)", "achieved": round(gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None,
                     "bytes_per_map": round(b_map, 1)},
        "plan_build_s": round(plan_s, 2),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ts * w, L, args.cpu_budget, positions=pos, label="two-site")
        out["cpu_baseline"]["value"] = round(out["cpu_baseline"]["value"], 2)
    print(json.dumps(out))


def bench_copy(S, args):
    """A plain device-to-device copy of 256 MiB, five times: the reference
    kernel of known size for scaling rocprofv3's FETCH_SIZE and WRITE_SIZE
    counters to bytes (run under --pmc; prints the bytes per copy)."""
    n = 256 << 20
    a, b = S.DeviceBuffer(n), S.DeviceBuffer(n)
    S.lib().sphr_memset_device(a.ptr, 1, n)
    for _ in range(5):
        if S.lib().sphr_memcpy_dtod(b.ptr, a.ptr, n):
            raise RuntimeError(S.last_error())
    S.synchronize()
    print(json.dumps({"workload": "copy", "bytes_per_copy": n, "copies": 5}))


# HBM bytes per launch of the fused sky kernel (float32, l_max 32, 1024
# segments): rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of the timed
# dispatches on this workload, scaled by the counters' reading of a 256 MiB
# copy (FETCH_SIZE reads half the copied bytes, WRITE_SIZE all of them);
# the summary with the commands is profiles/three_site_l32_traffic.md.
SKY_FUSED_F32_WRITE = 49048 * 1024
SKY_FUSED_F32_READ = 1781 * 1024 * 2
# the lag-domain kernel (float32, two pairs per pass) on the same workload:
# the series, the lag table and its register spills in, one row of partial
# sums per segment and the spills out
XCORR_F32_WRITE = 3264 * 1024
XCORR_F32_READ = 13063 * 1024 * 2


def sky_fused_traffic():
    return SKY_FUSED_F32_WRITE + SKY_FUSED_F32_READ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--segments", type=int, default=1024)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-budget", type=float, default=3.0)
    ap.add_argument("--workload", choices=["sky", "two_site", "localization", "jobs", "healpix", "td", "copy"], default="sky",
                    help="sky: the three-site cross-power maps (default); two_site: 2 sites, l_max 25, 16 Hann segments; "
                         "localization: the log-likelihood skies; "
                         "jobs: the uneven batch with a plan per job; healpix: RING maps + posterior of localization skies; "
                         "td: the time-domain correlator over a long stream")
    ap.add_argument("--events", type=int, default=8192)
    ap.add_argument("--jobs", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (segments are independent; "
              "see DESIGN.md for how the path would shard). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import sphradiometer_amd as S
    from sphradiometer_amd import inject
    S.lib()
    if not S.gpu_available():
        print("no GPU: " + S.last_error(), file=sys.stderr)
        sys.exit(1)
    if args.workload == "localization":
        bench_localization(S, args)
        return
    if args.workload == "jobs":
        bench_jobs(S, args)
        return
    if args.workload == "healpix":
        bench_healpix(S, args)
        return
    if args.workload == "td":
        bench_td(S, args)
        return
    if args.workload == "copy":
        bench_copy(S, args)
        return
    if args.workload == "two_site":
        bench_two_site(S, args)
        return

    L, N, nseg = 32, 1024, args.segments
    t0 = time.time()
    net = S.Network(inject.positions("ABC"), N, 1.0 / 4096)
    net.set_l(L)
    plan_s = time.time() - t0
    ts, _ = inject.three_site(nseg, N, seed=0)
    d_ts = S.DeviceBuffer.from_numpy(ts)
    print("plan built in %.2f s; data %s on device" % (plan_s, ts.shape), file=sys.stderr, flush=True)

    r64 = run_path(S, net, L, d_ts, nseg, S.FP64, args.steps, args.warmup)
    print("fp64: %.3f ms/step, stages %s" % (r64["ms_per_step"], ["%.4f" % t for t in r64["stage_ms"]]), file=sys.stderr, flush=True)
    rmx = run_path(S, net, L, d_ts, nseg, S.MIXED, args.steps, args.warmup)
    print("mixed: %.3f ms/step, stages %s" % (rmx["ms_per_step"], ["%.4f" % t for t in rmx["stage_ms"]]), file=sys.stderr, flush=True)

    pl = plain_fp64(S, net, ts, L, args.plain_budget, nseg)
    plain_err = float(np.linalg.norm(pl["alm"] - r64["alm"][pl["last"]]) / np.linalg.norm(r64["alm"][pl["last"]]))
    print("plain fp64 drop-in: %.1f maps/s over %d segments (batched fp64 vs plain a_lm rel err %.2g)"
          % (pl["value"], pl["segments"], plain_err), file=sys.stderr, flush=True)

    alm_err = np.linalg.norm(rmx["alm"] - r64["alm"], axis=1) / np.linalg.norm(r64["alm"], axis=1)
    map_err = float(np.linalg.norm(rmx["map"] - r64["map"]) / np.linalg.norm(r64["map"]))
    value = nseg / (rmx["ms_per_step"] * 1e-3)
    fp64_value = nseg / (r64["ms_per_step"] * 1e-3)

    names = ["xcorr", "none"] if rmx["info"].get("lag") else ["fft", "xproj"]
    names += ["sky_fused", "none"] if rmx["info"].get("fused") else ["alm_legendre", "phi_gemm"]
    sb = stage_bytes(rmx["info"], nseg, rmx["esz"])
    k = int(np.argmax(rmx["stage_ms"]))
    sb64 = stage_bytes(r64["info"], nseg, r64["esz"])
    k64 = int(np.argmax(r64["stage_ms"]))
    # the dominant (fused sky) kernel's algorithmic bytes: the a_lm and maps
    # it writes and the rotation rows it reads; the K-split partial sums it
    # reads are an intermediate of the pipeline, reported beside
    info, esz = rmx["info"], rmx["esz"]
    ncoef, nmesh = (L + 1) ** 2, (2 * L + 2) ** 2
    sky_out = nseg * (ncoef + nmesh) * 2 * esz + info["baselines"] * ncoef * 2 * esz
    partial_rd = (1 if info.get("lag") else info["k_splits"]) * nseg * info["columns"] * esz
    t_dom = rmx["stage_ms"][k] * 1e-3
    achieved = (sky_out if names[k] == "sky_fused" else sb[k]) / t_dom / 1e9
    # the whole step against the per-map bytes of the path (series in,
    # a_lm + map out, the delay-product table shared by the batch)
    b_map = 3 * N * 8 + ncoef * 2 * esz + nmesh * 2 * esz + info["table_bytes"] / nseg
    step_gbs = b_map * nseg / (rmx["ms_per_step"] * 1e-3) / 1e9
    # the lag-domain kernel (first stage): series and table in, partial
    # sums out, against HBM
    xcorr_line = None
    if info.get("lag"):
        xb = sb[0]
        xt = rmx["stage_ms"][0] * 1e-3
        xcorr_line = {"kernel": "xcorr (form %d)" % info.get("xcorr_form", 0), "algorithmic_bytes": xb,
                      "achieved": round(xb / xt / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(xb / xt / 1e9 / HBM_PEAK_GBS, 4),
                      "traffic": XCORR_F32_WRITE + XCORR_F32_READ if info.get("xcorr_form") == 4 and nseg == 1024 and L == 32 else None}
    # peak direction: mixed against float64 (both map-seeded + extremize)
    sep = np.degrees(np.arccos(np.clip(np.cos(rmx["peaks"][0]) * np.cos(r64["peaks"][0]) + np.sin(rmx["peaks"][0]) * np.sin(r64["peaks"][0])
                                       * np.cos(rmx["peaks"][1] - r64["peaks"][1]), -1.0, 1.0)))
    pk_val_err = np.abs(rmx["peaks"][2] - r64["peaks"][2]) / np.abs(r64["peaks"][2])

    out = {
        "metric": "sky maps/s, 3-det l_max=32: mixed-prec batched vs FP64 plain at <=1e-5 a_lm err",
        "value": round(value, 1),
        "unit": "sky maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(rmx["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: white N(0,1) float64 + one sine-Gaussian burst per segment (seed 0)",
        "config": {"workload": "three sites, %d segments x %d samples at 4096 Hz, sky l_max %d, a_lm + 66x66 map per segment" % (nseg, N, L),
                   "segments": nseg, "samples": N, "l_max": L, "detectors": 3, "baselines": 3},
        "fp64": {"value": round(fp64_value, 1), "ms_per_step": round(r64["ms_per_step"], 4),
                 "stage_ms": [round(t, 5) for t in r64["stage_ms"]],
                 "dominant": names[k64], "dominant_gbs": round(sb64[k64] / (r64["stage_ms"][k64] * 1e-3) / 1e9, 1)},
        "fp64_plain": {"value": round(pl["value"], 1), "segments": pl["segments"],
                       "path": "drop-in C API per segment: tseries_to_fseries x3, network_integrate_power_fd, to_mesh",
                       "alm_rel_err_vs_batched_fp64": plain_err},
        "mixed_vs_fp64_plain": round(value / pl["value"], 1),
        "mixed_vs_fp64": {"speedup": round(value / fp64_value, 3), "alm_rel_err_max": float(alm_err.max()),
                          "alm_rel_err_median": float(np.median(alm_err)), "map_rel_err": map_err,
                          "peak_sep_deg_median": float(np.median(sep)), "peak_sep_deg_p99": float(np.quantile(sep, 0.99)),
                          "peak_sep_deg_max": float(sep.max()), "peaks_within_0.01_deg": float(np.mean(sep <= 0.01)),
                          "peak_value_rel_err_max": float(pk_val_err.max())},
        "stage_ms": [round(t, 5) for t in rmx["stage_ms"]],
        "stages": names,
        "tuned": {"mixed": rmx["tuned"], "fp64": r64["tuned"], "info": rmx["info"]},
        "roofline": {"bound": "hbm", "kernel": names[k], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": sky_fused_traffic() if names[k] == "sky_fused" and nseg == 1024 and L == 32 else None,
                     "traffic_basis": "rocprofv3 FETCH_SIZE/WRITE_SIZE of the timed dispatches on this workload, copy-calibrated",
                     "algorithmic_bytes": sky_out if names[k] == "sky_fused" else sb[k],
                     "algorithmic_basis": "a_lm + maps written, rotation rows read",
                     "intermediate_read_bytes": partial_rd if names[k] == "sky_fused" else None,
                     "achieved_with_intermediates": round((sky_out + partial_rd) / t_dom / 1e9, 1) if names[k] == "sky_fused" else None},
        "xcorr_roofline": xcorr_line,
        "step_roofline": {"bytes_per_map": round(b_map, 1), "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4), "roofline_maps_per_s": round(HBM_PEAK_GBS * 1e9 / b_map, 1)},
        "plan_build_s": round(plan_s, 2),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ts, L, args.cpu_budget)
        out["cpu_baseline"]["value"] = round(out["cpu_baseline"]["value"], 2)
        out["vs_cpu_all_cores"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
