"""Benchmark of the MI355X acoustic finite-difference engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--shots S]

Default workload: reference-semantics RTM of a batch of 96 shots on a seeded
layered-and-faulted model of 751 x 2301 cells at 4 m (pml 30 -> 811 x 2361
padded cells), dt = 0.4 ms, nt = 5000, Ricker 12 Hz, migration velocity 10x
smoother than the true model.  The input records are forward-modeled from the
true model on the GPU before timing (inputs resident in HBM when the timed
region starts).  A step is one RTM of the whole batch: forward propagation
with boundary-strip saving, then source reconstruction + receiver
back-propagation + imaging, i.e. 3 wavefield propagations per shot.

Prints one JSON line: Gcell-updates/s over the timed steps, the dominant
kernel's HBM roofline fraction (HIP-event durations on the launch stream and
algorithmic bytes), and the reference's CPU path timed on this host on a
bounded sample of the same workload.
"""
import argparse
import contextlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s (batched shots) + LSRTM s/CG-iter vs plain FP32 at stated error"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of each kernel for the 96-shot workload, from
# rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated against a 1 GiB copy; the
# commands and counter values are in profiles/rtm_batch_hbm_traffic.md).
# Scaled by the number of shots in a launch.
MEASURED_TRAFFIC_96 = {"step_rtm_forward": 4.474e+09, "step_rtm_reverse": 8.846e+09, "rtm_image": 4.291e+09}


def rtm_workload(nshots, nt=5000):
    from seis_amd import models
    nx, nz, pml = 2301, 751, 30
    dx = dz = 4.0
    dt, fr = 0.0004, 12.0
    vtrue = models.layered_faulted(nx, nz, dx, dz, seed=1001)
    vmig = models.smooth_model(vtrue, 20.0)
    delay = int(round(1.0 / (fr * dt)))
    sxs = [12 + 24 * i for i in range(96)][:nshots]
    return dict(nx=nx, nz=nz, pml=pml, dx=dx, dz=dz, dt=dt, nt=nt, fr=fr, delay=delay, vtrue=vtrue, vmig=vmig,
                sxs=sxs, sz=2, gz=2, spread=240, rstep=3)


def spread_mask(w, sx):
    cols = np.arange(w["nx"])
    off = cols - sx
    return (np.abs(off) <= w["spread"] * w["rstep"]) & (off % w["rstep"] == 0)


def cpu_baseline(w, nt_sample, threads):
    """Reference RTM (compiled reference engine if present, else the oracle's
    restatement) of one shot with a shortened record, on `threads` host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    kind = "reference" if oracle_lib.have_reference() else "port"
    lib = oracle_lib.Reference() if kind == "reference" else oracle_lib.Oracle()
    lib.set_threads(threads)
    nx, nz, pml = w["nx"], w["nz"], w["pml"]
    wav = lib.ricker(w["dt"], w["fr"], w["delay"], nt_sample)
    gz = np.full(nx, w["gz"], np.int32)
    rec = np.zeros((nx, nt_sample), np.float32)
    sx = w["sxs"][len(w["sxs"]) // 2]
    rec[spread_mask(w, sx)] = np.random.default_rng(5).standard_normal((int(spread_mask(w, sx).sum()), nt_sample))
    t0 = time.perf_counter()
    with stdout_to_stderr():
        lib.rtm(w["dt"], w["dx"], w["dz"], nt_sample, 0, nx, nz, pml, sx, w["sz"], gz, wav, w["vmig"], rec)
    el = time.perf_counter() - t0
    cells = (nx + 2 * pml) * (nz + 2 * pml)
    updates = 3.0 * nt_sample * cells
    return dict(value=updates / el / 1e9, unit="Gcell-updates/s", cores=threads, kind=kind,
                sample="one shot of the same RTM workload with nt=%d (%.1f s of CPU work)" % (nt_sample, el),
                cpu_model=host_cpu_model())


def host_cpu_model():
    """Model name of the host CPU (first processor of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


@contextlib.contextmanager
def stdout_to_stderr():
    """The compiled reference prints its progress to stdout (fd 1); the
    bench's stdout carries only the JSON line."""
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        yield
    finally:
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)


def c1_single_shot(calls, with_cpu, threads):
    """The reference's CPU-runnable case: one-shot forward modeling of the
    3-layer model, 201 x 301 cells @ 10 m, pml 30, nt = 1500, through the
    single-shot host-buffer entry point (fd.h modeling), its time loop
    launched step by step."""
    import seis_amd
    nx, nz, pml, nt, dt, dx = 301, 201, 30, 1500, 0.001, 10.0
    vel = np.empty((nx, nz), np.float32)
    z = np.arange(nz) * dx
    vel[:] = np.where(z < 600.0, 1500.0, np.where(z < 1400.0, 2200.0, 3000.0)).astype(np.float32)
    gz = np.full(nx, 2, np.int32)
    wav = seis_amd.ricker(dt, 15.0, 67, nt)
    args = (dt, dx, dx, nt, 0, nx, nz, pml, 150, 2, gz, wav, vel)
    out = {"workload": "modeling, 1 shot, 201x301 cells @ 10 m, pml 30, nt=1500, 3-layer model, host buffers "
                       "in and out (fd.h modeling)"}
    rec = seis_amd.modeling(*args)  # warm-up (and the strip-length trial)
    t0 = time.perf_counter()
    for _ in range(calls):
        seis_amd.modeling(*args)
    out["ms_per_call"] = round((time.perf_counter() - t0) * 1e3 / calls, 3)
    out["strip_length"] = seis_amd.tuned_strip("forward", nx + 2 * pml, nz + 2 * pml, 1)
    cells = (nx + 2 * pml) * (nz + 2 * pml) * nt
    out["gcell_updates_per_s"] = round(cells / (out["ms_per_call"] * 1e-3) / 1e9, 2)
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        kind = "reference" if oracle_lib.have_reference() else "port"
        lib = oracle_lib.Reference() if kind == "reference" else oracle_lib.Oracle()
        lib.set_threads(threads)
        t0 = time.perf_counter()
        with stdout_to_stderr():
            ref = lib.modeling(*args)
        out["cpu_ms_per_call"] = round((time.perf_counter() - t0) * 1e3, 1)
        out["cpu"] = {"kind": kind, "cores": threads}
        out["equals_cpu"] = bool(np.array_equal(ref, rec))
    return out


def c3_born_adjoint(nshots, pair_shots, seed=1003):
    """Batched Born modeling with uneven shots: 500 x 1600 cells @ 5 m,
    pml 30, dt = 0.5 ms, record lengths drawn in 2000..4000 and spreads of
    200..1200 receivers, a sparse spike perturbation (2 % density, N(0,1)),
    every shot in one batch.  Then the BF16-snapshot Born / exact-adjoint pair
    on the first `pair_shots` shots with its FP64 dot-product test."""
    import ctypes
    import seis_amd
    from seis_amd import DeviceArray, models
    from seis_amd.batch import _layout, Grid, Shot
    from seis_amd._lib import check
    L = seis_amd.lib()
    nx, nz, pml, dx, dt, fr = 1600, 500, 30, 5.0, 0.0005, 15.0
    rng = np.random.default_rng(seed)
    # 1500-4500 m/s keeps v dt / dx at 0.45, inside the 8th-order scheme's
    # stability limit (0.55) with a margin
    v = models.layered_faulted(nx, nz, dx, dx, vmax=4500.0, seed=seed)
    dv = models.sparse_spikes(nx, nz, 0.02, seed)
    delay = int(round(1.0 / (fr * dt)))
    nts = rng.integers(2000, 4001, nshots)
    ngs = rng.integers(200, 1201, nshots)
    sxs = np.linspace(5, nx - 6, nshots).astype(int)
    shots = [Shot(sx=int(sx), sz=2, nt=int(n)) for sx, n in zip(sxs, nts)]
    grid = Grid(nx, nz, pml, dt, dx, dx)
    g = grid.c()
    arr, nw, ngz, nr = _layout(grid, shots)
    wav = np.concatenate([seis_amd.ricker(dt, fr, delay, int(n)) for n in nts])
    d_v, d_dv = DeviceArray.from_host(v), DeviceArray.from_host(dv)
    d_wav = DeviceArray.from_host(wav)
    d_gz = DeviceArray.from_host(np.full(nx * nshots, 2, np.int32))
    d_rec = DeviceArray((nr,), np.float32, zero=True)
    cells = (nx + 2 * pml) * (nz + 2 * pml)
    out = {"workload": "born_batch: %d shots in one batch, 500x1600 cells @ 5 m, pml 30, dt=0.5 ms, nt in "
                       "2000..4000, 200..1200 receivers, 2%% sparse spike perturbation" % nshots,
           "nt_range": [int(nts.min()), int(nts.max())]}
    # (the strip-length trial of this configuration runs inside: 12 launches)
    t0 = time.perf_counter()
    check(L.seis_born_batch(ctypes.byref(g), nshots, arr, d_v.ptr, d_dv.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr))
    seis_amd.synchronize()
    el = time.perf_counter() - t0
    out["born_s"] = round(el, 3)
    out["born_gcell_updates_per_s"] = round(2.0 * cells * float(nts.sum()) / el / 1e9, 1)
    log("C3: Born batch of %d shots in %.1f s" % (nshots, el))
    del d_rec
    # the Born / adjoint pair with BF16 snapshots on a subset
    sub = shots[:pair_shots]
    sarr, snw, sng, snr = _layout(grid, sub)
    mask = np.zeros(snr, np.float32)
    off = 0
    for s, ng in zip(sub, ngs[:pair_shots]):
        lo, hi = max(0, s.sx - int(ng) // 2), min(nx, s.sx + int(ng) // 2)
        mask[off:off + nx * s.nt].reshape(nx, s.nt)[lo:hi] = 1.0
        off += nx * s.nt
    sw = DeviceArray.from_host(np.concatenate([seis_amd.ricker(dt, fr, delay, s.nt) for s in sub]))
    sgz = DeviceArray.from_host(np.full(nx * pair_shots, 2, np.int32))
    plan = ctypes.c_void_p()
    t0 = time.perf_counter()
    check(L.seis_lsrtm_create(ctypes.byref(plan), ctypes.byref(g), pair_shots, sarr, d_v.ptr, sw.ptr, sgz.ptr, 1))
    dmask = DeviceArray.from_host(mask)
    check(L.seis_lsrtm_set_mask(plan, dmask.ptr))
    seis_amd.synchronize()
    out["pair_snapshot_s"] = round(time.perf_counter() - t0, 3)
    r = DeviceArray.from_host((rng.standard_normal(snr) * mask).astype(np.float32))
    lm = DeviceArray((snr,), np.float32, zero=True)
    ltr = DeviceArray((nx, nz), np.float32, zero=True)
    t0 = time.perf_counter()
    check(L.seis_lsrtm_born(plan, d_dv.ptr, lm.ptr))
    seis_amd.synchronize()
    t1 = time.perf_counter()
    check(L.seis_lsrtm_adjoint(plan, r.ptr, ltr.ptr))
    seis_amd.synchronize()
    t2 = time.perf_counter()
    def dot(pl, x, y, n):
        o = ctypes.c_double()
        check(L.seis_lsrtm_dot(pl, x.ptr, y.ptr, n, ctypes.byref(o)))
        return o.value

    def dot_test(pl, m, lm, r, ltr, nrec):
        a, b = dot(pl, lm, r, nrec), dot(pl, m, ltr, nx * nz)
        norms = np.sqrt(dot(pl, lm, lm, nrec) * dot(pl, r, r, nrec))
        return abs(a - b) / max(abs(a), abs(b), 1e-300), abs(a - b) / max(norms, 1e-300)

    rel, reln = dot_test(plan, d_dv, lm, r, ltr, snr)
    check(L.seis_lsrtm_destroy(plan))
    pair_cells = cells * float(sum(s.nt for s in sub))
    out["pair_shots"] = pair_shots
    out["pair_born_gcell_updates_per_s"] = round(pair_cells / (t1 - t0) / 1e9, 1)
    out["pair_adjoint_gcell_updates_per_s"] = round(pair_cells / (t2 - t1) / 1e9, 1)
    # |<L m, r> - <m, L^T r>| relative to the value and to ||L m|| ||r||: a
    # random r makes the two products cancel to ~1/sqrt(n) of the norms, so
    # FP32 rounding shows up far larger against the value
    out["dot_test_fp32_rel"] = float("%.3e" % rel)
    out["dot_test_fp32_rel_to_norms"] = float("%.3e" % reln)
    # the same test in FP64 (propagation, snapshots and products) on 4 shots
    n64 = min(4, pair_shots)
    sub64 = sub[:n64]
    a64, nw64, ng64, nr64 = _layout(grid, sub64)
    v64 = DeviceArray.from_host(v.astype(np.float64))
    w64 = DeviceArray.from_host(np.concatenate([seis_amd.ricker(dt, fr, delay, s.nt) for s in sub64]).astype(np.float64))
    p64 = ctypes.c_void_p()
    check(L.seis_lsrtm_create_f64(ctypes.byref(p64), ctypes.byref(g), n64, a64, v64.ptr, w64.ptr, sgz.ptr))
    m64 = DeviceArray.from_host(dv.astype(np.float64))
    r64 = DeviceArray.from_host((rng.standard_normal(nr64) * mask[:nr64]).astype(np.float64))
    lm64 = DeviceArray((nr64,), np.float64, zero=True)
    ltr64 = DeviceArray((nx, nz), np.float64, zero=True)
    check(L.seis_lsrtm_born(p64, m64.ptr, lm64.ptr))
    check(L.seis_lsrtm_adjoint(p64, r64.ptr, ltr64.ptr))
    rel64, reln64 = dot_test(p64, m64, lm64, r64, ltr64, nr64)
    check(L.seis_lsrtm_destroy(p64))
    out["dot_test_fp64_rel"] = float("%.3e" % rel64)
    out["dot_test_fp64_shots"] = n64
    out["dot_test_note"] = ("FP32: FP32 propagation, BF16 snapshots, FP64 inner products; FP64: every step in FP64 "
                            "(the stated bar, <= 1e-6)")
    log("C3: pair on %d shots; dot test FP32 %.2e (%.2e of the norms), FP64 %.2e" % (pair_shots, rel, reln, rel64))
    return out


def c5_sample(nshots, nt, seed=1005):
    """RTM throughput on the large grid: 2000 x 8000 cells @ 5 m, pml 30,
    layered-faulted model with a 4500 m/s salt body, dt = 0.5 ms, a short
    record of the workload's shots."""
    import ctypes
    import seis_amd
    from seis_amd import DeviceArray, models
    from seis_amd.batch import _layout, Grid, Shot, RTM_ISOLATE_BAND
    from seis_amd._lib import check
    L = seis_amd.lib()
    nx, nz, pml, dx, dt, fr = 8000, 2000, 30, 5.0, 0.0005, 12.0
    v = models.salt_body(models.layered_faulted(nx, nz, dx, dx, seed=seed), dx, dx, seed=seed)
    vm = models.smooth_model(v, 20.0)
    delay = int(round(1.0 / (fr * dt)))
    sxs = np.linspace(500, nx - 500, nshots).astype(int)
    shots = [Shot(sx=int(sx), sz=2, nt=nt) for sx in sxs]
    grid = Grid(nx, nz, pml, dt, dx, dx)
    g = grid.c()
    arr, nw, ngz, nr = _layout(grid, shots)
    d_v, d_vm = DeviceArray.from_host(v), DeviceArray.from_host(vm)
    d_wav = DeviceArray.from_host(np.tile(seis_amd.ricker(dt, fr, delay, nt), nshots))
    d_gz = DeviceArray.from_host(np.full(nx * nshots, 2, np.int32))
    d_rec = DeviceArray((nr,), np.float32, zero=True)
    d_img = DeviceArray((nshots * nx * nz,), np.float32, zero=True)
    check(L.seis_modeling_batch(ctypes.byref(g), nshots, arr, d_v.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr))
    seis_amd.synchronize()
    t0 = time.perf_counter()
    check(L.seis_rtm_batch_ex(ctypes.byref(g), nshots, arr, d_vm.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr, d_img.ptr,
                              RTM_ISOLATE_BAND))
    seis_amd.synchronize()
    el = time.perf_counter() - t0
    cells = (nx + 2 * pml) * (nz + 2 * pml)
    img = d_img.download()
    log("C5 sample: RTM of %d shots x %d steps in %.1f s" % (nshots, nt, el))
    return {"workload": "rtm_batch sample: %d shots, 2000x8000 cells @ 5 m, pml 30, nt=%d of the workload's 6000, "
                        "dt=0.5 ms, layered-faulted model with a salt body, isolated source band" % (nshots, nt),
            "s": round(el, 3), "s_per_shot_per_1000_steps": round(el / nshots * 1000.0 / nt, 3),
            "gcell_updates_per_s": round(3.0 * nt * cells * nshots / el / 1e9, 1),
            "image_finite": bool(np.isfinite(img).all())}


def plane_prepare_leg(nshots, ng, nt, n_p, with_cpu, seed=1006):
    """Plane-wave synthesis (planeWavePrepare): the workload's 96 shots of 480
    traces (every 3rd column of 2301), nt = 5000, phase-shifted and stacked
    for 8 ray parameters.  Gathers resident on the GPU before timing; the CPU
    baseline is the oracle's restatement of the reference loop on one shot and
    one ray parameter, scaled by shots x parameters."""
    import ctypes
    import seis_amd
    from seis_amd import DeviceArray
    from seis_amd._lib import check
    L = seis_amd.lib()
    nx, dx = 2301, 4.0
    rng = np.random.default_rng(seed)
    isx = np.linspace(300, nx - 300, nshots).astype(np.int32)
    sx = (isx.astype(np.float32) * np.float32(dx)).astype(np.float32)
    cols = []
    for s in range(nshots):
        c = (isx[s] - 3 * (ng // 2) + 3 * np.arange(ng)).astype(np.int32)
        c[(c < 0) | (c >= nx)] = -1
        cols.append(c)
    col = np.concatenate(cols)
    ngs = np.full(nshots, ng, np.int32)
    ntw = nt + 500
    wav = seis_amd.ricker(0.0004, 12.0, 208 + 500, ntw)
    d_data = DeviceArray((nshots * ng, nt), np.float32)
    chunk = 4 * ng
    first = None
    for r0 in range(0, nshots * ng, chunk):  # fill in pieces (bounded host memory)
        n = min(chunk, nshots * ng - r0)
        piece = rng.standard_normal((n, nt), dtype=np.float32)
        if first is None:
            first = piece[:ng].copy()
        check(L.seis_memcpy_htod(ctypes.c_void_p(d_data.ptr.value + r0 * nt * 4),
                                 piece.ctypes.data_as(ctypes.c_void_p), piece.nbytes))
    d_wav = DeviceArray.from_host(wav)
    d_sou = DeviceArray((n_p, nx, ntw), np.float32)
    d_rec = DeviceArray((n_p, nx, nt), np.float32)
    P = ctypes.c_void_p
    pmin, dp = np.float32(-2.0e-4), np.float32(5.0e-5)
    args = (nx, nt, ntw, nshots, n_p, float(pmin), float(dp), sx.ctypes.data_as(P), isx.ctypes.data_as(P),
            ngs.ctypes.data_as(P), col.ctypes.data_as(P), d_wav.ptr, d_data.ptr, d_sou.ptr, d_rec.ptr)
    check(L.seis_plane_prepare(*args))  # warm-up
    seis_amd.synchronize()
    seis_amd.profile_start()
    t0 = time.perf_counter()
    check(L.seis_plane_prepare(*args))
    seis_amd.synchronize()
    el = time.perf_counter() - t0
    st = seis_amd.profile_stop()
    out = {"workload": "planeWavePrepare: %d shots x %d traces, nt=%d, %d ray parameters, 2301 columns" %
                       (nshots, ng, nt, n_p),
           "s": round(el, 4), "trace_samples_per_s_per_parameter": round(nshots * ng * nt * n_p / el / 1e9, 2),
           "unit": "G trace-samples/s (x parameters)",
           "kernels": {k: {"launches": v["launches"], "ms": round(v["total_ms"], 3),
                           "GB/s": round(v["alg_bytes"] / max(v["total_ms"], 1e-9) / 1e6, 1)}
                       for k, v in st.items() if k.startswith("plane_")}}
    rec = d_rec.download()
    out["output_finite"] = bool(np.isfinite(rec).all())
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import Oracle
        o = Oracle()
        t0 = time.perf_counter()
        o.plane_prepare(nx, nt, ntw, 1, pmin, dp, sx[:1], isx[:1], ngs[:1], cols[0], wav, first)
        cpu1 = time.perf_counter() - t0
        est = cpu1 * nshots * n_p
        out["cpu"] = {"kind": "port", "cores": 1, "sample": "1 shot x %d traces x nt=%d, 1 ray parameter" % (ng, nt),
                      "sample_s": round(cpu1, 2), "est_s": round(est, 1), "gpu_speedup": round(est / el, 1)}
    log("plane prepare: %.3f s on the GPU" % el)
    return out


def snapshot_vs_recompute(nshots, nt, seed=1007):
    """The RTM workload's grid (751 x 2301 @ 4 m, pml 30, dt = 0.4 ms) imaged two
    ways for the same shots.
    * Recomputation: the reference's method. The forward pass keeps 4-layer
      boundary strips, and the source field is rebuilt backwards beside the
      receiver field (3 propagations).
    * Full snapshot storage: the forward pass stores the Born source of every
      step and model cell in BF16 with a per-step scale, and the reverse pass
      propagates the receiver field only and images against the snapshots
      (the exact adjoint of Born modeling, 2 propagations plus the snapshots'
      max-scale pass).
    The records are synthetic; the two images are different operators (FaQi
    cross-correlation and exact Born adjoint), so time and memory are compared,
    not images."""
    import ctypes
    import seis_amd
    from seis_amd import DeviceArray, models
    from seis_amd.batch import _layout, Grid, Shot, RTM_ISOLATE_BAND
    from seis_amd.lsrtm import LsrtmPlan, SNAP_BF16
    from seis_amd._lib import check
    L = seis_amd.lib()
    nx, nz, pml, dx, dt, fr = 2301, 751, 30, 4.0, 0.0004, 12.0
    vm = models.smooth_model(models.layered_faulted(nx, nz, dx, dx, seed=seed), 20.0)
    delay = int(round(1.0 / (fr * dt)))
    shots = [Shot(sx=int(sx), sz=2, nt=nt) for sx in np.linspace(300, nx - 300, nshots)]
    grid = Grid(nx, nz, pml, dt, dx, dx)
    g = grid.c()
    arr, nw, ngz, nr = _layout(grid, shots)
    wav = seis_amd.ricker(dt, fr, delay, nt)
    gzs = [np.full(nx, 2, np.int32)] * nshots
    d_vm = DeviceArray.from_host(vm)
    d_wav = DeviceArray.from_host(np.tile(wav, nshots))
    d_gz = DeviceArray.from_host(np.full(nx * nshots, 2, np.int32))
    rng = np.random.default_rng(seed)
    d_rec = DeviceArray.from_host(rng.standard_normal(nr, dtype=np.float32))
    d_img = DeviceArray((nshots * nx * nz,), np.float32, zero=True)
    seis_amd.synchronize()
    t0 = time.perf_counter()
    check(L.seis_rtm_batch_ex(ctypes.byref(g), nshots, arr, d_vm.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr, d_img.ptr,
                              RTM_ISOLATE_BAND))
    seis_amd.synchronize()
    t_rec = time.perf_counter() - t0
    del d_img
    seis_amd.profile_start()
    t0 = time.perf_counter()
    plan = LsrtmPlan(grid, shots, vm, [wav] * nshots, gzs, snapshot=SNAP_BF16)
    seis_amd.synchronize()
    t_snap = time.perf_counter() - t0
    t0 = time.perf_counter()
    img = plan.adjoint(d_rec)
    t_adj = time.perf_counter() - t0
    st = seis_amd.profile_stop()
    plan.close()
    cells = (nx + 2 * pml) * (nz + 2 * pml)
    strips_gb = 32.0 * nt * (nx + nz) / 1e9
    snap_gb = 2.0 * nx * nz * nt / 1e9
    log("C2 snapshot vs recompute: %.2f s vs %.2f + %.2f s for %d shots" % (t_rec, t_snap, t_adj, nshots))
    return {"workload": "rtm imaging, %d shots, 751x2301 cells @ 4 m, pml 30, nt=%d, dt=0.4 ms" % (nshots, nt),
            "recompute": {"method": "boundary strips + backward source reconstruction (reference; isolated band)",
                          "s": round(t_rec, 3), "s_per_shot": round(t_rec / nshots, 3),
                          "propagations": 3, "stored_gb_per_shot": round(strips_gb, 2)},
            "snapshots": {"method": "BF16 Born-source snapshot of every step, per-step max scale",
                          "forward_s": round(t_snap, 3), "reverse_s": round(t_adj, 3),
                          "s_per_shot": round((t_snap + t_adj) / nshots, 3), "propagations": 2,
                          "stored_gb_per_shot": round(snap_gb, 2),
                          "shots_per_288_gb": int(288.0 // snap_gb),
                          "kernels_ms": {k: round(v["total_ms"], 1) for k, v in st.items()}},
            "recompute_cells_per_s": round(3.0 * cells * nt * nshots / t_rec / 1e9, 1),
            "image_finite": bool(np.isfinite(img).all())}


def lsrtm_bench(nshots, nt, iters, err_shots, err_iters):
    """Seconds per CG iteration of least-squares RTM on the 376 x 1151 @ 8 m
    workload (the RTM model at half resolution, nt = 3000, dt = 0.8 ms, 400
    receivers per shot), with BF16 snapshots, beside the plain FP32 GPU path
    (per iteration one Born batch and one RTM batch, as the reference's
    forward_MPI + adjoint_MPI would), and the image error of BF16 and FP32
    snapshots against the FP64 GPU plan on a small batch."""
    import ctypes
    import seis_amd
    from seis_amd import DeviceArray, models
    from seis_amd.batch import _layout, Grid, Shot
    from seis_amd._lib import check
    L = seis_amd.lib()
    nx, nz, pml, dx, dt, fr = 1151, 376, 30, 8.0, 0.0008, 12.0
    v = np.ascontiguousarray(models.layered_faulted(2301, 751, 4.0, 4.0, seed=1001)[::2, ::2])
    v0 = models.smooth_model(v, 10.0)
    dv_true = (v - v0).astype(np.float32)
    delay = int(round(1.0 / (fr * dt)))
    grid = Grid(nx, nz, pml, dt, dx, dx)
    g = grid.c()
    all_sx = [4 + 9 * i for i in range(128)]

    def survey(sxs, nt_):
        shots = [Shot(sx=sx, sz=2, nt=nt_) for sx in sxs]
        arr, nw, ng, nr = _layout(grid, shots)
        mask = np.zeros((len(sxs), nx, nt_), np.float32)
        for j, sx in enumerate(sxs):
            mask[j, max(0, sx - 200):min(nx, sx + 200)] = 1.0
        return shots, arr, nr, mask.ravel()

    def born_data(shots, arr, nr, mask, vel, wav, gz, noise_seed):
        d_dv = DeviceArray.from_host(dv_true)
        d_rec = DeviceArray((nr,), np.float32, zero=True)
        check(L.seis_born_batch(ctypes.byref(g), len(shots), arr, vel.ptr, d_dv.ptr, wav.ptr, gz.ptr, d_rec.ptr))
        d = d_rec.download() * mask
        rms = float(np.sqrt(np.mean(d[mask > 0] ** 2)))
        d += (0.05 * rms * np.random.default_rng(noise_seed).standard_normal(d.shape)).astype(np.float32) * mask
        return d

    out = {"workload": "lsrtm_cg: %d of the 128 shots, 376x1151 cells @ 8 m, pml 30, nt=%d, dt=0.8 ms, "
                       "Ricker 12 Hz, 400 receivers per shot, Born data of the true perturbation + 5%% noise"
                       % (nshots, nt)}
    # --- timed: BF16-snapshot CGLS against the plain FP32 path
    sxs = all_sx[::2][:nshots]
    shots, arr, nr, mask = survey(sxs, nt)
    ns = len(shots)
    wav1 = seis_amd.ricker(dt, fr, delay, nt)
    d_v0 = DeviceArray.from_host(v0)
    d_v = DeviceArray.from_host(v)
    d_wav = DeviceArray.from_host(np.tile(wav1, ns))
    d_gz = DeviceArray.from_host(np.full(nx * ns, 2, np.int32))
    d_mask = DeviceArray.from_host(mask)
    dobs = DeviceArray.from_host(born_data(shots, arr, nr, mask, d_v0, d_wav, d_gz, 11))
    log("lsrtm: observed data for %d shots ready" % ns)
    t0 = time.perf_counter()
    plan = ctypes.c_void_p()
    check(L.seis_lsrtm_create(ctypes.byref(plan), ctypes.byref(g), ns, arr, d_v0.ptr, d_wav.ptr, d_gz.ptr, 1))
    seis_amd.synchronize()
    out["snapshot_s"] = round(time.perf_counter() - t0, 3)
    check(L.seis_lsrtm_set_mask(plan, d_mask.ptr))
    img = DeviceArray((nx, nz), np.float32, zero=True)
    hist = (ctypes.c_double * (2 * (iters + 1)))()
    done = ctypes.c_int(0)
    seis_amd.profile_start()
    t0 = time.perf_counter()
    check(L.seis_lsrtm_cg(plan, dobs.ptr, iters, 0.0, img.ptr, hist, ctypes.byref(done)))
    seis_amd.synchronize()
    el = time.perf_counter() - t0
    st = seis_amd.profile_stop()
    out["kernels"] = {k: {"launches": v["launches"], "ms": round(v["total_ms"], 2),
                          "GB/s": round(v["alg_bytes"] / max(v["total_ms"], 1e-9) / 1e6, 1)} for k, v in st.items()}
    # the same CG with the cross-correlation adjoint (the reference's imaging
    # condition on the snapshots, SEIS_LSRTM_ADJ_XCORR)
    check(L.seis_lsrtm_set_adjoint(plan, 1))
    img_x = DeviceArray((nx, nz), np.float32, zero=True)
    done_x = ctypes.c_int(0)
    seis_amd.profile_start()
    t0 = time.perf_counter()
    check(L.seis_lsrtm_cg(plan, dobs.ptr, iters, 0.0, img_x.ptr, None, ctypes.byref(done_x)))
    seis_amd.synchronize()
    el_x = time.perf_counter() - t0
    st_x = seis_amd.profile_stop()
    check(L.seis_lsrtm_destroy(plan))
    out["kernels_xcorr"] = {k: {"launches": v["launches"], "ms": round(v["total_ms"], 2)} for k, v in st_x.items()}
    out["s_per_cg_iter_bf16_xcorr"] = round(el_x / max(done_x.value, 1), 3)
    if done_x.value > 0 and "lsrtm_born" in st_x and "lsrtm_adjoint_xcorr" in st_x:
        l_s = st_x["lsrtm_born"]["total_ms"] / done_x.value / 1e3
        lt_s = st_x["lsrtm_adjoint_xcorr"]["total_ms"] / (done_x.value + 1) / 1e3
        out["s_per_cg_iter_bf16_xcorr_marginal"] = round(l_s + lt_s, 3)
        # the workload's 20 iterations: the initial L^T d spread over them
        out["s_per_cg_iter_bf16_xcorr_20"] = round((lt_s + 20 * (l_s + lt_s)) / 20, 3)
    log("lsrtm: %d CG iterations with the cross-correlation adjoint in %.1f s" % (done_x.value, el_x))
    h = np.array(hist[:]).reshape(-1, 2)[: done.value + 1]
    out["cg_iterations"] = done.value
    out["s_per_cg_iter_bf16"] = round(el / max(done.value, 1), 3)
    # one CGLS iteration is one L and one L^T; the run also holds the initial
    # L^T dobs, so the per-iteration cost without it:
    if done.value > 0 and "lsrtm_born" in st:
        adj = sum(st[k]["total_ms"] for k in ("lsrtm_adjoint", "lsrtm_adjoint_prep") if k in st)
        l_s, lt_s = st["lsrtm_born"]["total_ms"] / done.value / 1e3, adj / (done.value + 1) / 1e3
        out["s_per_cg_iter_bf16_marginal"] = round(l_s + lt_s, 3)
        out["s_per_cg_iter_bf16_20"] = round((lt_s + 20 * (l_s + lt_s)) / 20, 3)
    out["residual_rel"] = [round(float(np.sqrt(x / h[0, 0])), 5) for x in h[:, 0]]
    log("lsrtm: %d CG iterations (BF16 snapshots) in %.1f s" % (done.value, el))
    # plain FP32 GPU path: Born batch + RTM batch per iteration
    d_pert = DeviceArray.from_host(img.download())
    d_rec = DeviceArray((nr,), np.float32, zero=True)
    d_img = DeviceArray((ns * nx * nz,), np.float32, zero=True)
    t0 = time.perf_counter()
    check(L.seis_born_batch(ctypes.byref(g), ns, arr, d_v0.ptr, d_pert.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr))
    check(L.seis_rtm_batch_ex(ctypes.byref(g), ns, arr, d_v0.ptr, d_wav.ptr, d_gz.ptr, dobs.ptr, d_img.ptr, 1))
    seis_amd.synchronize()
    plain = time.perf_counter() - t0
    out["s_per_iter_plain_fp32"] = round(plain, 3)
    out["cells_per_propagation"] = float((nx + 2 * pml) * (nz + 2 * pml)) * nt * ns
    out["speedup_vs_plain_fp32"] = round(plain / out["s_per_cg_iter_bf16"], 2)
    if "s_per_cg_iter_bf16_marginal" in out:
        out["speedup_vs_plain_fp32_marginal"] = round(plain / out["s_per_cg_iter_bf16_marginal"], 2)
    out["speedup_vs_plain_fp32_xcorr"] = round(plain / out["s_per_cg_iter_bf16_xcorr"], 2)
    for key in ("", "_xcorr"):
        if "s_per_cg_iter_bf16%s_20" % key in out:
            out["speedup_vs_plain_fp32%s_20_iters" % key] = round(plain / out["s_per_cg_iter_bf16%s_20" % key], 2)
    if "s_per_cg_iter_bf16_xcorr_marginal" in out:
        out["speedup_vs_plain_fp32_xcorr_marginal"] = round(plain / out["s_per_cg_iter_bf16_xcorr_marginal"], 2)
    log("lsrtm: plain FP32 path, one iteration's Born + RTM in %.1f s" % plain)
    del d_img, d_rec, d_pert, dobs
    # --- image error of 16/32-bit snapshots against the FP64 plan (small batch)
    if err_shots > 0:
        esx = all_sx[32:32 + 8 * err_shots:8][:err_shots]
        eshots, earr, enr, emask = survey(esx, nt)
        ew = DeviceArray.from_host(np.tile(wav1, len(eshots)))
        egz = DeviceArray.from_host(np.full(nx * len(eshots), 2, np.int32))
        ed = born_data(eshots, earr, enr, emask, d_v0, ew, egz, 12)
        imgs = {}
        for kind, name, adj in ((0, "fp32", 0), (1, "bf16", 0), (2, "fp16", 0), (-1, "fp64", 0),
                                (1, "bf16_xcorr", 1), (-1, "fp64_xcorr", 1)):
            pl = ctypes.c_void_p()
            if kind < 0:
                dt64 = np.float64
                # the plan reads these during its lifetime: keep them referenced
                v64 = DeviceArray.from_host(v0.astype(dt64))
                w64 = DeviceArray.from_host(np.tile(wav1, len(eshots)).astype(dt64))
                check(L.seis_lsrtm_create_f64(ctypes.byref(pl), ctypes.byref(g), len(eshots), earr, v64.ptr, w64.ptr,
                                              egz.ptr))
            else:
                dt64 = np.float32
                check(L.seis_lsrtm_create(ctypes.byref(pl), ctypes.byref(g), len(eshots), earr, d_v0.ptr, ew.ptr,
                                          egz.ptr, kind))
            m = DeviceArray.from_host(emask.astype(dt64))
            check(L.seis_lsrtm_set_mask(pl, m.ptr))
            check(L.seis_lsrtm_set_adjoint(pl, adj))
            dd = DeviceArray.from_host(ed.astype(dt64))
            im = DeviceArray((nx, nz), dt64, zero=True)
            check(L.seis_lsrtm_cg(pl, dd.ptr, err_iters, 0.0, im.ptr, None, None))
            imgs[name] = im.download().astype(np.float64)
            check(L.seis_lsrtm_destroy(pl))
        ref = imgs["fp64"]
        out["image_rel_l2_vs_fp64"] = {k: float("%.3e" % (np.linalg.norm(imgs[k] - ref) / np.linalg.norm(ref)))
                                       for k in ("fp32", "bf16", "fp16")}
        refx = imgs["fp64_xcorr"]
        out["image_rel_l2_vs_fp64"]["bf16_xcorr"] = float("%.3e" % (np.linalg.norm(imgs["bf16_xcorr"] - refx)
                                                                    / np.linalg.norm(refx)))
        out["xcorr_vs_exact_fp64_image_rel"] = float("%.3e" % (np.linalg.norm(refx - ref) / np.linalg.norm(ref)))
        out["error_case"] = "%d shots, %d CG iterations, same grid and nt" % (len(eshots), err_iters)
        log("lsrtm: image errors vs FP64 %s" % out["image_rel_l2_vs_fp64"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--shots", type=int, default=96)
    ap.add_argument("--cpu-nt", type=int, default=900)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nt", type=int, default=5000,
                    help="record length; shorter runs keep every launch's size (for counter collection)")
    ap.add_argument("--lsrtm-shots", type=int, default=64, help="shots of the LSRTM timing (0 skips it)")
    ap.add_argument("--lsrtm-nt", type=int, default=3000)
    ap.add_argument("--lsrtm-iters", type=int, default=3)
    ap.add_argument("--c1-calls", type=int, default=10, help="timed calls of the C1 single-shot case (0 skips it)")
    ap.add_argument("--c3-shots", type=int, default=256, help="shots of the C3 Born batch (0 skips it)")
    ap.add_argument("--c3-pair-shots", type=int, default=16)
    ap.add_argument("--c5-shots", type=int, default=2, help="shots of the C5 RTM sample (0 skips it)")
    ap.add_argument("--c5-nt", type=int, default=500)
    ap.add_argument("--snap-shots", type=int, default=4,
                    help="shots of the snapshot-storage vs recomputation comparison (0 skips it)")
    ap.add_argument("--plane-shots", type=int, default=96, help="shots of the plane-wave synthesis leg (0 skips it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 8 device-to-device copies of 1 GiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the batch of shots is one GPU's share; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import seis_amd
    from seis_amd import DeviceArray
    from seis_amd._lib import check
    if args.copy_calibration:
        n = 1 << 28  # 1 GiB of float32
        a = DeviceArray((n,), np.float32, zero=True)
        b = DeviceArray((n,), np.float32, zero=True)
        for _ in range(8):
            check(seis_amd.lib().seis_copy_f32(b.ptr, a.ptr, n))
        seis_amd.synchronize()
        log("copied 8 x %d bytes device to device" % (4 * n))
        return
    from seis_amd.batch import _layout, Grid, Shot, RTM_ISOLATE_BAND
    import ctypes

    w = rtm_workload(args.shots, args.nt)
    if not args.no_cpu and args.cpu_nt <= w["delay"]:
        # the reference's engine exits the process when the wavelet delay
        # does not fit the record
        print("--cpu-nt must exceed the wavelet delay (%d samples)" % w["delay"], file=sys.stderr)
        sys.exit(2)
    nx, nz, pml, nt = w["nx"], w["nz"], w["pml"], w["nt"]
    grid = Grid(nx, nz, pml, w["dt"], w["dx"], w["dz"])
    shots = [Shot(sx=sx, sz=w["sz"], nt=nt, ndel=0, model=0) for sx in w["sxs"]]
    ns = len(shots)
    wav1 = seis_amd.ricker(w["dt"], w["fr"], w["delay"], nt)
    arr, nw, ng, nr = _layout(grid, shots)
    g = grid.c()
    d_vtrue = DeviceArray.from_host(w["vtrue"])
    d_vmig = DeviceArray.from_host(w["vmig"])
    d_wav = DeviceArray.from_host(np.tile(wav1, ns))
    d_gz = DeviceArray.from_host(np.full(nx * ns, w["gz"], np.int32))
    d_rec = DeviceArray((nr,), np.float32, zero=True)
    d_img = DeviceArray((ns * nx * nz,), np.float32, zero=True)
    L = seis_amd.lib()
    from seis_amd._lib import check

    # observed data: forward modeling through the true model, receivers limited
    # to each shot's 480-channel split spread
    t0 = time.perf_counter()
    check(L.seis_modeling_batch(ctypes.byref(g), ns, arr, d_vtrue.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr))
    rec = d_rec.download().reshape(ns, nx, nt)
    for j, s in enumerate(shots):
        rec[j][~spread_mask(w, s.sx)] = 0.0
    d_rec.upload(rec.ravel())
    prep_s = time.perf_counter() - t0
    log("observed records modeled for %d shots in %.1f s" % (ns, prep_s))

    # The reference's backward source reconstruction diverges in the absorbing
    # band on this model (inf/NaN images by nt = 5000, the reference included);
    # SEIS_RTM_ISOLATE_BAND keeps the band of that field at zero.  Same kernels
    # and work; the image equals the reference's beyond 34 cells of the edges.
    def step():
        check(L.seis_rtm_batch_ex(ctypes.byref(g), ns, arr, d_vmig.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr, d_img.ptr,
                                  RTM_ISOLATE_BAND))

    for i in range(args.warmup):
        step()
        log("warmup step %d/%d done" % (i + 1, args.warmup))
    seis_amd.synchronize()
    seis_amd.profile_start()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()  # returns when the batch has finished on the device
        log("timed step %d/%d done at %.1f s" % (i + 1, args.steps, time.perf_counter() - t0))
    seis_amd.synchronize()
    el = time.perf_counter() - t0
    stats = seis_amd.profile_stop()

    cells = (nx + 2 * pml) * (nz + 2 * pml)
    updates = 3.0 * nt * cells * ns * args.steps
    value = updates / el / 1e9
    dom = max(stats.items(), key=lambda kv: kv[1]["total_ms"])
    achieved = dom[1]["alg_bytes"] / (dom[1]["total_ms"] * 1e-3) / 1e9
    img = d_img.download()
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el * 1e3 / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered-faulted model; records forward-modeled on the GPU before timing)",
        "config": {"workload": "rtm_batch: %d shots, 751x2301 cells @ 4 m, pml 30, nt=%d, dt=0.4 ms, "
                               "Ricker 12 Hz, 480-channel split spread @ 12 m, reference imaging condition" % (ns, nt),
                   "shots": ns, "grid": [nz, nx], "padded_cells": cells, "nt": nt,
                   "propagations_per_shot": 3, "source_band": "isolated (SEIS_RTM_ISOLATE_BAND)",
                   "strip_length": seis_amd.tuned_strip("rtm", nx + 2 * w["pml"], nz + 2 * w["pml"], ns)},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (MEASURED_TRAFFIC_96[dom[0]] * ns / 96 if dom[0] in MEASURED_TRAFFIC_96 else None),
                     "avg_launch_ms": round(dom[1]["total_ms"] / dom[1]["launches"], 4),
                     "alg_bytes_per_launch": dom[1]["alg_bytes"] / dom[1]["launches"]},
        "kernels": {k: {"launches": v["launches"], "ms": round(v["total_ms"], 2),
                        "GB/s": round(v["alg_bytes"] / max(v["total_ms"], 1e-9) / 1e6, 1)}
                    for k, v in stats.items()},
        "prep_s": round(prep_s, 2),
        "image_finite": bool(np.isfinite(img).all()),
    }
    threads = min(16, len(os.sched_getaffinity(0)))
    if args.c3_shots > 0:
        del d_rec, d_img
        out["c3_born_adjoint"] = c3_born_adjoint(args.c3_shots, args.c3_pair_shots)
    if args.c5_shots > 0:
        out["c5_sample"] = c5_sample(args.c5_shots, args.c5_nt)
    if args.c1_calls > 0:
        log("C1 single-shot modeling, %d calls per mode" % args.c1_calls)
        out["c1_single_shot"] = c1_single_shot(args.c1_calls, not args.no_cpu, threads)
    if args.snap_shots > 0:
        out["snapshot_vs_recompute"] = snapshot_vs_recompute(args.snap_shots, args.nt)
    if args.plane_shots > 0:
        out["plane_prepare"] = plane_prepare_leg(args.plane_shots, 480, 5000, 8, not args.no_cpu)
    if args.lsrtm_shots > 0:
        out["lsrtm"] = lsrtm_bench(args.lsrtm_shots, args.lsrtm_nt, args.lsrtm_iters, 2, 3)
    if not args.no_cpu:
        log("timing the CPU baseline on %d threads" % threads)
        cb = cpu_baseline(w, args.cpu_nt, threads)
        cb["gpu_over_cpu"] = round(value / cb["value"], 1)
        out["cpu_baseline"] = cb
        ls = out.get("lsrtm")
        if ls and "s_per_cg_iter_bf16_marginal" in ls:
            # the reference's iteration is one Born (2 propagations) and one
            # RTM (3) per shot; priced at the measured reference rate
            cpu_iter = 5.0 * ls["cells_per_propagation"] / (cb["value"] * 1e9)
            ls["s_per_iter_cpu_reference_est"] = round(cpu_iter, 1)
            ls["speedup_vs_cpu_reference"] = round(cpu_iter / ls["s_per_cg_iter_bf16_marginal"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
