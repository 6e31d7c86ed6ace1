"""Benchmark of the MI355X acoustic finite-difference engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--shots-per-step S]

Headline workload: the large batched RTM survey.  2000 x 8000 cells @ 5 m,
pml 30, a seeded layered-and-faulted model (1500-5500 m/s, 40 dipping layers,
3 normal faults) with a 4500 m/s salt body inside a random smooth closed
contour; migration velocity 20 cells smoother.  512 shots between columns 500
and 7500, each recorded by 1000 receivers (every column from sx - 500 to
sx + 499, depth 10 m), dt = 0.5 ms, nt = 6000, Ricker 12 Hz.  Each shot runs
on its own window of the model, its spread widened by 50 columns on each side
(1100 columns, 1160 x 2060 padded cells), as the reference's getVel windows
every shot (lsrtm/inversion.cpp:552-598).  The records are forward-modeled
from the true model on the GPU before timing, so every input is resident in
HBM when the timed region starts.

A step is the reference-semantics RTM of `--shots-per-step` consecutive shots
of the survey at the full record length: forward propagation saving the
boundary strips, then source reconstruction, receiver back-propagation and
the imaging condition with FaQi, i.e. 3 wavefield propagations per shot, all
shots of the step in the same launches.  Step k takes shots S k .. S k + S - 1
(mod 512): with the default 26 shots per step the 20 timed steps image the
whole survey once.

Prints one JSON line: Gcell-updates/s over the timed steps, the dominant
kernel's HBM roofline fraction (HIP-event durations on the launch stream and
algorithmic bytes), the reference's CPU path timed on this host on a bounded
sample of the same workload, and the other workloads as extra objects.
"""
import argparse
import contextlib
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s (batched shots) + LSRTM s/CG-iter vs plain FP32 at stated error"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the headline's kernels (26 shots per launch, strip
# length 16, the tuner's usual choice on this launch), from rocprofv3 FETCH_SIZE
# / WRITE_SIZE calibrated against a 1 GiB copy; the commands and counter values
# are in profiles/c5_rtm_hbm_traffic.md.  step_rtm_reverse is the mean over the
# reverse launches (each but the first also images half of the interior
# columns), rtm_image the mean of the two imaging launches after the loop.
# Scaled by the shots of a launch.
MEASURED_TRAFFIC = {"shots": 26, "strip_length": 16,
                    "bytes": {"step_rtm_reverse": 3044062197, "step_rtm_forward": 1204190945,
                              "rtm_image": 601327872}}


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


def c5_survey(nt=6000):
    """The headline survey: model, migration model and per-shot windows."""
    from seis_amd import models
    nx, nz, pml, dx, dt, fr = 8000, 2000, 30, 5.0, 0.0005, 12.0
    t0 = time.perf_counter()
    v = models.salt_body(models.layered_faulted(nx, nz, dx, dx, seed=1005), dx, dx, seed=1005)
    vm = models.smooth_model(v, 20.0)
    log("C5 model built in %.1f s" % (time.perf_counter() - t0))
    sxs = [500 + int(round(i * 7000 / 511)) for i in range(512)]
    wins = [models.spread_window(nx, sx, sx - 500, sx + 499, 50) for sx in sxs]
    return dict(nx=nx, nz=nz, pml=pml, dx=dx, dt=dt, fr=fr, nt=nt, delay=int(round(1.0 / (fr * dt))), v=v, vm=vm,
                sxs=sxs, wins=wins, sz=2, gz=2)


def window_shots(grid, sxs, wins, nts, sz=2):
    """Shots on their windows [x0, x0 + nxw) of one model (views: model offset
    x0 nz), the packed arrays' offsets in shot order."""
    from seis_amd.batch import Shot, _layout
    shots = [Shot(sx=sx - x0, sz=sz, nt=int(nt), nx=n, x0=x0) for sx, (x0, n), nt in zip(sxs, wins, nts)]
    arr, nw, ng, nr = _layout(grid, shots, views=True)
    return shots, arr, nw, ng, nr


def sub_array(arr, idx):
    """seis_shot array of the shots idx of arr (offsets kept)."""
    from seis_amd._lib import seis_shot
    out = (seis_shot * len(idx))()
    for j, i in enumerate(idx):
        out[j] = arr[i]
    return out


def cpu_info():
    """lscpu facts of this host and the threads the CPU baseline may use."""
    info = {"model": host_cpu_model()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
            elif k == "CPU(s)":
                info["cpus"] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    info["affinity_cpus"] = aff
    info["threads_used"] = min(aff, int(omp)) if omp and omp.isdigit() else aff
    return info


def cpu_baseline(w, nt_sample, threads):
    """Reference RTM (the compiled reference engine if present, else the
    oracle's restatement) of one windowed shot of the headline survey with a
    shortened record, on `threads` host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    kind = "reference" if oracle_lib.have_reference() else "port"
    lib = oracle_lib.Reference() if kind == "reference" else oracle_lib.Oracle()
    lib.set_threads(threads)
    j = len(w["sxs"]) // 2
    x0, n = w["wins"][j]
    sx = w["sxs"][j] - x0
    wav = lib.ricker(w["dt"], w["fr"], w["delay"], nt_sample)
    gz = np.full(n, w["gz"], np.int32)
    rec = np.random.default_rng(5).standard_normal((n, nt_sample)).astype(np.float32)
    vel = np.ascontiguousarray(w["vm"][x0:x0 + n])
    t0 = time.perf_counter()
    with stdout_to_stderr():
        lib.rtm(w["dt"], w["dx"], w["dx"], nt_sample, 0, n, w["nz"], w["pml"], sx, w["sz"], gz, wav, vel, rec)
    el = time.perf_counter() - t0
    cells = (n + 2 * w["pml"]) * (w["nz"] + 2 * w["pml"])
    updates = 3.0 * nt_sample * cells
    return dict(value=updates / el / 1e9, unit="Gcell-updates/s", cores=threads, kind=kind,
                sample="RTM of one shot of the same survey on its %d-column window with nt=%d (%.1f s of CPU work)"
                       % (n, nt_sample, el))


def c2_rtm_batch(nt=5000):
    """The 751 x 2301 @ 4 m RTM workload: 96 shots every 96 m, 480 receivers
    at 12 m (every third column, split spread), each shot on its window
    (spread + 50 columns), dt = 0.4 ms, nt = 5000, Ricker 12 Hz, one batch."""
    import seis_amd
    from seis_amd import DeviceArray, models
    from seis_amd.batch import Grid, RTM_ISOLATE_BAND
    from seis_amd._lib import check
    L = seis_amd.lib()
    nx, nz, pml, dx, dt, fr = 2301, 751, 30, 4.0, 0.0004, 12.0
    vtrue = models.layered_faulted(nx, nz, dx, dx, seed=1001)
    vmig = models.smooth_model(vtrue, 20.0)
    delay = int(round(1.0 / (fr * dt)))
    sxs = [12 + 24 * i for i in range(96)]
    wins = []
    for sx in sxs:
        lo, hi = max(0, sx - 720), min(nx - 1, sx + 720)
        wins.append(models.spread_window(nx, sx, lo, hi, 50))
    grid = Grid(nx, nz, pml, dt, dx, dx)
    g = grid.c()
    shots, arr, nw, ng, nr = window_shots(grid, sxs, wins, [nt] * 96)
    d_vt, d_vm = DeviceArray.from_host(vtrue), DeviceArray.from_host(vmig)
    d_wav = DeviceArray.from_host(np.tile(seis_amd.ricker(dt, fr, delay, nt), 96))
    d_gz = DeviceArray.from_host(np.full(ng, 2, np.int32))
    d_rec = DeviceArray((nr,), np.float32, zero=True)
    nimg = sum(n for _, n in wins) * nz
    d_img = DeviceArray((nimg,), np.float32, zero=True)
    check(L.seis_modeling_batch(ctypes.byref(g), 96, arr, d_vt.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr))
    # receivers: every third column of the spread; the pad columns record nothing
    rec = d_rec.download()
    off = 0
    for sx, (x0, n) in zip(sxs, wins):
        r = rec[off:off + n * nt].reshape(n, nt)
        cols = np.arange(x0, x0 + n)
        r[~((np.abs(cols - sx) <= 720) & ((cols - sx) % 3 == 0))] = 0.0
        off += n * nt
    d_rec.upload(rec)
    del rec
    seis_amd.synchronize()
    seis_amd.profile_start()
    t0 = time.perf_counter()
    check(L.seis_rtm_batch_ex(ctypes.byref(g), 96, arr, d_vm.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr, d_img.ptr,
                              RTM_ISOLATE_BAND))
    seis_amd.synchronize()
    el = time.perf_counter() - t0
    st = seis_amd.profile_stop()
    cells = sum((n + 2 * pml) * (nz + 2 * pml) for _, n in wins)
    img = d_img.download()
    log("C2: RTM of 96 windowed shots in %.1f s" % el)
    return {"workload": "rtm_batch: 96 shots every 96 m on 751x2301 cells @ 4 m, 480 receivers @ 12 m per shot, "
                        "each shot on its window (spread + 50 columns, %d..%d columns), pml 30, nt=%d, dt=0.4 ms"
                        % (min(n for _, n in wins), max(n for _, n in wins), nt),
            "s": round(el, 3), "s_per_shot": round(el / 96, 4),
            "gcell_updates_per_s": round(3.0 * nt * cells / el / 1e9, 1),
            "kernels": {k: {"launches": v["launches"], "ms": round(v["total_ms"], 1),
                            "GB/s": round(v["alg_bytes"] / max(v["total_ms"], 1e-9) / 1e6, 1)} for k, v in st.items()},
            "image_finite": bool(np.isfinite(img).all())}


def c3_born_adjoint(nshots, pair_shots, seed=1003):
    """Batched Born modeling with uneven shots: 500 x 1600 cells @ 5 m,
    pml 30, dt = 0.5 ms, record lengths drawn in 2000..4000 and spreads of
    200..1200 receivers (every column, centred on the source), each shot on
    its window (spread + 50 columns), a sparse spike perturbation (2 %
    density, N(0,1)), every shot in one batch.  Then the BF16-snapshot Born /
    exact-adjoint pair on `pair_shots` of them with the dot-product test, and
    the same test in FP64 on 4."""
    import seis_amd
    from seis_amd import DeviceArray, models
    from seis_amd.batch import Grid
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
    wins = [models.spread_window(nx, int(sx), int(sx) - int(ng) // 2, int(sx) + (int(ng) - 1) // 2, 50)
            for sx, ng in zip(sxs, ngs)]
    grid = Grid(nx, nz, pml, dt, dx, dx)
    g = grid.c()
    shots, arr, nw, ng_off, nr = window_shots(grid, [int(s) for s in sxs], wins, nts)
    wav = np.concatenate([seis_amd.ricker(dt, fr, delay, int(n)) for n in nts])
    d_v, d_dv = DeviceArray.from_host(v), DeviceArray.from_host(dv)
    d_wav = DeviceArray.from_host(wav)
    d_gz = DeviceArray.from_host(np.full(ng_off, 2, np.int32))
    d_rec = DeviceArray((nr,), np.float32, zero=True)
    cells = [(n + 2 * pml) * (nz + 2 * pml) for _, n in wins]
    out = {"workload": "born_batch: %d shots in one batch, 500x1600 cells @ 5 m, pml 30, dt=0.5 ms, nt in "
                       "2000..4000, 200..1200 receivers, each shot on its window (%d..%d columns), 2%% sparse spike "
                       "perturbation" % (nshots, min(n for _, n in wins), max(n for _, n in wins)),
           "nt_range": [int(nts.min()), int(nts.max())]}
    # (the strip-length trial of this configuration runs inside: 12 launches)
    t0 = time.perf_counter()
    check(L.seis_born_batch(ctypes.byref(g), nshots, arr, d_v.ptr, d_dv.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr))
    seis_amd.synchronize()
    el = time.perf_counter() - t0
    out["born_s"] = round(el, 3)
    out["s_per_shot"] = round(el / nshots, 4)
    out["born_gcell_updates_per_s"] = round(2.0 * sum(c * float(n) for c, n in zip(cells, nts)) / el / 1e9, 1)
    log("C3: Born batch of %d windowed shots in %.1f s" % (nshots, el))
    del d_rec
    if pair_shots <= 0:
        return out
    # the Born / adjoint pair with BF16 snapshots on a subset (receivers: the
    # spread columns of each window)
    from seis_amd.lsrtm import LsrtmPlan, SNAP_BF16
    sub = list(range(pair_shots))
    sshots = [shots[i] for i in sub]
    mask = []
    for i in sub:
        x0, n = wins[i]
        cols = np.arange(x0, x0 + n)
        m = np.zeros((n, int(nts[i])), np.float32)
        m[np.abs(cols - sxs[i]) <= int(ngs[i]) // 2] = 1.0
        mask.append(m.ravel())
    mask = np.concatenate(mask)
    t0 = time.perf_counter()
    plan = LsrtmPlan(grid, sshots, v, [seis_amd.ricker(dt, fr, delay, s.nt) for s in sshots],
                     [np.full(s.nx, 2, np.int32) for s in sshots], snapshot=SNAP_BF16)
    plan.set_mask(mask)
    seis_amd.synchronize()
    out["pair_snapshot_s"] = round(time.perf_counter() - t0, 3)
    r = DeviceArray.from_host((rng.standard_normal(plan.nrec) * mask).astype(np.float32))
    t0 = time.perf_counter()
    lm = plan.born(d_dv)
    t1 = time.perf_counter()
    ltr = plan.adjoint(r)
    t2 = time.perf_counter()
    rr = r.download()
    a, b = float(np.dot(lm.astype(np.float64), rr)), float(np.sum(dv.astype(np.float64) * ltr))
    rel = abs(a - b) / max(abs(a), abs(b), 1e-300)
    reln = abs(a - b) / max(float(np.linalg.norm(lm) * np.linalg.norm(rr)), 1e-300)
    plan.close()
    pair_cells = sum(cells[i] * float(nts[i]) for i in sub)
    out["pair_shots"] = pair_shots
    out["pair_born_gcell_updates_per_s"] = round(pair_cells / (t1 - t0) / 1e9, 1)
    out["pair_adjoint_gcell_updates_per_s"] = round(pair_cells / (t2 - t1) / 1e9, 1)
    # |<L m, r> - <m, L^T r>| relative to the value and to ||L m|| ||r||: a
    # random r makes the two products cancel to ~1/sqrt(n) of the norms, so
    # FP32 rounding shows up far larger against the value
    out["dot_test_fp32_rel"] = float("%.3e" % rel)
    out["dot_test_fp32_rel_to_norms"] = float("%.3e" % reln)
    sub64 = sshots[:4]
    p64 = LsrtmPlan(grid, sub64, v.astype(np.float64), [seis_amd.ricker(dt, fr, delay, s.nt) for s in sub64],
                    [np.full(s.nx, 2, np.int32) for s in sub64], fp64=True)
    m64 = dv.astype(np.float64)
    r64 = np.random.default_rng(seed + 1).standard_normal(p64.nrec) * mask[:p64.nrec]
    a, b = float(np.dot(p64.born(m64), r64)), float(np.sum(m64 * p64.adjoint(r64)))
    p64.close()
    out["dot_test_fp64_rel"] = float("%.3e" % (abs(a - b) / max(abs(a), abs(b), 1e-300)))
    out["dot_test_fp64_shots"] = 4
    out["dot_test_note"] = ("FP32: FP32 propagation, BF16 snapshots, FP64 inner products; FP64: every step in FP64 "
                            "(the stated bar, <= 1e-6)")
    log("C3: pair on %d shots; dot test FP32 %.2e (%.2e of the norms), FP64 %s" % (pair_shots, rel, reln,
                                                                                   out["dot_test_fp64_rel"]))
    return out
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


def lsrtm_bench(iters, err_shots, plain_iters, xcorr_iters, single_shots=16):
    """Least-squares RTM on the 376 x 1151 @ 8 m workload (the RTM model at
    half resolution, nt = 3000, dt = 0.8 ms): all 128 shots, 400 receivers
    each (every column from sx - 200 to sx + 199), each shot on its window
    (spread + 40 columns); observed data = Born records of the true
    perturbation + 5 % Gaussian noise.  CGLS with BF16 snapshots and the
    exact adjoint for `iters` iterations, the same with the cross-correlation
    adjoint, the plain FP32 GPU path (per iteration one Born batch and one RTM
    batch of the same shots, as the reference's forward_MPI + adjoint_MPI),
    the same operators one shot per call (the plain FP32 single-shot GPU path,
    on `single_shots` of the shots, scaled to all), and, on `err_shots` of the
    shots, the image of every iteration with
    FP32 / BF16 / FP16 snapshots against the FP64 plan."""
    import seis_amd
    from seis_amd import DeviceArray, models
    from seis_amd.batch import Grid, RTM_ISOLATE_BAND
    from seis_amd.lsrtm import LsrtmPlan, SNAP_BF16, SNAP_FP16, SNAP_FP32
    from seis_amd._lib import check
    L = seis_amd.lib()
    nx, nz, pml, dx, dt, fr, nt = 1151, 376, 30, 8.0, 0.0008, 12.0, 3000
    v = np.ascontiguousarray(models.layered_faulted(2301, 751, 4.0, 4.0, seed=1001)[::2, ::2])
    v0 = models.smooth_model(v, 10.0)
    dv_true = (v - v0).astype(np.float32)
    delay = int(round(1.0 / (fr * dt)))
    grid = Grid(nx, nz, pml, dt, dx, dx)
    g = grid.c()
    all_sx = [4 + 9 * i for i in range(128)]
    wav1 = seis_amd.ricker(dt, fr, delay, nt)

    def survey(sxs):
        wins = [models.spread_window(nx, sx, max(0, sx - 200), min(nx - 1, sx + 199), 40) for sx in sxs]
        shots, arr, nw, ng, nr = window_shots(grid, sxs, wins, [nt] * len(sxs))
        mask = []
        for sx, (x0, n) in zip(sxs, wins):
            cols = np.arange(x0, x0 + n)
            m = np.zeros((n, nt), np.float32)
            m[(cols >= sx - 200) & (cols <= sx + 199)] = 1.0
            mask.append(m.ravel())
        return shots, arr, ng, nr, np.concatenate(mask), wins

    def born_data(shots, arr, ng, nr, mask, noise_seed):
        d_v0, d_dv = DeviceArray.from_host(v0), DeviceArray.from_host(dv_true)
        d_w = DeviceArray.from_host(np.tile(wav1, len(shots)))
        d_g = DeviceArray.from_host(np.full(ng, 2, np.int32))
        d_rec = DeviceArray((nr,), np.float32, zero=True)
        check(L.seis_born_batch(ctypes.byref(g), len(shots), arr, d_v0.ptr, d_dv.ptr, d_w.ptr, d_g.ptr, d_rec.ptr))
        d = d_rec.download() * mask
        rms = float(np.sqrt(np.mean(d[mask > 0].astype(np.float64) ** 2)))
        d += (0.05 * rms * np.random.default_rng(noise_seed).standard_normal(d.shape)).astype(np.float32) * mask
        return d

    shots, arr, ng, nr, mask, wins = survey(all_sx)
    ns = len(shots)
    cells = float(sum((n + 2 * pml) * (nz + 2 * pml) for _, n in wins)) * nt  # cell-updates per propagation
    out = {"workload": "lsrtm_cg: 128 shots, 376x1151 cells @ 8 m, pml 30, nt=%d, dt=0.8 ms, Ricker 12 Hz, 400 "
                       "receivers per shot, each shot on its window (%d..%d columns), Born data of the true "
                       "perturbation + 5%% noise" % (nt, min(n for _, n in wins), max(n for _, n in wins)),
           "cells_per_propagation": cells}
    dobs = born_data(shots, arr, ng, nr, mask, 11)
    log("lsrtm: observed data for %d shots ready" % ns)
    t0 = time.perf_counter()
    plan = LsrtmPlan(grid, shots, v0, [wav1] * ns, [np.full(s.nx, 2, np.int32) for s in shots], snapshot=SNAP_BF16)
    plan.set_mask(mask)
    seis_amd.synchronize()
    out["snapshot_s"] = round(time.perf_counter() - t0, 3)
    out["snapshot_gb"] = round(2.0 * sum(n for _, n in wins) * nz * nt / 1e9, 1)
    d_obs = DeviceArray.from_host(dobs)
    # the timed CG runs without the launch profiler (its event pair around
    # every launch adds about 5 % here, as the plain path below has none);
    # the kernel breakdown comes from a separate 2-iteration profiled run
    t0 = time.perf_counter()
    img, hist = plan.cg(d_obs, iters)
    el = time.perf_counter() - t0
    done = len(hist) - 1
    out["cg_iterations"] = done
    out["s_per_cg_iter_bf16"] = round(el / max(done, 1), 3)
    out["residual_rel"] = [round(float(np.sqrt(x / hist[0, 0])), 5) for x in hist[:, 0]]
    log("lsrtm: %d CG iterations (BF16 snapshots, exact adjoint) in %.1f s" % (done, el))
    # CGLS applies L^T once before its first iteration (s0 = L^T d); one L^T
    # alone, timed the same way, gives the steady per-iteration cost (one L and
    # one L^T plus the vector updates), the same work as one iteration of the
    # plain path below
    t0 = time.perf_counter()
    plan.adjoint(d_obs)
    el_adj = time.perf_counter() - t0
    out["s_adjoint_bf16"] = round(el_adj, 3)
    out["s_per_cg_iter_bf16_steady"] = round((el - el_adj) / max(done, 1), 3)
    seis_amd.profile_start()
    t0 = time.perf_counter()
    plan.cg(d_obs, 2)
    el_p = time.perf_counter() - t0
    st = seis_amd.profile_stop()
    out["kernels"] = {k: {"launches": v_["launches"], "ms": round(v_["total_ms"], 1),
                          "GB/s": round(v_["alg_bytes"] / max(v_["total_ms"], 1e-9) / 1e6, 1)} for k, v_ in st.items()}
    out["kernels_run"] = "2 CG iterations under the launch profiler, %.2f s" % el_p
    if xcorr_iters > 0:
        plan.set_adjoint(1)
        t0 = time.perf_counter()
        img_x, hist_x = plan.cg(d_obs, xcorr_iters)
        el_x = time.perf_counter() - t0
        out["s_per_cg_iter_bf16_xcorr"] = round(el_x / max(len(hist_x) - 1, 1), 3)
        out["xcorr_iterations"] = len(hist_x) - 1
        log("lsrtm: %d CG iterations with the cross-correlation adjoint in %.1f s" % (len(hist_x) - 1, el_x))
    plan.close()
    # plain FP32 GPU path, per iteration: Born batch of the current image and
    # RTM batch of the residual, over the same windows
    d_v0 = DeviceArray.from_host(v0)
    d_w = DeviceArray.from_host(np.tile(wav1, ns))
    d_g = DeviceArray.from_host(np.full(ng, 2, np.int32))
    d_pert = DeviceArray.from_host(img)
    d_rec = DeviceArray((nr,), np.float32, zero=True)
    d_img = DeviceArray((sum(n for _, n in wins) * nz,), np.float32, zero=True)
    t0 = time.perf_counter()
    for _ in range(plain_iters):
        check(L.seis_born_batch(ctypes.byref(g), ns, arr, d_v0.ptr, d_pert.ptr, d_w.ptr, d_g.ptr, d_rec.ptr))
        check(L.seis_rtm_batch_ex(ctypes.byref(g), ns, arr, d_v0.ptr, d_w.ptr, d_g.ptr, d_obs.ptr, d_img.ptr,
                                  RTM_ISOLATE_BAND))
    seis_amd.synchronize()
    plain = (time.perf_counter() - t0) / plain_iters
    # the plain FP32 single-shot GPU path: the same Born and RTM operators with
    # one shot per call, shot after shot (one iteration's work for
    # `single_shots` shots spread over the survey, scaled to all of them)
    single = None
    if single_shots > 0:
        idx = list(range(0, ns, max(1, ns // single_shots)))[:single_shots]
        t0 = time.perf_counter()
        for i in idx:
            one = sub_array(arr, [i])
            check(L.seis_born_batch(ctypes.byref(g), 1, one, d_v0.ptr, d_pert.ptr, d_w.ptr, d_g.ptr, d_rec.ptr))
            check(L.seis_rtm_batch_ex(ctypes.byref(g), 1, one, d_v0.ptr, d_w.ptr, d_g.ptr, d_obs.ptr, d_img.ptr,
                                      RTM_ISOLATE_BAND))
        seis_amd.synchronize()
        single = (time.perf_counter() - t0) * ns / len(idx)
    del d_img, d_rec, d_pert, d_obs
    out["s_per_iter_plain_fp32"] = round(plain, 3)
    out["plain_iterations_timed"] = plain_iters
    out["speedup_vs_plain_fp32"] = round(plain / out["s_per_cg_iter_bf16"], 2)
    out["speedup_vs_plain_fp32_steady"] = round(plain / out["s_per_cg_iter_bf16_steady"], 2)
    if "s_per_cg_iter_bf16_xcorr" in out:
        out["speedup_vs_plain_fp32_xcorr"] = round(plain / out["s_per_cg_iter_bf16_xcorr"], 2)
    log("lsrtm: plain FP32 path, %.2f s per iteration (Born + RTM batches)" % plain)
    if single is not None:
        out["s_per_iter_plain_fp32_single_shot"] = round(single, 3)
        out["plain_single_shot_sample"] = "%d of the %d shots, one Born and one RTM call each, scaled to all" % (
            len(idx), ns)
        out["speedup_vs_plain_fp32_single_shot"] = round(single / out["s_per_cg_iter_bf16"], 2)
        log("lsrtm: plain FP32 single-shot path, %.2f s per iteration" % single)
    # --- image error per iteration against the FP64 plan, on a subset
    if err_shots > 0:
        esx = all_sx[4::128 // err_shots][:err_shots]
        eshots, earr, eng, enr, emask, ewins = survey(esx)
        ed = born_data(eshots, earr, eng, enr, emask, 12)
        egz = [np.full(s.nx, 2, np.int32) for s in eshots]
        its = {}
        hists = {}
        for name, kind in (("fp64", None), ("fp32", SNAP_FP32), ("bf16", SNAP_BF16), ("fp16", SNAP_FP16)):
            if kind is None:
                p = LsrtmPlan(grid, eshots, v0.astype(np.float64), [wav1.astype(np.float64)] * len(eshots), egz,
                              fp64=True)
                dd = ed.astype(np.float64)
                p.set_mask(emask.astype(np.float64))
            else:
                p = LsrtmPlan(grid, eshots, v0, [wav1] * len(eshots), egz, snapshot=kind)
                dd = ed
                p.set_mask(emask)
            _, h, it = p.cg(dd, iters, iterates=True)
            p.close()
            its[name] = it.astype(np.float64)
            hists[name] = h
        ref = its["fp64"]
        n = min(len(ref), *(len(its[k]) for k in ("fp32", "bf16", "fp16")))
        err = {k: [float("%.3e" % (np.linalg.norm(its[k][i] - ref[i]) / max(np.linalg.norm(ref[i]), 1e-300)))
                   for i in range(n)] for k in ("fp32", "bf16", "fp16")}
        out["image_rel_l2_vs_fp64_per_iter"] = err
        out["image_rel_l2_vs_fp64_final"] = {k: e[-1] for k, e in err.items()}
        out["residual_rel_fp64_per_iter"] = [round(float(np.sqrt(x / hists["fp64"][0, 0])), 5)
                                             for x in hists["fp64"][:, 0]]
        out["error_case"] = "%d of the 128 shots, %d CG iterations, same grid, windows and nt" % (len(eshots), n)
        log("lsrtm: final image errors vs FP64 %s" % out["image_rel_l2_vs_fp64_final"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots-per-step", type=int, default=26)
    ap.add_argument("--nt", type=int, default=6000,
                    help="record length of the headline survey; shorter runs keep every launch's size "
                         "(for counter collection)")
    ap.add_argument("--cpu-nt", type=int, default=600)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lsrtm-iters", type=int, default=20, help="CG iterations of the LSRTM leg (0 skips it)")
    ap.add_argument("--lsrtm-xcorr-iters", type=int, default=20)
    ap.add_argument("--lsrtm-plain-iters", type=int, default=1)
    ap.add_argument("--lsrtm-err-shots", type=int, default=8)
    ap.add_argument("--lsrtm-single-shots", type=int, default=16,
                    help="shots of the plain FP32 single-shot timing (scaled to all 128; 0 skips it)")
    ap.add_argument("--c1-calls", type=int, default=10, help="timed calls of the C1 single-shot case (0 skips it)")
    ap.add_argument("--c2", type=int, default=1, help="run the 96-shot RTM batch on the 751 x 2301 grid (0 skips)")
    ap.add_argument("--c3-shots", type=int, default=256, help="shots of the C3 Born batch (0 skips it)")
    ap.add_argument("--c3-pair-shots", type=int, default=16)
    ap.add_argument("--snap-shots", type=int, default=4,
                    help="shots of the snapshot-storage vs recomputation comparison (0 skips it)")
    ap.add_argument("--plane-shots", type=int, default=96, help="shots of the plane-wave synthesis leg (0 skips it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 8 device-to-device copies of 1 GiB (HBM counter calibration)")
    ap.add_argument("--no-tune", action="store_true",
                    help="default strip length (16 columns) instead of the tuner's choice, for counter runs, "
                         "whose serialised launches would mislead the tuner's timing")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the batch of shots is one GPU's share; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import seis_amd
    from seis_amd import DeviceArray
    from seis_amd._lib import check
    if args.no_tune:
        seis_amd.set_tuning(False)
    if args.copy_calibration:
        n = 1 << 28  # 1 GiB of float32
        a = DeviceArray((n,), np.float32, zero=True)
        b = DeviceArray((n,), np.float32, zero=True)
        for _ in range(8):
            check(seis_amd.lib().seis_copy_f32(b.ptr, a.ptr, n))
        seis_amd.synchronize()
        log("copied 8 x %d bytes device to device" % (4 * n))
        return
    from seis_amd.batch import Grid, RTM_ISOLATE_BAND

    w = c5_survey(args.nt)
    if not args.no_cpu and args.cpu_nt <= w["delay"]:
        # the reference's engine exits the process when the wavelet delay
        # does not fit the record
        print("--cpu-nt must exceed the wavelet delay (%d samples)" % w["delay"], file=sys.stderr)
        sys.exit(2)
    nx, nz, pml, nt = w["nx"], w["nz"], w["pml"], w["nt"]
    grid = Grid(nx, nz, pml, w["dt"], w["dx"], w["dx"])
    g = grid.c()
    ns = len(w["sxs"])
    shots, arr, nw, ng, nr = window_shots(grid, w["sxs"], w["wins"], [nt] * ns, w["sz"])
    wav1 = seis_amd.ricker(w["dt"], w["fr"], w["delay"], nt)
    L = seis_amd.lib()
    d_vtrue = DeviceArray.from_host(w["v"])
    d_vmig = DeviceArray.from_host(w["vm"])
    d_wav = DeviceArray.from_host(np.tile(wav1, ns))
    d_gz = DeviceArray.from_host(np.full(ng, w["gz"], np.int32))
    d_rec = DeviceArray((nr,), np.float32, zero=True)
    # observed data: forward modeling of every shot through the true model
    t0 = time.perf_counter()
    check(L.seis_modeling_batch(ctypes.byref(g), ns, arr, d_vtrue.ptr, d_wav.ptr, d_gz.ptr, d_rec.ptr))
    seis_amd.synchronize()
    prep_s = time.perf_counter() - t0
    log("observed records modeled for %d shots in %.1f s" % (ns, prep_s))
    del d_vtrue
    S = args.shots_per_step
    cells_of = [(n + 2 * pml) * (nz + 2 * pml) for _, n in w["wins"]]
    img_elems = max(n for _, n in w["wins"]) * nz * S
    d_img = DeviceArray((img_elems,), np.float32, zero=True)

    # The reference's backward source reconstruction diverges in the absorbing
    # band on long records in fast media (inf/NaN images, the reference
    # included); SEIS_RTM_ISOLATE_BAND keeps the band of that field at zero.
    # Same kernels and work; the image equals the reference's beyond 34 cells
    # of the edges.
    def step(k):
        idx = [(k * S + j) % ns for j in range(S)]
        check(L.seis_rtm_batch_ex(ctypes.byref(g), S, sub_array(arr, idx), d_vmig.ptr, d_wav.ptr, d_gz.ptr,
                                  d_rec.ptr, d_img.ptr, RTM_ISOLATE_BAND))
        return 3.0 * nt * sum(cells_of[i] for i in idx)

    for i in range(args.warmup):
        step(i)
        log("warmup step %d/%d done" % (i + 1, args.warmup))
    seis_amd.synchronize()
    seis_amd.profile_start()
    t0 = time.perf_counter()
    updates = 0.0
    for i in range(args.steps):
        updates += step(i)  # returns when the step's shots have finished on the device
        log("timed step %d/%d done at %.1f s" % (i + 1, args.steps, time.perf_counter() - t0))
    seis_amd.synchronize()
    el = time.perf_counter() - t0
    stats = seis_amd.profile_stop()
    value = updates / el / 1e9
    dom = max(stats.items(), key=lambda kv: kv[1]["total_ms"])
    achieved = dom[1]["alg_bytes"] / (dom[1]["total_ms"] * 1e-3) / 1e9
    img = d_img.download()
    s_per_shot = el / (args.steps * S)
    traffic = MEASURED_TRAFFIC["bytes"].get(dom[0])
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
        "data": "synthetic (seeded layered-faulted model with a salt body; records forward-modeled on the GPU "
                "before timing)",
        "config": {"workload": "rtm_survey (the large batched RTM/LSRTM stress case): 512 shots, 2000x8000 cells "
                               "@ 5 m with a salt body, pml 30, nt=%d, dt=0.5 ms, Ricker 12 Hz, 1000 receivers "
                               "per shot, each shot on its 1100-column window, %d shots per step (step k images "
                               "shots %dk..%dk+%d mod 512), reference imaging condition" % (nt, S, S, S, S - 1),
                   "shots": ns, "shots_per_step": S, "grid": [nz, nx], "window_padded_cells": cells_of[0],
                   "nt": nt, "propagations_per_shot": 3, "source_band": "isolated (SEIS_RTM_ISOLATE_BAND)",
                   "strip_length": seis_amd.tuned_strip("rtm", max(n for _, n in w["wins"]) + 2 * pml,
                                                        nz + 2 * pml, S)},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic * S / MEASURED_TRAFFIC["shots"] if traffic else None,
                     "traffic_strip_length": MEASURED_TRAFFIC["strip_length"],
                     "avg_launch_ms": round(dom[1]["total_ms"] / dom[1]["launches"], 4),
                     "alg_bytes_per_launch": dom[1]["alg_bytes"] / dom[1]["launches"]},
        "kernels": {k: {"launches": v["launches"], "ms": round(v["total_ms"], 2),
                        "GB/s": round(v["alg_bytes"] / max(v["total_ms"], 1e-9) / 1e6, 1)}
                    for k, v in stats.items()},
        "survey": {"s_per_shot": round(s_per_shot, 4), "s_for_512_shots": round(s_per_shot * ns, 1),
                   "shot_rtms_timed": args.steps * S},
        "prep_s": round(prep_s, 2),
        "image_finite": bool(np.isfinite(img).all()),
    }
    del d_rec, d_img, d_wav, d_gz, d_vmig
    ci = cpu_info()
    threads = ci["threads_used"]
    if args.c2:
        out["c2_rtm_batch"] = c2_rtm_batch()
    if args.c3_shots > 0:
        out["c3_born_adjoint"] = c3_born_adjoint(args.c3_shots, args.c3_pair_shots)
    if args.lsrtm_iters > 0:
        out["lsrtm"] = lsrtm_bench(args.lsrtm_iters, args.lsrtm_err_shots, args.lsrtm_plain_iters,
                                   args.lsrtm_xcorr_iters, args.lsrtm_single_shots)
    if args.c1_calls > 0:
        log("C1 single-shot modeling, %d calls per mode" % args.c1_calls)
        out["c1_single_shot"] = c1_single_shot(args.c1_calls, not args.no_cpu, threads)
    if args.snap_shots > 0:
        out["snapshot_vs_recompute"] = snapshot_vs_recompute(args.snap_shots, 5000)
    if args.plane_shots > 0:
        out["plane_prepare"] = plane_prepare_leg(args.plane_shots, 480, 5000, 8, not args.no_cpu)
    if not args.no_cpu:
        log("timing the CPU baseline on %d threads" % threads)
        cb = cpu_baseline(w, args.cpu_nt, threads)
        cb["gpu_over_cpu"] = round(value / cb["value"], 1)
        cb["cpu"] = ci
        out["cpu_baseline"] = cb
        ls = out.get("lsrtm")
        if ls:
            # the reference's iteration is one Born (2 propagations) and one
            # RTM (3) per shot; priced at the measured reference rate
            cpu_iter = 5.0 * ls["cells_per_propagation"] / (cb["value"] * 1e9)
            ls["s_per_iter_cpu_reference_est"] = round(cpu_iter, 1)
            ls["speedup_vs_cpu_reference"] = round(cpu_iter / ls["s_per_cg_iter_bf16"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
