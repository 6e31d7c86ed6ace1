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
MEASURED_TRAFFIC_96 = {"step_rtm_forward": 3.989e9, "step_rtm_reverse": 8.761e9, "rtm_image": 2.821e9}


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
    lib.rtm(w["dt"], w["dx"], w["dz"], nt_sample, 0, nx, nz, pml, sx, w["sz"], gz, wav, w["vmig"], rec)
    el = time.perf_counter() - t0
    cells = (nx + 2 * pml) * (nz + 2 * pml)
    updates = 3.0 * nt_sample * cells
    return dict(value=updates / el / 1e9, unit="Gcell-updates/s", cores=threads, kind=kind,
                sample="one shot of the same RTM workload with nt=%d (%.1f s of CPU work)" % (nt_sample, el))


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


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
                   "propagations_per_shot": 3, "source_band": "isolated (SEIS_RTM_ISOLATE_BAND)"},
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
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        log("timing the CPU baseline on %d threads" % threads)
        cb = cpu_baseline(w, args.cpu_nt, threads)
        cb["gpu_over_cpu"] = round(value / cb["value"], 1)
        out["cpu_baseline"] = cb
    print(json.dumps(out))


if __name__ == "__main__":
    main()
