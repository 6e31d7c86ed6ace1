"""Write the golden vectors of the 2D acoustic engine from the compiled reference.

Run in the container that holds the reference tree, after `make -C oracle ref`:

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py harmonic   # only harmonic.npz
    python tests/golden/make_golden.py c1_rtm     # only c1_rtm.npz (after c1_modeling.npz)
    python tests/golden/make_golden.py c1_born_illum  # only c1_born_illum.npz (after c1_rtm.npz)
    python tests/golden/make_golden.py edge_grids     # only edge_grids.npz

It calls the reference's own modeling / forward / adjoint / illum and the
plane-wave forwordPlane / adjointPlane / illumPlane (src/util/fd.h:3-11)
through oracle/_ref/libseisref.so and stores inputs and outputs as .npz
fixtures next to this script.  The fixtures are data only; they pin the CPU
oracle (tests/test_oracle_golden.py) and, through it, the HIP engine.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import Reference  # noqa: E402


def layered(nx, nz, dz, interfaces, speeds):
    """[nx][nz] velocity with horizontal interfaces at depths `interfaces` (m)."""
    z = np.arange(nz) * dz
    col = np.full(nz, speeds[0], np.float32)
    for zi, v in zip(interfaces, speeds[1:]):
        col[z >= zi] = v
    return np.repeat(col[None, :], nx, axis=0).astype(np.float32)


def small_case(seed, nx, nz, pml, nt, ndel, dt=0.001, dx=10.0, dz=10.0):
    rng = np.random.default_rng(seed)
    vel = layered(nx, nz, dz, [nz * dz / 3, 2 * nz * dz / 3], [1500.0, 2200.0, 3000.0])
    vel += rng.uniform(-60, 60, vel.shape).astype(np.float32)
    gz = rng.integers(0, 3, nx).astype(np.int32)
    sx, sz = int(rng.integers(nx // 4, 3 * nx // 4)), int(rng.integers(0, 4))
    fr = 15.0
    wav = None  # filled by the reference's own Ricker
    dv = (rng.standard_normal((nx, nz)) * (rng.random((nx, nz)) < 0.05) * 50).astype(np.float32)
    rec_in = rng.standard_normal((nx, nt)).astype(np.float32)
    return dict(nx=nx, nz=nz, pml=pml, nt=nt, ndel=ndel, dt=dt, dx=dx, dz=dz, sx=sx, sz=sz, fr=fr,
                delay=25, vel=vel, gz=gz, dv=dv, rec_in=rec_in, wav=wav)


def run_case(ref, c, what):
    args = (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"], c["gz"],
            c["wav"], c["vel"])
    out = {}
    if "modeling" in what:
        out["rec"] = ref.modeling(*args)
    if "born" in what:
        out["born"] = ref.born(*args, c["dv"])
    if "rtm" in what:
        out["img"] = ref.rtm(*args, c["rec_in"])
    if "illum" in what:
        out["ss"] = ref.illum(c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"],
                              c["sz"], c["wav"], c["vel"])
    return out


def plane_case(ref, seed, nx, nz, pml, nt, sz, gz, dip, dt=0.001, dx=10.0, dz=10.0):
    """Plane-wave case: the line source's column ix fires the reference
    Ricker delayed by dip * ix samples (sou is [nx][nt], ndel = 0)."""
    rng = np.random.default_rng(seed)
    vel = layered(nx, nz, dz, [nz * dz / 2], [1800.0, 2600.0])
    vel += rng.uniform(-60, 60, vel.shape).astype(np.float32)
    fr, delay = 20.0, 25
    wav = ref.ricker(dt, fr, delay, nt)
    sou = np.zeros((nx, nt), np.float32)
    for ix in range(nx):
        sh = int(dip * ix)
        sou[ix, sh:] = wav[: nt - sh]
    dv = (rng.standard_normal((nx, nz)) * (rng.random((nx, nz)) < 0.05) * 50).astype(np.float32)
    rec_in = rng.standard_normal((nx, nt)).astype(np.float32)
    return dict(nx=nx, nz=nz, pml=pml, nt=nt, ndel=0, dt=dt, dx=dx, dz=dz, sx=0, sz=sz, gz_row=gz, fr=fr,
                delay=delay, vel=vel, sou=sou, dv=dv, rec_in=rec_in)


def save(name, c, out, note):
    meta = {k: c[k] for k in ("nx", "nz", "pml", "nt", "ndel", "dt", "dx", "dz", "sx", "sz", "fr", "delay")}
    if "gz_row" in c:
        meta["gz_row"] = c["gz_row"]
    meta["note"] = note
    arrays = {k: v for k, v in c.items() if isinstance(v, np.ndarray)}
    arrays.update(out)
    np.savez_compressed(os.path.join(HERE, name), meta=json.dumps(meta), **arrays)
    print(name, {k: v.shape for k, v in arrays.items()})


def harmonic_cases(ref):
    """harmonic (InitInversion.cpp:141-151) of seeded traces: lengths around
    the filter's half length (100) and past 1024 samples, phase angles kx*x of
    zero, of both signs and beyond 2 pi, and a trace of signed zeros."""
    rng = np.random.default_rng(31)
    lengths = [1, 5, 99, 100, 101, 201, 250, 1030, 2500, 64]
    angles = [(0.0, 1500.0), (2.5e-4, 3210.0), (-4.0e-4, 1999.0), (1.1e-3, 7525.0), (3.3e-4, -600.0),
              (-1.0e-3, 40.0), (5.0e-4, 0.0), (2.0e-3, 4711.0), (7.7e-5, 12345.0), (6.0e-4, 2100.0)]
    xs, out, n, kx, x0 = [], [], [], [], []
    for i, (m, (k, x)) in enumerate(zip(lengths, angles)):
        tr = rng.standard_normal(m).astype(np.float32)
        if i == len(lengths) - 1:
            tr = np.where(np.arange(m) % 2 == 0, np.float32(-0.0), np.float32(0.0)).astype(np.float32)
        xs.append(tr)
        out.append(ref.harmonic(tr, k, x))
        n.append(m)
        kx.append(k)
        x0.append(x)
    arrays = dict(x=np.concatenate(xs), out=np.concatenate(out), n=np.array(n, np.int32),
                  kx=np.array(kx, np.float32), xs=np.array(x0, np.float32))
    meta = {"note": "harmonic of the reference (cos(kx x) in - sin(kx x) H[in], 201-tap Hilbert); "
                    "traces concatenated, lengths in n"}
    np.savez_compressed(os.path.join(HERE, "harmonic.npz"), meta=json.dumps(meta), **arrays)
    print("harmonic.npz", {k: v.shape for k, v in arrays.items()})


def c1_rtm(ref):
    """The reference's adjoint (RTM, fd.cpp:592-719) on the one-shot 201 x 301
    grid, nt = 1500: the record of c1_modeling.npz imaged with a smoothed
    model (only the image is stored; the inputs are rebuilt from that fixture)."""
    with np.load(os.path.join(HERE, "c1_modeling.npz"), allow_pickle=False) as f:
        c = {k: f[k] for k in f.files}
    meta = json.loads(str(c.pop("meta")))
    vel = c["vel"].astype(np.float64)
    k = np.array([1, 4, 6, 4, 1], np.float64) / 16.0
    for _ in range(3):  # a smooth migration model: separable binomial passes with edge replication
        vel = np.apply_along_axis(lambda a: np.convolve(np.pad(a, 2, mode="edge"), k, mode="valid"), 0, vel)
        vel = np.apply_along_axis(lambda a: np.convolve(np.pad(a, 2, mode="edge"), k, mode="valid"), 1, vel)
    vmig = vel.astype(np.float32)
    t0 = time.time()
    img = ref.rtm(meta["dt"], meta["dx"], meta["dz"], meta["nt"], meta["ndel"], meta["nx"], meta["nz"], meta["pml"],
                  meta["sx"], meta["sz"], c["gz"], c["wav"], vmig, c["rec"])
    print("C1 RTM %.2f s" % (time.time() - t0))
    np.savez_compressed(os.path.join(HERE, "c1_rtm.npz"),
                        meta=json.dumps({"note": "reference adjoint (RTM) of the c1_modeling record with a smoothed "
                                                 "model; inputs in c1_modeling.npz"}),
                        vmig=vmig, img=img)


def c1_born_illum(ref):
    """Born records (fd.cpp:81-177) and source illumination (fd.cpp:275-339) of
    the reference on the one-shot 201 x 301 grid, nt = 1500.  The perturbation
    is the true model minus the smoothed one of c1_rtm.npz.  The Born record
    (1.8 MB) is kept as the SHA-256 of its float32 bytes plus its norm; the
    illumination image is kept whole."""
    import hashlib
    with np.load(os.path.join(HERE, "c1_modeling.npz"), allow_pickle=False) as f:
        c = {k: f[k] for k in f.files}
    meta = json.loads(str(c.pop("meta")))
    with np.load(os.path.join(HERE, "c1_rtm.npz"), allow_pickle=False) as f:
        vmig = f["vmig"]
    dv = (c["vel"] - vmig).astype(np.float32)
    a = (meta["dt"], meta["dx"], meta["dz"], meta["nt"], meta["ndel"], meta["nx"], meta["nz"], meta["pml"],
         meta["sx"], meta["sz"])
    born = ref.born(*a, c["gz"], c["wav"], vmig, dv)
    ss = ref.illum(*a, c["wav"], vmig)
    np.savez_compressed(os.path.join(HERE, "c1_born_illum.npz"),
                        meta=json.dumps({"note": "reference Born record (sha256 of float32 bytes, norm) and "
                                                 "illumination of the one-shot grid; dv = vel - vmig of c1_rtm.npz"}),
                        born_sha256=np.frombuffer(hashlib.sha256(np.ascontiguousarray(born).tobytes()).digest(),
                                                  np.uint8),
                        born_norm=np.array(np.linalg.norm(born.astype(np.float64))), ss=ss)
    print("c1_born_illum.npz written")


def edge_cases(ref):
    """Grids at the edges of the engine's layout: no absorbing band (pml 0),
    and padded grids too small for an interior rectangle, where every cell
    takes the per-cell path.  Modeling, Born, RTM and illumination of each."""
    rng = np.random.default_rng(41)
    arrays = {}
    cases = [(20, 16, 0), (4, 5, 2), (6, 4, 3), (9, 8, 0)]
    for n, (nx, nz, pml) in enumerate(cases):
        vel = (2000 + 500 * rng.random((nx, nz))).astype(np.float32)
        gz = rng.integers(0, nz, nx).astype(np.int32)
        dv = (rng.standard_normal((nx, nz)) * 20).astype(np.float32)
        nt, ndel = 60, 2
        wav = ref.ricker(0.001, 30.0, 10 + ndel, nt + ndel)
        a = (0.001, 10.0, 10.0, nt, ndel, nx, nz, pml, nx // 2, nz // 2, gz, wav, vel)
        rec = ref.modeling(*a)
        arrays.update({"vel%d" % n: vel, "gz%d" % n: gz, "dv%d" % n: dv, "wav%d" % n: wav, "rec%d" % n: rec,
                       "born%d" % n: ref.born(*a, dv), "img%d" % n: ref.rtm(*a, rec),
                       "ss%d" % n: ref.illum(*a[:10], wav, vel)})
    meta = {"note": "edge grids (pml 0; no interior rectangle): modeling, born, rtm (of the modeled record), illum",
            "cases": cases, "dt": 0.001, "dx": 10.0, "dz": 10.0, "nt": 60, "ndel": 2}
    np.savez_compressed(os.path.join(HERE, "edge_grids.npz"), meta=json.dumps(meta), **arrays)
    print("edge_grids.npz", len(arrays), "arrays")


def main():
    ref = Reference()
    ref.set_threads(os.cpu_count() or 1)
    if sys.argv[1:] == ["harmonic"]:
        harmonic_cases(ref)
        return
    if sys.argv[1:] == ["c1_rtm"]:
        c1_rtm(ref)
        return
    if sys.argv[1:] == ["c1_born_illum"]:
        c1_born_illum(ref)
        return
    if sys.argv[1:] == ["edge_grids"]:
        edge_cases(ref)
        return
    harmonic_cases(ref)
    c1_rtm(ref)
    c1_born_illum(ref)
    edge_cases(ref)
    cases = [
        ("small_a.npz", small_case(11, 24, 20, 6, 80, 5)),
        ("small_b.npz", small_case(12, 40, 33, 10, 150, 0)),
        ("small_c.npz", small_case(13, 31, 45, 12, 120, 7)),
        ("small_d.npz", small_case(14, 64, 48, 16, 200, 3)),
    ]
    for name, c in cases:
        c["wav"] = ref.ricker(c["dt"], c["fr"], c["delay"] + c["ndel"], c["nt"] + c["ndel"])
        out = run_case(ref, c, ("modeling", "born", "rtm", "illum"))
        save(name, c, out, "seeded small grid; modeling, born, rtm and illum of the reference engine")

    # Plane-wave engine (fd.h:6-7, 11): forwordPlane, illumPlane, adjointPlane.
    for name, args in (("plane_a.npz", (21, 36, 30, 10, 120, 1, 2, 0.35)),
                       ("plane_b.npz", (22, 50, 41, 14, 170, 0, 3, -0.0))):
        c = plane_case(ref, *args)
        a = (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sz"], c["gz_row"],
             c["vel"], c["sou"])
        out = {"born": ref.born_plane(*a, c["dv"]), "ss": ref.illum_plane(*a), "img": ref.adjoint_plane(*a, c["rec_in"])}
        save(name, c, out, "plane-wave line source; forwordPlane, illumPlane and adjointPlane of the reference engine")

    # Plain forward modeling, 1 shot: 201 x 301 at 10 m, 3 layers 1500/2200/3000 m/s with
    # interfaces at 600 m and 1400 m, dt = 1 ms, nt = 1500, pml 30, Ricker 15 Hz (delay 67
    # samples, no extra pre-roll), source at (z = 20 m, x = 1500 m), 301 receivers at z = 20 m.
    nx, nz = 301, 201
    c1 = dict(nx=nx, nz=nz, pml=30, nt=1500, ndel=0, dt=0.001, dx=10.0, dz=10.0, sx=150, sz=2, fr=15.0,
              delay=67)
    c1["vel"] = layered(nx, nz, 10.0, [600.0, 1400.0], [1500.0, 2200.0, 3000.0])
    c1["gz"] = np.full(nx, 2, np.int32)
    c1["wav"] = ref.ricker(c1["dt"], c1["fr"], c1["delay"] + c1["ndel"], c1["nt"] + c1["ndel"])
    t0 = time.time()
    rec = ref.modeling(c1["dt"], c1["dx"], c1["dz"], c1["nt"], c1["ndel"], nx, nz, c1["pml"], c1["sx"], c1["sz"],
                       c1["gz"], c1["wav"], c1["vel"])
    print("C1 modeling %.2f s" % (time.time() - t0))
    save("c1_modeling.npz", c1, {"rec": rec}, "single-shot FP32 forward modeling of the 3-layer model")


if __name__ == "__main__":
    main()
