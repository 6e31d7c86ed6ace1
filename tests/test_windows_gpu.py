"""Shots of different window widths in one launch, and batches cut into
chunks by device memory.

The reference computes every shot on its own window of the model (getVel,
lsrtm/inversion.cpp:552-598); here shots of different widths share the same
launches.  Every shot's record, Born record, image and illumination must be
bit-identical to the single-shot call on its window (which is itself pinned
to the oracle), whatever else is in the launch and however the batch is cut
into chunks."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _mixed_case(seed=41):
    from seis_amd import ricker
    from seis_amd.batch import Grid, Shot
    rng = np.random.default_rng(seed)
    nz, pml, dt, dx = 40, 12, 0.001, 10.0
    widths = [48, 23, 70, 48, 5, 31]          # 5: a padded grid with no interior rectangle
    nts = [90, 130, 60, 111, 70, 95]
    ndels = [0, 5, 11, 2, 0, 3]
    shots, vels, dvs, wavs, gzs = [], [], [], [], []
    for w, nt, nd in zip(widths, nts, ndels):
        shots.append(Shot(sx=int(rng.integers(0, w)), sz=int(rng.integers(0, 4)), nt=nt, ndel=nd, nx=w))
        vels.append((1800.0 + rng.uniform(0, 600, (w, nz))).astype(np.float32))
        dvs.append((rng.standard_normal((w, nz)) * (rng.random((w, nz)) < 0.05) * 40).astype(np.float32))
        wavs.append(ricker(dt, 20.0, 20 + nd, nt + nd))
        gzs.append(rng.integers(0, nz, w).astype(np.int32))
    grid = Grid(max(widths), nz, pml, dt, dx, dx)
    return grid, shots, vels, dvs, wavs, gzs


def _single_args(grid, s, v, wav, gz):
    return (grid.dt, grid.dx, grid.dz, s.nt, s.ndel, s.nx, grid.nz, grid.pml, s.sx, s.sz, gz, wav, v)


def test_mixed_widths_equal_single_shots(engine, oracle):
    """Modeling, Born, RTM and illumination of six shots of widths 5..70 in
    one batch, each against the single-shot entry point on its window, and
    one shot against the oracle."""
    from seis_amd.batch import born_batch, illum_batch, modeling_batch, rtm_batch
    print("case windows_mixed_widths", flush=True)
    grid, shots, vels, dvs, wavs, gzs = _mixed_case()
    vel = np.concatenate([v.ravel() for v in vels])
    dv = np.concatenate([d.ravel() for d in dvs])
    recs = modeling_batch(grid, shots, vel, wavs, gzs)
    borns = born_batch(grid, shots, vel, dv, wavs, gzs)
    imgs = rtm_batch(grid, shots, vel, wavs, gzs, recs)
    ills = illum_batch(grid, shots, vel, wavs)
    for j, s in enumerate(shots):
        a = _single_args(grid, s, vels[j], wavs[j], gzs[j])
        assert np.array_equal(recs[j], engine.modeling(*a)), j
        assert np.array_equal(borns[j], engine.forward(*a, dvs[j])), j
        assert np.array_equal(imgs[j], engine.adjoint(*a, recs[j])), j
        ia = a[:10] + (wavs[j], vels[j])
        assert np.array_equal(ills[j], engine.illum(*ia)), j
    a = _single_args(grid, shots[2], vels[2], wavs[2], gzs[2])
    assert np.array_equal(recs[2], oracle.modeling(*a))
    assert rel_l2(imgs[2], oracle.rtm(*a, recs[2])) <= 1e-4


def test_chunked_batches_equal_whole(engine):
    """A forced small memory budget cuts the batch into several chunks (and
    a cap of shots per launch into more); the outputs equal the uncut run
    bit for bit."""
    from seis_amd.batch import (PLAN_MODEL, PLAN_RTM, batch_plan, born_batch, modeling_batch, rtm_batch,
                                set_batch_limits)
    print("case windows_chunked", flush=True)
    grid, shots, vels, dvs, wavs, gzs = _mixed_case(43)
    vel = np.concatenate([v.ravel() for v in vels])
    dv = np.concatenate([d.ravel() for d in dvs])
    whole = modeling_batch(grid, shots, vel, wavs, gzs)
    bwhole = born_batch(grid, shots, vel, dv, wavs, gzs)
    iwhole = rtm_batch(grid, shots, vel, wavs, gzs, whole)
    # budgets of one to three shots' working sets: the widest shot (70
    # columns) holds 4 fields of 102 x 128 floats for modeling, 13 plus its
    # strips for RTM
    bmodel, brtm = 400_000, 1_200_000
    assert len(batch_plan(PLAN_MODEL, grid, shots, bmodel)) >= 3
    assert len(batch_plan(PLAN_RTM, grid, shots, brtm)) >= 3
    try:
        set_batch_limits(bmodel, 0)
        cut = modeling_batch(grid, shots, vel, wavs, gzs)
        set_batch_limits(2 * bmodel, 0)
        bcut = born_batch(grid, shots, vel, dv, wavs, gzs)
        set_batch_limits(brtm, 0)
        icut = rtm_batch(grid, shots, vel, wavs, gzs, whole)
        set_batch_limits(0, 2)  # two shots per launch
        icut2 = rtm_batch(grid, shots, vel, wavs, gzs, whole)
    finally:
        set_batch_limits(0, 0)
    for j in range(len(shots)):
        assert np.array_equal(cut[j], whole[j]), j
        assert np.array_equal(bcut[j], bwhole[j]), j
        assert np.array_equal(icut[j], iwhole[j]), j
        assert np.array_equal(icut2[j], iwhole[j]), j


def test_shot_larger_than_budget_is_an_error(engine):
    from seis_amd._lib import SeisError
    from seis_amd.batch import modeling_batch, set_batch_limits
    print("case windows_budget_error", flush=True)
    grid, shots, vels, dvs, wavs, gzs = _mixed_case(44)
    vel = np.concatenate([v.ravel() for v in vels])
    try:
        set_batch_limits(4096, 0)
        with pytest.raises(SeisError, match="more than the"):
            modeling_batch(grid, shots, vel, wavs, gzs)
    finally:
        set_batch_limits(0, 0)


def test_lsrtm_windows(engine):
    """An LSRTM plan over windowed shots: L of a shot equals Born modeling on
    its window of the model with its window of the perturbation; L^T sums the
    shots' window gradients into the image; the FP64 dot test holds."""
    from seis_amd import ricker
    from seis_amd.batch import Grid, Shot, born_batch
    from seis_amd.lsrtm import LsrtmPlan, SNAP_FP32
    print("case lsrtm_windows", flush=True)
    rng = np.random.default_rng(45)
    nx, nz, pml, dt, dx = 90, 36, 10, 0.001, 10.0
    grid = Grid(nx, nz, pml, dt, dx, dx)
    v = (1900.0 + 500.0 * np.linspace(0, 1, nz)[None, :] + rng.uniform(0, 80, (nx, nz))).astype(np.float32)
    pert = (rng.standard_normal((nx, nz)) * (rng.random((nx, nz)) < 0.1) * 30).astype(np.float32)
    wins = [(0, 40), (25, 50), (60, 30), (0, 90)]  # (x0, nx): overlapping, at both ends, full width
    shots = [Shot(sx=w // 2, sz=2, nt=120, nx=w, x0=x0) for x0, w in wins]
    wavs = [ricker(dt, 22.0, 25, 120) for _ in shots]
    gzs = [np.full(w, 1, np.int32) for _, w in wins]
    plan = LsrtmPlan(grid, shots, v, wavs, gzs, snapshot=SNAP_FP32)
    try:
        lm = plan.born(pert)
        d = rng.standard_normal(plan.nrec).astype(np.float32)
        img = plan.adjoint(d)
    finally:
        plan.close()
    off = 0
    for (x0, w), s in zip(wins, shots):
        one = Grid(w, nz, pml, dt, dx, dx)
        ref = born_batch(one, [Shot(sx=s.sx, sz=2, nt=120)], v[x0:x0 + w], pert[x0:x0 + w], wavs[:1],
                         [np.full(w, 1, np.int32)])[0]
        assert np.array_equal(lm[off:off + w * 120].reshape(w, 120), ref), (x0, w)
        off += w * 120
    # L^T: each shot alone on its window, placed into the image, summed in shot order
    acc = np.zeros((nx, nz), np.float32)
    off = 0
    for (x0, w), s in zip(wins, shots):
        p1 = LsrtmPlan(Grid(w, nz, pml, dt, dx, dx), [Shot(sx=s.sx, sz=2, nt=120)], v[x0:x0 + w], wavs[:1],
                       [np.full(w, 1, np.int32)], snapshot=SNAP_FP32)
        try:
            g1 = p1.adjoint(d[off:off + w * 120])
        finally:
            p1.close()
        part = np.zeros((nx, nz), np.float32)
        part[x0:x0 + w] = g1
        acc = acc + part
        off += w * 120
    assert np.array_equal(img, acc)
    # FP64 dot test on the windowed plan
    p64 = LsrtmPlan(grid, shots, v.astype(np.float64), wavs, gzs, fp64=True)
    try:
        m = pert.astype(np.float64)
        r = rng.standard_normal(p64.nrec)
        a, b = float(np.dot(p64.born(m), r)), float(np.sum(m * p64.adjoint(r)))
    finally:
        p64.close()
    assert abs(a - b) / max(abs(a), abs(b)) <= 1e-6
