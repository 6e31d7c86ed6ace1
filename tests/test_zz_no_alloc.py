"""seis_modeling_batch_static: modeling of a batch in the library's static
device scratch (no allocation), against seis_modeling_batch on the same
inputs, on two calls in a row, the second with a smaller batch; and its
rejection of a shot larger than the scratch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(nshots, nx, nz, seed):
    from seis_amd import ricker
    from seis_amd.batch import Grid, Shot, _layout
    rng = np.random.default_rng(seed)
    grid = Grid(nx, nz, 14, 0.001, 10.0, 10.0)
    shots = [Shot(sx=int(rng.integers(0, nx)), sz=2, nt=int(rng.integers(80, 160)), ndel=int(rng.integers(0, 6)))
             for _ in range(nshots)]
    vel = (1800.0 + rng.uniform(0, 900, (nx, nz))).astype(np.float32)
    wav = np.concatenate([ricker(0.001, 20.0, 25 + s.ndel, s.nt + s.ndel) for s in shots])
    gz = rng.integers(0, nz, nx * nshots).astype(np.int32)
    arr, nw, ng, nr = _layout(grid, shots)
    return grid, shots, arr, vel, wav, gz, nr


def _both(case):
    from seis_amd import DeviceArray, lib, synchronize
    from seis_amd._lib import check
    grid, shots, arr, vel, wav, gz, nr = case
    L = lib()
    g = grid.c()
    d_v, d_w, d_g = DeviceArray.from_host(vel), DeviceArray.from_host(wav), DeviceArray.from_host(gz)
    r1 = DeviceArray((nr,), np.float32, zero=True)
    r2 = DeviceArray((nr,), np.float32, zero=True)
    check(L.seis_modeling_batch(ctypes.byref(g), len(shots), arr, d_v.ptr, d_w.ptr, d_g.ptr, r1.ptr))
    check(L.seis_modeling_batch_static(ctypes.byref(g), len(shots), arr, d_v.ptr, d_w.ptr, d_g.ptr, r2.ptr))
    synchronize()
    return r1.download(), r2.download()


def test_static_modeling_equals_allocating_entry_point(engine):
    print("case no_alloc two calls: 5 shots 120x90, then 2 shots 60x50", flush=True)
    a, b = _both(_case(5, 120, 90, 3))
    assert np.array_equal(a, b) and np.abs(a).max() > 0
    a, b = _both(_case(2, 60, 50, 4))  # smaller: the scratch keeps the first call's contents
    assert np.array_equal(a, b) and np.abs(a).max() > 0


def test_static_modeling_chunks_within_scratch(engine):
    """More shots than the 64 MiB scratch holds at once (39 of these):
    the call runs them in consecutive chunks in the scratch (39 per chunk)."""
    print("case no_alloc chunked: 70 shots 300x200", flush=True)
    a, b = _both(_case(70, 300, 200, 5))
    assert np.array_equal(a, b)


def test_static_modeling_rejects_shot_beyond_scratch(engine):
    from seis_amd import lib
    from seis_amd._lib import SeisError, check
    grid, shots, arr, vel, wav, gz, nr = _case(1, 3000, 2000, 6)
    g = grid.c()
    with pytest.raises(SeisError, match="static scratch"):
        check(lib().seis_modeling_batch_static(ctypes.byref(g), 1, arr, None, None, None, None))
