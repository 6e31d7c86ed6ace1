"""Cutting a batch into chunks that fit a device-memory budget (host-side
arithmetic of the batch calls; no GPU needed).

A shot holds nf padded fields of (NX + 8) x ld floats (ld = NZ + 4 + a top
margin that aligns row max(pml, 4), rounded up to 64) for the run, plus, for RTM, its boundary strips of nt * 8 * (nx +
nz) floats, plus its descriptor.  The cut must be consecutive, cover every
shot, keep each chunk within the budget, and be greedy: the next shot would
not have fitted.  Also the host-side checks of the batch wrappers."""
import numpy as np
import pytest

from seis_amd.batch import (PLAN_BORN, PLAN_ILLUM, PLAN_MODEL, PLAN_RTM, PLAN_RTM_PLANE, Grid, Shot, batch_plan,
                            born_batch, modeling_batch, pack, rtm_batch)
from seis_amd._lib import SeisError

# fields per shot: three levels per propagated field, the band's stencil values,
# and for RTM the padded image and four U buffers of the imaging condition
FIELDS = {PLAN_MODEL: 4, PLAN_BORN: 8, PLAN_RTM: 13, PLAN_ILLUM: 5, PLAN_RTM_PLANE: 13}
DESC_MAX = 512  # the descriptor is a few hundred bytes; the checks hold for any size up to this


def _fixed_bytes(mode, grid, s):
    """Fields and strips of one shot (its descriptor excluded)."""
    nx = s.nx if s.nx > 0 else grid.nx
    NX, NZ = nx + 2 * grid.pml, grid.nz + 2 * grid.pml
    m = max(grid.pml, 4) % 64          # top margin: row max(pml, 4) on a 64-float boundary
    mz = (64 - m) % 64
    mz = mz + 64 if mz < 4 else mz
    ld = (mz + NZ + 4 + 63) // 64 * 64
    b = (NX + 8) * ld * FIELDS[mode] * 4
    nst = s.nt if mode == PLAN_RTM else s.nt + s.ndel if mode == PLAN_RTM_PLANE else 0
    st = nst * 8 * (nx + grid.nz) * 4
    return b + (st + 255) // 256 * 256


def _shots(seed, n=40):
    rng = np.random.default_rng(seed)
    return [Shot(sx=1, sz=1, nt=int(rng.integers(50, 900)), ndel=int(rng.integers(0, 30)),
                 nx=int(rng.integers(8, 400))) for _ in range(n)]


@pytest.mark.parametrize("mode", [PLAN_MODEL, PLAN_BORN, PLAN_RTM, PLAN_ILLUM])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_chunks_fit_and_are_greedy(mode, seed):
    grid = Grid(400, 150, 20, 0.001, 10.0, 10.0)
    shots = _shots(seed)
    fixed = [_fixed_bytes(mode, grid, s) for s in shots]
    rng = np.random.default_rng(100 + seed)
    for budget in [int(max(fixed) * f) + 1024 for f in (1.2, 2.5, 7.0)] + [int(rng.integers(max(fixed) + 1024, sum(fixed)))]:
        chunks = batch_plan(mode, grid, shots, budget)
        assert chunks[0][0] == 0 and chunks[-1][1] == len(shots)
        for (a, b), (c, _) in zip(chunks, chunks[1:]):
            assert b == c and b > a
        for a, b in chunks:
            assert sum(fixed[a:b]) + 256 <= budget
            if b < len(shots):  # greedy: one more shot would not have fitted
                assert sum(fixed[a:b + 1]) + (b + 1 - a) * DESC_MAX + 256 > budget


def test_shot_cap_and_one_chunk():
    grid = Grid(400, 150, 20, 0.001, 10.0, 10.0)
    shots = _shots(7, 25)
    assert batch_plan(PLAN_RTM, grid, shots, 1 << 62) == [(0, 25)]
    assert batch_plan(PLAN_RTM, grid, shots, 1 << 62, max_shots=8) == [(0, 8), (8, 16), (16, 24), (24, 25)]
    assert batch_plan(PLAN_MODEL, grid, [], 1 << 30) == []


def test_rtm_strips_dominate_at_survey_scale():
    """The 2000 x 8000 workload's RTM shots on 1100-column windows with
    nt = 6000: about 0.7 GB each, so 512 of them need several chunks of a
    288 GB part, while modeling fits in one."""
    grid = Grid(8000, 2000, 30, 0.0005, 5.0, 5.0)
    shots = [Shot(sx=550, sz=2, nt=6000, nx=1100) for _ in range(512)]
    one = _fixed_bytes(PLAN_RTM, grid, shots[0])
    assert 0.6e9 < one < 0.8e9
    chunks = batch_plan(PLAN_RTM, grid, shots, 270 * 10**9)
    assert len(chunks) == 2
    assert all(b - a <= 270 * 10**9 // one for a, b in chunks)
    assert batch_plan(PLAN_MODEL, grid, shots, 270 * 10**9) == [(0, 512)]


def test_shot_larger_than_budget_is_rejected():
    grid = Grid(400, 150, 20, 0.001, 10.0, 10.0)
    with pytest.raises(SeisError, match="more than the"):
        batch_plan(PLAN_RTM, grid, _shots(9, 3), 1 << 20)


def test_plane_shots_span_the_grid():
    grid = Grid(64, 40, 10, 0.001, 10.0, 10.0)
    with pytest.raises(SeisError, match="plane-wave"):
        batch_plan(PLAN_RTM_PLANE, grid, [Shot(sx=0, sz=1, nt=10, nx=32)], 1 << 30)
    assert batch_plan(PLAN_RTM_PLANE, grid, [Shot(sx=0, sz=1, nt=10)], 1 << 30) == [(0, 1)]


def test_wrapper_rejects_short_inputs():
    """Host-side checks of the batch wrappers: a short wavelet, a gz row of
    the wrong length, a model stack missing a model or window, a record of
    the wrong size."""
    grid = Grid(30, 20, 8, 0.001, 10.0, 10.0)
    shots = [Shot(sx=3, sz=1, nt=50, ndel=4), Shot(sx=5, sz=1, nt=40, model=1)]
    ok_w = [np.zeros(54, np.float32), np.zeros(40, np.float32)]
    ok_g = [np.zeros(30, np.int32)] * 2
    vel = np.full((2, 30, 20), 2000.0, np.float32)
    with pytest.raises(ValueError, match="wavelet"):
        pack(grid, shots, [np.zeros(50, np.float32), ok_w[1]], ok_g)
    with pytest.raises(ValueError, match="gz"):
        pack(grid, shots, ok_w, [np.zeros(29, np.int32), ok_g[1]])
    with pytest.raises(ValueError, match="vel has"):
        modeling_batch(grid, shots, vel[:1], ok_w, ok_g)
    with pytest.raises(ValueError, match="dv has"):
        born_batch(grid, shots, vel, vel[:1], ok_w, ok_g)
    with pytest.raises(ValueError, match="record"):
        rtm_batch(grid, shots, vel, ok_w, ok_g, [np.zeros((30, 50), np.float32), np.zeros((30, 39), np.float32)])
    win = [Shot(sx=3, sz=1, nt=50, nx=12), Shot(sx=3, sz=1, nt=50, nx=17)]
    with pytest.raises(ValueError, match="vel has"):
        modeling_batch(grid, win, np.zeros(12 * 20 + 16 * 20, np.float32), [ok_w[0]] * 2,
                       [np.zeros(12, np.int32), np.zeros(17, np.int32)])
