"""The two forms of a step launch give identical results.

A step is either two launches (the interior sweep plus the band cells'
stencil values, then the band kernel's one-way update and blend) or one
launch in which every band cell recomputes its neighbours' stencil values.
The tuner picks between them by timing on small grids, so the parity cases
are run here with each form forced (tuning modes 0 and 2) against the same
golden vectors and oracle results as the default runs.
"""
import numpy as np
import pytest

import test_gpu_parity as parity
import test_lsrtm_gpu as lsrtm
from conftest import load_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[0, 2], ids=["band_launch", "one_launch"])
def sweep_mode(request, engine):
    engine.set_tuning(request.param)
    print("tuning mode", request.param, flush=True)
    yield request.param
    engine.set_tuning(1)


def test_small_fixtures(engine, sweep_mode):
    for name in parity.SMALL:
        parity.test_modeling_matches_reference(engine, name)
        parity.test_born_matches_reference(engine, name)
        parity.test_rtm_matches_reference(engine, name)
        parity.test_illum_matches_reference(engine, name)


def test_plane_fixtures_and_batches(engine, oracle, sweep_mode):
    for name in parity.PLANE:
        parity.test_plane_wave_matches_reference(engine, name)
    for pml in (2, 10):
        parity.test_plane_wave_batch_matches_oracle(engine, oracle, pml)


def test_batches_and_edge_grids(engine, oracle, sweep_mode):
    parity.test_batch_equals_single_shots(engine, oracle)
    parity.test_rtm_isolated_band_equals_reference_interior(oracle)
    parity.test_rtm_odd_record_lengths_match_oracle(engine, oracle)
    for pml in (2, 10):
        parity.test_illum_batch_matches_oracle(engine, oracle, pml)
    parity.test_edge_grids_match_reference(engine)


def test_c1_modeling_and_rtm(engine, sweep_mode):
    parity.test_c1_modeling_matches_reference(engine)
    parity.test_c1_rtm_matches_reference(engine)


def test_lsrtm_snapshot_modes(engine, oracle, sweep_mode):
    """Snapshot creation, Born from snapshots and the cross-correlation
    adjoint go through the step launches; the exact adjoint keeps its own."""
    lsrtm.test_born_fp32_matches_reference_born(engine, oracle)
    lsrtm.test_xcorr_adjoint_matches_oracle(engine, oracle)
    lsrtm.test_dot_product_fp64(engine, oracle)


def test_tuner_choice_on_c1(engine):
    """Tuned (mode 1): the C1 grid gets a choice among the small-grid
    candidates and the record stays bit-exact."""
    engine.set_tuning(1)
    c = load_fixture("c1_modeling.npz")
    rec = engine.modeling(*parity._args(c))
    assert np.array_equal(rec, c["rec"])
    ch = engine.tuned_sweep("forward", int(c["nx"]) + 2 * int(c["pml"]), int(c["nz"]) + 2 * int(c["pml"]), 1)
    print("c1 sweep choice", ch, flush=True)
    assert ch is not None
    bx, one = ch
    assert bx in ((1, 2, 4, 8) if one else (1, 2, 4, 8, 16))
