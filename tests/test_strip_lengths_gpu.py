"""Every strip length of the interior sweep gives identical results.

The tuner picks the strip length by timing: on small grids (such as every
golden-vector case) it tries strips of 1, 2 and 4 columns besides 8 and 16,
so the default parity runs mostly take short strips.  Here the parity cases
run again with tuning off (strips of 16 columns) against the same golden
vectors and oracle results.
"""
import numpy as np
import pytest

import test_gpu_parity as parity
import test_lsrtm_gpu as lsrtm
from conftest import load_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture
def strip16(engine):
    engine.set_tuning(0)
    print("tuning off: strips of 16 columns", flush=True)
    yield
    engine.set_tuning(1)


def test_small_fixtures(engine, strip16):
    for name in parity.SMALL:
        parity.test_modeling_matches_reference(engine, name)
        parity.test_born_matches_reference(engine, name)
        parity.test_rtm_matches_reference(engine, name)
        parity.test_illum_matches_reference(engine, name)


def test_plane_fixtures_and_batches(engine, oracle, strip16):
    for name in parity.PLANE:
        parity.test_plane_wave_matches_reference(engine, name)
    for pml in (2, 10):
        parity.test_plane_wave_batch_matches_oracle(engine, oracle, pml)


def test_batches_and_edge_grids(engine, oracle, strip16):
    parity.test_batch_equals_single_shots(engine, oracle)
    parity.test_rtm_isolated_band_equals_reference_interior(oracle)
    parity.test_rtm_odd_record_lengths_match_oracle(engine, oracle)
    for pml in (2, 10):
        parity.test_illum_batch_matches_oracle(engine, oracle, pml)
    parity.test_edge_grids_match_reference(engine)


def test_c1_modeling_and_rtm(engine, strip16):
    parity.test_c1_modeling_matches_reference(engine)
    parity.test_c1_rtm_matches_reference(engine)


def test_lsrtm_snapshot_modes(engine, oracle, strip16):
    lsrtm.test_born_fp32_matches_reference_born(engine, oracle)
    lsrtm.test_xcorr_adjoint_matches_oracle(engine, oracle)
    lsrtm.test_dot_product_fp64(engine, oracle)


def test_tuner_choice_on_c1(engine):
    """Tuned: the C1 grid gets one of the small-grid candidates and the record
    stays bit-exact."""
    engine.set_tuning(1)
    c = load_fixture("c1_modeling.npz")
    rec = engine.modeling(*parity._args(c))
    assert np.array_equal(rec, c["rec"])
    bx = engine.tuned_strip("forward", int(c["nx"]) + 2 * int(c["pml"]), int(c["nz"]) + 2 * int(c["pml"]), 1)
    print("c1 strip length", bx, flush=True)
    assert bx in (1, 2, 4, 8, 16)
