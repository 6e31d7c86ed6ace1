"""The CPU oracle against the golden vectors of the compiled reference engine.

The fixtures (tests/golden/*.npz) hold the outputs of the reference's own
modeling / forward / adjoint / illum (src/util/fd.h:3-5, 10), written by
tests/golden/make_golden.py.  The float32 oracle keeps the reference's operation
order, so it must reproduce them bit for bit.
"""
import numpy as np
import pytest

from conftest import fixture_names, load_fixture, rel_l2

SMALL = [n for n in fixture_names() if n.startswith("small_")]


def _args(c):
    return (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"], c["gz"],
            c["wav"], c["vel"])


@pytest.mark.parametrize("name", SMALL)
def test_modeling_bitexact(oracle, name):
    c = load_fixture(name)
    rec = oracle.modeling(*_args(c))
    assert np.array_equal(rec, c["rec"]), name


@pytest.mark.parametrize("name", SMALL)
def test_born_bitexact(oracle, name):
    c = load_fixture(name)
    rec = oracle.born(*_args(c), c["dv"])
    assert np.array_equal(rec, c["born"]), name


@pytest.mark.parametrize("name", SMALL)
def test_rtm_bitexact(oracle, name):
    c = load_fixture(name)
    img = oracle.rtm(*_args(c), c["rec_in"])
    assert np.array_equal(img, c["img"]), name


def _illum_args(c):
    return (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"], c["wav"],
            c["vel"])


@pytest.mark.parametrize("name", SMALL)
def test_illum_bitexact(oracle, name):
    c = load_fixture(name)
    ss = oracle.illum(*_illum_args(c))
    assert np.array_equal(ss, c["ss"]), name


@pytest.mark.parametrize("name", SMALL)
def test_illum_is_sum_of_squared_modeling_samples(oracle, name):
    """Independent of the illumination code: at each receiver cell the
    illumination is the float32 running sum of the squared record samples of
    steps it > ndel (sample index itp = it - ndel >= 1), because modeling
    records the same field after the same boundary update.  Pins the strict
    inequality, the time alignment and the interior crop."""
    c = load_fixture(name)
    rec = oracle.modeling(*_args(c))
    ss = oracle.illum(*_illum_args(c))
    for ix in range(c["nx"]):
        acc = np.float32(0)
        for v in rec[ix, 1:]:
            acc = np.float32(acc + np.float32(v * v))
        assert ss[ix, c["gz"][ix]] == acc, ix
    # a sum from it >= ndel would add rec[:, 0]^2: the fixtures with a
    # pre-roll (ndel > 0) have nonzero first samples, so the loop above sees it
    assert c["ndel"] == 0 or np.count_nonzero(rec[:, 0]) > 0


def test_c1_modeling_bitexact(oracle):
    """Single-shot FP32 forward modeling on the 201 x 301 three-layer model."""
    c = load_fixture("c1_modeling.npz")
    rec = oracle.modeling(*_args(c))
    assert np.array_equal(rec, c["rec"])


def test_c1_fp64_oracle_close(oracle):
    """The FP64 instance agrees with the FP32 reference to FP32 rounding level."""
    c = load_fixture("c1_modeling.npz")
    a = list(_args(c))
    a[11] = c["wav"].astype(np.float64)
    a[12] = c["vel"].astype(np.float64)
    rec64 = oracle.modeling(*a, dtype=np.float64)
    err = rel_l2(c["rec"], rec64)
    assert err < 1e-4, err


def test_fixture_wavelet_is_ricker(oracle):
    c = load_fixture("c1_modeling.npz")
    w = oracle.ricker(c["dt"], c["fr"], c["delay"] + c["ndel"], c["nt"] + c["ndel"])
    assert np.array_equal(w, c["wav"])


PLANE = [n for n in fixture_names() if n.startswith("plane_")]


def _plane_args(c):
    return (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sz"], c["gz_row"],
            c["vel"], c["sou"])


@pytest.mark.parametrize("name", PLANE)
def test_plane_wave_bitexact(oracle, name):
    """forwordPlane / illumPlane / adjointPlane (fd.h:6-7, 11)."""
    c = load_fixture(name)
    a = _plane_args(c)
    assert np.array_equal(oracle.born_plane(*a, c["dv"]), c["born"]), name
    assert np.array_equal(oracle.illum_plane(*a), c["ss"]), name
    assert np.array_equal(oracle.adjoint_plane(*a, c["rec_in"]), c["img"]), name


def test_line_source_in_one_column_is_the_point_source(oracle):
    """Independent of the plane-wave code: a line source that fires in one
    column only is the point source there, so Born records and illumination
    equal the (golden-pinned) point-source ones with all receivers at one
    depth."""
    c = load_fixture("small_c.npz")
    nx, nt, ndel = c["nx"], c["nt"], c["ndel"]
    sou = np.zeros((nx, nt + ndel), np.float32)
    sou[c["sx"]] = c["wav"][: nt + ndel]
    gzr = 2
    gz = np.full(nx, gzr, np.int32)
    pa = (c["dt"], c["dx"], c["dz"], nt, ndel, nx, c["nz"], c["pml"], c["sz"], gzr, c["vel"], sou)
    pt = (c["dt"], c["dx"], c["dz"], nt, ndel, nx, c["nz"], c["pml"], c["sx"], c["sz"], gz, c["wav"], c["vel"])
    assert np.array_equal(oracle.born_plane(*pa, c["dv"]), oracle.born(*pt, c["dv"]))
    assert np.array_equal(oracle.illum_plane(*pa), oracle.illum(*_illum_args(c)))
