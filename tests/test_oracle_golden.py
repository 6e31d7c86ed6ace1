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


def _harmonic_cases():
    c = load_fixture("harmonic.npz")
    off = np.concatenate([[0], np.cumsum(c["n"])])
    return [(c["x"][off[i]:off[i + 1]], c["kx"][i], c["xs"][i], c["out"][off[i]:off[i + 1]])
            for i in range(len(c["n"]))]


def test_harmonic_bitexact(oracle):
    """harmonic (InitInversion.cpp:141-151) against the compiled reference,
    bit for bit including signed zeros."""
    for x, kx, xs, want in _harmonic_cases():
        got = oracle.harmonic(x, kx, xs)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (len(x), kx, xs)


def test_harmonic_special_angles(oracle):
    """kx*x = 0 leaves the trace as it is; kx*x = pi/2 gives -H[trace]."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal(300).astype(np.float32)
    assert np.array_equal(oracle.harmonic(x, 0.0, 1234.0), x)
    assert np.array_equal(oracle.harmonic(x, 1e-3, 0.0), x)
    h = oracle.hilbert_apply(100, x, np.float64)
    got = oracle.harmonic(x.astype(np.float64), np.pi / 2, 1.0, np.float64)
    assert np.abs(got + h).max() < 1e-12 * np.abs(h).max()


def _plane_survey(seed, nx=23, nt=140, ns=5):
    """Shots with uneven trace counts, traces outside the model (-1), two
    traces of one shot in one column, and two shots at one source column."""
    rng = np.random.default_rng(seed)
    sx = np.array([50.0, 410.0, 410.0, 170.0, 90.0][:ns], np.float32)
    isx = np.array([1, 9, 9, 4, 2][:ns], np.int32)
    cols, traces = [], []
    for s in range(ns):
        ng = int(rng.integers(3, 19))
        c = rng.integers(-1, nx, ng).astype(np.int32)
        if ng > 3:
            c[2] = c[3] = 7
        cols.append(c)
        traces.append(rng.standard_normal((ng, nt)).astype(np.float32))
    return sx, isx, cols, traces


def test_plane_prepare_zero_ray_parameter_stacks(oracle):
    """p = 0: every phase shift is the identity, so sou stacks the wavelet at
    the source columns and rec stacks each shot's traces by column (a later
    trace of a shot replacing an earlier one), shots summed in order."""
    nx, nt = 23, 140
    sx, isx, cols, traces = _plane_survey(8, nx, nt)
    wav = np.random.default_rng(9).standard_normal(nt + 11).astype(np.float32)
    sou, rec = oracle.plane_prepare(nx, nt, wav.shape[0], 1, 0.0, 0.0, sx, isx, [len(c) for c in cols],
                                    np.concatenate(cols), wav, np.concatenate(traces))
    ws = np.zeros((nx, wav.shape[0]), np.float32)
    rs = np.zeros((nx, nt), np.float32)
    for s in range(len(sx)):
        ws[isx[s]] = ws[isx[s]] + wav
        placed = np.zeros((nx, nt), np.float32)
        for g, c in enumerate(cols[s]):
            if c >= 0:
                placed[c] = traces[s][g]
        rs = placed + rs
    assert np.array_equal(sou[0], ws)
    assert np.array_equal(rec[0], rs)


def test_plane_prepare_single_shot_is_harmonic(oracle):
    """One shot and one ray parameter: the source column holds harmonic of the
    wavelet (a golden trace of the compiled reference, so its expected output
    is the reference's), each trace's column the harmonic of the trace."""
    nx, nt = 17, 260
    wav, p, xs, want = _harmonic_cases()[7]
    rng = np.random.default_rng(12)
    col = np.array([3, -1, 5, 11, 16], np.int32)
    tr = rng.standard_normal((5, nt)).astype(np.float32)
    sou, rec = oracle.plane_prepare(nx, nt, wav.shape[0], 1, p, 0.5, [xs], [6], [5], col, wav, tr)
    assert np.array_equal(sou[0, 6], want)
    assert not np.any(np.delete(sou[0], 6, axis=0))
    for g, c in enumerate(col):
        if c >= 0:
            assert np.array_equal(rec[0, c], oracle.harmonic(tr[g], p, xs))
    assert not np.any(rec[0, [0, 1, 2, 4, 6, 7, 8, 9, 10, 12, 13, 14, 15]])


def test_plane_prepare_ray_parameters_independent(oracle):
    """Slice k of an np-parameter call equals a one-parameter call at pmin + k*dp."""
    nx, nt = 23, 140
    sx, isx, cols, traces = _plane_survey(10, nx, nt)
    wav = np.random.default_rng(2).standard_normal(nt + 5).astype(np.float32)
    args = (sx, isx, [len(c) for c in cols], np.concatenate(cols), wav, np.concatenate(traces))
    pmin, dp = np.float32(-2.0e-4), np.float32(1.5e-4)
    sou, rec = oracle.plane_prepare(nx, nt, wav.shape[0], 4, pmin, dp, *args)
    for k in range(4):
        s1, r1 = oracle.plane_prepare(nx, nt, wav.shape[0], 1, np.float32(pmin + np.float32(k) * dp), 0.0, *args)
        assert np.array_equal(sou[k], s1[0]) and np.array_equal(rec[k], r1[0]), k


def test_plane_prepare_sums_each_shots_phase_shift(oracle):
    """Several shots at one nonzero ray parameter: each shot's wavelet and
    traces are shifted with that shot's own source x, then summed in order."""
    nx, nt = 23, 140
    sx, isx, cols, traces = _plane_survey(14, nx, nt)
    wav = np.random.default_rng(4).standard_normal(nt + 3).astype(np.float32)
    p = np.float32(6.1e-4)
    sou, rec = oracle.plane_prepare(nx, nt, wav.shape[0], 1, p, 0.0, sx, isx, [len(c) for c in cols],
                                    np.concatenate(cols), wav, np.concatenate(traces))
    ws = np.zeros((nx, wav.shape[0]), np.float32)
    rs = np.zeros((nx, nt), np.float32)
    for s in range(len(sx)):
        ws[isx[s]] = ws[isx[s]] + oracle.harmonic(wav, p, sx[s])
        placed = np.zeros((nx, nt), np.float32)
        for g, c in enumerate(cols[s]):
            if c >= 0:
                placed[c] = oracle.harmonic(traces[s][g], p, sx[s])
        rs = placed + rs
    assert np.array_equal(sou[0], ws) and np.array_equal(rec[0], rs)


def test_c1_rtm_bitexact(oracle):
    """The reference's adjoint (RTM) on the one-shot 201 x 301 grid, nt = 1500,
    imaging the modeled record with a smoothed model."""
    c = load_fixture("c1_modeling.npz")
    g = load_fixture("c1_rtm.npz")
    img = oracle.rtm(c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"],
                     c["gz"], c["wav"], g["vmig"], c["rec"])
    assert np.array_equal(img, g["img"])


def _c1_born_illum_inputs():
    c = load_fixture("c1_modeling.npz")
    vmig = load_fixture("c1_rtm.npz")["vmig"]
    a = (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"])
    return c, vmig, (c["vel"] - vmig).astype(np.float32), a


def test_c1_born_and_illum_bitexact(oracle):
    """The reference's Born record (by its SHA-256) and illumination on the
    one-shot 201 x 301 grid, nt = 1500."""
    import hashlib
    c, vmig, dv, a = _c1_born_illum_inputs()
    g = load_fixture("c1_born_illum.npz")
    born = oracle.born(*a, c["gz"], c["wav"], vmig, dv)
    assert hashlib.sha256(np.ascontiguousarray(born).tobytes()).digest() == g["born_sha256"].tobytes()
    assert np.array_equal(oracle.illum(*a, c["wav"], vmig), g["ss"])


def _edge_cases():
    c = load_fixture("edge_grids.npz")
    for n, (nx, nz, pml) in enumerate(c["cases"]):
        a = (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], nx, nz, pml, nx // 2, nz // 2, c["gz%d" % n],
             c["wav%d" % n], c["vel%d" % n])
        yield n, (nx, nz, pml), a, c


def test_edge_grids_bitexact(oracle):
    """pml 0, and padded grids without an interior rectangle: modeling, Born,
    RTM and illumination bit for bit against the compiled reference."""
    for n, shape, a, c in _edge_cases():
        assert np.array_equal(oracle.modeling(*a), c["rec%d" % n]), shape
        assert np.array_equal(oracle.born(*a, c["dv%d" % n]), c["born%d" % n]), shape
        assert np.array_equal(oracle.rtm(*a, c["rec%d" % n]), c["img%d" % n]), shape
        assert np.array_equal(oracle.illum(*a[:10], c["wav%d" % n], c["vel%d" % n]), c["ss%d" % n]), shape
