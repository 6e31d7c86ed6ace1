"""GPU parity of the HIP engine against the CPU oracle and the golden vectors.

The engine keeps the reference's FP32 operation order and is compiled without
contraction, so modeling, Born and RTM are expected to match the reference's
FP32 results bit for bit; the asserted bars are the stated tolerances
(shot gathers relative L2 <= 1e-5, images <= 1e-4) plus a bit-exact check
reported separately.
"""
import numpy as np
import pytest

from conftest import fixture_names, load_fixture, rel_l2

pytestmark = pytest.mark.gpu

SMALL = [n for n in fixture_names() if n.startswith("small_")]
GATHER_TOL = 1e-5
IMAGE_TOL = 1e-4


def _args(c):
    return (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"], c["gz"],
            c["wav"], c["vel"])


@pytest.mark.parametrize("name", SMALL)
def test_modeling_matches_reference(engine, name):
    c = load_fixture(name)
    print("case", name, flush=True)
    rec = engine.modeling(*_args(c))
    assert rel_l2(rec, c["rec"]) <= GATHER_TOL
    assert np.array_equal(rec, c["rec"])


@pytest.mark.parametrize("name", SMALL)
def test_born_matches_reference(engine, name):
    c = load_fixture(name)
    print("case", name, flush=True)
    rec = engine.forward(*_args(c), c["dv"])
    assert rel_l2(rec, c["born"]) <= GATHER_TOL
    assert np.array_equal(rec, c["born"])


@pytest.mark.parametrize("name", SMALL)
def test_rtm_matches_reference(engine, name):
    c = load_fixture(name)
    print("case", name, flush=True)
    img = engine.adjoint(*_args(c), c["rec_in"])
    assert rel_l2(img, c["img"]) <= IMAGE_TOL
    assert np.array_equal(img, c["img"])


def test_c1_modeling_matches_reference(engine):
    print("case c1_modeling", flush=True)
    c = load_fixture("c1_modeling.npz")
    rec = engine.modeling(*_args(c))
    assert rel_l2(rec, c["rec"]) <= GATHER_TOL
    assert np.array_equal(rec, c["rec"])


def test_c1_born_and_rtm_match_oracle(engine, oracle):
    """C1 geometry, Born with a sparse spike reflectivity and RTM of the C1 record."""
    print("case c1_born_rtm", flush=True)
    c = load_fixture("c1_modeling.npz")
    rng = np.random.default_rng(21)
    dv = (rng.standard_normal(c["vel"].shape) * (rng.random(c["vel"].shape) < 0.02) * 50).astype(np.float32)
    a = _args(c)
    ref = oracle.born(*a, dv)
    got = engine.forward(*a, dv)
    assert rel_l2(got, ref) <= GATHER_TOL
    assert np.array_equal(got, ref)
    ref = oracle.rtm(*a, c["rec"])
    got = engine.adjoint(*a, c["rec"])
    assert rel_l2(got, ref) <= IMAGE_TOL
    assert np.array_equal(got, ref)


def test_batch_equals_single_shots(engine, oracle):
    """Uneven record lengths and pre-rolls in one batch, each shot as if alone."""
    print("case batch_uneven", flush=True)
    rng = np.random.default_rng(7)
    nx, nz, pml = 48, 40, 12
    dt, dx, dz = 0.001, 10.0, 10.0
    g = engine.Grid(nx, nz, pml, dt, dx, dz)
    vels = np.stack([np.full((nx, nz), 1800.0, np.float32) + rng.uniform(0, 600, (nx, nz)).astype(np.float32)
                     for _ in range(2)])
    dvs = (rng.standard_normal(vels.shape) * (rng.random(vels.shape) < 0.05) * 40).astype(np.float32)
    shots = [engine.Shot(sx=int(rng.integers(0, nx)), sz=int(rng.integers(0, 4)), nt=int(nt), ndel=int(nd),
                         model=int(mi)) for nt, nd, mi in [(90, 0, 0), (130, 5, 1), (60, 11, 0), (110, 2, 1)]]
    wavs = [oracle.ricker(dt, 20.0, 20 + s.ndel, s.nt + s.ndel) for s in shots]
    gzs = [rng.integers(0, nz, nx).astype(np.int32) for _ in shots]
    recs = engine.modeling_batch(g, shots, vels, wavs, gzs)
    borns = engine.born_batch(g, shots, vels, dvs, wavs, gzs)
    imgs = engine.rtm_batch(g, shots, vels, wavs, gzs, recs)
    for j, s in enumerate(shots):
        a = (dt, dx, dz, s.nt, s.ndel, nx, nz, pml, s.sx, s.sz, gzs[j], wavs[j], vels[s.model])
        assert np.array_equal(recs[j], oracle.modeling(*a)), j
        assert np.array_equal(borns[j], oracle.born(*a, dvs[s.model])), j
        assert rel_l2(imgs[j], oracle.rtm(*a, recs[j])) <= IMAGE_TOL, j


def test_rtm_isolated_band_equals_reference_interior(oracle):
    """SEIS_RTM_ISOLATE_BAND holds the reconstructed source field at zero in
    the absorbing band.  The 4-layer strips shield the interior wavefields, so
    the image can differ from the reference's only through the strip cells and
    the 61-tap column Hilbert filter: nowhere beyond 4 + 30 cells of the model
    edge, where it must be bit-identical; the whole image is bit-identical
    to the oracle's isolated-band RTM."""
    from seis_amd import ricker
    from seis_amd.batch import Grid, Shot, rtm_batch
    print("case rtm_isolated_band 100x90 nt=400", flush=True)
    nx, nz, pml, nt = 100, 90, 12, 400
    dt, dx, dz = 0.001, 10.0, 10.0
    rng = np.random.default_rng(31)
    v = (1800.0 + 1500.0 * np.linspace(0, 1, nz)[None, :] + 60.0 * rng.standard_normal((nx, nz))).astype(np.float32)
    gz = np.full(nx, 2, np.int32)
    wav = ricker(dt, 20.0, 50, nt)
    rec = rng.standard_normal((nx, nt)).astype(np.float32)
    ref = oracle.rtm(dt, dx, dz, nt, 0, nx, nz, pml, 50, 2, gz, wav, v, rec)
    grid = Grid(nx, nz, pml, dt, dx, dz)
    img = rtm_batch(grid, [Shot(sx=50, sz=2, nt=nt)], v, [wav], [gz], [rec], isolate_band=True)[0]
    m = 34
    assert np.isfinite(img).all()
    assert np.array_equal(img[m:nx - m, m:nz - m], ref[m:nx - m, m:nz - m])
    assert not np.array_equal(img, ref)  # the edge rows do see the change
    # the whole image, edges included, against the oracle's isolated variant
    iso = oracle.rtm(dt, dx, dz, nt, 0, nx, nz, pml, 50, 2, gz, wav, v, rec, isolate=True)
    assert np.array_equal(img, iso)


def test_rtm_odd_record_lengths_match_oracle(engine, oracle):
    """The imaging pass applies two reverse steps at a time: shots whose
    record ends inside a pair (odd nt next to even nt) and a batch whose
    longest record is odd (a final single step) image as the reference does."""
    from seis_amd.batch import Grid, Shot, rtm_batch
    print("case rtm_odd_nt", flush=True)
    rng = np.random.default_rng(17)
    nx, nz, pml, dt = 40, 36, 8, 0.001
    g = Grid(nx, nz, pml, dt, 10.0, 10.0)
    v = (1900.0 + rng.uniform(0, 500, (nx, nz))).astype(np.float32)
    shots = [Shot(sx=int(rng.integers(0, nx)), sz=1, nt=nt, ndel=nd) for nt, nd in [(101, 0), (64, 3), (77, 6)]]
    wavs = [oracle.ricker(dt, 25.0, 15 + s.ndel, s.nt + s.ndel) for s in shots]
    gzs = [rng.integers(0, 4, nx).astype(np.int32) for _ in shots]
    recs = [rng.standard_normal((nx, s.nt)).astype(np.float32) for s in shots]
    imgs = rtm_batch(g, shots, v, wavs, gzs, recs)
    for j, s in enumerate(shots):
        ref = oracle.rtm(dt, 10.0, 10.0, s.nt, s.ndel, nx, nz, pml, s.sx, s.sz, gzs[j], wavs[j], v, recs[j])
        assert np.array_equal(imgs[j], ref), j


@pytest.mark.parametrize("name", SMALL)
def test_illum_matches_reference(engine, name):
    """Source illumination (fd.h:10 illum) against the reference's golden
    vectors, bit for bit."""
    c = load_fixture(name)
    print("case illum", name, flush=True)
    ss = engine.illum(c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"],
                      c["wav"], c["vel"])
    assert rel_l2(ss, c["ss"]) <= IMAGE_TOL
    assert np.array_equal(ss, c["ss"])


@pytest.mark.parametrize("pml", [2, 12])
def test_illum_batch_matches_oracle(engine, oracle, pml):
    """Batched illumination of shots with uneven record lengths and
    pre-rolls, each as the oracle gives it alone; pml 2 puts model cells in
    the ring path."""
    from seis_amd.batch import Grid, Shot, illum_batch
    print("case illum_batch pml=%d" % pml, flush=True)
    rng = np.random.default_rng(40 + pml)
    nx, nz, dt = 46, 38, 0.001
    g = Grid(nx, nz, pml, dt, 10.0, 10.0)
    v = (1800.0 + rng.uniform(0, 900, (nx, nz))).astype(np.float32)
    shots = [Shot(sx=int(rng.integers(0, nx)), sz=int(rng.integers(0, nz)), nt=nt, ndel=nd)
             for nt, nd in [(120, 0), (75, 9), (160, 4)]]
    wavs = [oracle.ricker(dt, 20.0, 20 + s.ndel, s.nt + s.ndel) for s in shots]
    ss = illum_batch(g, shots, v, wavs)
    for j, s in enumerate(shots):
        ref = oracle.illum(dt, 10.0, 10.0, s.nt, s.ndel, nx, nz, pml, s.sx, s.sz, wavs[j], v)
        assert np.array_equal(ss[j], ref), j


PLANE = [n for n in fixture_names() if n.startswith("plane_")]


@pytest.mark.parametrize("name", PLANE)
def test_plane_wave_matches_reference(engine, name):
    """forwordPlane / illumPlane / adjointPlane (fd.h:6-7, 11) against the
    reference's golden vectors, bit for bit."""
    c = load_fixture(name)
    print("case plane", name, flush=True)
    a = (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sz"], c["gz_row"], c["vel"],
         c["sou"])
    assert np.array_equal(engine.forward_plane(*a, c["dv"]), c["born"])
    assert np.array_equal(engine.illum_plane(*a), c["ss"])
    img = engine.adjoint_plane(*a, c["rec_in"])
    assert rel_l2(img, c["img"]) <= IMAGE_TOL
    assert np.array_equal(img, c["img"])


@pytest.mark.parametrize("pml", [2, 10])
def test_plane_wave_batch_matches_oracle(engine, oracle, pml):
    """Plane-wave shots of uneven length and pre-roll in one batch, each as
    the oracle gives it alone; pml 2 puts model cells in the ring path."""
    from seis_amd.batch import Grid, Shot, born_plane_batch, illum_plane_batch, rtm_plane_batch
    print("case plane_batch pml=%d" % pml, flush=True)
    rng = np.random.default_rng(60 + pml)
    nx, nz, dt = 42, 34, 0.001
    g = Grid(nx, nz, pml, dt, 10.0, 10.0)
    v = (1900.0 + rng.uniform(0, 700, (nx, nz))).astype(np.float32)
    dv = (rng.standard_normal((nx, nz)) * (rng.random((nx, nz)) < 0.05) * 40).astype(np.float32)
    specs = [(110, 0, 1, 2), (75, 4, 0, 3), (96, 9, 2, 1)]
    shots = [Shot(sx=0, sz=sz, nt=nt, ndel=nd) for nt, nd, sz, _ in specs]
    gzr = [r for _, _, _, r in specs]
    sous = []
    for s in shots:
        w = oracle.ricker(dt, 25.0, 15, s.nt + s.ndel)
        u = np.zeros((nx, s.nt + s.ndel), np.float32)
        for ix in range(nx):
            sh = int(0.4 * ix)
            u[ix, sh:] = w[: s.nt + s.ndel - sh]
        sous.append(u)
    recs = [rng.standard_normal((nx, s.nt)).astype(np.float32) for s in shots]
    born = born_plane_batch(g, shots, v, sous, gzr, dv)
    ss = illum_plane_batch(g, shots, v, sous)
    img = rtm_plane_batch(g, shots, v, sous, gzr, recs)
    for j, s in enumerate(shots):
        a = (dt, 10.0, 10.0, s.nt, s.ndel, nx, nz, pml, s.sz, gzr[j], v, sous[j])
        assert np.array_equal(born[j], oracle.born_plane(*a, dv)), j
        assert np.array_equal(ss[j], oracle.illum_plane(*a)), j
        assert np.array_equal(img[j], oracle.adjoint_plane(*a, recs[j])), j


def test_c1_rtm_matches_reference(engine):
    """fd.h adjoint (RTM) on the one-shot 201 x 301 grid, nt = 1500, against the
    compiled reference's image (tests/golden/c1_rtm.npz)."""
    c = load_fixture("c1_modeling.npz")
    g = load_fixture("c1_rtm.npz")
    print("case c1_rtm 201x301 nt=1500", flush=True)
    img = engine.adjoint(c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"],
                         c["gz"], c["wav"], g["vmig"], c["rec"])
    err = rel_l2(img, g["img"])
    print("c1 rtm rel_l2 %.3e bitexact %s" % (err, np.array_equal(img, g["img"])))
    assert err <= IMAGE_TOL
    assert np.array_equal(img, g["img"])


def test_c1_born_and_illum_match_reference(engine):
    """fd.h forward (Born) and illum on the one-shot 201 x 301 grid, nt = 1500,
    against the compiled reference (Born by the SHA-256 of its record)."""
    import hashlib
    c = load_fixture("c1_modeling.npz")
    vmig = load_fixture("c1_rtm.npz")["vmig"]
    g = load_fixture("c1_born_illum.npz")
    a = (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"])
    print("case c1_born_illum 201x301 nt=1500", flush=True)
    born = engine.forward(*a, c["gz"], c["wav"], vmig, (c["vel"] - vmig).astype(np.float32))
    assert abs(np.linalg.norm(born.astype(np.float64)) - float(g["born_norm"])) <= GATHER_TOL * float(g["born_norm"])
    assert hashlib.sha256(np.ascontiguousarray(born).tobytes()).digest() == g["born_sha256"].tobytes()
    ss = engine.illum(*a, c["wav"], vmig)
    assert rel_l2(ss, g["ss"]) <= IMAGE_TOL
    assert np.array_equal(ss, g["ss"])


def test_edge_grids_match_reference(engine):
    """pml 0 (no absorbing band) and padded grids too small for an interior
    rectangle (every cell on the per-cell path), against the compiled
    reference: modeling, Born, RTM and illumination bit for bit."""
    c = load_fixture("edge_grids.npz")
    for n, (nx, nz, pml) in enumerate(c["cases"]):
        print("case edge grid", nx, nz, pml, flush=True)
        a = (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], nx, nz, pml, nx // 2, nz // 2)
        g, w, v = c["gz%d" % n], c["wav%d" % n], c["vel%d" % n]
        assert np.array_equal(engine.modeling(*a, g, w, v), c["rec%d" % n]), (nx, nz, pml)
        assert np.array_equal(engine.forward(*a, g, w, v, c["dv%d" % n]), c["born%d" % n]), (nx, nz, pml)
        assert np.array_equal(engine.adjoint(*a, g, w, v, c["rec%d" % n]), c["img%d" % n]), (nx, nz, pml)
        assert np.array_equal(engine.illum(*a, w, v), c["ss%d" % n]), (nx, nz, pml)
