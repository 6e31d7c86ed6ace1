"""GPU parity of plane-wave synthesis (planeWavePrepare, lsrtm/inversion.cpp:
808-912, and its harmonic, util/InitInversion.cpp:141-151).

Integer placement and FP32 sums in the reference's order: every output is
checked bit for bit (as raw bits, so signed zeros count) against the golden
vectors of the compiled reference and against the oracle; at full survey
size, through properties that fix the result exactly.
"""
import numpy as np
import pytest

from conftest import load_fixture

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_harmonic_matches_reference(engine):
    c = load_fixture("harmonic.npz")
    off = np.concatenate([[0], np.cumsum(c["n"])])
    for i in range(len(c["n"])):
        x = c["x"][off[i]:off[i + 1]]
        print("harmonic case", i, "n", len(x), flush=True)
        got = engine.harmonic(x, c["kx"][i], c["xs"][i])
        assert _bits_equal(got, c["out"][off[i]:off[i + 1]]), i


def _survey(seed, nx, nt, ns, ngmax):
    """Uneven trace counts, traces outside the model (-1), two traces of a shot
    in one column, and shots sharing a source column."""
    rng = np.random.default_rng(seed)
    isx = rng.integers(0, nx, ns).astype(np.int32)
    if ns > 2:
        isx[2] = isx[1]
    sx = (isx * np.float32(12.5) + np.float32(300.0)).astype(np.float32)
    cols, traces = [], []
    for s in range(ns):
        ng = int(rng.integers(1, ngmax + 1))
        c = rng.integers(-1, nx, ng).astype(np.int32)
        if ng > 3:
            c[3] = c[2]
        cols.append(c)
        traces.append(rng.standard_normal((ng, nt)).astype(np.float32))
    return sx, isx, cols, traces


@pytest.mark.parametrize("seed,nx,nt,ntw,ns,np_", [(1, 37, 1100, 1137, 7, 11), (2, 20, 90, 90, 3, 1),
                                                   (3, 64, 1025, 1300, 1, 8)])
def test_plane_prepare_matches_oracle(engine, oracle, seed, nx, nt, ntw, ns, np_):
    print("plane prepare case", seed, nx, nt, ntw, ns, np_, flush=True)
    sx, isx, cols, traces = _survey(seed, nx, nt, ns, 40)
    wav = np.random.default_rng(seed + 100).standard_normal(ntw).astype(np.float32)
    pmin, dp = np.float32(-3.0e-4), np.float32(7.5e-5)
    sou, rec = engine.plane_prepare(nx, nt, np_, pmin, dp, sx, isx, traces, cols, wav)
    osou, orec = oracle.plane_prepare(nx, nt, ntw, np_, pmin, dp, sx, isx, [len(c) for c in cols],
                                      np.concatenate(cols), wav, np.concatenate(traces))
    assert _bits_equal(sou, osou)
    assert _bits_equal(rec, orec)


def test_plane_prepare_no_shots(engine):
    sou, rec = engine.plane_prepare(9, 50, 3, 0.0, 1e-4, np.zeros(0, np.float32), np.zeros(0, np.int32), [], [],
                                    np.ones(60, np.float32))
    assert sou.shape == (3, 9, 60) and rec.shape == (3, 9, 50)
    assert not sou.any() and not rec.any()


def test_plane_prepare_full_survey_properties(engine):
    """96 shots of 480 traces, 2000 samples, 1440 columns, 8 ray parameters:
    the p = 0 slice is the plain stack by column (every phase shift is the
    identity), and slice k equals a one-parameter call at pmin + k dp."""
    nx, nt, ntw, ns, ng = 1440, 2000, 2100, 96, 480
    rng = np.random.default_rng(77)
    isx = np.sort(rng.integers(0, nx, ns)).astype(np.int32)
    sx = (isx * np.float32(12.5)).astype(np.float32)
    cols, traces = [], []
    for s in range(ns):
        c = (isx[s] - 240 + 3 * np.arange(ng)).astype(np.int32)
        c[(c < 0) | (c >= nx)] = -1
        cols.append(c)
        traces.append(rng.standard_normal((ng, nt), dtype=np.float32))
    wav = rng.standard_normal(ntw, dtype=np.float32)
    from seis_amd.device import DeviceArray
    ddata = DeviceArray.from_host(np.concatenate(traces))
    print("full survey: np=8 call", flush=True)
    dp = np.float32(1.0e-4)
    sou, rec = engine.plane_prepare(nx, nt, 8, 0.0, dp, sx, isx, ddata, cols, wav)
    ws = np.zeros((nx, ntw), np.float32)
    rs = np.zeros((nx, nt), np.float32)
    for s in range(ns):
        ws[isx[s]] = ws[isx[s]] + wav
        placed = np.zeros((nx, nt), np.float32)
        ok = cols[s] >= 0
        placed[cols[s][ok]] = traces[s][ok]
        rs = placed + rs
    assert _bits_equal(sou[0], ws)
    assert _bits_equal(rec[0], rs)
    for k in (1, 5, 7):
        print("full survey: one-parameter call k =", k, flush=True)
        s1, r1 = engine.plane_prepare(nx, nt, 1, np.float32(np.float32(k) * dp), 0.0, sx, isx, ddata, cols, wav)
        assert _bits_equal(sou[k], s1[0]) and _bits_equal(rec[k], r1[0]), k
