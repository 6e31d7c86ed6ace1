"""The oracle's exact transpose of the Born operator.

born_adjoint has no counterpart in the reference (the reference's `adjoint`
is an imaging operator, not a transpose), so it is pinned by invariants:
  * the dot-product test <L dv, r> = <dv, L^T r> in FP64 (to 1e-12 here;
    the bar for the engine is 1e-6), on grids whose absorbing band, corners,
    ring and low-order edge all carry signal;
  * the FP32 instance agrees with the FP64 one to FP32 rounding;
  * Born is the derivative of modeling: (m(v + e dv) - m(v - e dv)) / 2e -> L dv.
"""
import numpy as np
import pytest

from conftest import load_fixture, rel_l2


def _case(seed, nx=22, nz=19, pml=5, nt=70, ndel=3):
    rng = np.random.default_rng(seed)
    vel = (1800 + 900 * rng.random((nx, nz))).astype(np.float64)
    gz = rng.integers(0, nz, nx).astype(np.int32)
    dt, dx, dz = 0.001, 10.0, 9.0
    sx, sz = int(rng.integers(0, nx)), int(rng.integers(0, nz))
    return dict(dt=dt, dx=dx, dz=dz, nt=nt, ndel=ndel, nx=nx, nz=nz, pml=pml, sx=sx, sz=sz, gz=gz, vel=vel, rng=rng)


def _args(c, wav):
    return (c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"], c["gz"],
            wav, c["vel"])


@pytest.mark.parametrize("seed,shape", [(1, (22, 19, 5)), (2, (13, 16, 7)), (3, (30, 12, 4))])
def test_dot_product_fp64(oracle, seed, shape):
    nx, nz, pml = shape
    c = _case(seed, nx, nz, pml)
    rng = c["rng"]
    wav = oracle.ricker(c["dt"], 30.0, 15 + c["ndel"], c["nt"] + c["ndel"], dtype=np.float64)
    dv = rng.standard_normal((nx, nz))
    r = rng.standard_normal((nx, c["nt"]))
    Ld = oracle.born(*_args(c, wav), dv, dtype=np.float64)
    Ltr = oracle.born_adjoint(*_args(c, wav), r, dtype=np.float64)
    lhs, rhs = float(np.sum(Ld * r)), float(np.sum(dv * Ltr))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def test_adjoint_fp32_close_to_fp64(oracle):
    c = _case(4)
    rng = c["rng"]
    wav64 = oracle.ricker(c["dt"], 30.0, 15 + c["ndel"], c["nt"] + c["ndel"], dtype=np.float64)
    r = rng.standard_normal((c["nx"], c["nt"]))
    g64 = oracle.born_adjoint(*_args(c, wav64), r, dtype=np.float64)
    c32 = dict(c, vel=c["vel"].astype(np.float32))
    g32 = oracle.born_adjoint(*_args(c32, wav64.astype(np.float32)), r.astype(np.float32))
    assert rel_l2(g32, g64) < 1e-4


def test_born_is_derivative_of_modeling(oracle):
    c = load_fixture("small_b.npz")
    a = [c["dt"], c["dx"], c["dz"], c["nt"], c["ndel"], c["nx"], c["nz"], c["pml"], c["sx"], c["sz"], c["gz"],
         c["wav"].astype(np.float64)]
    v = c["vel"].astype(np.float64)
    dv = c["dv"].astype(np.float64)
    # Born linearizes only the wave operator: it leaves out the velocity in the
    # source amplitude vt2 (fd.cpp:51) and in the edge-replicated absorbing band
    # (setVel copies the first/last model row and column).  Keep dv off those.
    dv[c["sx"], c["sz"]] = 0.0
    dv[0, :] = dv[-1, :] = dv[:, 0] = dv[:, -1] = 0.0
    born = oracle.born(*a, v, dv, dtype=np.float64)
    errs = []
    for eps in (1e-2, 1e-3):
        fd = (oracle.modeling(*a, v + eps * dv, dtype=np.float64) -
              oracle.modeling(*a, v - eps * dv, dtype=np.float64)) / (2 * eps)
        errs.append(rel_l2(fd, born))
    # second-order convergence of the central difference to the Born record
    assert errs[1] < 1e-6 and errs[1] < errs[0], errs


def _interior_case():
    """Source, perturbation and live receivers well inside the model, pml 20,
    and a record short enough (70 steps, at most 14 cells of travel at
    2000 m/s) that no wave reaches the absorbing band."""
    nx = nz = 64
    pml, nt, ndel = 20, 70, 0
    rng = np.random.default_rng(21)
    vel = (1900 + 200 * rng.random((nx, nz))).astype(np.float64)
    gz = np.full(nx, 32, np.int32)
    dv = np.zeros((nx, nz))
    dv[26:38, 26:38] = rng.standard_normal((12, 12))
    r = np.zeros((nx, nt))
    r[24:40] = rng.standard_normal((16, nt))
    c = dict(dt=0.001, dx=10.0, dz=10.0, nt=nt, ndel=ndel, nx=nx, nz=nz, pml=pml, sx=32, sz=30, gz=gz, vel=vel)
    return c, dv, r


def test_xcorr_adjoint_is_the_transpose_away_from_the_band(oracle):
    """Where no wave reaches the band, the cross-correlation adjoint (forward
    operator run backwards, rec * vt2 injected, grad += u S / vt2) is the
    exact transpose: same image as born_adjoint, and the dot test with born
    holds, in FP64.  A one-step shift of the imaging time would break both."""
    c, dv, r = _interior_case()
    wav = oracle.ricker(c["dt"], 25.0, 40, c["nt"] + c["ndel"], dtype=np.float64)
    gx = oracle.born_adjoint_xcorr(*_args(c, wav), r, dtype=np.float64)
    ge = oracle.born_adjoint(*_args(c, wav), r, dtype=np.float64)
    assert np.abs(gx).max() > 0
    assert np.linalg.norm(gx - ge) <= 1e-10 * np.linalg.norm(ge)
    Ld = oracle.born(*_args(c, wav), dv, dtype=np.float64)
    lhs, rhs = float(np.sum(Ld * r)), float(np.sum(dv * gx))
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def test_xcorr_adjoint_detects_a_time_shift(oracle):
    """The same comparison with the record delayed by one sample fails by far:
    the agreement above pins the time alignment, not just the amplitude."""
    c, dv, r = _interior_case()
    wav = oracle.ricker(c["dt"], 25.0, 40, c["nt"] + c["ndel"], dtype=np.float64)
    ge = oracle.born_adjoint(*_args(c, wav), r, dtype=np.float64)
    shifted = np.zeros_like(r)
    shifted[:, 1:] = r[:, :-1]
    gs = oracle.born_adjoint_xcorr(*_args(c, wav), shifted, dtype=np.float64)
    assert np.linalg.norm(gs - ge) > 1e-2 * np.linalg.norm(ge)


def test_xcorr_adjoint_differs_only_through_the_band(oracle):
    """With waves in the band the two adjoints differ, but stay close."""
    c = _case(5)
    rng = c["rng"]
    wav = oracle.ricker(c["dt"], 30.0, 15 + c["ndel"], c["nt"] + c["ndel"], dtype=np.float64)
    r = rng.standard_normal((c["nx"], c["nt"]))
    gx = oracle.born_adjoint_xcorr(*_args(c, wav), r, dtype=np.float64)
    ge = oracle.born_adjoint(*_args(c, wav), r, dtype=np.float64)
    err = np.linalg.norm(gx - ge) / np.linalg.norm(ge)
    assert 1e-8 < err < 0.5, err
