"""Closed-form and brute-force checks of the oracle's building blocks.

These pin what the golden vectors leave open and fail on the plausible slips
(a wrong coefficient, sign, index or taper):
  * the 8th-order stencil is exact for polynomials of degree <= 9 and reduces
    to the 3-point operator in the 4-cell edge band (fd.cpp:1226-1271);
  * Ricker and Hilbert taps follow their formulas (wavelet.cpp:7-18,
    hilbert.h:59-70) and the Hilbert FIR is a convolution with them;
  * the sponge weights and velocity padding match a brute-force restatement
    of InitInversion.cpp:11-40, 92-113.
"""
import numpy as np
import pytest

from conftest import rel_l2


def test_stencil_polynomial_exactness(oracle):
    nx, nz, pml = 20, 18, 0
    dx, dz, dt = 2.0, 3.0, 0.01
    x = np.arange(nx)[:, None] * dx
    z = np.arange(nz)[None, :] * dz
    # degree-8 polynomial in x and z: the 9-point second derivative is exact
    px = (x / 10.0) ** 8 - 0.5 * (x / 10.0) ** 3
    pz = (z / 10.0) ** 7 + 0.25 * (z / 10.0) ** 2
    u = px + pz
    d2 = 8 * 7 * (x / 10.0) ** 6 / 100.0 - 0.5 * 6 * (x / 10.0) / 100.0 + 7 * 6 * (z / 10.0) ** 5 / 100.0 + 0.5 / 100.0
    vel = np.full((nx, nz), 100.0)
    prv = np.zeros_like(u)
    nxt = oracle.stencil_step(vel, prv, u, pml, dt, dx, dz, dtype=np.float64)
    vt2 = (dt * 100.0) ** 2
    expect = vt2 * d2 + 2 * u
    inner = (slice(4, nx - 4), slice(4, nz - 4))
    assert np.allclose(nxt[inner], expect[inner], rtol=1e-10, atol=1e-9)
    # the edge band uses the 3-point operator
    ix, iz = 2, 9
    uxx = (u[ix + 1, iz] + u[ix - 1, iz] - 2 * u[ix, iz]) / dx ** 2
    uzz = (u[ix, iz + 1] + u[ix, iz - 1] - 2 * u[ix, iz]) / dz ** 2
    assert np.isclose(nxt[ix, iz], vt2 * (uxx + uzz) + 2 * u[ix, iz], rtol=1e-12)
    # the outer ring is not written by the stencil
    assert np.all(nxt[0, :] == 0) and np.all(nxt[:, -1] == 0)


def test_stencil_symbol_at_nyquist(oracle):
    """Coefficient sum at k = pi: -205/72 - 2(8/5 + 1/5 + 8/315 + 1/560)... = -6.5016."""
    n = 40
    u = np.zeros((n, n))
    u[:, :] = ((-1.0) ** np.arange(n))[None, :]  # checkerboard in z only
    vel = np.ones((n, n))
    nxt = oracle.stencil_step(vel, np.zeros_like(u), u, 0, 1.0, 1.0, 1.0, dtype=np.float64)
    sym = (nxt[10, 10] - 2 * u[10, 10]) / u[10, 10]
    c = [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560]
    expect = c[0] + 2 * sum(ck * (-1) ** k for k, ck in enumerate(c) if k > 0)
    assert abs(sym - expect) < 1e-12 and abs(expect + 6.50158730) < 1e-6


def test_ricker_formula(oracle):
    dt, fr, nd, n = 0.001, 15.0, 67, 300
    w = oracle.ricker(dt, fr, nd, n, dtype=np.float64)
    t = (np.arange(n) - nd) * dt
    a = 3.1415926 ** 2 * fr ** 2
    assert np.allclose(w, (1 - 2 * a * t * t) * np.exp(-a * t * t), rtol=0, atol=1e-14)
    assert w[nd] == 1.0 and np.argmax(w) == nd


def test_hilbert_taps_formula(oracle):
    half = 30
    h = oracle.hilbert_taps(half, dtype=np.float64)
    i = np.arange(1, half + 1)
    taper = 0.54 + 0.46 * np.cos(3.1415926 * i / half)
    hi = taper * (i % 2) * 2.0 / (3.1415926 * i)
    assert h[half] == 0
    assert np.allclose(h[half + 1:], hi, atol=1e-15)
    assert np.allclose(h[:half][::-1], -hi, atol=1e-15)  # antisymmetric
    assert np.all(h[half + 2::2] == 0)  # even taps vanish


def test_hilbert_apply_is_convolution(oracle):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(97)
    half = 30
    h = oracle.hilbert_taps(half, dtype=np.float64)
    y = oracle.hilbert_apply(half, x, dtype=np.float64)
    full = np.convolve(x, h)  # full[j + half] = sum_i x[i] h[j - i + half]
    assert np.allclose(y, full[half:half + x.size], atol=1e-12)
    # a cosine becomes (approximately) a sine in the middle of a long trace
    n = 2048
    t = np.arange(n)
    c = np.cos(2 * np.pi * 0.1 * t)
    s = oracle.hilbert_apply(half, c, dtype=np.float64)
    mid = slice(200, n - 200)
    assert rel_l2(s[mid], np.sin(2 * np.pi * 0.1 * t)[mid]) < 0.02


@pytest.mark.parametrize("nx,nz,pml", [(5, 4, 3), (12, 9, 6), (30, 20, 10)])
def test_sponge_weights_bruteforce(oracle, nx, nz, pml):
    w = oracle.sponge_weights(nz, nx, pml, dtype=np.float64)
    NX, NZ = nx + 2 * pml, nz + 2 * pml
    for ix in range(NX):
        for iz in range(NZ):
            if ix < pml:
                e = ix / pml
            elif ix >= pml + nx:
                e = (NX - 1 - ix) / pml
            elif iz < pml:
                e = iz / pml
            elif iz >= pml + nz:
                e = (NZ - 1 - iz) / pml
            else:
                e = 1.0
            assert abs(w[ix, iz] - e) < 1e-12, (ix, iz)


def test_pad_velocity_is_clamped_model(oracle):
    rng = np.random.default_rng(5)
    v = rng.random((7, 5))
    pml = 3
    vp = oracle.pad_velocity(v, pml, dtype=np.float64)
    for ix in range(v.shape[0] + 2 * pml):
        for iz in range(v.shape[1] + 2 * pml):
            e = v[min(max(ix - pml, 0), v.shape[0] - 1), min(max(iz - pml, 0), v.shape[1] - 1)]
            assert vp[ix, iz] == e


def test_isolated_band_rtm_matches_reference_away_from_edges(oracle):
    """The oracle's isolated-band RTM (the engine's SEIS_RTM_ISOLATE_BAND
    variant) zeroes the reconstructed source field outside the model area.
    The 4-layer strips shield the interior wavefields from that change, so
    the image equals the reference's RTM beyond 4 + 30 cells of the model
    edge (strip layer + Hilbert half length) and differs at the edges."""
    rng = np.random.default_rng(5)
    nx, nz, pml, nt, dt = 80, 76, 10, 220, 0.001
    v = (1800.0 + 900.0 * np.linspace(0, 1, nz)[None, :] + 40.0 * rng.standard_normal((nx, nz))).astype(np.float32)
    gz = np.full(nx, 2, np.int32)
    wav = oracle.ricker(dt, 20.0, 40, nt)
    rec = rng.standard_normal((nx, nt)).astype(np.float32)
    a = (dt, 10.0, 10.0, nt, 0, nx, nz, pml, 40, 2, gz, wav, v, rec)
    ref = oracle.rtm(*a)
    iso = oracle.rtm(*a, isolate=True)
    m = 34
    assert np.isfinite(iso).all()
    assert np.array_equal(iso[m:nx - m, m:nz - m], ref[m:nx - m, m:nz - m])
    assert not np.array_equal(iso, ref)
