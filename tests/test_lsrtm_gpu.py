"""GPU LSRTM operator pair: L (Born from snapshots) and its exact transpose.

Checks, against the CPU oracle and the mathematics:
  * L with FP32 snapshots reproduces the reference Born records bit for bit
    (the snapshot holds vt2*lap*2/v^3 exactly as fd.cpp:134-139 forms it);
  * L^T matches the oracle's scatter-form transpose (FP64 plan to 1e-12,
    FP32 plan to 1e-4);
  * the Born/adjoint dot-product test holds to 1e-6 in FP64 (and is
    reported for FP32 and BF16 snapshots);
  * BF16 / FP16 snapshot records stay within 1e-2 of the FP32 ones.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _setup(engine, oracle, seed=3, nx=40, nz=30, pml=8, shots=((90, 2), (120, 0), (70, 5))):
    rng = np.random.default_rng(seed)
    dt, dx, dz = float(np.float32(0.001)), 10.0, 10.0  # the C ABI carries dt in FP32
    vel = (2000 + 800 * rng.random((nx, nz))).astype(np.float32)
    g = engine.Grid(nx, nz, pml, dt, dx, dz)
    sh = [engine.Shot(sx=int(rng.integers(0, nx)), sz=int(rng.integers(0, 3)), nt=nt, ndel=nd) for nt, nd in shots]
    wavs = [oracle.ricker(dt, 25.0, 18 + s.ndel, s.nt + s.ndel) for s in sh]
    gzs = [rng.integers(0, nz, nx).astype(np.int32) for _ in sh]
    return rng, g, sh, vel, wavs, gzs


def _split(g, shots, rec):
    out, off = [], 0
    for s in shots:
        n = g.nx * s.nt
        out.append(rec[off:off + n].reshape(g.nx, s.nt))
        off += n
    return out


def _cgls_reorth(A, d, k):
    """Textbook CGLS on the explicit matrix A from x = 0, with every new
    gradient s = A^T r Gram-Schmidt-orthogonalised against the earlier ones
    (as the GPU solver does); history of (||r||^2, ||s||^2)."""
    x = np.zeros(A.shape[1])
    r = d.copy()
    s = A.T @ r
    basis = [s / np.linalg.norm(s)]
    p = s.copy()
    gamma = s @ s
    hist = [(r @ r, gamma)]
    for _ in range(k):
        q = A @ p
        alpha = gamma / (q @ q)
        x += alpha * p
        r -= alpha * q
        s = A.T @ r
        for u in basis:
            s = s - (s @ u) * u
        basis.append(s / np.linalg.norm(s))
        gnew = s @ s
        p = s + (gnew / gamma) * p
        gamma = gnew
        hist.append((r @ r, gamma))
    return x, np.array(hist)


def test_born_fp32_matches_reference_born(engine, oracle):
    from seis_amd.lsrtm import LsrtmPlan, SNAP_FP32
    print("case lsrtm_born_fp32", flush=True)
    rng, g, shots, vel, wavs, gzs = _setup(engine, oracle)
    dv = (rng.standard_normal((g.nx, g.nz)) * 30).astype(np.float32)
    plan = LsrtmPlan(g, shots, vel, wavs, gzs, snapshot=SNAP_FP32)
    recs = _split(g, shots, plan.born(dv))
    for j, s in enumerate(shots):
        ref = oracle.born(g.dt, g.dx, g.dz, s.nt, s.ndel, g.nx, g.nz, g.pml, s.sx, s.sz, gzs[j], wavs[j], vel, dv)
        assert rel_l2(recs[j], ref) <= 1e-5
        assert np.array_equal(recs[j], ref), j


def test_adjoint_matches_oracle_transpose(engine, oracle):
    from seis_amd.lsrtm import LsrtmPlan, SNAP_FP32
    print("case lsrtm_adjoint", flush=True)
    rng, g, shots, vel, wavs, gzs = _setup(engine, oracle, seed=4)
    r = rng.standard_normal(sum(g.nx * s.nt for s in shots))
    rs = _split(g, shots, r)
    ref = sum(oracle.born_adjoint(g.dt, g.dx, g.dz, s.nt, s.ndel, g.nx, g.nz, g.pml, s.sx, s.sz, gzs[j],
                                  wavs[j].astype(np.float64), vel.astype(np.float64), rs[j], dtype=np.float64)
              for j, s in enumerate(shots))
    p64 = LsrtmPlan(g, shots, vel.astype(np.float64), [w.astype(np.float64) for w in wavs], gzs, fp64=True)
    got = p64.adjoint(r)
    assert rel_l2(got, ref) <= 1e-12, rel_l2(got, ref)
    p32 = LsrtmPlan(g, shots, vel, wavs, gzs, snapshot=SNAP_FP32)
    got32 = p32.adjoint(r.astype(np.float32))
    assert rel_l2(got32, ref) <= 1e-4


def test_dot_product_fp64(engine, oracle):
    from seis_amd.lsrtm import LsrtmPlan
    print("case lsrtm_dot", flush=True)
    rng, g, shots, vel, wavs, gzs = _setup(engine, oracle, seed=5)
    plan = LsrtmPlan(g, shots, vel.astype(np.float64), [w.astype(np.float64) for w in wavs], gzs, fp64=True)
    x = rng.standard_normal((g.nx, g.nz))
    y = rng.standard_normal(plan.nrec)
    Lx = plan.born(x)
    Lty = plan.adjoint(y)
    lhs, rhs = plan.dot(Lx, y), plan.dot(x.ravel(), Lty.ravel())
    err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print("dot test fp64 rel err %.3e" % err)
    assert err <= 1e-6


@pytest.mark.parametrize("kind,tol", [(1, 1e-2), (2, 2e-3)])
def test_16bit_snapshots_close_to_fp32(engine, oracle, kind, tol):
    from seis_amd.lsrtm import LsrtmPlan, SNAP_FP32
    print("case lsrtm_snap_kind", kind, flush=True)
    rng, g, shots, vel, wavs, gzs = _setup(engine, oracle, seed=6)
    dv = (rng.standard_normal((g.nx, g.nz)) * 30).astype(np.float32)
    ref = LsrtmPlan(g, shots, vel, wavs, gzs, snapshot=SNAP_FP32).born(dv)
    got = LsrtmPlan(g, shots, vel, wavs, gzs, snapshot=kind).born(dv)
    assert rel_l2(got, ref) <= tol


def test_cgls_matches_textbook_cgls_on_explicit_operator(engine, oracle):
    """The GPU CGLS (FP64 plan) against CGLS written out in numpy on the
    explicit matrix of the same operator (columns L e_j from the plan's own
    Born), with the gradients re-orthogonalised on both sides, and L^T
    against that matrix's transpose; after as many iterations as unknowns the
    residual is the least-squares minimum."""
    from seis_amd.lsrtm import LsrtmPlan
    print("case cgls_explicit 6x5 nt=60", flush=True)
    nx, nz, pml, nt = 6, 5, 4, 60
    rng = np.random.default_rng(8)
    dt, dx, dz = float(np.float32(0.001)), 10.0, 10.0
    vel = 2000.0 + 500.0 * rng.random((nx, nz))
    g = engine.Grid(nx, nz, pml, dt, dx, dz)
    shots = [engine.Shot(sx=2, sz=1, nt=nt)]
    plan = LsrtmPlan(g, shots, vel, [oracle.ricker(dt, 30.0, 15, nt)], [np.zeros(nx, np.int32)], fp64=True)
    n = nx * nz
    Lm = np.stack([plan.born(np.eye(n)[j].reshape(nx, nz)) for j in range(n)], axis=1)
    d = rng.standard_normal(plan.nrec)
    # the transpose of the explicit matrix
    lt = plan.adjoint(d).ravel()
    assert np.linalg.norm(lt - Lm.T @ d) <= 1e-12 * np.linalg.norm(Lm.T @ d)
    # textbook CGLS (re-orthogonalised gradients) from x = 0
    k = 10
    x, ref = _cgls_reorth(Lm, d, k)
    img, hist = plan.cg(d, niter=k)
    assert hist.shape == (k + 1, 2)
    assert np.allclose(hist, ref, rtol=1e-9, atol=0)
    # the two sides sum their dot products in different orders
    assert np.linalg.norm(img.ravel() - x) <= 1e-7 * np.linalg.norm(x)
    # n iterations come down to the least-squares minimum (never below it;
    # finite-precision CG on this conditioning stays a few 1e-6 above it)
    img_n, hist_n = plan.cg(d, niter=n)
    xls = np.linalg.lstsq(Lm, d, rcond=None)[0]
    rmin = np.sum((d - Lm @ xls) ** 2)
    assert rmin * (1 - 1e-10) <= hist_n[-1, 0] <= rmin * (1 + 1e-4)
    assert np.all(np.diff(hist_n[:, 0]) <= 1e-12 * hist_n[0, 0])  # CGLS residuals never increase
    plan.close()


def test_xcorr_adjoint_matches_oracle(engine, oracle):
    """The cross-correlation adjoint (the reference's imaging condition on the
    stored snapshots) against the oracle's restatement: the FP32 plan with FP32
    snapshots bit for bit (summed over shots in shot order), the FP64 plan to
    1e-12."""
    from seis_amd.lsrtm import LsrtmPlan, SNAP_FP32, ADJ_XCORR
    print("case lsrtm_adjoint_xcorr", flush=True)
    rng, g, shots, vel, wavs, gzs = _setup(engine, oracle, seed=6)
    r = rng.standard_normal(sum(g.nx * s.nt for s in shots))
    rs = _split(g, shots, r)
    a = lambda j, s, dt: (g.dt, g.dx, g.dz, s.nt, s.ndel, g.nx, g.nz, g.pml, s.sx, s.sz, gzs[j])  # noqa: E731
    ref32 = sum(oracle.born_adjoint_xcorr(*a(j, s, None), wavs[j], vel, rs[j].astype(np.float32))
                for j, s in enumerate(shots))
    p32 = LsrtmPlan(g, shots, vel, wavs, gzs, snapshot=SNAP_FP32)
    p32.set_adjoint(ADJ_XCORR)
    got32 = p32.adjoint(r.astype(np.float32))
    print("xcorr fp32 rel %.3e bitexact %s" % (rel_l2(got32, ref32), np.array_equal(got32, ref32)))
    assert rel_l2(got32, ref32) <= 1e-6
    assert np.array_equal(got32, ref32)
    ref64 = sum(oracle.born_adjoint_xcorr(*a(j, s, None), wavs[j].astype(np.float64), vel.astype(np.float64), rs[j],
                                          dtype=np.float64) for j, s in enumerate(shots))
    p64 = LsrtmPlan(g, shots, vel.astype(np.float64), [w.astype(np.float64) for w in wavs], gzs, fp64=True)
    p64.set_adjoint(ADJ_XCORR)
    assert rel_l2(p64.adjoint(r), ref64) <= 1e-12


def test_xcorr_cg_with_bf16_snapshots(engine, oracle):
    """CGLS with the cross-correlation adjoint and BF16 snapshots: the data
    residual falls and the image stays within 1e-3 of the FP64 plan run the same
    way; the exact adjoint stays the default."""
    from seis_amd.lsrtm import LsrtmPlan, SNAP_BF16, ADJ_XCORR
    print("case lsrtm_xcorr_cg", flush=True)
    rng, g, shots, vel, wavs, gzs = _setup(engine, oracle, seed=7)
    dv = np.zeros((g.nx, g.nz), np.float32)
    dv[10:30, 12:24] = rng.standard_normal((20, 12)).astype(np.float32) * 40
    p64 = LsrtmPlan(g, shots, vel.astype(np.float64), [w.astype(np.float64) for w in wavs], gzs, fp64=True)
    d = p64.born(dv.astype(np.float64))
    p64.set_adjoint(ADJ_XCORR)
    x64, h64 = p64.cg(d, niter=4)
    pb = LsrtmPlan(g, shots, vel, wavs, gzs, snapshot=SNAP_BF16)
    pb.set_adjoint(ADJ_XCORR)
    xb, hb = pb.cg(d.astype(np.float32), niter=4)
    assert hb.shape == (5, 2)
    assert hb[-1, 0] < 0.5 * hb[0, 0]
    err = rel_l2(xb, x64)
    print("xcorr cg bf16 vs fp64 image rel %.3e, residual %.3e -> %.3e" % (err, hb[0, 0], hb[-1, 0]))
    assert err <= 1e-3


def test_weighted_cgls_matches_textbook_on_explicit_operator(engine, oracle):
    """CG with an image-space preconditioner W (seis_lsrtm_set_weight) against
    textbook CGLS on the explicit matrix L diag(W), returning W x; a weight of
    ones gives the unweighted run bit for bit."""
    from seis_amd.lsrtm import LsrtmPlan
    print("case weighted_cgls 6x5 nt=60", flush=True)
    nx, nz, pml, nt = 6, 5, 4, 60
    rng = np.random.default_rng(9)
    dt, dx, dz = float(np.float32(0.001)), 10.0, 10.0
    vel = 2000.0 + 500.0 * rng.random((nx, nz))
    g = engine.Grid(nx, nz, pml, dt, dx, dz)
    shots = [engine.Shot(sx=3, sz=1, nt=nt)]
    plan = LsrtmPlan(g, shots, vel, [oracle.ricker(dt, 30.0, 15, nt)], [np.zeros(nx, np.int32)], fp64=True)
    n = nx * nz
    Lm = np.stack([plan.born(np.eye(n)[j].reshape(nx, nz)) for j in range(n)], axis=1)
    d = rng.standard_normal(plan.nrec)
    k = 8
    x0, h0 = plan.cg(d, niter=k)
    plan.set_weight(np.ones((nx, nz)))
    x1, h1 = plan.cg(d, niter=k)
    assert np.array_equal(x0, x1) and np.array_equal(h0, h1)
    w = 0.5 + rng.random(n)
    plan.set_weight(w.reshape(nx, nz))
    xw, hw = plan.cg(d, niter=k)
    A = Lm * w[None, :]
    y, ref = _cgls_reorth(A, d, k)
    assert np.allclose(hw, ref, rtol=1e-9, atol=0)
    assert np.linalg.norm(xw.ravel() - w * y) <= 1e-6 * np.linalg.norm(w * y)
    plan.set_weight(None)
    x2, _ = plan.cg(d, niter=k)
    assert np.array_equal(x0, x2)
    plan.close()


def test_cg_iterates_history(engine, oracle):
    """seis_lsrtm_cg_history: the image after iteration i equals the image of
    a run of i iterations, bit for bit, and the last one the returned image."""
    from seis_amd.lsrtm import LsrtmPlan, SNAP_FP32
    print("case lsrtm_cg_history", flush=True)
    rng, g, shots, vel, wavs, gzs = _setup(engine, oracle, seed=10)
    plan = LsrtmPlan(g, shots, vel, wavs, gzs, snapshot=SNAP_FP32)
    d = rng.standard_normal(plan.nrec).astype(np.float32)
    img, hist, its = plan.cg(d, 5, iterates=True)
    assert its.shape == (5, g.nx, g.nz) and np.array_equal(its[-1], img)
    for k in (1, 3):
        ik, _ = plan.cg(d, k)
        assert np.array_equal(its[k - 1], ik)
    plan.close()


def test_fp32_cg_stays_with_fp64_over_20_iterations(engine, oracle):
    """On an ill-conditioned case (layered-faulted model, 4 shots, data with
    5 % noise) plain CG in FP32 leaves the FP64 iterates by several percent
    after about 10 iterations, FP64 itself leaves the exact-arithmetic ones
    later; with the re-orthogonalised gradients the FP32 plan (FP32 and BF16
    snapshots) stays within 1e-4 / 1e-3 of the FP64 plan at every iteration."""
    from seis_amd.batch import Grid, Shot
    from seis_amd.lsrtm import LsrtmPlan, SNAP_BF16, SNAP_FP32
    from seis_amd.models import layered_faulted, smooth_model
    print("case lsrtm_cg_precision 120x60 nt=500 20 iterations", flush=True)
    nx, nz, pml, dx, dt, nt = 120, 60, 12, 8.0, 0.0008, 500
    v = np.ascontiguousarray(layered_faulted(240, 120, 4.0, 4.0, seed=1001)[::2, ::2])
    v0 = smooth_model(v, 6.0)
    dv = (v - v0).astype(np.float64)
    g = Grid(nx, nz, pml, dt, dx, dx)
    shots = [Shot(sx=sx, sz=2, nt=nt) for sx in (20, 50, 80, 100)]
    wav = oracle.ricker(dt, 20.0, 50, nt)
    gzs = [np.full(nx, 2, np.int32)] * 4
    p64 = LsrtmPlan(g, shots, v0.astype(np.float64), [wav.astype(np.float64)] * 4, gzs, fp64=True)
    d = p64.born(dv)
    d = d + 0.05 * np.sqrt(np.mean(d ** 2)) * np.random.default_rng(1).standard_normal(d.shape)
    _, h64, it64 = p64.cg(d, 20, iterates=True)
    p64.close()
    for kind, tol in ((SNAP_FP32, 1e-4), (SNAP_BF16, 1e-3)):
        p = LsrtmPlan(g, shots, v0, [wav] * 4, gzs, snapshot=kind)
        _, h, it = p.cg(d.astype(np.float32), 20, iterates=True)
        p.close()
        errs = [rel_l2(it[i], it64[i]) for i in range(20)]
        print("kind %d: image error per iteration %s" % (kind, ["%.1e" % e for e in errs]))
        assert max(errs) <= tol
        assert h[-1, 0] < 0.5 * h[0, 0]
