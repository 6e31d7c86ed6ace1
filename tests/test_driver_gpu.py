"""End-to-end parameter-file driver on the GPU (seis_run): shot records, Born
records and images written in the reference's file layout, compared with the
glue restatement (tests/glue_oracle.py) driving the CPU oracle."""
import os

import numpy as np
import pytest

import glue_oracle as G
from conftest import rel_l2

pytestmark = pytest.mark.gpu

NX, NZ, NT = 70, 40, 300


def _survey(tmp):
    rng = np.random.default_rng(17)
    v = np.empty((NX, NZ), np.float32)
    v[:, :] = 1800.0
    v[:, 15:] = 2400.0
    v[:, 28:] = 3000.0
    v += rng.uniform(-40, 40, v.shape).astype(np.float32)
    v0 = np.full_like(v, 2100.0)
    lines = []
    for sx in (150.0, 350.0, 560.0):
        for gx in np.arange(sx - 150.0, sx + 151.0, 10.0):
            if 0.0 <= gx <= (NX - 1) * 10.0:
                lines.append("%12.2f%12.2f%12.2f%12.2f" % (sx, 20.0, gx, 10.0))
    ng = {}
    for ln in lines:
        ng[ln.split()[0]] = ng.get(ln.split()[0], 0) + 1
    (tmp / "coords.txt").write_text("\n".join(lines) + "\n")
    v.tofile(tmp / "v.bin")
    v0.tofile(tmp / "v0.bin")
    (tmp / "WDIR").mkdir()
    (tmp / "CSG").mkdir()
    text = """nthread 1
ns 3
ngmax %d
nx %d
nz %d
nt %d
npml 12
delay 20
delaycal 0
maxiter 3
lpad 4
rpad 4
dx 10.0
dz 10.0
dt 0.001
fr 25
velfx 0.0
sbegin 0
send 2
precsg %s/CSG/CSG
wdir %s/WDIR/
v0file %s/v0.bin
vfile %s/v.bin
coordfile %s/coords.txt
""" % (max(ng.values()), NX, NZ, NT, tmp, tmp, tmp, tmp, tmp)
    (tmp / "param.txt").write_text(text)
    shots = []
    for sx in sorted({float(ln.split()[0]) for ln in lines}):
        sel = [ln.split() for ln in lines if float(ln.split()[0]) == sx]
        shots.append((sx, 20.0, [float(s[2]) for s in sel], [float(s[3]) for s in sel]))
    P = dict(dx=10.0, dz=10.0, dt=0.001, fr=25.0, nx=NX, nz=NZ, nt=NT, lpad=4, rpad=4, velfx=0.0, delay=20, npml=12)
    return P, shots, v, v0


def _window_case(P, shot, model):
    sx, sz, gx, gz = shot
    fx, nxw, cols = G.window(P, sx, gx)
    igz = G.receiver_gz(P, fx, nxw, gx, gz)
    isx, isz = G.source_cell(P, fx, sx, sz)
    return fx, nxw, np.ascontiguousarray(model[cols]), igz, isx, isz


def test_modeling_rtm_lsrtm_files(engine, oracle, tmp_path):
    from seis_amd import driver
    print("case driver_end_to_end", flush=True)
    P, shots, v, v0 = _survey(tmp_path)
    param = tmp_path / "param.txt"
    wav = oracle.ricker(P["dt"], P["fr"], P["delay"], NT)
    driver.run(param, "modeling")
    csgs = []
    for is_, shot in enumerate(shots):
        fx, nxw, vw, igz, isx, isz = _window_case(P, shot, v)
        rec = oracle.modeling(P["dt"], P["dx"], P["dz"], NT, 0, nxw, NZ, P["npml"], isx, isz, igz, wav, vw)
        csg = G.reorder(P, fx, nxw, shot[2], NT, rec=rec, to_gather=True)
        csg = G.remove_direct(P, csg, v, shot[0], shot[1], shot[2], shot[3])
        got = np.fromfile(tmp_path / "CSG" / ("CSG%d.dat" % is_), np.float32).reshape(len(shot[2]), NT)
        assert np.array_equal(got, csg), is_
        csgs.append(got)
    driver.run(param, "rtm")
    img = np.fromfile(tmp_path / "WDIR" / "RTM_IMAGE.dat", np.float32).reshape(NX, NZ)
    ref = np.zeros((NX, NZ), np.float32)
    for is_, shot in enumerate(shots):
        fx, nxw, vw, igz, isx, isz = _window_case(P, shot, v0)
        csg = G.remove_direct(P, csgs[is_], v0, shot[0], shot[1], shot[2], shot[3])
        rec = G.reorder(P, fx, nxw, shot[2], NT, csg=csg, to_gather=False)
        w_img = oracle.rtm(P["dt"], P["dx"], P["dz"], NT, 0, nxw, NZ, P["npml"], isx, isz, igz, wav, vw, rec)
        part = G.to_global(P, w_img, fx, 0.0, NX, NZ)
        assert os.path.exists(tmp_path / "WDIR" / ("CSG_IMAGE%d.dat" % is_))
        ref = part + ref
    assert rel_l2(img, ref) <= 1e-4
    driver.run(param, "lsrtm")
    ls = np.fromfile(tmp_path / "WDIR" / "LSRTM_IMAGE.dat", np.float32).reshape(NX, NZ)
    assert np.isfinite(ls).all() and np.abs(ls).max() > 0
    # the same CG with the cross-correlation adjoint: a nearby image
    driver.run(param, "lsrtm_xcorr", out_file=tmp_path / "WDIR" / "LSRTM_XCORR.dat")
    lx = np.fromfile(tmp_path / "WDIR" / "LSRTM_XCORR.dat", np.float32).reshape(NX, NZ)
    assert np.isfinite(lx).all() and rel_l2(lx, ls) < 0.5
    # illumination-preconditioned: a different, finite image
    driver.run(param, "lsrtm_illum", out_file=tmp_path / "WDIR" / "LSRTM_ILLUM.dat")
    li = np.fromfile(tmp_path / "WDIR" / "LSRTM_ILLUM.dat", np.float32).reshape(NX, NZ)
    assert np.isfinite(li).all() and np.abs(li).max() > 0 and not np.array_equal(li, ls)


def test_illum_files(engine, oracle, tmp_path):
    """seis_run illum (illum_MPI): per-shot windowed illumination through the
    v0file background, moved to the global grid -> <wdir>ILLUM<is>.dat, and
    their sum in shot order, against the glue restatement over the oracle."""
    from seis_amd import driver
    print("case driver_illum", flush=True)
    P, shots, v, v0 = _survey(tmp_path)
    param = tmp_path / "param.txt"
    wav = oracle.ricker(P["dt"], P["fr"], P["delay"], NT)
    driver.run(param, "illum")
    total = np.fromfile(tmp_path / "WDIR" / "ILLUM.dat", np.float32).reshape(NX, NZ)
    ref = np.zeros((NX, NZ), np.float32)
    for is_, shot in enumerate(shots):
        fx, nxw, vw, igz, isx, isz = _window_case(P, shot, v0)
        ss = oracle.illum(P["dt"], P["dx"], P["dz"], NT, 0, nxw, NZ, P["npml"], isx, isz, wav, vw)
        part = G.to_global(P, ss, fx, 0.0, NX, NZ)
        got = np.fromfile(tmp_path / "WDIR" / ("ILLUM%d.dat" % is_), np.float32).reshape(NX, NZ)
        assert np.array_equal(got, part), is_
        ref = part + ref
    assert np.array_equal(total, ref)


def test_plane_wave_files(engine, oracle, tmp_path):
    """seis_run plane_prepare / rtm_plane / illum_plane (planeWavePrepare,
    adjointPlane_MPI, illumPlane_MPI) with a mask: the PLANE_S_NEW /
    PLANE_R_NEW files bit for bit against the oracle's planeWavePrepare, the
    per-parameter images and illuminations against the oracle's adjointPlane /
    illumPlane, and their sums in parameter order."""
    from seis_amd import driver
    print("case driver_plane", flush=True)
    P, shots, v, v0 = _survey(tmp_path)
    param = tmp_path / "param.txt"
    n_p, dp, pmin = 3, 1.0e-4, -1.0e-4
    mask = np.ones((NX, NZ), np.float32)
    mask[:, :6] = 0.0
    mask.tofile(tmp_path / "mask.bin")
    text = param.read_text().replace("maxiter 3\n", "maxiter 3\nmask 1\nplaneTag 1\n")
    text += "maskfile %s/mask.bin\nnp %d\ndp %g\npmin %g\n" % (tmp_path, n_p, dp, pmin)
    param.write_text(text)
    rng = np.random.default_rng(5)
    gathers = []
    for is_, shot in enumerate(shots):
        g = rng.standard_normal((len(shot[2]), NT)).astype(np.float32)
        g.tofile(tmp_path / "CSG" / ("CSG%d.dat" % is_))
        gathers.append(g)
    driver.run(param, "plane_prepare")
    wav = oracle.ricker(P["dt"], P["fr"], P["delay"], NT)
    isx, cols = G.plane_geometry(P, [s[0] for s in shots], [s[2] for s in shots])
    sou, rec = oracle.plane_prepare(NX, NT, NT, n_p, np.float32(pmin), np.float32(dp),
                                    np.array([s[0] for s in shots], np.float32), isx, [len(c) for c in cols],
                                    np.concatenate(cols), wav, np.concatenate(gathers))
    for k in range(n_p):
        gs = np.fromfile(tmp_path / "WDIR" / ("PLANE_S_NEW%d.dat" % k), np.float32).reshape(NX, NT)
        gr = np.fromfile(tmp_path / "WDIR" / ("PLANE_R_NEW%d.dat" % k), np.float32).reshape(NX, NT)
        assert np.array_equal(gs, sou[k]) and np.array_equal(gr, rec[k]), k
    sz, gz = G.plane_rows(P, shots[0][1], shots[0][3][0])
    driver.run(param, "rtm_plane")
    driver.run(param, "illum_plane")
    img = np.fromfile(tmp_path / "WDIR" / "PLANE_IMAGE.dat", np.float32).reshape(NX, NZ)
    ss = np.fromfile(tmp_path / "WDIR" / "ILLUM_PLANE.dat", np.float32).reshape(NX, NZ)
    img_sum = np.zeros((NX, NZ), np.float32)
    ss_sum = np.zeros((NX, NZ), np.float32)
    for k in range(n_p):
        a = (P["dt"], P["dx"], P["dz"], NT, 0, NX, NZ, P["npml"], sz, gz, v0, sou[k])
        want_img = oracle.adjoint_plane(*a, rec[k]) * mask
        want_ss = oracle.illum_plane(*a)
        got_img = np.fromfile(tmp_path / "WDIR" / ("IMAGE%d.dat" % k), np.float32).reshape(NX, NZ)
        got_ss = np.fromfile(tmp_path / "WDIR" / ("ILLUM_PLANE%d.dat" % k), np.float32).reshape(NX, NZ)
        assert rel_l2(got_img, want_img) <= 1e-4, k
        assert rel_l2(got_ss, want_ss) <= 1e-5, k
        assert not got_img[:, :6].any()
        img_sum = got_img + img_sum
        ss_sum = got_ss + ss_sum
    assert np.array_equal(img, img_sum) and np.array_equal(ss, ss_sum)
