"""Restatement of the inversion driver's per-shot glue (test infrastructure).

Follows lsrtm/inversion.cpp of the reference with its float/double promotions
(numpy float32 scalars for C float, Python floats for C double):
  window        getVel       inversion.cpp:552-598
  receiver_gz   getIgz       inversion.cpp:600-625
  reorder       swapReord    inversion.cpp:626-643
  to_global     swapModel    inversion.cpp:644-665
  remove_direct removeDirect inversion.cpp:699-752
  source_cell                inversion.cpp:186-187
  plane_geometry planeWavePrepare's columns  inversion.cpp:845-893
  plane_rows     the plane-wave source / receiver rows  inversion.cpp:331-332
The reference's inversion.cpp needs MPI (absent here), so it cannot be run;
these are pinned by the hand-computed cases in tests/test_driver_glue.py.
"""
import numpy as np

f32 = np.float32


def window(p, sx, gx):
    dx = f32(p["dx"])
    gx = np.asarray(gx, np.float32)
    gx_min, gx_max = f32(gx.min()), f32(gx.max())
    x_min, x_max = min(f32(sx), gx_min), max(f32(sx), gx_max)
    x_min_pad = f32(x_min - f32(f32(p["lpad"]) * dx))
    x_max_pad = f32(x_max + f32(f32(p["rpad"]) * dx))
    velfx = f32(p["velfx"])
    x_min_vel = f32(velfx + dx)
    x_max_vel = f32(f32(velfx + f32(f32(p["nx"] - 1) * dx)) - dx)
    fx = x_min_pad
    nxw = int(f32(f32(f32(x_max_pad - x_min_pad) / dx) + f32(1)))
    cols = []
    for ix in range(nxw):
        xcor = f32(fx + f32(f32(ix) * dx))
        if x_min_vel <= xcor <= x_max_vel:
            xloc = int(f32(f32(xcor - velfx) / dx))
        elif xcor < x_min_vel:
            xloc = int(f32(f32(x_min_vel - velfx) / dx))
        else:
            xloc = int(f32(f32(x_max_vel - velfx) / dx))
        cols.append(xloc)
    return fx, nxw, np.array(cols)


def _edges(fx, nxw, dx):
    x_min = f32(float(fx) - 0.001 * float(dx))
    x_max = f32(float(f32(fx + f32(f32(nxw - 1) * dx))) + 0.001 * float(dx))
    return x_min, x_max


def receiver_gz(p, fx, nxw, gx, gz):
    dx, dz = f32(p["dx"]), f32(p["dz"])
    x_min, x_max = _edges(fx, nxw, dx)
    igz = np.zeros(nxw, np.int32)
    for x, z in zip(np.asarray(gx, np.float32), np.asarray(gz, np.float32)):
        if x_min <= x <= x_max:
            xloc = int(f32(f32(f32(x - x_min) / dx) + f32(0.1)))
            zloc = int(f32(f32(f32(z - f32(0)) / dz) + f32(0.1)))
            igz[xloc] = zloc
    return igz


def source_cell(p, fx, sx, sz):
    ix = int(f32(f32(f32(f32(sx) - fx) / f32(p["dx"])) + f32(0.0001)))
    iz = int((float(f32(sz)) - 0.0) / float(f32(p["dz"])) + float(f32(0.0001)))
    return ix, iz


def reorder(p, fx, nxw, gx, nt, csg=None, rec=None, to_gather=True):
    dx = f32(p["dx"])
    x_min, x_max = _edges(fx, nxw, dx)
    if to_gather:
        out = np.zeros((len(gx), nt), np.float32)
    else:
        out = np.zeros((nxw, nt), np.float32)
    for ig, g in enumerate(np.asarray(gx, np.float32)):
        if x_min <= g <= x_max:
            xloc = int(f32(f32(f32(g - x_min) / dx) + f32(0.1)))
            if to_gather:
                out[ig] = rec[xloc]
            else:
                out[xloc] = csg[ig]
    return out


def to_global(p, m, mfx, dfx, nxd, nz):
    dx = f32(p["dx"])
    nxm = m.shape[0]
    x_min, x_max = _edges(mfx, nxm, dx)
    d = np.zeros((nxd, nz), np.float32)
    for ix in range(4, nxd - 4):
        xcor = f32(f32(dfx) + f32(f32(ix - 1) * dx))
        if x_min <= xcor <= x_max:
            xloc = int(f32(f32(f32(xcor - x_min) / dx) + f32(0.0001)))
            if 4 <= xloc <= nxm - 4:
                d[ix, 4:nz - 4] += m[xloc, 4:nz - 4]
    return d


def remove_direct(p, csg, v, sx, sz, gx, gz):
    csg = csg.copy()
    nt, nx = p["nt"], p["nx"]
    dx, dz, dt, fr = f32(p["dx"]), f32(p["dz"]), f32(p["dt"]), f32(p["fr"])
    velfx = f32(p["velfx"])
    delay = p["delay"]
    isx = int(f32(f32(f32(f32(sx) - velfx) / dx) + f32(0.0001)))
    for ig in range(len(gx)):
        t = f32(0)
        g, zg = f32(gx[ig]), f32(gz[ig])
        igx = int(f32(f32(f32(g - velfx) / dx) + f32(0.0001)))
        igz = int(f32(f32(f32(zg - f32(0)) / dz) + f32(0.0001)))
        adddelay = int(abs(f32(f32(f32(zg - f32(sz)) / dz) + f32(0.0001)))) + 10
        rng = range(igx, isx) if igx <= isx else range(igx, isx, -1)
        for i in rng:
            if 0 <= i < nx:
                t = f32(t + f32(dx / f32(v[i, igz])))
            elif igx < 0:
                t = f32(t + f32(dx / f32(v[0, igz])))
            elif igx >= nx:
                t = f32(t + f32(dx / f32(v[nx - 1, igz])))
        s = f32(f32(f32(t / dt) + f32(delay)) + f32(adddelay))
        tail = float(f32(f32(f32(f32(1) / fr) / dt) * f32(0.5))) * 5.0
        it = int((float(s) + tail) + float(f32(0.0001)))
        it = min(it, nt - 1)
        it = max(it, 1)
        csg[ig, :it] = 0
    return csg


def plane_geometry(p, sxs, gxs):
    """Source column of every shot and model column of every trace (-1
    outside the model) of planeWavePrepare, over the whole model."""
    dx, velfx = f32(p["dx"]), f32(p["velfx"])
    isx = [int(f32(f32(f32(f32(sx) - velfx) / dx) + f32(0.001))) for sx in sxs]
    x_min, x_max = _edges(velfx, p["nx"], dx)
    cols = []
    for gx in gxs:
        c = []
        for g in np.asarray(gx, np.float32):
            c.append(int(f32(f32(f32(g - x_min) / dx) + f32(0.1))) if x_min <= g <= x_max else -1)
        cols.append(np.array(c, np.int32))
    return np.array(isx, np.int32), cols


def plane_rows(p, sz0, gz0):
    dz = float(f32(p["dz"]))
    return int((float(f32(sz0)) - 0.0) / dz + float(f32(0.0001))), int((float(f32(gz0)) - 0.0) / dz + float(f32(0.0001)))
