"""Batched, device-resident shot processing over the C ABI.

A batch shares one model grid (nx, nz, pml, dt, dx, dz); every shot has its own
source cell, record length nt, pre-roll ndel, wavelet, receiver depths, model
(by offset into the batch's model array) and record.  This replaces the
reference's per-shot MPI worker loops (lsrtm/inversion.cpp:141-206, 365-551):
all shots advance together, one launch per time step.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib, seis_grid, seis_shot
from .device import DeviceArray


@dataclass
class Grid:
    nx: int
    nz: int
    pml: int
    dt: float
    dx: float
    dz: float

    def c(self):
        return seis_grid(self.nx, self.nz, self.pml, self.dt, self.dx, self.dz)


@dataclass
class Shot:
    sx: int
    sz: int
    nt: int
    ndel: int = 0
    model: int = 0  # index of this shot's model in the batch model stack
    nx: int = 0     # columns of the shot's own window (0: the grid's nx)
    x0: int = 0     # LSRTM plans: first image column of the window


def width(grid, s):
    """Columns of a shot's grid: its window, or the grid's nx."""
    return s.nx if s.nx > 0 else grid.nx


def _layout(grid, shots, views=False):
    """Offsets of each shot's model / wavelet / gz row / record in packed
    batch arrays.  Models are stacked by index for full-width shots; with
    windowed shots (Shot.nx > 0) each shot takes its window at its own slot
    of the packed window models, or, with views=True, its window is the
    column range [x0, x0 + nx) of one [nx][nz] model (offset x0 nz)."""
    arr = (seis_shot * len(shots))()
    wav_off = gz_off = rec_off = mod_off = 0
    windowed = any(s.nx > 0 for s in shots)
    for j, s in enumerate(shots):
        nx = width(grid, s)
        if views:
            model = s.x0 * grid.nz
        else:
            model = mod_off if windowed else s.model * grid.nx * grid.nz
        arr[j] = seis_shot(s.sx, s.sz, s.nt, s.ndel, model, wav_off, gz_off, rec_off, s.nx, s.x0)
        wav_off += s.nt + s.ndel
        gz_off += nx
        rec_off += nx * s.nt
        mod_off += nx * grid.nz
    return arr, wav_off, gz_off, rec_off


def _dev(a, dtype):
    if isinstance(a, DeviceArray):
        return a, None
    d = DeviceArray.from_host(np.ascontiguousarray(a, dtype))
    return d, d


def pack(grid, shots, wavs, gzs):
    """Concatenate per-shot wavelets and receiver-depth rows in batch order.
    Every wavelet needs nt + ndel samples and every gz row one entry per
    column of its shot; anything else is a ValueError (the engine reads
    device memory at offsets derived from these lengths)."""
    if len(wavs) != len(shots) or len(gzs) != len(shots):
        raise ValueError("one wavelet and one gz row per shot")
    for j, (w, g, s) in enumerate(zip(wavs, gzs, shots)):
        if np.size(w) < s.nt + s.ndel:
            raise ValueError("shot %d: wavelet has %d samples, needs nt + ndel = %d" % (j, np.size(w), s.nt + s.ndel))
        if np.size(g) != width(grid, s):
            raise ValueError("shot %d: gz has %d entries, needs %d (one per column)" % (j, np.size(g), width(grid, s)))
    wav = np.concatenate([np.asarray(w, np.float32).ravel()[: s.nt + s.ndel] for w, s in zip(wavs, shots)])
    gz = np.concatenate([np.asarray(g, np.int32).ravel() for g in gzs])
    return wav, gz


def _check_models(grid, shots, vel, name="vel"):
    """A host model array must hold every model the shots index: the
    (max index + 1) stacked models of full-width shots, or each windowed
    shot's own [nx][nz] window in shot order."""
    if isinstance(vel, DeviceArray):
        return
    if any(s.nx > 0 for s in shots):
        need = sum(width(grid, s) * grid.nz for s in shots)
    else:
        need = (max(s.model for s in shots) + 1) * grid.nx * grid.nz
    if np.size(vel) < need:
        raise ValueError("%s has %d values, the shots need %d" % (name, np.size(vel), need))


def split_records(grid, shots, rec):
    """Per-shot [nx][nt] views of a packed record array."""
    out, off = [], 0
    for s in shots:
        n = width(grid, s) * s.nt
        out.append(rec[off:off + n].reshape(width(grid, s), s.nt))
        off += n
    return out


def split_images(grid, shots, img):
    """Per-shot images of a packed image array: [nshots][nx][nz] when every
    shot has the grid's width, else a list of [nx_s][nz] arrays."""
    if all(width(grid, s) == grid.nx for s in shots):
        return img.reshape(len(shots), grid.nx, grid.nz)
    out, off = [], 0
    for s in shots:
        n = width(grid, s) * grid.nz
        out.append(img[off:off + n].reshape(width(grid, s), grid.nz))
        off += n
    return out


def set_batch_limits(nbytes=0, shots=0):
    """Cap the device working set of one batch chunk (bytes, 0: free memory
    less a reserve) and the shots per launch (0: no cap)."""
    check(lib().seis_set_batch_limits(int(nbytes), int(shots)))


PLAN_MODEL, PLAN_BORN, PLAN_RTM, PLAN_ILLUM, PLAN_RTM_PLANE = 0, 1, 2, 3, 4


def batch_plan(mode, grid, shots, budget, max_shots=0):
    """Chunks [start, end) a batch call would run for `budget` bytes (host only)."""
    arr, _, _, _ = _layout(grid, shots)
    starts = (ctypes.c_int * (len(shots) + 1))()
    n = ctypes.c_int(0)
    g = grid.c()
    check(lib().seis_batch_plan(int(mode), ctypes.byref(g), len(shots), arr, int(budget), int(max_shots), starts,
                                ctypes.byref(n)))
    return [(starts[i], starts[i + 1]) for i in range(n.value)]


RTM_ISOLATE_BAND = 1  # SEIS_RTM_ISOLATE_BAND of include/seis.h


def _run(fn, grid, shots, vel, extra, wav, gz, out_elems, extra_dtype=np.float32, rec_in=None, flags=0):
    arr, nw, ng, nr = _layout(grid, shots)
    _check_models(grid, shots, vel)
    if extra is not None:
        _check_models(grid, shots, extra, "dv")
    keep = []
    dvel, k = _dev(vel, np.float32); keep.append(k)
    dwav, k = _dev(wav, np.float32); keep.append(k)
    dgz, k = _dev(gz, np.int32); keep.append(k)
    g = grid.c()
    nimg = sum(width(grid, s) for s in shots) * grid.nz
    if fn == "rtm":
        if not isinstance(rec_in, DeviceArray) and np.size(rec_in) != nr:
            raise ValueError("records have %d samples, the shots need %d" % (np.size(rec_in), nr))
        drec, k = _dev(rec_in, np.float32); keep.append(k)
        dimg = DeviceArray((nimg,), np.float32, zero=True)
        check(lib().seis_rtm_batch_ex(ctypes.byref(g), len(shots), arr, dvel.ptr, dwav.ptr, dgz.ptr, drec.ptr,
                                      dimg.ptr, int(flags)))
        return split_images(grid, shots, dimg.download())
    drec = DeviceArray((nr,), np.float32, zero=True)
    if fn == "model":
        check(lib().seis_modeling_batch(ctypes.byref(g), len(shots), arr, dvel.ptr, dwav.ptr, dgz.ptr, drec.ptr))
    else:
        ddv, k = _dev(extra, np.float32); keep.append(k)
        check(lib().seis_born_batch(ctypes.byref(g), len(shots), arr, dvel.ptr, ddv.ptr, dwav.ptr, dgz.ptr,
                                    drec.ptr))
    return split_records(grid, shots, drec.download())


def modeling_batch(grid, shots, vel, wavs, gzs):
    """Forward records of every shot; vel is [nmodels][nx][nz] (or [nx][nz]),
    or, with windowed shots (Shot.nx > 0), each shot's window in shot order."""
    wav, gz = pack(grid, shots, wavs, gzs)
    return _run("model", grid, shots, vel, None, wav, gz, None)


def born_batch(grid, shots, vel, dv, wavs, gzs):
    """Born records of every shot; dv indexed like vel."""
    wav, gz = pack(grid, shots, wavs, gzs)
    return _run("born", grid, shots, vel, dv, wav, gz, None)


def rtm_batch(grid, shots, vel, wavs, gzs, recs, isolate_band=False):
    """Reference-semantics RTM image of every shot: [nshots][nx][nz] (a list
    of [nx_s][nz] with windowed shots).
    isolate_band holds the reconstructed source field at zero in the absorbing
    band (SEIS_RTM_ISOLATE_BAND), which keeps long records in fast media
    finite; the image then differs from the reference's only near the edges."""
    wav, gz = pack(grid, shots, wavs, gzs)
    for j, (r, s) in enumerate(zip(recs, shots)):
        if np.size(r) != width(grid, s) * s.nt:
            raise ValueError("shot %d: record has %d samples, needs nx * nt = %d" % (j, np.size(r), width(grid, s) * s.nt))
    rec = np.concatenate([np.ascontiguousarray(r, np.float32).ravel() for r in recs])
    return _run("rtm", grid, shots, vel, None, wav, gz, None, rec_in=rec,
                flags=RTM_ISOLATE_BAND if isolate_band else 0)


def illum_batch(grid, shots, vel, wavs):
    """Source illumination of every shot: [nshots][nx][nz] (illum_MPI's
    per-shot ILLUM<is>, which it then sums)."""
    for j, (w, s) in enumerate(zip(wavs, shots)):
        if np.size(w) < s.nt + s.ndel:
            raise ValueError("shot %d: wavelet has %d samples, needs nt + ndel = %d" % (j, np.size(w), s.nt + s.ndel))
    wav = np.concatenate([np.asarray(w, np.float32).ravel()[: s.nt + s.ndel] for w, s in zip(wavs, shots)])
    arr, nw, ng, nr = _layout(grid, shots)
    _check_models(grid, shots, vel)
    dvel, kv = _dev(vel, np.float32)
    dwav, kw = _dev(wav, np.float32)
    dss = DeviceArray((sum(width(grid, s) for s in shots) * grid.nz,), np.float32, zero=True)
    g = grid.c()
    check(lib().seis_illum_batch(ctypes.byref(g), len(shots), arr, dvel.ptr, dwav.ptr, dss.ptr))
    return split_images(grid, shots, dss.download())


def _plane_batch(fn, grid, shots, vel, sous, gz_rows, extra=None, recs=None):
    arr, nw, ng, nr = _layout(grid, shots)
    sou = np.concatenate([np.ascontiguousarray(u, np.float32).ravel() for u in sous])
    assert sou.shape[0] == nw * grid.nx
    # the layout's wavelet offsets count samples; sou blocks hold nx of them
    for j in range(len(shots)):
        arr[j].wav *= grid.nx
    gz = np.concatenate([np.full(grid.nx, r, np.int32) for r in gz_rows])
    keep = []
    dvel, k = _dev(vel, np.float32); keep.append(k)
    dsou, k = _dev(sou, np.float32); keep.append(k)
    dgz, k = _dev(gz, np.int32); keep.append(k)
    g = grid.c()
    n = len(shots)
    if fn == "born":
        ddv, k = _dev(extra, np.float32); keep.append(k)
        drec = DeviceArray((nr,), np.float32, zero=True)
        check(lib().seis_born_plane_batch(ctypes.byref(g), n, arr, dvel.ptr, ddv.ptr, dsou.ptr, dgz.ptr, drec.ptr))
        return split_records(grid, shots, drec.download())
    dout = DeviceArray((n * grid.nx * grid.nz,), np.float32, zero=True)
    if fn == "illum":
        check(lib().seis_illum_plane_batch(ctypes.byref(g), n, arr, dvel.ptr, dsou.ptr, dout.ptr))
    else:
        rec = np.concatenate([np.ascontiguousarray(r, np.float32).ravel() for r in recs])
        drec, k = _dev(rec, np.float32); keep.append(k)
        check(lib().seis_rtm_plane_batch(ctypes.byref(g), n, arr, dvel.ptr, dsou.ptr, dgz.ptr, drec.ptr, dout.ptr))
    return dout.download().reshape(n, grid.nx, grid.nz)


def born_plane_batch(grid, shots, vel, sous, gz_rows, dv):
    """Plane-wave Born records of every shot (sous[j]: [nx][nt+ndel] fired at
    row shots[j].sz; receivers at row gz_rows[j])."""
    return _plane_batch("born", grid, shots, vel, sous, gz_rows, extra=dv)


def illum_plane_batch(grid, shots, vel, sous):
    """Plane-wave illumination of every shot: [nshots][nx][nz]."""
    return _plane_batch("illum", grid, shots, vel, sous, [0] * len(shots))


def rtm_plane_batch(grid, shots, vel, sous, gz_rows, recs):
    """Plane-wave images of every shot (adjointPlane): [nshots][nx][nz]."""
    return _plane_batch("rtm", grid, shots, vel, sous, gz_rows, recs=recs)
