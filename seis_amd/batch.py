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


def _layout(grid, shots):
    """Offsets of each shot's wavelet / gz row / record in packed batch arrays."""
    arr = (seis_shot * len(shots))()
    wav_off = gz_off = rec_off = 0
    for j, s in enumerate(shots):
        arr[j] = seis_shot(s.sx, s.sz, s.nt, s.ndel, s.model * grid.nx * grid.nz, wav_off, gz_off, rec_off)
        wav_off += s.nt + s.ndel
        gz_off += grid.nx
        rec_off += grid.nx * s.nt
    return arr, wav_off, gz_off, rec_off


def _dev(a, dtype):
    if isinstance(a, DeviceArray):
        return a, None
    d = DeviceArray.from_host(np.ascontiguousarray(a, dtype))
    return d, d


def pack(grid, shots, wavs, gzs):
    """Concatenate per-shot wavelets and receiver-depth rows in batch order."""
    wav = np.concatenate([np.asarray(w, np.float32)[: s.nt + s.ndel] for w, s in zip(wavs, shots)])
    gz = np.concatenate([np.asarray(g, np.int32) for g in gzs])
    assert gz.shape[0] == grid.nx * len(shots)
    return wav, gz


def split_records(grid, shots, rec):
    """Per-shot [nx][nt] views of a packed record array."""
    out, off = [], 0
    for s in shots:
        n = grid.nx * s.nt
        out.append(rec[off:off + n].reshape(grid.nx, s.nt))
        off += n
    return out


RTM_ISOLATE_BAND = 1  # SEIS_RTM_ISOLATE_BAND of include/seis.h


def _run(fn, grid, shots, vel, extra, wav, gz, out_elems, extra_dtype=np.float32, rec_in=None, flags=0):
    arr, nw, ng, nr = _layout(grid, shots)
    keep = []
    dvel, k = _dev(vel, np.float32); keep.append(k)
    dwav, k = _dev(wav, np.float32); keep.append(k)
    dgz, k = _dev(gz, np.int32); keep.append(k)
    g = grid.c()
    if fn == "rtm":
        drec, k = _dev(rec_in, np.float32); keep.append(k)
        dimg = DeviceArray((len(shots) * grid.nx * grid.nz,), np.float32, zero=True)
        check(lib().seis_rtm_batch_ex(ctypes.byref(g), len(shots), arr, dvel.ptr, dwav.ptr, dgz.ptr, drec.ptr,
                                      dimg.ptr, int(flags)))
        return dimg.download().reshape(len(shots), grid.nx, grid.nz)
    drec = DeviceArray((nr,), np.float32, zero=True)
    if fn == "model":
        check(lib().seis_modeling_batch(ctypes.byref(g), len(shots), arr, dvel.ptr, dwav.ptr, dgz.ptr, drec.ptr))
    else:
        ddv, k = _dev(extra, np.float32); keep.append(k)
        check(lib().seis_born_batch(ctypes.byref(g), len(shots), arr, dvel.ptr, ddv.ptr, dwav.ptr, dgz.ptr,
                                    drec.ptr))
    return split_records(grid, shots, drec.download())


def modeling_batch(grid, shots, vel, wavs, gzs):
    """Forward records of every shot; vel is [nmodels][nx][nz] (or [nx][nz])."""
    wav, gz = pack(grid, shots, wavs, gzs)
    return _run("model", grid, shots, vel, None, wav, gz, None)


def born_batch(grid, shots, vel, dv, wavs, gzs):
    """Born records of every shot; dv indexed like vel."""
    wav, gz = pack(grid, shots, wavs, gzs)
    return _run("born", grid, shots, vel, dv, wav, gz, None)


def rtm_batch(grid, shots, vel, wavs, gzs, recs, isolate_band=False):
    """Reference-semantics RTM image of every shot: [nshots][nx][nz].
    isolate_band holds the reconstructed source field at zero in the absorbing
    band (SEIS_RTM_ISOLATE_BAND), which keeps long records in fast media
    finite; the image then differs from the reference's only near the edges."""
    wav, gz = pack(grid, shots, wavs, gzs)
    rec = np.concatenate([np.ascontiguousarray(r, np.float32).ravel() for r in recs])
    return _run("rtm", grid, shots, vel, None, wav, gz, None, rec_in=rec,
                flags=RTM_ISOLATE_BAND if isolate_band else 0)


def illum_batch(grid, shots, vel, wavs):
    """Source illumination of every shot: [nshots][nx][nz] (illum_MPI's
    per-shot ILLUM<is>, which it then sums)."""
    wav = np.concatenate([np.asarray(w, np.float32)[: s.nt + s.ndel] for w, s in zip(wavs, shots)])
    arr, nw, ng, nr = _layout(grid, shots)
    dvel, kv = _dev(vel, np.float32)
    dwav, kw = _dev(wav, np.float32)
    dss = DeviceArray((len(shots) * grid.nx * grid.nz,), np.float32, zero=True)
    g = grid.c()
    check(lib().seis_illum_batch(ctypes.byref(g), len(shots), arr, dvel.ptr, dwav.ptr, dss.ptr))
    return dss.download().reshape(len(shots), grid.nx, grid.nz)


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
