"""Plane-wave synthesis of shot gathers on the GPU.

``harmonic`` is the reference's phase shift of one trace
(src/util/InitInversion.cpp:141-151) and ``plane_prepare`` its
planeWavePrepare (src/lsrtm/inversion.cpp:808-912): for every ray parameter
p_k = pmin + k*dp, the plane-wave source sou_k [nx][ntw] and record
rec_k [nx][nt] that the plane-wave engine (``seis_amd.fd.adjoint_plane``,
``illum_plane``, ``forward_plane``) takes.
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .device import DeviceArray


def harmonic(x, kx, xs):
    """cos(kx*xs) x - sin(kx*xs) H[x], H the 201-tap Hilbert filter (float32)."""
    x = np.ascontiguousarray(x, np.float32)
    if x.ndim != 1:
        raise ValueError("harmonic takes one trace")
    out = np.zeros_like(x)
    check(lib().seis_harmonic(x.ctypes.data, out.ctypes.data, x.shape[0], float(kx), float(xs)))
    return out


def plane_prepare(nx, nt, np_, pmin, dp, sx, isx, traces, cols, wav, download=True):
    """planeWavePrepare for np_ ray parameters.

    sx[s]      source x of shot s (metres); isx[s] its model column;
    traces[s]  the gather of shot s, [ng_s][nt] float32 (host array or DeviceArray
               of the concatenated gathers, then cols gives the split);
    cols[s]    model column of each trace of shot s (-1: outside the model);
    wav        the source wavelet [ntw].
    Returns sou [np_][nx][ntw] and rec [np_][nx][nt] (numpy, or DeviceArrays
    when download is False).
    """
    sx = np.ascontiguousarray(sx, np.float32)
    isx = np.ascontiguousarray(isx, np.int32)
    ns = sx.shape[0]
    cols = [np.ascontiguousarray(c, np.int32).reshape(-1) for c in cols]
    ng = np.array([c.shape[0] for c in cols], np.int32)
    col = np.concatenate(cols) if ns else np.zeros(0, np.int32)
    wav = np.ascontiguousarray(wav, np.float32)
    ntw = wav.shape[0]
    if np_ == 0:
        return np.zeros((0, nx, ntw), np.float32), np.zeros((0, nx, nt), np.float32)
    if isinstance(traces, DeviceArray):
        ddata = traces
    else:
        data = np.concatenate([np.ascontiguousarray(t, np.float32).reshape(-1, nt) for t in traces]) if ns \
            else np.zeros((0, nt), np.float32)
        if data.shape[0] != int(ng.sum()):
            raise ValueError("traces and cols disagree on the trace counts")
        ddata = DeviceArray.from_host(data if data.size else np.zeros(1, np.float32))
    dwav = DeviceArray.from_host(wav)
    sou = DeviceArray((np_, nx, ntw), np.float32)
    rec = DeviceArray((np_, nx, nt), np.float32)
    P = ctypes.c_void_p
    check(lib().seis_plane_prepare(nx, nt, ntw, ns, np_, float(pmin), float(dp), sx.ctypes.data_as(P),
                                   isx.ctypes.data_as(P), ng.ctypes.data_as(P), col.ctypes.data_as(P), dwav.ptr,
                                   ddata.ptr, sou.ptr, rec.ptr))
    if not download:
        return sou, rec
    return sou.download(), rec.download()
