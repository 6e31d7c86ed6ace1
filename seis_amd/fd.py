"""Host mirror of the reference engine entry points (src/util/fd.h:3-5, 10).

Same argument order and meaning as the reference; arrays are numpy arrays in
the reference layouts (vel/dv [nx][nz], rec [nx][nt], wav [nt+ndel], gz [nx]).
Unlike the reference, which returns a freshly allocated float** (fd.cpp:27,79),
these return new numpy arrays.  The work runs on the GPU through the C ABI.
"""
import ctypes

import numpy as np

from ._lib import check, lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _prep(nt, ndel, nx, nz, gz, wav, vel):
    gz = np.ascontiguousarray(gz, np.int32)
    wav = np.ascontiguousarray(wav, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    if gz.shape != (nx,):
        raise ValueError("gz must have nx entries")
    if wav.shape[0] < nt + ndel:
        raise ValueError("wav must have nt + ndel samples")
    if vel.shape != (nx, nz):
        raise ValueError("vel must be [nx][nz]")
    return gz, wav, vel


def modeling(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel):
    """Forward shot record rec[nx][nt] (fd.cpp:15-80)."""
    gz, wav, vel = _prep(nt, ndel, nx, nz, gz, wav, vel)
    rec = np.zeros((nx, nt), np.float32)
    check(lib().seis_modeling(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, _p(gz), _p(wav), _p(vel), _p(rec)))
    return rec


def forward(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, dv):
    """Born (scattered) shot record rec[nx][nt] for perturbation dv (fd.cpp:81-177)."""
    gz, wav, vel = _prep(nt, ndel, nx, nz, gz, wav, vel)
    dv = np.ascontiguousarray(dv, np.float32)
    if dv.shape != (nx, nz):
        raise ValueError("dv must be [nx][nz]")
    rec = np.zeros((nx, nt), np.float32)
    check(lib().seis_born(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, _p(gz), _p(wav), _p(vel), _p(dv), _p(rec)))
    return rec


def adjoint(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, rec):
    """RTM image img[nx][nz] of record rec[nx][nt] (fd.cpp:592-719)."""
    gz, wav, vel = _prep(nt, ndel, nx, nz, gz, wav, vel)
    rec = np.ascontiguousarray(rec, np.float32)
    if rec.shape != (nx, nt):
        raise ValueError("rec must be [nx][nt]")
    img = np.zeros((nx, nz), np.float32)
    check(lib().seis_rtm(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, _p(gz), _p(wav), _p(vel), _p(rec), _p(img)))
    return img


def illum(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, wav, vel):
    """Source illumination ss[nx][nz] = sum over steps it > ndel of u^2
    (fd.cpp:275-339; no receivers)."""
    wav = np.ascontiguousarray(wav, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    if wav.shape[0] < nt + ndel:
        raise ValueError("wav must have nt + ndel samples")
    if vel.shape != (nx, nz):
        raise ValueError("vel must be [nx][nz]")
    ss = np.zeros((nx, nz), np.float32)
    check(lib().seis_illum(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, _p(wav), _p(vel), _p(ss)))
    return ss


def _plane_prep(nt, ndel, nx, nz, vel, sou):
    vel = np.ascontiguousarray(vel, np.float32)
    sou = np.ascontiguousarray(sou, np.float32)
    if vel.shape != (nx, nz):
        raise ValueError("vel must be [nx][nz]")
    if sou.shape != (nx, nt + ndel):
        raise ValueError("sou must be [nx][nt + ndel]")
    return vel, sou


def forward_plane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou, dv):
    """Born record rec[nx][nt] of the plane-wave line source sou[nx][nt+ndel]
    fired at row sz, receivers at row gz (forwordPlane, fd.cpp:178-274)."""
    vel, sou = _plane_prep(nt, ndel, nx, nz, vel, sou)
    dv = np.ascontiguousarray(dv, np.float32)
    rec = np.zeros((nx, nt), np.float32)
    check(lib().seis_born_plane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, _p(vel), _p(sou), _p(dv), _p(rec)))
    return rec


def illum_plane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou):
    """Illumination ss[nx][nz] of the line source (illumPlane, fd.cpp:720-786)."""
    vel, sou = _plane_prep(nt, ndel, nx, nz, vel, sou)
    ss = np.zeros((nx, nz), np.float32)
    check(lib().seis_illum_plane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, _p(vel), _p(sou), _p(ss)))
    return ss


def adjoint_plane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou, rec):
    """Plane-wave image img[nx][nz] of record rec[nx][nt] (adjointPlane,
    fd.cpp:787-921)."""
    vel, sou = _plane_prep(nt, ndel, nx, nz, vel, sou)
    rec = np.ascontiguousarray(rec, np.float32)
    if rec.shape != (nx, nt):
        raise ValueError("rec must be [nx][nt]")
    img = np.zeros((nx, nz), np.float32)
    check(lib().seis_rtm_plane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, _p(vel), _p(sou), _p(rec), _p(img)))
    return img
