"""Least-squares RTM plans over the C ABI (include/seis.h, seis_lsrtm_*).

The reference declares its LSRTM operator and CG solver (conope_MPI,
cgsolver_MPI, lsrtm/inversion.h:81,87) without defining them.  A plan holds
the per-step Born-source snapshots of the background for a batch of shots and
applies L (Born records of an image perturbation), its exact transpose L^T and
CGLS on min ||L x - d||^2, all on the GPU.
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .batch import _layout
from .device import DeviceArray

SNAP_FP32, SNAP_BF16, SNAP_FP16 = 0, 1, 2
ADJ_EXACT, ADJ_XCORR = 0, 1  # exact transpose / the reference's cross-correlation imaging


class LsrtmPlan:
    """A plan over shots of one model vel [nx][nz]; a shot with Shot.nx > 0
    propagates on its window, columns [Shot.x0, Shot.x0 + Shot.nx) of the
    model, and its records are [Shot.nx][nt]."""

    def __init__(self, grid, shots, vel, wavs, gzs, snapshot=SNAP_BF16, fp64=False):
        self.grid, self.shots, self.fp64 = grid, shots, fp64
        self.dtype = np.float64 if fp64 else np.float32
        # every shot's velocity is its window [x0, x0 + nx) of the one model
        arr, nw, ng, nr = _layout(grid, shots, views=True)
        self._arr = arr
        self.nrec = nr
        self.model = (grid.nx, grid.nz)
        wav = np.concatenate([np.asarray(w)[: s.nt + s.ndel] for w, s in zip(wavs, shots)]).astype(self.dtype)
        gz = np.concatenate([np.asarray(g, np.int32) for g in gzs])
        self._vel = DeviceArray.from_host(np.ascontiguousarray(vel, self.dtype))
        self._wav = DeviceArray.from_host(wav)
        self._gz = DeviceArray.from_host(gz)
        g = grid.c()
        h = ctypes.c_void_p()
        if fp64:
            check(lib().seis_lsrtm_create_f64(ctypes.byref(h), ctypes.byref(g), len(shots), arr, self._vel.ptr,
                                              self._wav.ptr, self._gz.ptr))
        else:
            check(lib().seis_lsrtm_create(ctypes.byref(h), ctypes.byref(g), len(shots), arr, self._vel.ptr,
                                          self._wav.ptr, self._gz.ptr, int(snapshot)))
        self.handle = h

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().seis_lsrtm_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _dev(self, a, shape):
        if isinstance(a, DeviceArray):
            return a
        return DeviceArray.from_host(np.ascontiguousarray(a, self.dtype).reshape(shape))

    def born(self, pert):
        """Packed Born records L pert (host array of nrec samples)."""
        dp = self._dev(pert, self.model)
        rec = DeviceArray((self.nrec,), self.dtype, zero=True)
        check(lib().seis_lsrtm_born(self.handle, dp.ptr, rec.ptr))
        return rec.download()

    def adjoint(self, rec):
        """Image L^T rec (host [nx][nz])."""
        dr = self._dev(rec, (self.nrec,))
        img = DeviceArray(self.model, self.dtype, zero=True)
        check(lib().seis_lsrtm_adjoint(self.handle, dr.ptr, img.ptr))
        return img.download()

    def dot(self, a, b):
        da, db = self._dev(a, (np.size(a),)), self._dev(b, (np.size(b),))
        out = ctypes.c_double()
        check(lib().seis_lsrtm_dot(self.handle, da.ptr, db.ptr, int(np.size(a)), ctypes.byref(out)))
        return out.value

    def set_adjoint(self, kind):
        """ADJ_EXACT (the exact transpose, default) or ADJ_XCORR (the
        reference's imaging condition on the snapshots) for adjoint() and cg()."""
        check(lib().seis_lsrtm_set_adjoint(self.handle, int(kind)))

    def set_weight(self, weight):
        """Image-space preconditioner W ([nx][nz]) of cg(): CGLS on L W, image
        W x; None clears."""
        if weight is None:
            self._weight = None
            check(lib().seis_lsrtm_set_weight(self.handle, None))
            return
        self._weight = self._dev(weight, self.model)
        check(lib().seis_lsrtm_set_weight(self.handle, self._weight.ptr))

    def set_mask(self, mask):
        """Record mask (packed like the records) applied in L and L^T; None clears."""
        if mask is None:
            self._mask = None
            check(lib().seis_lsrtm_set_mask(self.handle, None))
            return
        self._mask = self._dev(mask, (self.nrec,))
        check(lib().seis_lsrtm_set_mask(self.handle, self._mask.ptr))

    def cg(self, dobs, niter, rtol=0.0, iterates=False):
        """CGLS image and history [(||r||^2, ||L^T r||^2)] per completed
        iteration; with iterates=True also the image after every iteration
        ([iterations][nx][nz])."""
        dd = self._dev(dobs, (self.nrec,))
        img = DeviceArray(self.model, self.dtype, zero=True)
        hist = (ctypes.c_double * (2 * (niter + 1)))()
        done = ctypes.c_int(0)
        its = DeviceArray((max(niter, 1),) + self.model, self.dtype, zero=True) if iterates else None
        check(lib().seis_lsrtm_cg_history(self.handle, dd.ptr, int(niter), float(rtol), img.ptr, hist,
                                          ctypes.byref(done), its.ptr if its is not None else None))
        h = np.array(hist[:]).reshape(-1, 2)[: done.value + 1]
        if iterates:
            return img.download(), h, its.download()[: done.value]
        return img.download(), h
