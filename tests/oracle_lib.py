"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, where it has
been built, to the compiled reference engine (oracle/_ref/libseisref.so).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use these as checkers / CPU baselines.  Arrays follow the
reference layouts: models [nx][nz] (z fastest), records [nx][nt] (t fastest).
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libseisref.so")

_P = ctypes.c_void_p
_I = ctypes.c_int


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _Lib:
    def __init__(self, path, prefix, typed):
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        self.typed = typed  # oracle has _f32/_f64 variants, reference is float only

    def _fn(self, name, dtype):
        if self.typed:
            suf = "_f32" if dtype == np.float32 else "_f64"
            return getattr(self.lib, self.prefix + name + suf)
        if dtype != np.float32:
            raise TypeError("reference engine is float32 only")
        return getattr(self.lib, self.prefix + name)

    def set_threads(self, n):
        getattr(self.lib, self.prefix + "set_threads")(_I(int(n)))

    @staticmethod
    def _scalar(dtype):
        return ctypes.c_float if dtype == np.float32 else ctypes.c_double

    def ricker(self, dt, fr, num_delay, n, dtype=np.float32):
        out = np.zeros(n, dtype)
        S = self._scalar(dtype)
        self._fn("ricker", dtype)(S(dt), S(fr), _I(num_delay), _I(n), _ptr(out))
        return out

    def _geom_call(self, name, dtype, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz):
        S = self._scalar(dtype)
        gz = np.ascontiguousarray(gz, np.int32)
        assert gz.shape == (nx,)
        return self._fn(name, dtype), [S(dt), S(dx), S(dz), _I(nt), _I(ndel), _I(nx), _I(nz), _I(pml),
                                       _I(sx), _I(sz), _ptr(gz)], gz

    def modeling(self, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, dtype=np.float32):
        f, args, keep = self._geom_call("modeling", dtype, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz)
        wav = np.ascontiguousarray(wav, dtype)
        vel = np.ascontiguousarray(vel, dtype)
        assert wav.shape[0] >= nt + ndel and vel.shape == (nx, nz)
        rec = np.zeros((nx, nt), dtype)
        f(*args, _ptr(wav), _ptr(vel), _ptr(rec))
        return rec

    def born(self, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, dv, dtype=np.float32):
        f, args, keep = self._geom_call("born", dtype, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz)
        wav = np.ascontiguousarray(wav, dtype)
        vel = np.ascontiguousarray(vel, dtype)
        dv = np.ascontiguousarray(dv, dtype)
        assert vel.shape == (nx, nz) and dv.shape == (nx, nz)
        rec = np.zeros((nx, nt), dtype)
        f(*args, _ptr(wav), _ptr(vel), _ptr(dv), _ptr(rec))
        return rec

    def rtm(self, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, rec, dtype=np.float32, isolate=False):
        """isolate: the engine's isolated-band variant (source field held at
        zero outside the model area); the oracle library only."""
        f, args, keep = self._geom_call("rtm_isolated" if isolate else "rtm", dtype, dt, dx, dz, nt, ndel, nx, nz,
                                        pml, sx, sz, gz)
        wav = np.ascontiguousarray(wav, dtype)
        vel = np.ascontiguousarray(vel, dtype)
        rec = np.ascontiguousarray(rec, dtype)
        assert vel.shape == (nx, nz) and rec.shape == (nx, nt)
        img = np.zeros((nx, nz), dtype)
        f(*args, _ptr(wav), _ptr(vel), _ptr(rec), _ptr(img))
        return img


    def illum(self, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, wav, vel, dtype=np.float32):
        """Source illumination ss[nx][nz] (fd.h:10; no receivers)."""
        S = self._scalar(dtype)
        wav = np.ascontiguousarray(wav, dtype)
        vel = np.ascontiguousarray(vel, dtype)
        assert wav.shape[0] >= nt + ndel and vel.shape == (nx, nz)
        ss = np.zeros((nx, nz), dtype)
        self._fn("illum", dtype)(S(dt), S(dx), S(dz), _I(nt), _I(ndel), _I(nx), _I(nz), _I(pml), _I(sx), _I(sz),
                                 _ptr(wav), _ptr(vel), _ptr(ss))
        return ss


    # plane-wave engine (fd.h:6-7, 11): sou is [nx][nt+ndel], receivers at row gz
    def _plane_args(self, dtype, dt, dx, dz, nt, ndel, nx, nz, pml, sz, vel, sou):
        S = self._scalar(dtype)
        vel = np.ascontiguousarray(vel, dtype)
        sou = np.ascontiguousarray(sou, dtype)
        assert vel.shape == (nx, nz) and sou.shape == (nx, nt + ndel)
        head = [S(dt), S(dx), S(dz), _I(nt), _I(ndel), _I(nx), _I(nz), _I(pml), _I(sz)]
        return head, vel, sou

    def born_plane(self, dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou, dv, dtype=np.float32):
        head, vel, sou = self._plane_args(dtype, dt, dx, dz, nt, ndel, nx, nz, pml, sz, vel, sou)
        dv = np.ascontiguousarray(dv, dtype)
        rec = np.zeros((nx, nt), dtype)
        self._fn("born_plane", dtype)(*head, _I(gz), _ptr(vel), _ptr(sou), _ptr(dv), _ptr(rec))
        return rec

    def illum_plane(self, dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou, dtype=np.float32):
        head, vel, sou = self._plane_args(dtype, dt, dx, dz, nt, ndel, nx, nz, pml, sz, vel, sou)
        ss = np.zeros((nx, nz), dtype)
        extra = [] if self.typed else [_I(gz)]  # the reference's illumPlane takes (and ignores) gz
        self._fn("illum_plane", dtype)(*head, *extra, _ptr(vel), _ptr(sou), _ptr(ss))
        return ss

    def adjoint_plane(self, dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou, rec, dtype=np.float32):
        head, vel, sou = self._plane_args(dtype, dt, dx, dz, nt, ndel, nx, nz, pml, sz, vel, sou)
        rec = np.ascontiguousarray(rec, dtype)
        assert rec.shape == (nx, nt)
        img = np.zeros((nx, nz), dtype)
        self._fn("adjoint_plane", dtype)(*head, _I(gz), _ptr(vel), _ptr(sou), _ptr(rec), _ptr(img))
        return img

    def harmonic(self, x, kx, xs, dtype=np.float32):
        """cos(kx xs) x - sin(kx xs) H[x] with the 201-tap Hilbert filter."""
        S = self._scalar(dtype)
        x = np.ascontiguousarray(x, dtype)
        out = np.zeros_like(x)
        self._fn("harmonic", dtype)(_ptr(x), _ptr(out), _I(x.shape[0]), S(kx), S(xs))
        return out


class Oracle(_Lib):
    def __init__(self, path=ORACLE_SO):
        super().__init__(path, "oracle_", True)

    def born_adjoint(self, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, rec, dtype=np.float32):
        """grad[nx][nz] = L^T rec, the exact transpose of born()."""
        f, args, keep = self._geom_call("born_adjoint", dtype, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz)
        wav = np.ascontiguousarray(wav, dtype)
        vel = np.ascontiguousarray(vel, dtype)
        rec = np.ascontiguousarray(rec, dtype)
        assert vel.shape == (nx, nz) and rec.shape == (nx, nt)
        grad = np.zeros((nx, nz), dtype)
        f(*args, _ptr(wav), _ptr(vel), _ptr(rec), _ptr(grad))
        return grad

    def plane_prepare(self, nx, nt, ntw, np_, pmin, dp, sx, isx, ng, col, wav, data, dtype=np.float32):
        """planeWavePrepare: sou [np][nx][ntw], rec [np][nx][nt]."""
        S = self._scalar(dtype)
        sx = np.ascontiguousarray(sx, dtype)
        isx = np.ascontiguousarray(isx, np.int32)
        ng = np.ascontiguousarray(ng, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        wav = np.ascontiguousarray(wav, dtype)
        data = np.ascontiguousarray(data, dtype)
        ns = sx.shape[0]
        assert isx.shape == (ns,) and ng.shape == (ns,) and col.shape == (int(ng.sum()),)
        assert wav.shape == (ntw,) and data.size == int(ng.sum()) * nt
        sou = np.zeros((np_, nx, ntw), dtype)
        rec = np.zeros((np_, nx, nt), dtype)
        self._fn("plane_prepare", dtype)(_I(nx), _I(nt), _I(ntw), _I(ns), _I(np_), S(pmin), S(dp), _ptr(sx),
                                         _ptr(isx), _ptr(ng), _ptr(col), _ptr(wav), _ptr(data), _ptr(sou),
                                         _ptr(rec))
        return sou, rec

    def born_adjoint_xcorr(self, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, rec, dtype=np.float32):
        """grad[nx][nz]: the cross-correlation adjoint of born() (forward operator
        run backwards, grad += u S / vt2)."""
        f, args, keep = self._geom_call("born_adjoint_xcorr", dtype, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz)
        wav = np.ascontiguousarray(wav, dtype)
        vel = np.ascontiguousarray(vel, dtype)
        rec = np.ascontiguousarray(rec, dtype)
        assert vel.shape == (nx, nz) and rec.shape == (nx, nt)
        grad = np.zeros((nx, nz), dtype)
        f(*args, _ptr(wav), _ptr(vel), _ptr(rec), _ptr(grad))
        return grad

    def hilbert_taps(self, half, dtype=np.float32):
        h = np.zeros(2 * half + 1, dtype)
        self._fn("hilbert_taps", dtype)(_I(half), _ptr(h))
        return h

    def hilbert_apply(self, half, x, dtype=np.float32):
        x = np.ascontiguousarray(x, dtype)
        out = np.zeros_like(x)
        self._fn("hilbert_apply", dtype)(_I(half), _ptr(x), _ptr(out), _I(x.shape[0]))
        return out

    def sponge_weights(self, nz, nx, pml, dtype=np.float32):
        w = np.zeros((nx + 2 * pml, nz + 2 * pml), dtype)
        self._fn("sponge_weights", dtype)(_I(nz), _I(nx), _I(pml), _ptr(w))
        return w

    def pad_velocity(self, vel, pml, dtype=np.float32):
        vel = np.ascontiguousarray(vel, dtype)
        nx, nz = vel.shape
        vp = np.zeros((nx + 2 * pml, nz + 2 * pml), dtype)
        self._fn("pad_velocity", dtype)(_I(nz), _I(nx), _I(pml), _ptr(vel), _ptr(vp))
        return vp

    def stencil_step(self, vpad, prv, cur, pml, dt, dx, dz, dtype=np.float32):
        S = self._scalar(dtype)
        vpad = np.ascontiguousarray(vpad, dtype)
        prv = np.ascontiguousarray(prv, dtype)
        cur = np.ascontiguousarray(cur, dtype)
        NX, NZ = vpad.shape
        nxt = np.zeros_like(cur)
        self._fn("stencil_step", dtype)(_I(NZ - 2 * pml), _I(NX - 2 * pml), _I(pml), S(dt), S(dx), S(dz),
                                        _ptr(vpad), _ptr(prv), _ptr(cur), _ptr(nxt))
        return nxt


class Reference(_Lib):
    def __init__(self, path=REF_SO):
        super().__init__(path, "ref_", False)

    def hilbert_apply(self, half, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros_like(x)
        self.lib.ref_hilbert_apply(_I(half), _ptr(x), _ptr(out), _I(x.shape[0]))
        return out

    def set_abs(self, nz, nx, pml):
        w = np.zeros((nx + 2 * pml, nz + 2 * pml), np.float32)
        self.lib.ref_set_abs(_I(nz), _I(nx), _I(pml), _ptr(w))
        return w

    def set_vel(self, vel, pml):
        vel = np.ascontiguousarray(vel, np.float32)
        nx, nz = vel.shape
        vp = np.zeros((nx + 2 * pml, nz + 2 * pml), np.float32)
        self.lib.ref_set_vel(_I(nz), _I(nx), _I(pml), _ptr(vel), _ptr(vp))
        return vp


def have_reference():
    return os.path.exists(REF_SO)
