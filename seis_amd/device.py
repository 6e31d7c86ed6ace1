"""Device buffers managed through the engine's C ABI (no torch needed)."""
import ctypes

import numpy as np

from ._lib import check, lib


def device_count():
    n = ctypes.c_int(0)
    rc = lib().seis_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    check(lib().seis_synchronize())


def set_tuning(on):
    """Strip-length autotuning of the interior sweeps: on (the default) or
    off (strip length 16).  Results are identical either way."""
    check(lib().seis_set_tuning(int(on)))


def tuned_strip(kind, nx_padded, nz_padded, nshots):
    """The tuner's strip length for a configuration, or None when none was
    made."""
    v = ctypes.c_int(0)
    check(lib().seis_tuned_strip(kind.encode(), int(nx_padded), int(nz_padded), int(nshots), ctypes.byref(v)))
    return v.value or None


def profile_start():
    """Start per-launch kernel timing (HIP events on the launch stream)."""
    check(lib().seis_profile_start())


def profile_stop():
    """Stop kernel timing; returns {kernel: dict(launches, total_ms, alg_bytes, cell_updates)}."""
    from ._lib import seis_kernel_stat
    stats = (seis_kernel_stat * 64)()
    n = ctypes.c_int(0)
    check(lib().seis_profile_stop(stats, 64, ctypes.byref(n)))
    out = {}
    for s in stats[: min(n.value, 64)]:
        out[s.name.decode()] = dict(launches=s.launches, total_ms=s.total_ms, alg_bytes=s.alg_bytes,
                                    cell_updates=s.cell_updates)
    return out


class DeviceArray:
    """A typed device allocation holding one numpy-shaped array."""

    def __init__(self, shape, dtype=np.float32, zero=False):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().seis_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p
        if zero:
            check(lib().seis_memset(self.ptr, 0, self.nbytes))

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype if dtype is not None else a.dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().seis_memcpy_htod(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes))

    def download(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().seis_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes))
        return out

    def zero(self):
        check(lib().seis_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().seis_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
