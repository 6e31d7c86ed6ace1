"""Loader and ctypes signatures of the engine's C ABI (include/seis.h)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(HERE, "libseis_amd.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_SZ = ctypes.c_size_t


class SeisError(RuntimeError):
    """A call into the engine returned an error code."""


class seis_grid(ctypes.Structure):
    _fields_ = [("nx", _I), ("nz", _I), ("pml", _I), ("dt", _F), ("dx", _F), ("dz", _F)]


class seis_shot(ctypes.Structure):
    _fields_ = [("sx", _I), ("sz", _I), ("nt", _I), ("ndel", _I), ("model", ctypes.c_longlong),
                ("wav", ctypes.c_longlong), ("gz", ctypes.c_longlong), ("rec", ctypes.c_longlong),
                ("nx", _I), ("x0", _I)]


class seis_kernel_stat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 40), ("launches", _I), ("total_ms", ctypes.c_double),
                ("alg_bytes", ctypes.c_double), ("cell_updates", ctypes.c_double)]


_SIGS = {
    "seis_last_error": (ctypes.c_char_p, []),
    "seis_device_count": (_I, [ctypes.POINTER(_I)]),
    "seis_set_device": (_I, [_I]),
    "seis_malloc": (_I, [ctypes.POINTER(_P), _SZ]),
    "seis_free": (_I, [_P]),
    "seis_memcpy_htod": (_I, [_P, _P, _SZ]),
    "seis_memcpy_dtoh": (_I, [_P, _P, _SZ]),
    "seis_memset": (_I, [_P, _I, _SZ]),
    "seis_copy_f32": (_I, [_P, _P, ctypes.c_longlong]),
    "seis_set_tuning": (_I, [_I]),
    "seis_tuned_strip": (_I, [ctypes.c_char_p, _I, _I, _I, ctypes.POINTER(_I)]),
    "seis_synchronize": (_I, []),
    "seis_modeling": (_I, [_F, _F, _F, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "seis_born": (_I, [_F, _F, _F, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "seis_rtm": (_I, [_F, _F, _F, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "seis_illum": (_I, [_F, _F, _F, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P]),
    "seis_born_plane": (_I, [_F, _F, _F, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "seis_illum_plane": (_I, [_F, _F, _F, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P]),
    "seis_rtm_plane": (_I, [_F, _F, _F, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "seis_born_plane_batch": (_I, [ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot), _P, _P, _P, _P, _P]),
    "seis_illum_plane_batch": (_I, [ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot), _P, _P, _P]),
    "seis_rtm_plane_batch": (_I, [ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot), _P, _P, _P, _P, _P]),
    "seis_illum_batch": (_I, [ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot), _P, _P, _P]),
    "seis_modeling_batch": (_I, [ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot), _P, _P, _P, _P]),
    "seis_modeling_batch_static": (_I, [ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot), _P, _P, _P, _P]),
    "seis_born_batch": (_I, [ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot), _P, _P, _P, _P, _P]),
    "seis_rtm_batch": (_I, [ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot), _P, _P, _P, _P, _P]),
    "seis_rtm_batch_ex": (_I, [ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot), _P, _P, _P, _P, _P, _I]),
    "seis_lsrtm_create": (_I, [ctypes.POINTER(_P), ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot),
                               _P, _P, _P, _I]),
    "seis_lsrtm_create_f64": (_I, [ctypes.POINTER(_P), ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot),
                                   _P, _P, _P]),
    "seis_lsrtm_destroy": (_I, [_P]),
    "seis_lsrtm_born": (_I, [_P, _P, _P]),
    "seis_lsrtm_adjoint": (_I, [_P, _P, _P]),
    "seis_lsrtm_dot": (_I, [_P, _P, _P, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double)]),
    "seis_lsrtm_set_mask": (_I, [_P, _P]),
    "seis_lsrtm_set_adjoint": (_I, [_P, _I]),
    "seis_lsrtm_set_weight": (_I, [_P, _P]),
    "seis_lsrtm_cg": (_I, [_P, _P, _I, ctypes.c_double, _P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I)]),
    "seis_lsrtm_cg_history": (_I, [_P, _P, _I, ctypes.c_double, _P, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(_I), _P]),
    "seis_ricker": (_I, [_F, _F, _I, _I, _P]),
    "seis_harmonic": (_I, [_P, _P, _I, _F, _F]),
    "seis_plane_prepare": (_I, [_I, _I, _I, _I, _I, _F, _F, _P, _P, _P, _P, _P, _P, _P, _P]),
    "seis_write_su": (_I, [ctypes.c_char_p, _I, _I, _P]),
    "seis_su_split": (_I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
    "seis_fd_free": (None, [_P]),
    "seis_run": (_I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
    "seis_read_params": (_I, [ctypes.c_char_p, ctypes.c_char_p, _I]),
    "seis_read_coords": (_I, [ctypes.c_char_p, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), _P, _P, _P]),
    "seis_set_batch_limits": (_I, [_SZ, _I]),
    "seis_batch_plan": (_I, [_I, ctypes.POINTER(seis_grid), _I, ctypes.POINTER(seis_shot), _SZ, _I,
                             ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "seis_profile_start": (_I, []),
    "seis_profile_stop": (_I, [ctypes.POINTER(seis_kernel_stat), _I, ctypes.POINTER(_I)]),
}

_lib = None


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise ImportError(
                "seis_amd: %s is missing; build it with __graft_entry__.build() "
                "(make -C seis_amd/csrc)" % library_path)
        L = ctypes.CDLL(library_path)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().seis_last_error().decode(errors="replace")
        raise SeisError("seis error %d: %s" % (rc, msg))
    return rc


def exported_symbols():
    """Names declared in include/seis.h that the library must export."""
    return list(_SIGS)
