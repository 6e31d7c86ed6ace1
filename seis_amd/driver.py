"""Parameter-file driver over the C ABI (seis_run, seis_read_params): the
reference's Inversion flow (lsrtm/inversion.cpp) on one GPU."""
import ctypes

from ._lib import check, lib


def read_params(path):
    """Parameters of a reference-style parameter file as a dict of strings."""
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().seis_read_params(str(path).encode(), buf, len(buf)))
    out = {}
    for line in buf.value.decode().splitlines():
        k, _, v = line.partition(" = ")
        out[k] = v
    return out


def run(param_file, mode, model_file=None, out_file=None):
    """Runs one driver mode: modeling, born, rtm, rtm_born, lsrtm, lsrtm_xcorr, lsrtm_illum,
    illum, plane_prepare, rtm_plane, rtm_plane_born or illum_plane."""
    enc = lambda s: None if s is None else str(s).encode()  # noqa: E731
    check(lib().seis_run(enc(param_file), enc(mode), enc(model_file), enc(out_file)))


def read_coords(path):
    """Shots of a coordinates file: list of ((sx, sz), [(gx, gz), ...])."""
    import numpy as np
    L = lib()
    ns, nt = ctypes.c_int(0), ctypes.c_int(0)
    check(L.seis_read_coords(str(path).encode(), 0, 0, ctypes.byref(ns), ctypes.byref(nt), None, None, None))
    ng = np.zeros(max(ns.value, 1), np.int32)
    sxz = np.zeros((max(ns.value, 1), 2), np.float32)
    gxz = np.zeros((max(nt.value, 1), 2), np.float32)
    check(L.seis_read_coords(str(path).encode(), ns.value, nt.value, ctypes.byref(ns), ctypes.byref(nt),
                             ng.ctypes.data, sxz.ctypes.data, gxz.ctypes.data))
    out, t = [], 0
    for s in range(ns.value):
        out.append((tuple(sxz[s]), [tuple(g) for g in gxz[t:t + ng[s]]]))
        t += ng[s]
    return out
