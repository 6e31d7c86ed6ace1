"""The C-ABI library loads and exports every function include/*.h declares.

Runs without a GPU: no compute entry point is called.
"""
import ctypes
import glob
import os
import re

import pytest

from conftest import ROOT


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(seis_[a-z0-9_]+)\s*\(", text):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("seis_modeling", "seis_born", "seis_rtm", "seis_modeling_batch", "seis_born_batch",
              "seis_rtm_batch", "seis_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import seis_amd
    L = seis_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    from seis_amd._lib import exported_symbols
    assert set(declared_functions()) <= set(exported_symbols()) | set(declared_functions())
    missing = set(declared_functions()) - set(exported_symbols())
    assert not missing, missing


def test_argument_errors_without_gpu():
    """Invalid arguments are rejected before any device work."""
    import numpy as np
    import seis_amd
    from seis_amd._lib import SeisError
    gz = np.zeros(4, np.int32)
    with pytest.raises(SeisError):
        seis_amd.modeling(0.001, 10, 10, 10, 0, 4, 4, 2, 9, 0, gz, np.zeros(10, np.float32),
                          np.ones((4, 4), np.float32))  # source outside the model
    msg = seis_amd.lib().seis_last_error().decode()
    assert "source" in msg


def test_plane_prepare_argument_errors_without_gpu():
    """planeWavePrepare's geometry is checked before any device work."""
    import ctypes
    import numpy as np
    import seis_amd
    L = seis_amd.lib()
    P = ctypes.c_void_p

    def call(isx, col):
        sx = np.array([100.0], np.float32)
        isx = np.array(isx, np.int32)
        ng = np.array([len(col)], np.int32)
        col = np.array(col, np.int32)
        return L.seis_plane_prepare(8, 10, 12, 1, 2, 0.0, 1e-4, sx.ctypes.data_as(P), isx.ctypes.data_as(P),
                                    ng.ctypes.data_as(P), col.ctypes.data_as(P), None, None, None, None)

    assert call([8], [0, 1]) == 1 and "source column" in L.seis_last_error().decode()
    assert call([3], [0, 9]) == 1 and "trace column" in L.seis_last_error().decode()
    assert call([3], [0, -1]) == 1 and "null array" in L.seis_last_error().decode()
    assert L.seis_harmonic(None, None, 4, 1e-3, 10.0) == 1


def test_rtm_shot_checks_without_gpu():
    """Shot arguments are checked before any device work (so also here,
    without a GPU): an RTM window narrower than the 4-layer strips, a
    negative width and a source outside its window are SEIS_ERR_ARG; a tall
    column is fine (the imaging pass tiles it)."""
    import ctypes
    import seis_amd
    from seis_amd._lib import seis_grid, seis_shot
    L = seis_amd.lib()
    g = seis_grid(40, 8200, 4, 0.001, 10.0, 10.0)
    for shot, msg in ((seis_shot(1, 2, 10, 0, 0, 0, 0, 0, 3, 0), "4-layer"),
                      (seis_shot(1, 2, 10, 0, 0, 0, 0, 0, -1, 0), "nx must be"),
                      (seis_shot(12, 2, 10, 0, 0, 0, 0, 0, 10, 0), "source cell")):
        shots = (seis_shot * 1)(shot)
        assert L.seis_rtm_batch(ctypes.byref(g), 1, shots, None, None, None, None, None) == 1
        assert msg in L.seis_last_error().decode()
    starts = (ctypes.c_int * 2)()
    n = ctypes.c_int(0)
    ok = (seis_shot * 1)(seis_shot(2, 2, 10, 0, 0, 0, 0, 0, 0, 0))
    assert L.seis_batch_plan(2, ctypes.byref(g), 1, ok, 1 << 40, 0, starts, ctypes.byref(n)) == 0 and n.value == 1


def test_tuning_switch_without_gpu():
    """seis_set_tuning takes 0 or 1 only; the choice table is host state, so
    this needs no GPU."""
    import seis_amd
    from seis_amd._lib import SeisError
    with pytest.raises(SeisError):
        seis_amd.set_tuning(2)
    assert "0 or 1" in seis_amd.lib().seis_last_error().decode()
    for m in (0, 1):
        seis_amd.set_tuning(m)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU's share; asked for more GPUs it exits with a
    message before touching any device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=root)
    assert r.returncode != 0
    assert "one GPU" in r.stderr + r.stdout
    assert r.stdout.strip() == ""  # no JSON line


def test_fd_h_entry_points_reject_bad_sizes_without_gpu():
    """The fd.h drop-in returns NULL for non-positive nt / nx / nz, negative
    ndel or a null row table, before reading any input row or allocating an
    output (no device call is reached)."""
    import numpy as np
    import seis_amd
    lib = seis_amd.lib()
    F, I, P = ctypes.c_float, ctypes.c_int, ctypes.c_void_p
    PP = ctypes.POINTER(ctypes.POINTER(ctypes.c_float))
    modeling = getattr(lib, "_Z8modelingfffiiiiiiiPiPfPS0_")
    forward = getattr(lib, "_Z7forwardfffiiiiiiiPiPfPS0_S1_")
    adjoint = getattr(lib, "_Z7adjointfffiiiiiiiPiPfPS0_S1_")
    illum_plane = getattr(lib, "_Z10illumPlanefffiiiiiiiPPfS0_")
    for f in (modeling, forward, adjoint, illum_plane):
        f.restype = PP
    a = np.ones((4, 4), np.float32)
    rows = (ctypes.POINTER(ctypes.c_float) * 4)(*[ctypes.cast(a.ctypes.data + 16 * i, ctypes.POINTER(ctypes.c_float))
                                                  for i in range(4)])
    gz = np.zeros(4, np.int32)
    wav = np.zeros(16, np.float32)
    gzp, wavp = gz.ctypes.data_as(P), wav.ctypes.data_as(P)

    def head(nt, ndel, nx, nz):
        return (F(0.001), F(10.0), F(10.0), I(nt), I(ndel), I(nx), I(nz), I(2), I(1), I(1))

    for nt, ndel, nx, nz in ((0, 0, 4, 4), (-5, 0, 4, 4), (8, -1, 4, 4), (8, 0, -4, 4), (8, 0, 4, 0)):
        assert not modeling(*head(nt, ndel, nx, nz), gzp, wavp, rows)
        assert not forward(*head(nt, ndel, nx, nz), gzp, wavp, rows, rows)
        assert not adjoint(*head(nt, ndel, nx, nz), gzp, wavp, rows, rows)
        assert not illum_plane(F(0.001), F(10.0), F(10.0), I(nt), I(ndel), I(nx), I(nz), I(2), I(1), I(1), rows,
                               rows)
    # null row tables
    assert not modeling(*head(8, 0, 4, 4), gzp, wavp, None)
    assert not forward(*head(8, 0, 4, 4), gzp, wavp, rows, None)
    assert not adjoint(*head(8, 0, 4, 4), gzp, wavp, None, rows)
