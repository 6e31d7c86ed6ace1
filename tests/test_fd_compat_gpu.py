"""The fd.h-signature entry points (include/seis_fd.h) on the GPU: the exact
C++ symbols a reference build links against in place of util/fd.cpp's
modeling / forward / adjoint / illum (fd.h:3-5, 10), called with the reference's
MyAlloc-style float** inputs, against the reference's golden vectors."""
import ctypes

import numpy as np
import pytest

from conftest import load_fixture

pytestmark = pytest.mark.gpu

F, I, P = ctypes.c_float, ctypes.c_int, ctypes.c_void_p
PP = ctypes.POINTER(ctypes.POINTER(ctypes.c_float))


def _table(a):
    """float** row table into one contiguous [n2][n1] block (MyAlloc layout)."""
    a = np.ascontiguousarray(a, np.float32)
    rows = (ctypes.POINTER(ctypes.c_float) * a.shape[0])()
    base = a.ctypes.data
    for i in range(a.shape[0]):
        rows[i] = ctypes.cast(base + 4 * a.shape[1] * i, ctypes.POINTER(ctypes.c_float))
    return rows, a


def _result(ptr, n2, n1):
    assert ptr, "engine call returned NULL"
    return np.ctypeslib.as_array(ptr[0], shape=(n2 * n1,)).reshape(n2, n1).copy()


@pytest.mark.parametrize("name", ["small_a.npz", "small_b.npz", "small_c.npz", "small_d.npz"])
def test_fd_h_entry_points_match_reference(engine, name):
    lib = engine.lib()
    c = load_fixture(name)
    print("case fd.h", name, flush=True)
    sym = {"modeling": "_Z8modelingfffiiiiiiiPiPfPS0_", "forward": "_Z7forwardfffiiiiiiiPiPfPS0_S1_",
           "adjoint": "_Z7adjointfffiiiiiiiPiPfPS0_S1_", "illum": "_Z5illumfffiiiiiiiPfPS_"}
    fns = {}
    for k, mangled in sym.items():
        f = getattr(lib, mangled)
        f.restype = PP
        fns[k] = f
    head = (F(c["dt"]), F(c["dx"]), F(c["dz"]), I(c["nt"]), I(c["ndel"]), I(c["nx"]), I(c["nz"]), I(c["pml"]),
            I(c["sx"]), I(c["sz"]))
    gz = np.ascontiguousarray(c["gz"], np.int32)
    wav = np.ascontiguousarray(c["wav"], np.float32)
    vel_t, vel_a = _table(c["vel"])
    keep = [gz, wav, vel_a]
    gzp, wavp = gz.ctypes.data_as(P), wav.ctypes.data_as(P)
    rec = fns["modeling"](*head, gzp, wavp, vel_t)
    assert np.array_equal(_result(rec, c["nx"], c["nt"]), c["rec"])
    lib.seis_fd_free(rec)
    dv_t, dv_a = _table(c["dv"])
    born = fns["forward"](*head, gzp, wavp, vel_t, dv_t)
    assert np.array_equal(_result(born, c["nx"], c["nt"]), c["born"])
    lib.seis_fd_free(born)
    rin_t, rin_a = _table(c["rec_in"])
    img = fns["adjoint"](*head, gzp, wavp, vel_t, rin_t)
    assert np.array_equal(_result(img, c["nx"], c["nz"]), c["img"])
    lib.seis_fd_free(img)
    ss = fns["illum"](*head, wavp, vel_t)
    assert np.array_equal(_result(ss, c["nx"], c["nz"]), c["ss"])
    lib.seis_fd_free(ss)
    del keep, dv_a, rin_a


@pytest.mark.parametrize("name", ["plane_a.npz", "plane_b.npz"])
def test_fd_h_plane_entry_points_match_reference(engine, name):
    """forwordPlane / adjointPlane / illumPlane (fd.h:6-7, 11) through their
    mangled symbols with float** tables."""
    lib = engine.lib()
    c = load_fixture(name)
    print("case fd.h plane", name, flush=True)
    fns = {}
    for k, mangled in (("born", "_Z12forwordPlanefffiiiiiiiPPfS0_S0_"), ("rtm", "_Z12adjointPlanefffiiiiiiiPPfS0_S0_"),
                       ("illum", "_Z10illumPlanefffiiiiiiiPPfS0_")):
        f = getattr(lib, mangled)
        f.restype = PP
        fns[k] = f
    head = (F(c["dt"]), F(c["dx"]), F(c["dz"]), I(c["nt"]), I(c["ndel"]), I(c["nx"]), I(c["nz"]), I(c["pml"]),
            I(c["sz"]), I(c["gz_row"]))
    vel_t, vel_a = _table(c["vel"])
    sou_t, sou_a = _table(c["sou"])
    dv_t, dv_a = _table(c["dv"])
    rin_t, rin_a = _table(c["rec_in"])
    out = fns["born"](*head, vel_t, sou_t, dv_t)
    assert np.array_equal(_result(out, c["nx"], c["nt"]), c["born"])
    lib.seis_fd_free(out)
    out = fns["illum"](*head, vel_t, sou_t)
    assert np.array_equal(_result(out, c["nx"], c["nz"]), c["ss"])
    lib.seis_fd_free(out)
    out = fns["rtm"](*head, vel_t, sou_t, rin_t)
    assert np.array_equal(_result(out, c["nx"], c["nz"]), c["img"])
    lib.seis_fd_free(out)
    del vel_a, sou_a, dv_a, rin_a
