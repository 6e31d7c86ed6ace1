"""Ricker source wavelet (reference src/util/wavelet.cpp:7-18), computed by
the engine library's host code (seis_ricker) with the reference's float
arithmetic."""
import ctypes

import numpy as np

from ._lib import check, lib


def ricker(dt, fr, num_delay, n):
    """w[i] = (1 - 2 a t^2) exp(-a t^2), a = pi^2 fr^2 (pi = 3.1415926),
    t = (i - num_delay) dt; float32, n samples."""
    out = np.zeros(int(n), np.float32)
    check(lib().seis_ricker(float(dt), float(fr), int(num_delay), int(n), out.ctypes.data_as(ctypes.c_void_p)))
    return out
