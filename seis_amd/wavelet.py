"""Ricker source wavelet (reference src/util/wavelet.cpp:7-18)."""
import numpy as np


def ricker(dt, fr, num_delay, n):
    """w[i] = (1 - 2 a t^2) exp(-a t^2), a = pi^2 fr^2 with pi = 3.1415926,
    t = (i - num_delay) dt, evaluated in float32 like the reference."""
    a = np.float32(3.1415926 * 3.1415926 * float(np.float32(fr)) * float(np.float32(fr)))
    i = np.arange(n)
    t = (i - num_delay).astype(np.float32) * np.float32(1.0) * np.float32(dt)
    t2 = t * t
    return ((np.float32(1.0) - np.float32(2.0) * a * t2) * np.exp(-a * t2)).astype(np.float32)
