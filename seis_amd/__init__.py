"""seis_amd: MI355X (gfx950) engine for the 2D acoustic finite-difference path.

The compute lives in the HIP library ``seis_amd/libseis_amd.so`` behind the C
ABI of ``include/seis.h``.  This package is a thin host-side mirror of the
reference's engine interface (bbsun/Seis src/util/fd.h): ``modeling``,
``forward`` (Born) and ``adjoint`` (RTM) with the reference's arguments and
array layouts, plus batched, device-resident variants.  There is no CPU
fallback: if the library is missing or no GPU is present, calls raise.
"""
from ._lib import SeisError, lib, library_path  # noqa: F401
from .device import (DeviceArray, device_count, profile_start, profile_stop, set_tuning, synchronize,  # noqa: F401
                     tuned_strip)
from .fd import adjoint, adjoint_plane, forward, forward_plane, illum, illum_plane, modeling  # noqa: F401
from .batch import Grid, Shot, born_batch, illum_batch, modeling_batch, rtm_batch  # noqa: F401
from .plane import harmonic, plane_prepare  # noqa: F401
from .wavelet import ricker  # noqa: F401
