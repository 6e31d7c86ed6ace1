"""SU interop of the reference: writeSu (src/util/io.cpp:8-58) and the
get_info utility (zzz/get_info/get_info.c), host-side through the C ABI."""
import numpy as np

from ._lib import check, lib


def write_su(path, data):
    """Writes data [ntraces][nsamples] (or one trace [nsamples]) as writeSu
    does: the 2D form for a 2D array, the one-trace form for a 1D array."""
    a = np.ascontiguousarray(data, np.float32)
    if a.ndim == 1:
        n1, n2 = a.shape[0], 0
    elif a.ndim == 2:
        n2, n1 = a.shape
    else:
        raise ValueError("data must be 1D or 2D")
    check(lib().seis_write_su(str(path).encode(), n1, n2, a.ctypes.data))


def su_split(su_file, info_file, coord_file, directory, key="fldr"):
    """Splits an SU file into <directory>CSG<i>.dat gathers plus the
    coordinates and statistics files, as get_info does (key: fldr, ep, cdp
    or sx).  `directory` is used as a prefix, like get_info's dir argument."""
    check(lib().seis_su_split(str(su_file).encode(), str(info_file).encode(), str(coord_file).encode(),
                              str(directory).encode(), key.encode()))
