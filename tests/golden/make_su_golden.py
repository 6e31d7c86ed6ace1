"""Golden vectors of the reference's SU interop, from the reference compiled
from its own sources (oracle/_ref, built by `make -C oracle ref`):

* writeSu (src/util/io.cpp:8-58) through the harness's ref_write_su, for a
  block of traces and for the one-trace form;
* the get_info utility (zzz/get_info/get_info.c), compiled by the same
  recipe, run on a synthetic SU file of three shots with uneven trace counts,
  grouped by fldr and by sx.

Run in the container that holds the reference tree:

    python tests/golden/make_su_golden.py

Everything stored is data: the input SU bytes and the output files' bytes.
"""
import ctypes
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

# byte offsets of the SU trace header fields written here
OFF = dict(tracl=0, tracr=4, fldr=8, tracf=12, ep=16, cdp=20, cdpt=24, offset=36, gelev=40, selev=44, sx=72, sy=76,
           gx=80, gy=84, ns=114, dt=116)


def su_trace(fields, data):
    h = bytearray(240)
    for k, v in fields.items():
        fmt = "<H" if k in ("ns", "dt") else "<i"
        h[OFF[k]:OFF[k] + (2 if k in ("ns", "dt") else 4)] = np.array(v, fmt[1:]).astype(fmt).tobytes()
    return bytes(h) + np.asarray(data, "<f4").tobytes()


def synthetic_su(rng, nt=8):
    """Three shots (fldr 101, 102, 103) of 4, 6 and 5 traces, receivers on
    both sides of the source, elevations set."""
    out = []
    tr = 0
    for shot, (ntr, sx) in enumerate([(4, 1500), (6, 1750), (5, 2010)]):
        for j in range(ntr):
            tr += 1
            gx = sx - 60 + 30 * j
            f = dict(tracl=tr, tracr=tr, fldr=101 + shot, tracf=j + 1, ep=7 + shot, cdp=(sx + gx) // 2, cdpt=j + 1,
                     offset=gx - sx, gelev=5 + j, selev=12 + shot, sx=sx, sy=40, gx=gx, gy=35 + j, ns=nt, dt=500)
            out.append(su_trace(f, rng.standard_normal(nt).astype(np.float32)))
    return b"".join(out)


def main():
    lib = ctypes.CDLL(os.path.join(REF_DIR, "libseisref.so"))
    get_info = os.path.join(REF_DIR, "get_info")
    rng = np.random.default_rng(77)
    arrays = {}
    with tempfile.TemporaryDirectory() as tmp:
        block = rng.standard_normal((5, 7)).astype(np.float32)
        one = rng.standard_normal(9).astype(np.float32)
        p2, p1 = os.path.join(tmp, "block.su"), os.path.join(tmp, "one.su")
        lib.ref_write_su(p2.encode(), 7, 5, block.ctypes.data_as(ctypes.c_void_p))
        lib.ref_write_su(p1.encode(), 9, 0, one.ctypes.data_as(ctypes.c_void_p))
        arrays.update(write_block=block, write_one=one,
                      write_block_su=np.frombuffer(open(p2, "rb").read(), np.uint8),
                      write_one_su=np.frombuffer(open(p1, "rb").read(), np.uint8))
        su = synthetic_su(rng)
        arrays["split_input"] = np.frombuffer(su, np.uint8)
        for key in ("fldr", "sx"):
            d = os.path.join(tmp, key)
            os.mkdir(d)
            src = os.path.join(d, "in.su")
            open(src, "wb").write(su)
            subprocess.run([get_info, src, os.path.join(d, "info.txt"), os.path.join(d, "coord.txt"), d + "/", key],
                           check=True, stdout=subprocess.DEVNULL)
            arrays["split_%s_info" % key] = np.frombuffer(open(os.path.join(d, "info.txt"), "rb").read(), np.uint8)
            arrays["split_%s_coord" % key] = np.frombuffer(open(os.path.join(d, "coord.txt"), "rb").read(), np.uint8)
            i = 0
            while os.path.exists(os.path.join(d, "CSG%d.dat" % i)):
                arrays["split_%s_csg%d" % (key, i)] = np.frombuffer(open(os.path.join(d, "CSG%d.dat" % i), "rb").read(),
                                                                    np.uint8)
                i += 1
            print(key, "gathers:", i)
    np.savez_compressed(os.path.join(HERE, "su_interop.npz"), **arrays)
    print({k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    sys.exit(main())
