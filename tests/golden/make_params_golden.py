"""Write golden cases of the parameter-file grammar and the coordinates
reader from the compiled reference.

Run in the container that holds the reference tree, after `make -C oracle ref`:

    python tests/golden/make_params_golden.py

For every parameter text below it runs the reference's own ParserFromFile
(util/parser.cpp) in the sequence of Inversion::readParamFile
(lsrtm/inversion.cpp:38-79) on a paramLSRTM (lsrtm/datastr.h:51-75), and for
every coordinates text the reference's Coords readers (util/coord.cpp:22-176),
through oracle/_ref/libseisref.so (ref_read_params, ref_coords).  It stores
the texts and the parsed results as params_coords.json next to this script:
data only, which pins the driver's parser and coordinates reader
(tests/test_params_golden.py).  Only texts that hold every required key are
used: on a missing key the reference's error path passes a std::string
through printf's varargs (util/debug.cpp:13-16), which aborts.
"""
import ctypes
import json
import os
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libseisref.so")
REF_SRC = "/root/reference/src"

BASE = """nthread   4
ns        3
ngmax     40
nx        60       ! model columns
nz        50
nt        400
npml      20
delay     10
delaycal  0
lpad 5
rpad 6
dx 10.0
dz 10.0
dt 0.001
fr 25
velfx 0.0
sbegin 0
send 2
precsg     /tmp/seis/CSG/CSG
wdir       /tmp/seis/WDIR/
v0file     /tmp/seis/v0.bin
coordfile /tmp/seis/coords.txt
"""

PARAM_CASES = {
    "base": BASE,
    # a key in a comment, a key line without a value, the first valued line wins
    "comments_first_value": "! nx 999\nnx\nnx 77\nnx 88\n" + BASE.replace("nx        60       ! model columns\n", ""),
    # optional keys missing take their defaults; signs, exponents, CRLF endings
    "defaults_signs_crlf": BASE.replace("npml      20\n", "").replace("delay     10\n", "")
    .replace("velfx 0.0", "velfx -12.5e1").replace("dt 0.001", "dt +1e-3").replace("\n", "\r\n"),
    # a tab after an optional key: the parser looks for "key " and falls back to the default
    "tab_optional": BASE.replace("npml      20", "npml\t20").replace("lpad 5", "lpad\t5"),
    # mask = 1 requires maskfile; planeTag = 1 reads np, dp, pmin; a float value for an int key
    "mask_plane": BASE + "mask 1\nmaskfile /tmp/seis/mask.bin\nplaneTag 1\nnp 81\ndp 0.5\npmin -21\nmaxiter 7.9\n",
    # a string whose value starts after digits and punctuation
    "string_scan": BASE.replace("wdir       /tmp/seis/WDIR/", "wdir  = ./run2/WDIR/"),
}


def ref_params(lib, text):
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False, newline="") as f:
        f.write(text)
        path = f.name
    try:
        buf = ctypes.create_string_buffer(1 << 16)
        lib.ref_read_params(path.encode(), buf, len(buf))
    finally:
        os.unlink(path)
    out = {}
    for line in buf.value.decode().splitlines():
        k, _, v = line.partition(" = ")
        out[k] = v
    return out


def get_info_lines(shots, jitter):
    """Coordinates in get_info's %12.2f x 4 format (zzz/get_info/get_info.c:109-152)."""
    lines = []
    for (sx, sz, gxs, gz), j in zip(shots, jitter):
        for gx in gxs:
            lines.append("%12.2f%12.2f%12.2f%12.2f" % (sx + j, sz, gx, gz))
    return lines


SHOTS = [(150.0, 20.0, [0.0 + 10 * i for i in range(31)], 10.0),
         (350.0, 20.0, [200.0 + 10 * i for i in range(31)], 10.0),
         (560.0, 25.0, [410.0 + 10 * i for i in range(26)], 12.5)]

COORD_CASES = {
    "get_info": "\n".join(get_info_lines(SHOTS, [0.37, 1.62, 0.05])) + "\n",
    "no_final_newline": "\n".join(get_info_lines(SHOTS, [0.0, 0.0, 0.0])),
    "single_trace_shots_spaces": "100 5 90 5\n200\t5\t210\t5\n300  5   290 6\n300 5 310 6\n",
}


def ref_coords(lib, text):
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False, newline="") as f:
        f.write(text)
        path = f.name
    try:
        ng = (ctypes.c_int * 64)()
        sxz = (ctypes.c_float * 128)()
        gxz = (ctypes.c_float * 4096)()
        ns = lib.ref_coords(path.encode(), 64, 2048, ng, sxz, gxz)
    finally:
        os.unlink(path)
    out, t = [], 0
    for s in range(ns):
        out.append({"s": [sxz[2 * s], sxz[2 * s + 1]],
                    "g": [[gxz[2 * (t + g)], gxz[2 * (t + g) + 1]] for g in range(ng[s])]})
        t += ng[s]
    return out


def main():
    lib = ctypes.CDLL(REF_SO)
    lib.ref_read_params.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    lib.ref_coords.restype = ctypes.c_int
    lib.ref_coords.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p]
    cases = dict(PARAM_CASES)
    # The reference's own parameter files lack sbegin and send, which
    # readParamFile requires (datastr.h:56-57): appended here.
    for name in ("param_marm.txt2", "param_simple.txt2"):
        with open(os.path.join(REF_SRC, name)) as f:
            cases[name] = f.read() + "sbegin 0\nsend 1\n"
    doc = {"params": {k: {"text": v, "ref": ref_params(lib, v)} for k, v in cases.items()},
           "coords": {k: {"text": v, "ref": ref_coords(lib, v)} for k, v in COORD_CASES.items()},
           "source": "reference ParserFromFile (util/parser.cpp) in readParamFile's order and Coords "
                     "(util/coord.cpp), compiled from source (oracle/Makefile ref)"}
    with open(os.path.join(HERE, "params_coords.json"), "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print("wrote params_coords.json: %d parameter texts, %d coordinate texts" % (len(cases), len(COORD_CASES)))


if __name__ == "__main__":
    main()
