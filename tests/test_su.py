"""SU interop against the reference compiled from its own sources: writeSu
(io.cpp:8-58) and the get_info utility (zzz/get_info/get_info.c), whose
outputs are stored byte for byte in tests/golden/su_interop.npz
(tests/golden/make_su_golden.py).  Host code only: runs without a GPU."""
import os

import numpy as np
import pytest

from conftest import HERE

GOLD = np.load(os.path.join(HERE, "golden", "su_interop.npz"), allow_pickle=False)


def test_write_su_block_matches_reference(tmp_path):
    from seis_amd.su import write_su
    p = tmp_path / "b.su"
    write_su(p, GOLD["write_block"])
    assert p.read_bytes() == GOLD["write_block_su"].tobytes()


def test_write_su_one_trace_matches_reference(tmp_path):
    """The one-trace form, including its dt of 1000000 truncated to 16 bits."""
    from seis_amd.su import write_su
    p = tmp_path / "o.su"
    write_su(p, GOLD["write_one"])
    got = p.read_bytes()
    assert got == GOLD["write_one_su"].tobytes()
    assert int.from_bytes(got[116:118], "little") == 1000000 % 65536


@pytest.mark.parametrize("key", ["fldr", "sx"])
def test_su_split_matches_get_info(tmp_path, key):
    """Gathers, coordinates (with get_info's rand() jitter of the source x)
    and the statistics table, byte for byte."""
    from seis_amd.su import su_split
    src = tmp_path / "in.su"
    src.write_bytes(GOLD["split_input"].tobytes())
    su_split(src, tmp_path / "info.txt", tmp_path / "coord.txt", str(tmp_path) + "/", key)
    assert (tmp_path / "info.txt").read_bytes() == GOLD["split_%s_info" % key].tobytes()
    assert (tmp_path / "coord.txt").read_bytes() == GOLD["split_%s_coord" % key].tobytes()
    n = len([k for k in GOLD.files if k.startswith("split_%s_csg" % key)])
    assert n == 3
    for i in range(n):
        assert (tmp_path / ("CSG%d.dat" % i)).read_bytes() == GOLD["split_%s_csg%d" % (key, i)].tobytes(), i
    assert not (tmp_path / ("CSG%d.dat" % n)).exists()


def test_split_gathers_are_the_input_traces(tmp_path):
    """Independent of get_info: concatenated, the gathers are the input's
    trace samples in order (headers stripped)."""
    from seis_amd.su import su_split
    raw = GOLD["split_input"].tobytes()
    nt = int.from_bytes(raw[114:116], "little")
    rec = 240 + 4 * nt
    samples = b"".join(raw[i + 240:i + rec] for i in range(0, len(raw), rec))
    src = tmp_path / "in.su"
    src.write_bytes(raw)
    su_split(src, tmp_path / "i.txt", tmp_path / "c.txt", str(tmp_path) + "/", "fldr")
    got = b"".join((tmp_path / ("CSG%d.dat" % i)).read_bytes() for i in range(3))
    assert got == samples
