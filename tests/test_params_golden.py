"""The parameter-file grammar and the coordinates reader against the
reference's own ParserFromFile and Coords, compiled from source, on the cases
of tests/golden/params_coords.json (written by make_params_golden.py):
comments, a key line without a value, optional keys and defaults, signs,
exponents, CRLF endings, a tab after a key, mask and plane-wave keys, the
reference's own param_marm / param_simple files; coordinates in get_info's
format with and without a final newline, and a first shot of one trace.
Floats must agree to the bit, strings and counts exactly."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DOC = json.load(open(os.path.join(HERE, "golden", "params_coords.json")))


@pytest.mark.parametrize("name", sorted(DOC["params"]))
def test_params_match_reference(tmp_path, name):
    from seis_amd.driver import read_params
    case = DOC["params"][name]
    f = tmp_path / "param.txt"
    f.write_bytes(case["text"].encode())
    got = read_params(f)
    for key, ref in case["ref"].items():
        assert key in got, key
        if key in ("dx", "dz", "dt", "fr", "velfx", "dp", "pmin"):
            assert np.float32(got[key]) == np.float32(ref), (key, got[key], ref)
        else:
            assert got[key] == ref, (key, got[key], ref)


@pytest.mark.parametrize("name", sorted(DOC["coords"]))
def test_coords_match_reference(tmp_path, name):
    from seis_amd.driver import read_coords
    case = DOC["coords"][name]
    f = tmp_path / "coords.txt"
    f.write_bytes(case["text"].encode())
    got = read_coords(f)
    ref = case["ref"]
    assert len(got) == len(ref)
    for (s, g), r in zip(got, ref):
        assert list(np.float32(s)) == list(np.float32(r["s"]))
        assert len(g) == len(r["g"])
        assert np.array_equal(np.array(g, np.float32).reshape(-1, 2), np.array(r["g"], np.float32).reshape(-1, 2))
