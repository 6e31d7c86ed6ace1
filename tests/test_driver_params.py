"""Parameter files in the reference's grammar (util/parser.cpp:16-168) with
the keys, defaults and optional flags of paramLSRTM (lsrtm/datastr.h:51-75)."""
import pytest

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


def _parse(tmp_path, text):
    from seis_amd.driver import read_params
    f = tmp_path / "param.txt"
    f.write_text(text)
    return read_params(f)


def test_values_and_defaults(tmp_path):
    p = _parse(tmp_path, BASE)
    assert p["nx"] == "60" and p["nz"] == "50" and p["nt"] == "400" and p["npml"] == "20"
    assert float(p["dt"]) == pytest.approx(0.001) and float(p["fr"]) == 25.0
    assert p["precsg"] == "/tmp/seis/CSG/CSG" and p["wdir"] == "/tmp/seis/WDIR/"
    assert p["maxiter"] == "30"      # DEFAULT_MAXITER, datastr.h:9
    assert p["mask"] == "0" and p["planeTag"] == "0"
    assert p["vfile"] == ""          # optional string


def test_comment_and_first_valued_line_win(tmp_path):
    text = "! nx 999\nnx\nnx 77\nnx 88\n" + BASE.replace("nx        60       ! model columns\n", "")
    p = _parse(tmp_path, text)
    assert p["nx"] == "77"


def test_int_value_truncates_like_stoi(tmp_path):
    p = _parse(tmp_path, BASE.replace("nt        400", "nt        400.7"))
    assert p["nt"] == "400"


def test_missing_mandatory_key(tmp_path):
    from seis_amd._lib import SeisError
    with pytest.raises(SeisError, match="key sbegin not found"):
        _parse(tmp_path, BASE.replace("sbegin 0\n", ""))


def test_key_without_value(tmp_path):
    from seis_amd._lib import SeisError
    with pytest.raises(SeisError, match="key nz value missing"):
        _parse(tmp_path, BASE.replace("nz        50", "nz   "))


def test_mask_requires_maskfile(tmp_path):
    from seis_amd._lib import SeisError
    with pytest.raises(SeisError, match="maskfile"):
        _parse(tmp_path, BASE + "mask 1\n")
    p = _parse(tmp_path, BASE + "mask 1\nmaskfile /tmp/seis/mask.bin\n")
    assert p["maskfile"] == "/tmp/seis/mask.bin"


def test_plane_wave_keys_only_with_planetag(tmp_path):
    from seis_amd._lib import SeisError
    with pytest.raises(SeisError, match="key np not found"):
        _parse(tmp_path, BASE + "planeTag 1\n")
    p = _parse(tmp_path, BASE + "planeTag 1\nnp 81\ndp 0.5\npmin -21\n")
    assert p["planeTag"] == "1"


def test_unreadable_file():
    from seis_amd.driver import read_params
    from seis_amd._lib import SeisError
    with pytest.raises(SeisError, match="can not open"):
        read_params("/nonexistent/param.txt")
