"""The .ini run input: keys, quoting, switches and error codes of the
reference's GetPrivateProfileString reads (CUDA_EGS.cu:936-1480), checked on
two of the reference's own run inputs (tests/golden/ini, copied verbatim)."""
import os

import pytest

from gegs_amd import runinput

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ini")


def test_five_field_head_input():
    r = runinput.read_run_input(os.path.join(GOLD, "TestHeAbDo_5Fields_S.ini"))
    assert r.output_prefix == "V12306" and r.output_directory == "output"
    assert r.histories == 100_000_000
    assert (r.gpu_number, r.gpu_id) == (1, 0)
    assert r.pegs_file == "700icru.pegs4dat" and r.default_medium == "AIR700ICRU"
    assert (r.phsp_file_number, r.phase_space_file, r.phsp_z) == (4, "VarNoJaws4", 16.44)
    assert r.secjaws and r.secjaws_file == "VarianMLC.egsinp"
    assert r.mlc and r.mlcinfo_file == "HEAD\\HeadMLC_5Fields.mlcinfo"
    assert not r.block and not r.wedge
    assert r.use_ctcreate and r.xyz_boundaries == [-8.265, 8.085, -24.195, -7.845, -21.45, -13.351]
    assert r.ct_voxel_sizes == [0.15, 0.15, 0.15]
    assert r.change_coordinate
    assert r.isocenter == [-0.09, -16.02, -17.63] and r.source_to_isocenter == 100.0
    assert r.dose_calibration and r.histories_per_cgy == 6.020e13


def test_calibration_input_uses_egsphant():
    r = runinput.read_run_input(os.path.join(GOLD, "TestCalibration.ini"))
    assert not r.use_ctcreate and r.egsphant_file == "Calibration.egsphant"
    assert not r.change_coordinate
    assert r.histories == 50_000_000


def _write(tmp_path, text):
    p = tmp_path / "run.ini"
    p.write_text(text)
    return str(p)


BASE = """[General]
output prefix = "x"
histories = 1e6
[Data]
data directory = "d"
egsinp file = "a.egsinp"
pegs file = "p"
photon xsections = "si"
atomic ff file = "ff"
default medium = AIR700ICRU
[Phase space]
phsp file number = 1
phase space file = "ph"
z position = 10
[Phantom]
EGSPHANT FILE = "w.egsphant"
isocenter location = 0 0 0
source to isocenter = 100
"""


def test_parse_semantics(tmp_path):
    r = runinput.read_run_input(_write(tmp_path, BASE))
    assert r.egsphant_file == "w.egsphant"          # key names are case-insensitive
    assert r.output_directory == "."                # missing directory -> current
    assert not r.mlc and r.mlcinfo_file is None     # module switch absent -> off


@pytest.mark.parametrize("edit,code", [
    (("histories = 1e6", ""), 2013),
    (("histories = 1e6", "histories = many"), 2014),
    (("histories = 1e6", "histories = -5"), 2015),
    (("pegs file = \"p\"", ""), 2012),
    (("phsp file number = 1", "phsp file number = one"), 5014),
    (("isocenter location = 0 0 0", "isocenter location = 0 0"), 2011),
    (("[Phantom]", "[MLC]\nmodule MLC = On\n[Phantom]"), 2011),
])
def test_error_codes(tmp_path, edit, code):
    with pytest.raises(runinput.RunInputError) as e:
        runinput.read_run_input(_write(tmp_path, BASE.replace(*edit)))
    assert e.value.code == code


def test_missing_file():
    with pytest.raises(runinput.RunInputError) as e:
        runinput.read_run_input("/nonexistent/run.ini")
    assert e.value.code == 1001
