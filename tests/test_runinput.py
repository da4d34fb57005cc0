"""The .ini run input: keys, quoting, switches and error codes of the
reference's GetPrivateProfileString reads (CUDA_EGS.cu:936-1480), checked on
two of the reference's own run inputs (tests/golden/ini, copied verbatim)."""
import os

import numpy as np
import pytest

from gegs_amd import runinput

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ini")


def test_five_field_head_input():
    r = runinput.read_run_input(os.path.join(GOLD, "TestHeAbDo_5Fields_S.ini"))
    assert r.output_prefix == "V12306" and r.output_directory == "output"
    assert r.histories == 100_000_000
    assert (r.gpu_number, r.gpu_id) == (1, 0)
    assert r.pegs_file == "700icru.pegs4dat" and r.default_medium == "AIR700ICRU"
    f = lambda v: float(np.float32(v))              # sscanf "%f" into a float
    assert (r.phsp_file_number, r.phase_space_file, r.phsp_z) == (4, "VarNoJaws4", f(16.44))
    assert r.secjaws and r.secjaws_file == "VarianMLC.egsinp"
    assert r.mlc and r.mlcinfo_file == "HEAD\\HeadMLC_5Fields.mlcinfo"
    assert not r.block and not r.wedge
    assert r.use_ctcreate and r.xyz_boundaries == [f(v) for v in (-8.265, 8.085, -24.195, -7.845, -21.45, -13.351)]
    assert r.ct_voxel_sizes == [f(0.15)] * 3
    assert r.change_coordinate
    assert r.isocenter == [f(-0.09), f(-16.02), f(-17.63)] and r.source_to_isocenter == 100.0
    assert r.dose_calibration and r.histories_per_cgy == f(6.020e13)


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
[SecJaws]
secjaws file = "j.egsinp"
[MLC]
mlcinfo file = "f.mlcinfo"
[Phantom]
EGSPHANT FILE = "w.egsphant"
isocenter location = 0 0 0
source to isocenter = 100
"""


def test_parse_semantics(tmp_path):
    r = runinput.read_run_input(_write(tmp_path, BASE))
    assert r.egsphant_file == "w.egsphant"          # key names are case-insensitive
    assert r.output_directory == "."                # missing directory -> current
    assert not r.mlc and r.mlcinfo_file == "f.mlcinfo"     # module switch absent -> off, the list still read
    assert not r.secjaws and r.secjaws_file == "j.egsinp"
    assert r.change_coordinate                      # default Yes (CUDA_EGS.cu:1402-1410)
    assert r.dose_unit == np.float32(1.6e-8) and not r.dose_calibration
    assert not (r.output_factor or r.interpolate or r.absolute_dose or r.relative_dose)


def test_reference_defaults_and_exact_switches(tmp_path):
    text = BASE.replace('output prefix = "x"\n', "").replace("[Phantom]", "[Phantom]\nchange coordinate = no")
    text = text.replace("[MLC]", "[MLC]\nmodule MLC = on")
    r = runinput.read_run_input(_write(tmp_path, text))
    assert r.output_prefix == "GEGS" and r.warnings       # default prefix with a warning
    assert r.change_coordinate                      # only "No" turns it off
    assert not r.mlc                                # "on" is not "On"
    r = runinput.read_run_input(_write(tmp_path, BASE.replace("[Phantom]", "[Phantom]\nchange coordinate = No")))
    assert not r.change_coordinate


def test_sscanf_prefixes_and_binary32(tmp_path):
    text = BASE.replace("histories = 1e6", "histories = 123456789 per MU").replace(
        "z position = 10", "z position = 16.44cm")
    r = runinput.read_run_input(_write(tmp_path, text))
    assert r.histories == 123456792                 # "%f" into a float, then truncated to ulong
    assert r.phsp_z == float(np.float32(16.44))


def test_output_sections(tmp_path):
    text = BASE + """[Calibration]
dose calibration = Yes
histories per cGy = 6.020E+13
[Output factor]
use output factor = Yes
output factor file = "of.txt"
[Interpolation]
do interpolate = Yes
starting point = -10.00 -26.00 -21.60
xyz dimensions = 512 512 56
3d voxel sizes = 0.0390625 0.0390625 0.15
[Dose output]
absolute dose = Yes
dose position = -0.09 -16.02 -17.63
dose measured = 200.8
relative dose = Yes
starting point = -2.75 -0.1 -2.75
xyz dimensions = 23 1 23
3d voxel sizes = 0.25 0.0 0.25
PTW729/PTW1000 = PTW1000
gamma index (mm,%) = 3 3
norm to Max/Center = Center
measured data = "PTW1000.txt"
"""
    r = runinput.read_run_input(_write(tmp_path, text.replace("isocenter location = 0 0 0",
                                                              "isocenter location = 1 2 3")))
    f = np.float32
    assert r.dose_unit == f(f(f(1.6e-8) * f(6.02e13)) / f(1e6))
    assert r.output_factor and r.output_factor_file == "of.txt"
    assert r.interp_dimen == [512, 512, 56]
    assert r.interp_start == [float(f(f(-10.0) + f(0.0390625) / f(2))), float(f(f(-26.0) + f(0.0390625) / f(2))),
                              float(f(f(-21.6) + f(0.15) / f(2)))]
    assert r.dose_position == [float(f(-0.09)), float(f(-16.02)), float(f(-17.63))] and r.dose_measured == float(f(200.8))
    assert r.reldose_start == [float(f(f(-2.75) + f(1))), float(f(f(-0.1) + f(2))), float(f(f(-2.75) + f(3)))]
    assert r.reldose_dimen == [23, 1, 23] and r.array_type == 2 and r.norm_type == 2
    assert r.gamma_len == float(f(f(3) * f(0.1))) and r.gamma_dose == float(f(f(3) * f(0.01)))
    assert r.measured_data == "PTW1000.txt"


@pytest.mark.parametrize("edit,code", [
    (("histories = 1e6", ""), 2013),
    (("histories = 1e6", "histories = many"), 2014),
    (("histories = 1e6", "histories = -5"), 2015),
    (("pegs file = \"p\"", ""), 2012),
    (("phsp file number = 1", "phsp file number = one"), 5014),
    (("isocenter location = 0 0 0", "isocenter location = 0 0"), 2011),
    (('mlcinfo file = "f.mlcinfo"', ""), 2011),
    (('secjaws file = "j.egsinp"', ""), 2011),
    (("[Phantom]", "[Output factor]\nuse output factor = Yes\n[Phantom]"), 2012),
    (("[Phantom]", "[Interpolation]\ndo interpolate = Yes\nstarting point = 0 0\n[Phantom]"), 2011),
])
def test_error_codes(tmp_path, edit, code):
    with pytest.raises(runinput.RunInputError) as e:
        runinput.read_run_input(_write(tmp_path, BASE.replace(*edit)))
    assert e.value.code == code


def test_missing_file():
    with pytest.raises(runinput.RunInputError) as e:
        runinput.read_run_input("/nonexistent/run.ini")
    assert e.value.code == 1001
