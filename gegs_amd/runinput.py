"""The reference's run input: a Windows-style ``.ini`` file read key by key.

The reference reads its input with ``GetPrivateProfileString`` (CUDA_EGS.cu:
936-1480): sections in brackets, ``key = value`` lines, values optionally in
double quotes, a missing key returned as the string ``<<missing>>``; a missing
or unparsable required key ends the run with an error code (``error()``,
output.c:51-89).  ``read_run_input`` keeps those keys, defaults and codes and
raises ``RunInputError`` with the code instead of exiting.  The transport
uses the [General] history count, the [Phase space] source, the [SecJaws] /
[MLC] module switches and files, and the [Phantom] geometry; the other
sections are parsed and kept for a host that runs the rest of the
reference's pipeline.
"""
from __future__ import annotations

import dataclasses
import os
import re

import numpy as np

MISSING = "<<missing>>"


class RunInputError(ValueError):
    def __init__(self, code, msg):
        super().__init__("error %d: %s" % (code, msg))
        self.code = code


def parse_ini(text):
    """{section: {key: value}} with GetPrivateProfileString's matching:
    section and key names compare case-insensitively, surrounding blanks are
    dropped, and one pair of enclosing double quotes is removed from a value.
    A repeated key keeps its first value."""
    out = {}
    sec = None
    for raw in text.splitlines():
        line = raw.strip()
        if not line or line[0] in ";#":
            continue
        if line.startswith("[") and "]" in line:
            sec = line[1:line.index("]")].strip().lower()
            out.setdefault(sec, {})
            continue
        if sec is None or "=" not in line:
            continue
        k, v = line.split("=", 1)
        k = k.strip().lower()
        v = v.strip()
        if len(v) >= 2 and v[0] == '"' and v[-1] == '"':
            v = v[1:-1]
        out[sec].setdefault(k, v)
    return out


def get(ini, section, key, default=MISSING):
    return ini.get(section.lower(), {}).get(key.lower(), default)


def _req(ini, section, key, code, path):
    v = get(ini, section, key)
    if v == MISSING:
        raise RunInputError(code, 'Property "%s" is not specified in the section "%s" in the input file "%s".'
                            % (key, section, path))
    return v


_NUM = re.compile(r"\s*[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")
_INT = re.compile(r"\s*[-+]?\d+")


def _scan(v, n, kind):
    """sscanf of n blank-separated "%f" (kind float, rounded to binary32) or
    "%d" (kind int) conversions: each takes the longest number at its
    position; None when fewer than n convert."""
    out, pos = [], 0
    rx = _INT if kind is int else _NUM
    for _ in range(n):
        m = rx.match(v, pos)
        if not m:
            return None
        out.append(int(m.group()) if kind is int else float(np.float32(float(m.group()))))
        pos = m.end()
    return out


def _floats(v, n, code, what, path):
    xs = _scan(v, n, float)
    if xs is None:
        raise RunInputError(code, 'Could not parse the %s "%s" in the input file "%s".' % (what, v, path))
    return xs


def _ints(v, n, code, what, path):
    xs = _scan(v, n, int)
    if xs is None:
        raise RunInputError(code, 'Could not parse the %s "%s" in the input file "%s".' % (what, v, path))
    return xs


def _switch(v, on):
    """The reference's switches compare the whole value, case included."""
    return v == on


@dataclasses.dataclass
class RunInput:
    path: str
    output_prefix: str
    output_directory: str
    histories: int                   # histories per MU: the "histories" value rounded to binary32, truncated
    gpu_number: int
    gpu_id: int
    data_directory: str
    egsinp_file: str
    pegs_file: str
    photon_xsections: str
    atomic_ff_file: str
    default_medium: str
    phsp_file_number: int
    phase_space_file: str
    phsp_z: float
    secjaws: bool
    secjaws_file: str
    mlc: bool
    mlcinfo_file: str
    block: bool
    wedge: bool
    use_ctcreate: bool
    ct_ramps_file: str | None
    dicom_file_list: str | None
    xyz_boundaries: list | None
    ct_voxel_sizes: list | None
    change_coordinate: bool
    egsphant_file: str | None
    isocenter: list
    source_to_isocenter: float
    dose_calibration: bool
    histories_per_cgy: float | None
    sections: dict
    histories_input: float = 0.0     # the "histories" value as binary32 (histories per MU)
    dose_unit: float = 1.6e-8        # MeV/g -> cGy, times histories per cGy / histories per MU with calibration
    output_factor: bool = False
    output_factor_file: str | None = None
    interpolate: bool = False
    interp_start: list | None = None      # starting point + half a voxel (CUDA_EGS.cu:1557-1559)
    interp_dimen: list | None = None
    interp_sizes: list | None = None
    absolute_dose: bool = False
    dose_position: list | None = None
    dose_measured: float | None = None
    relative_dose: bool = False
    reldose_start: list | None = None     # starting point + isocentre (CUDA_EGS.cu:1615-1617)
    reldose_dimen: list | None = None
    reldose_sizes: list | None = None
    array_type: int = 0                   # 1 PTW729, 2 PTW1000, 0 neither
    norm_type: int = 0                    # 1 Max, 2 Center, 0 none
    gamma_len: float | None = None        # cm
    gamma_dose: float | None = None       # fraction
    measured_data: str | None = None
    warnings: list = dataclasses.field(default_factory=list)


DOSE_UNIT = np.float32(1.6e-8)           # CUDA_EGS.h:536, MeV/g -> cGy


def read_run_input(path):
    """Parse a run input file with the reference's keys, defaults and error
    codes (CUDA_EGS.cu:936-1680).  File names are kept as written (the
    reference prefixes the data directory when it opens them)."""
    if not os.path.exists(path):
        raise RunInputError(1001, 'The input file "%s" does not exist.' % path)
    with open(path, "r", errors="replace") as f:
        ini = parse_ini(f.read())
    warn = []
    prefix = get(ini, "General", "output prefix")
    if prefix == MISSING:
        prefix = "GEGS"
        warn.append("No output prefix was specified. Default value is used.")
    outdir = get(ini, "General", "output directory")
    if outdir in (MISSING, ""):
        outdir = "."
        warn.append("No output directory was specified. Current Working Directory is used.")
    gpu_no = get(ini, "General", "GPU number")
    gpu_id = get(ini, "General", "GPU id")
    if gpu_no == MISSING:
        gpu_no = 1
    else:
        g = _scan(gpu_no, 1, int)
        if g is None:
            raise RunInputError(2004, 'The given GPU number "%s" could not be parsed as a 4-byte integer.' % gpu_no)
        gpu_no = g[0]
    if gpu_id == MISSING:
        gpu_id = 0
    else:
        g = _scan(gpu_id, 1, int)
        if g is None:
            raise RunInputError(2004, 'The given GPU id "%s" could not be parsed as a 4-byte integer.' % gpu_id)
        gpu_id = g[0]
    h = _req(ini, "General", "histories", 2013, path)
    hv = _scan(h, 1, float)
    if hv is None:
        raise RunInputError(2014, 'Could not parse the number of histories "%s".' % h)
    hist = hv[0]
    if hist <= 0:
        raise RunInputError(2015, "The number of histories must be positive, but %g was specified." % hist)
    data_dir = _req(ini, "Data", "data directory", 2017, path)
    egsinp = _req(ini, "Data", "egsinp file", 2011, path)
    pegs = _req(ini, "Data", "pegs file", 2012, path)
    xs = _req(ini, "Data", "photon xsections", 2019, path)
    ff = _req(ini, "Data", "atomic ff file", 2020, path)
    dmed = _req(ini, "Data", "default medium", 2013, path)
    nphsp = _req(ini, "Phase space", "phsp file number", 5013, path)
    nv = _scan(nphsp, 1, int)
    if nv is None:
        raise RunInputError(5014, 'Could not parse phsp file number "%s".' % nphsp)
    nphsp = nv[0]
    phsp = _req(ini, "Phase space", "phase space file", 2012, path)
    z = _req(ini, "Phase space", "z position", 5013, path)
    zv = _scan(z, 1, float)
    if zv is None:
        raise RunInputError(5014, 'Could not parse the z position "%s".' % z)
    z = zv[0]
    secjaws = _switch(get(ini, "SecJaws", "module SecJaws"), "On")
    secjaws_file = _req(ini, "SecJaws", "secjaws file", 2011, path)        # read with the jaws off too
    mlc = _switch(get(ini, "MLC", "module MLC"), "On")
    mlc_file = _req(ini, "MLC", "mlcinfo file", 2011, path)               # the field list, always
    block = _switch(get(ini, "Block", "module Block"), "On")
    wedge = _switch(get(ini, "Wedge", "module Wedge"), "On")
    if block:
        _req(ini, "Block", "block file", 2011, path)
    if wedge:
        _req(ini, "Wedge", "wedge file", 2011, path)
    use_ct = _switch(get(ini, "CTcreate", "use ctcreate"), "Yes")
    ramps = dlist = bounds = vox = egsphant = None
    if use_ct:
        ramps = _req(ini, "CTcreate", "CT ramps file", 2012, path)
        dlist = get(ini, "CTcreate", "dicom file list")
        if dlist == MISSING:
            dlist = _req(ini, "CTcreate", "dicom directory", 2012, path)
        bounds = _floats(_req(ini, "CTcreate", "xyz boundaries", 2011, path), 6, 2011,
                         "boundaries of the subset of CT", path)
        vox = _floats(_req(ini, "CTcreate", "3d voxel sizes", 2011, path), 3, 2011,
                      "voxel dimensions of CT", path)
    else:
        egsphant = _req(ini, "Phantom", "egsphant file", 2011, path)
    change = get(ini, "Phantom", "change coordinate") != "No"               # default Yes
    iso = _floats(_req(ini, "Phantom", "isocenter location", 2011, path), 3, 2011, "isocenter location", path)
    sid = _floats(_req(ini, "Phantom", "source to isocenter", 2011, path), 1, 2011, "source to isocenter", path)[0]
    f32 = np.float32
    cal = _switch(get(ini, "Calibration", "dose calibration"), "Yes")
    hpc = None
    dose_unit = DOSE_UNIT
    if cal:
        v = _req(ini, "Calibration", "histories per cGy", 5013, path)
        hv = _scan(v, 1, float)
        if hv is None:
            raise RunInputError(5014, 'Could not parse histories per cGy "%s".' % v)
        hpc = hv[0]
        dose_unit = f32(f32(dose_unit * f32(hpc)) / f32(hist))          # CUDA_EGS.cu:1478, in float
    r = RunInput(path, prefix, outdir, int(hist), gpu_no, gpu_id, data_dir, egsinp, pegs, xs, ff, dmed,
                 nphsp, phsp, z, secjaws, secjaws_file, mlc, mlc_file, block, wedge, use_ct, ramps, dlist,
                 bounds, vox, change, egsphant, iso, sid, cal, hpc, ini)
    r.histories_input = hist
    r.dose_unit = float(dose_unit)
    r.warnings = warn
    r.output_factor = _switch(get(ini, "Output factor", "use output factor"), "Yes")
    if r.output_factor:
        r.output_factor_file = _req(ini, "Output factor", "output factor file", 2012, path)
    r.interpolate = _switch(get(ini, "Interpolation", "do interpolate"), "Yes")
    if r.interpolate:
        st = _floats(_req(ini, "Interpolation", "starting point", 2011, path), 3, 2011,
                     "starting point of interpolation", path)
        r.interp_dimen = _ints(_req(ini, "Interpolation", "xyz dimensions", 2011, path), 3, 2011,
                               "xyz dimensions of interpolation", path)
        r.interp_sizes = _floats(_req(ini, "Interpolation", "3d voxel sizes", 2011, path), 3, 2011,
                                 "3d voxel sizes of interpolation", path)
        r.interp_start = [float(f32(f32(a) + f32(b) / f32(2))) for a, b in zip(st, r.interp_sizes)]
    r.absolute_dose = _switch(get(ini, "Dose output", "absolute dose"), "Yes")
    if r.absolute_dose:
        r.dose_position = _floats(_req(ini, "Dose output", "dose position", 2011, path), 3, 2011,
                                  "dose position of absolute dose", path)
        r.dose_measured = _floats(_req(ini, "Dose output", "dose measured", 2011, path), 1, 2011,
                                  "dose measured of absolute dose", path)[0]
    r.relative_dose = _switch(get(ini, "Dose output", "relative dose"), "Yes")
    if r.relative_dose:
        st = _floats(_req(ini, "Dose output", "starting point", 2011, path), 3, 2011,
                     "starting point of relative dose", path)
        r.reldose_start = [float(f32(f32(a) + f32(b))) for a, b in zip(st, iso)]
        r.reldose_dimen = _ints(_req(ini, "Dose output", "xyz dimensions", 2011, path), 3, 2011,
                                "xyz dimensions of relative dose", path)
        r.reldose_sizes = _floats(_req(ini, "Dose output", "3d voxel sizes", 2011, path), 3, 2011,
                                  "3d voxel sizes of relative dose", path)
        a = get(ini, "Dose output", "PTW729/PTW1000")
        r.array_type = 1 if a == "PTW729" else (2 if a == "PTW1000" else 0)
        nt = get(ini, "Dose output", "norm to Max/Center")
        r.norm_type = 1 if nt == "Max" else (2 if nt == "Center" else 0)
        g = _floats(_req(ini, "Dose output", "gamma index (mm,%)", 2011, path), 2, 2011,
                    "gamma index paremeters", path)
        r.gamma_len = float(f32(f32(g[0]) * f32(0.1)))
        r.gamma_dose = float(f32(f32(g[1]) * f32(0.01)))
        r.measured_data = _req(ini, "Dose output", "measured data", 2012, path)
    return r
