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


def _floats(v, n, code, what, path):
    try:
        xs = [float(t) for t in v.split()]
    except ValueError:
        xs = []
    if len(xs) != n:
        raise RunInputError(code, 'Could not parse the %s "%s" in the input file "%s".' % (what, v, path))
    return xs


def _switch(v, on):
    return v != MISSING and v.strip().lower() == on.lower()


@dataclasses.dataclass
class RunInput:
    path: str
    output_prefix: str
    output_directory: str
    histories: int
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
    secjaws_file: str | None
    mlc: bool
    mlcinfo_file: str | None
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


def read_run_input(path):
    """Parse a run input file with the reference's keys, defaults and error
    codes (CUDA_EGS.cu:936-1480)."""
    if not os.path.exists(path):
        raise RunInputError(1001, 'The input file "%s" does not exist.' % path)
    with open(path, "r", errors="replace") as f:
        ini = parse_ini(f.read())
    prefix = get(ini, "General", "output prefix")
    if prefix == MISSING:
        raise RunInputError(2001, 'Property "output prefix" is not specified in the section "General".')
    outdir = get(ini, "General", "output directory")
    if outdir in (MISSING, ""):
        outdir = "."
    gpu_no = get(ini, "General", "GPU number")
    gpu_id = get(ini, "General", "GPU id")
    try:
        gpu_no = 1 if gpu_no == MISSING else int(gpu_no)
    except ValueError:
        raise RunInputError(2004, 'The given GPU number "%s" could not be parsed as a 4-byte integer.' % gpu_no)
    try:
        gpu_id = 0 if gpu_id == MISSING else int(gpu_id)
    except ValueError:
        raise RunInputError(2004, 'The given GPU id "%s" could not be parsed as a 4-byte integer.' % gpu_id)
    h = _req(ini, "General", "histories", 2013, path)
    try:
        hist = float(h)
    except ValueError:
        raise RunInputError(2014, 'Could not parse the number of histories "%s".' % h)
    if hist <= 0:
        raise RunInputError(2015, "The number of histories must be positive, but %g was specified." % hist)
    data_dir = _req(ini, "Data", "data directory", 2017, path)
    egsinp = _req(ini, "Data", "egsinp file", 2011, path)
    pegs = _req(ini, "Data", "pegs file", 2012, path)
    xs = _req(ini, "Data", "photon xsections", 2019, path)
    ff = _req(ini, "Data", "atomic ff file", 2020, path)
    dmed = _req(ini, "Data", "default medium", 2013, path)
    nphsp = _req(ini, "Phase space", "phsp file number", 5013, path)
    try:
        nphsp = int(nphsp)
    except ValueError:
        raise RunInputError(5014, 'Could not parse phsp file number "%s".' % nphsp)
    phsp = _req(ini, "Phase space", "phase space file", 2012, path)
    z = _req(ini, "Phase space", "z position", 5013, path)
    try:
        z = float(z)
    except ValueError:
        raise RunInputError(5014, 'Could not parse the z position "%s".' % z)
    secjaws = _switch(get(ini, "SecJaws", "module SecJaws"), "On")
    mlc = _switch(get(ini, "MLC", "module MLC"), "On")
    block = _switch(get(ini, "Block", "module Block"), "On")
    wedge = _switch(get(ini, "Wedge", "module Wedge"), "On")
    secjaws_file = _req(ini, "SecJaws", "secjaws file", 2011, path) if secjaws else None
    mlc_file = _req(ini, "MLC", "mlcinfo file", 2011, path) if mlc else None
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
    change = _switch(get(ini, "Phantom", "change coordinate"), "Yes")
    iso = _floats(_req(ini, "Phantom", "isocenter location", 2011, path), 3, 2011, "isocenter location", path)
    sid = _floats(_req(ini, "Phantom", "source to isocenter", 2011, path), 1, 2011, "source to isocenter", path)[0]
    cal = _switch(get(ini, "Calibration", "dose calibration"), "Yes")
    hpc = None
    if cal:
        v = _req(ini, "Calibration", "histories per cGy", 5013, path)
        try:
            hpc = float(v)
        except ValueError:
            raise RunInputError(5014, 'Could not parse histories per cGy "%s".' % v)
    return RunInput(path, prefix, outdir, int(hist), gpu_no, gpu_id, data_dir, egsinp, pegs, xs, ff, dmed,
                    nphsp, phsp, z, secjaws, secjaws_file, mlc, mlc_file, block, wedge, use_ct, ramps, dlist,
                    bounds, vox, change, egsphant if egsphant != MISSING else None, iso, sid, cal, hpc, ini)
