"""ctypes binding of include/gegs_host.h: the run's file formats as the C++
library reads and writes them (egsphant, egsphsp, media-table and head files,
the batches of a field, the dose file).  Host only: no GPU is touched."""
from __future__ import annotations

import ctypes

import numpy as np

from . import hatch, head as headmod, native
from . import phantom as ph
from . import source as srcmod

NAME_LEN = 64
MXMED = 6
HOST_FIX_Q4 = 1
HOST_FIX_Q15 = 2
HOST_FIX_Q16 = 4
PHSP_REFUSED = 1

HOST_SYMBOLS = [
    "gegs_host_last_error", "gegs_media_file_write", "gegs_media_file_read", "gegs_egsphant_read",
    "gegs_egsphant_free", "gegs_voxel_mass", "gegs_region_table", "gegs_phsp_open", "gegs_phsp_total",
    "gegs_phsp_header_path", "gegs_phsp_read", "gegs_phsp_close", "gegs_field_schedule", "gegs_write_dose",
    "gegs_head_file_write", "gegs_head_file_read", "gegs_head_free",
    "gegs_input_last_error", "gegs_read_secjaws", "gegs_read_dynvmlc", "gegs_read_block", "gegs_read_wedge",
    "gegs_build_block_wedge", "gegs_read_mlcinfo", "gegs_read_leaf_file", "gegs_read_output_factors",
    "gegs_field_output_factor", "gegs_host_bounds", "gegs_interp_dose", "gegs_write_interp_dose", "gegs_abs_dose",
    "gegs_reldose_points", "gegs_reldose_compare", "gegs_write_relative_dose",
    "gegs_ct_last_error", "gegs_dicom_read_series", "gegs_ct_volume_free", "gegs_read_ct_ramps",
    "gegs_ct_to_medium", "gegs_ctcreate", "gegs_egsphant_write_ctcreate",
]
NOT_FOUND = 1


class CEgsphant(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32), ("nmed", ctypes.c_uint32),
                ("names", ctypes.c_void_p), ("xb", ctypes.POINTER(ctypes.c_float)),
                ("yb", ctypes.POINTER(ctypes.c_float)), ("zb", ctypes.POINTER(ctypes.c_float)),
                ("media", ctypes.POINTER(ctypes.c_int32)), ("density", ctypes.POINTER(ctypes.c_float))]


class CCtVolume(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int32 * 3), ("voxel", ctypes.c_float * 3), ("offset", ctypes.c_float * 3),
                ("hu", ctypes.POINTER(ctypes.c_int16))]


class CCtRamp(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * NAME_LEN), ("ct_upper", ctypes.c_int32), ("density_lower", ctypes.c_float),
                ("density_upper", ctypes.c_float)]


class CSchedule(ctypes.Structure):
    _fields_ = [("histories", ctypes.c_uint64), ("n_repeat", ctypes.c_uint32), ("batch_size", ctypes.c_uint32),
                ("n_batch", ctypes.c_uint64), ("last_batch", ctypes.c_uint64), ("n_loads", ctypes.c_uint64),
                ("n_dropped", ctypes.c_uint64), ("dropped_records", ctypes.c_uint64), ("records", ctypes.c_uint64),
                ("repeat_counts", ctypes.c_uint64 * 256)]


class CHead(ctypes.Structure):
    _fields_ = [("has_mlc", ctypes.c_uint32), ("has_jaws", ctypes.c_uint32), ("has_block_wedge", ctypes.c_uint32),
                ("n_sets", ctypes.c_uint32), ("mlc", headmod.Mlc), ("jaws", headmod.Jaws),
                ("block_wedge", headmod.BlockWedge), ("leaf_positions", ctypes.POINTER(ctypes.c_float)),
                ("jaw_positions", ctypes.POINTER(ctypes.c_float))]


class CMlcField(ctypes.Structure):
    _fields_ = [("angle", ctypes.c_double), ("x1", ctypes.c_double), ("x2", ctypes.c_double),
                ("y1", ctypes.c_double), ("y2", ctypes.c_double), ("mu", ctypes.c_double),
                ("leaf_file", ctypes.c_char * 512)]


class HostError(ValueError):
    def __init__(self, code, msg):
        super().__init__("error %d: %s" % (code, msg))
        self.code = code


_L = None


def lib():
    global _L
    if _L is not None:
        return _L
    L = native.load()
    vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
    for name in HOST_SYMBOLS:
        if getattr(L, name, None) is None:
            raise RuntimeError("symbol %s missing from %s" % (name, native.LIB_PATH))
    L.gegs_host_last_error.restype = ctypes.c_char_p
    L.gegs_media_file_write.argtypes = [ctypes.c_char_p, vp, vp]
    L.gegs_media_file_read.argtypes = [ctypes.c_char_p, vp, vp]
    L.gegs_egsphant_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(CEgsphant)]
    L.gegs_egsphant_free.argtypes = [ctypes.POINTER(CEgsphant)]
    L.gegs_voxel_mass.argtypes = [ctypes.POINTER(CEgsphant), vp]
    L.gegs_region_table.argtypes = [ctypes.POINTER(CEgsphant), vp, vp, vp, ctypes.c_uint16, vp]
    L.gegs_phsp_open.argtypes = [ctypes.c_char_p, u32, u32, ctypes.POINTER(vp)]
    L.gegs_phsp_total.argtypes = [vp]
    L.gegs_phsp_total.restype = u64
    L.gegs_phsp_header_path.argtypes = [vp, u32]
    L.gegs_phsp_header_path.restype = ctypes.c_char_p
    L.gegs_phsp_read.argtypes = [vp, u64, u64, ctypes.c_float, vp]
    L.gegs_phsp_close.argtypes = [vp]
    L.gegs_field_schedule.argtypes = [u64, u64, u32, ctypes.POINTER(CSchedule), vp, u64]
    L.gegs_write_dose.argtypes = [ctypes.c_char_p, u32, u32, u32, vp, vp, ctypes.c_float, vp, ctypes.c_char_p,
                                  ctypes.c_size_t]
    L.gegs_head_file_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(CHead)]
    L.gegs_head_file_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(CHead)]
    L.gegs_head_free.argtypes = [ctypes.POINTER(CHead)]
    L.gegs_input_last_error.restype = ctypes.c_char_p
    L.gegs_read_secjaws.argtypes = [ctypes.c_char_p, vp, vp]
    L.gegs_read_dynvmlc.argtypes = [ctypes.c_char_p, vp, vp]
    L.gegs_read_block.argtypes = [ctypes.c_char_p, vp, vp]
    L.gegs_read_wedge.argtypes = [ctypes.c_char_p, vp, vp, vp, vp]
    L.gegs_build_block_wedge.argtypes = [vp, ctypes.c_float, vp, u32, vp]
    L.gegs_read_mlcinfo.argtypes = [ctypes.c_char_p, vp, u32, vp]
    L.gegs_read_leaf_file.argtypes = [ctypes.c_char_p, vp]
    L.gegs_read_output_factors.argtypes = [ctypes.c_char_p, vp, vp, u32, vp]
    L.gegs_field_output_factor.argtypes = [ctypes.c_double] * 4 + [vp, vp, u32, vp]
    L.gegs_ct_last_error.restype = ctypes.c_char_p
    L.gegs_dicom_read_series.argtypes = [ctypes.c_char_p, ctypes.POINTER(CCtVolume)]
    L.gegs_ct_volume_free.argtypes = [ctypes.POINTER(CCtVolume)]
    L.gegs_ct_volume_free.restype = None
    L.gegs_read_ct_ramps.argtypes = [ctypes.c_char_p, vp, vp, vp]
    L.gegs_ct_to_medium.argtypes = [vp, ctypes.c_size_t, vp, u32, ctypes.c_int32, ctypes.c_int, vp, vp]
    L.gegs_ctcreate.argtypes = [ctypes.c_char_p, vp, vp, vp, vp, u32, ctypes.c_int32, ctypes.c_int,
                                ctypes.POINTER(CEgsphant)]
    L.gegs_egsphant_write_ctcreate.argtypes = [ctypes.c_char_p, ctypes.POINTER(CEgsphant)]
    _L = L
    return L


def _check(rc):
    if rc:
        raise HostError(rc, lib().gegs_host_last_error().decode())


def _check_in(rc):
    """Status of a head-input reader: None for GEGS_NOT_FOUND, raises on errors."""
    if rc == NOT_FOUND:
        return None
    if rc:
        raise HostError(rc, lib().gegs_input_last_error().decode())
    return True


def _names_block(names, n=MXMED):
    buf = ctypes.create_string_buffer(n * NAME_LEN)
    for i, s in enumerate(names):
        b = s.encode()
        if len(b) >= NAME_LEN:
            raise ValueError("medium name %r longer than %d bytes" % (s, NAME_LEN - 1))
        ctypes.memmove(ctypes.addressof(buf) + i * NAME_LEN, b, len(b))
    return buf


def _decode_names(buf, n):
    return [bytes(buf[i * NAME_LEN:(i + 1) * NAME_LEN]).split(b"\0", 1)[0].decode() for i in range(n)]


# ------------------------------------------------------------------ media-table file
def write_media_file(path, tables):
    """The media tables (hatch.MediaTables) and their names in one file."""
    _check(lib().gegs_media_file_write(path.encode(), ctypes.addressof(tables.c), _names_block(tables.names)))


def read_media_file(path):
    """(hatch.CMediaTables, names) of a media-table file."""
    c = hatch.CMediaTables()
    names = ctypes.create_string_buffer(MXMED * NAME_LEN)
    _check(lib().gegs_media_file_read(path.encode(), ctypes.addressof(c), names))
    return c, _decode_names(names.raw, c.nmed)


# ------------------------------------------------------------------ egsphant, masses, regions
class Egsphant:
    """A phantom read by the C++ reader; arrays copied out, C memory freed."""

    def __init__(self, path):
        c = CEgsphant()
        _check(lib().gegs_egsphant_read(path.encode(), ctypes.byref(c)))
        self._adopt(c)

    @classmethod
    def _owning(cls, c):
        self = cls.__new__(cls)
        self._c = None
        self._adopt(c)
        return self

    def _adopt(self, c):
        try:
            self.nx, self.ny, self.nz, nmed = c.nx, c.ny, c.nz, c.nmed
            raw = ctypes.string_at(c.names, nmed * NAME_LEN)
            self.media_names = _decode_names(raw, nmed)
            self.xb = np.ctypeslib.as_array(c.xb, (self.nx + 1,)).copy()
            self.yb = np.ctypeslib.as_array(c.yb, (self.ny + 1,)).copy()
            self.zb = np.ctypeslib.as_array(c.zb, (self.nz + 1,)).copy()
            n = self.nx * self.ny * self.nz
            self.media_idx = np.ctypeslib.as_array(c.media, (n,)).copy()
            self.density = np.ctypeslib.as_array(c.density, (n,)).copy()
            self.voxel_mass = np.zeros(n, dtype=np.float32)
            _check(lib().gegs_voxel_mass(ctypes.byref(c), self.voxel_mass.ctypes.data))
            self._c = c
        except Exception:
            lib().gegs_egsphant_free(ctypes.byref(c))
            raise

    def region_table(self, tables_c, names, other_media, flags=ph.DEFAULT_FLAGS):
        out = np.zeros(self.nx * self.ny * self.nz + ph.VAC_AND_OTH, dtype=ph.REGION_DTYPE)
        _check(lib().gegs_region_table(ctypes.byref(self._c), ctypes.addressof(tables_c), _names_block(names),
                                       _names_block(other_media, 5), int(flags), out.ctypes.data))
        return out

    def close(self):
        if self._c is not None:
            lib().gegs_egsphant_free(ctypes.byref(self._c))
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ egsphsp
class PhaseSpace:
    """gegs_phsp: the run's phase-space files through the C++ reader."""

    def __init__(self, base, n_files, fixes=0):
        h = ctypes.c_void_p()
        _check(lib().gegs_phsp_open(base.encode(), int(n_files), int(fixes), ctypes.byref(h)))
        self.h = h
        self.total = int(lib().gegs_phsp_total(h))
        self.header_paths = []
        i = 0
        while True:
            p = lib().gegs_phsp_header_path(h, i)
            if p is None:
                break
            self.header_paths.append(p.decode())
            i += 1

    def read(self, istart, length, z):
        """Decoded source records (source.SOURCE_DTYPE), or None where the
        reference's read refuses."""
        out = np.zeros(int(length), dtype=srcmod.SOURCE_DTYPE)
        rc = lib().gegs_phsp_read(self.h, int(istart), int(length), float(z), out.ctypes.data)
        if rc == PHSP_REFUSED:
            return None
        _check(rc)
        return out

    def close(self):
        if self.h:
            lib().gegs_phsp_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ batches, dose file, head file
def field_schedule(histories, total_records, fixes=0):
    """(CSchedule, loads as an (n, 2) array of first record, count)."""
    s = CSchedule()
    _check(lib().gegs_field_schedule(int(histories), int(total_records), int(fixes), ctypes.byref(s), None, 0))
    loads = np.zeros((s.n_loads, 2), dtype=np.uint64)
    _check(lib().gegs_field_schedule(int(histories), int(total_records), int(fixes), ctypes.byref(s),
                                     loads.ctypes.data if s.n_loads else None, s.n_loads))
    return s, loads


def write_dose(root, nx, ny, nz, edep, mass, dose_unit):
    """(dose float32, file name) as the C++ writer makes them."""
    edep = np.ascontiguousarray(edep, dtype=np.float64)
    mass = np.ascontiguousarray(mass, dtype=np.float32)
    out = np.zeros(nx * ny * nz, dtype=np.float32)
    fname = ctypes.create_string_buffer(len(root) + 64)
    _check(lib().gegs_write_dose(root.encode(), nx, ny, nz, edep.ctypes.data, mass.ctypes.data,
                                 float(np.float32(dose_unit)), out.ctypes.data, fname, len(fname)))
    return out, fname.value.decode()


def write_head_file(path, mlc=None, jaws=None, block_wedge=None, leaf_positions=None, jaw_positions=None):
    h = CHead()
    keep = []
    n_sets = 0
    if mlc is not None:
        lp = np.ascontiguousarray(leaf_positions, dtype=np.float32).reshape(-1, 120)
        keep.append(lp)
        h.has_mlc, h.mlc = 1, mlc
        h.leaf_positions = lp.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        n_sets = len(lp)
    if jaws is not None:
        jp = np.ascontiguousarray(jaw_positions, dtype=np.float32).reshape(-1, 4)
        keep.append(jp)
        if mlc is not None and len(jp) != n_sets:
            raise ValueError("leaf and jaw position sets differ in number")
        h.has_jaws, h.jaws = 1, jaws
        h.jaw_positions = jp.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        n_sets = len(jp)
    if block_wedge is not None:
        h.has_block_wedge, h.block_wedge = 1, block_wedge
    h.n_sets = n_sets
    _check(lib().gegs_head_file_write(path.encode(), ctypes.byref(h)))


def read_head_file(path):
    """dict(mlc, jaws, block_wedge, leaf_positions, jaw_positions); absent
    parts are None."""
    h = CHead()
    _check(lib().gegs_head_file_read(path.encode(), ctypes.byref(h)))
    try:
        out = dict(mlc=None, jaws=None, block_wedge=None, leaf_positions=None, jaw_positions=None)
        if h.has_mlc:
            out["mlc"] = headmod.Mlc.from_buffer_copy(h.mlc)
            out["leaf_positions"] = np.ctypeslib.as_array(h.leaf_positions, (h.n_sets, 120)).copy()
        if h.has_jaws:
            out["jaws"] = headmod.Jaws.from_buffer_copy(h.jaws)
            out["jaw_positions"] = np.ctypeslib.as_array(h.jaw_positions, (h.n_sets, 4)).copy()
        if h.has_block_wedge:
            out["block_wedge"] = headmod.BlockWedge.from_buffer_copy(h.block_wedge)
        return out
    finally:
        lib().gegs_head_free(ctypes.byref(h))


# ------------------------------------------------------------------ head inputs of a run
def read_secjaws(path):
    """(head.Jaws, medium) of the SECJAWS component, or None."""
    j = headmod.Jaws()
    med = ctypes.create_string_buffer(NAME_LEN)
    if _check_in(lib().gegs_read_secjaws(path.encode(), ctypes.addressof(j), med)) is None:
        return None
    return j, med.value.decode()


def read_dynvmlc(path):
    """(head.Mlc built from the DYNVMLC component, leaf medium), or None."""
    m = headmod.Mlc()
    med = ctypes.create_string_buffer(NAME_LEN)
    if _check_in(lib().gegs_read_dynvmlc(path.encode(), ctypes.addressof(m), med)) is None:
        return None
    return m, med.value.decode()


def read_block(path):
    c = np.zeros(10, dtype=np.float32)
    med = ctypes.create_string_buffer(NAME_LEN)
    if _check_in(lib().gegs_read_block(path.encode(), c.ctypes.data, med)) is None:
        return None
    return c, med.raw.split(b"\0", 1)[0].decode()


def read_wedge(path):
    rmax = ctypes.c_float()
    layers = np.zeros(24, dtype=np.float32)
    n = ctypes.c_uint32()
    med = ctypes.create_string_buffer(NAME_LEN)
    if _check_in(lib().gegs_read_wedge(path.encode(), ctypes.byref(rmax), layers.ctypes.data, ctypes.byref(n),
                                       med)) is None:
        return None
    return rmax.value, layers[:6 * n.value].reshape(-1, 6).copy(), med.raw.split(b"\0", 1)[0].decode()


def build_block_wedge(block=None, wedge=None):
    """head.BlockWedge from block values (10) and/or wedge (rmax, layers)."""
    bw = headmod.BlockWedge()
    b = None if block is None else np.ascontiguousarray(block, dtype=np.float32)
    rmax, layers, n = 0.0, None, 0
    if wedge is not None:
        rmax = float(wedge[0])
        layers = np.ascontiguousarray(wedge[1], dtype=np.float32).reshape(-1)
        n = len(layers) // 6
    _check_in(lib().gegs_build_block_wedge(b.ctypes.data if b is not None else None, rmax,
                                           layers.ctypes.data if layers is not None else None, n,
                                           ctypes.addressof(bw)))
    return bw


def read_mlcinfo(path):
    n = ctypes.c_uint32()
    _check_in(lib().gegs_read_mlcinfo(path.encode(), None, 0, ctypes.byref(n)))
    arr = (CMlcField * max(1, n.value))()
    _check_in(lib().gegs_read_mlcinfo(path.encode(), ctypes.addressof(arr), n.value, ctypes.byref(n)))
    return [dict(angle=f.angle, x1=f.x1, x2=f.x2, y1=f.y1, y2=f.y2, mu=f.mu, leaf_file=f.leaf_file.decode())
            for f in arr[:n.value]]


def read_leaf_file(path):
    out = np.zeros(120, dtype=np.float32)
    _check_in(lib().gegs_read_leaf_file(path.encode(), out.ctypes.data))
    return out


def read_output_factors(path):
    n = ctypes.c_uint32()
    _check_in(lib().gegs_read_output_factors(path.encode(), None, None, 0, ctypes.byref(n)))
    sizes = np.zeros(max(1, n.value), dtype=np.float32)
    factors = np.zeros(max(1, n.value), dtype=np.float32)
    _check_in(lib().gegs_read_output_factors(path.encode(), sizes.ctypes.data, factors.ctypes.data, n.value,
                                             ctypes.byref(n)))
    return sizes[:n.value], factors[:n.value]


def field_output_factor(x1, x2, y1, y2, sizes, factors):
    sizes = np.ascontiguousarray(sizes, dtype=np.float32)
    factors = np.ascontiguousarray(factors, dtype=np.float32)
    of = ctypes.c_float()
    if _check_in(lib().gegs_field_output_factor(float(x1), float(x2), float(y1), float(y2), sizes.ctypes.data,
                                                factors.ctypes.data, len(sizes), ctypes.byref(of))) is None:
        return None
    return np.float32(of.value)


# ------------------------------------------------------------------ dose outputs
def _dose_outputs_lib():
    L = lib()
    vp, u32 = ctypes.c_void_p, ctypes.c_uint32
    L.gegs_host_bounds.argtypes = [vp, u32, vp]
    L.gegs_interp_dose.argtypes = [vp, u32, u32, u32, vp, vp, vp, vp, vp, vp, vp]
    L.gegs_abs_dose.argtypes = [vp, u32, u32, u32, vp, vp, vp, vp, ctypes.c_int, vp, vp, vp]
    return L


def host_bounds(b):
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = np.zeros_like(b)
    _check(_dose_outputs_lib().gegs_host_bounds(b.ctypes.data, len(b) - 1, out.ctypes.data))
    return out


def interp_dose(dose, xc, yc, zc, start, sizes, dimen):
    """The interpolated dose [dimen_z, dimen_y, dimen_x] of dose [nz, ny, nx]
    on the host coordinates xc, yc, zc (host_bounds)."""
    d = np.ascontiguousarray(dose, dtype=np.float32)
    nz, ny, nx = d.shape
    xc, yc, zc = (np.ascontiguousarray(b, dtype=np.float32) for b in (xc, yc, zc))
    st = np.ascontiguousarray(start, dtype=np.float32)
    sz = np.ascontiguousarray(sizes, dtype=np.float32)
    dm = np.ascontiguousarray(dimen, dtype=np.uint32)
    out = np.zeros((dm[2], dm[1], dm[0]), dtype=np.float32)
    _check(_dose_outputs_lib().gegs_interp_dose(d.ctypes.data, nx, ny, nz, xc.ctypes.data, yc.ctypes.data,
                                                zc.ctypes.data, st.ctypes.data, sz.ctypes.data, dm.ctypes.data,
                                                out.ctypes.data))
    return out


def abs_dose(dose, xc, yc, zc, pos, fix_q9=False):
    """(value, (i, j, k), voxel dose) as postproc.abs_dose."""
    d = np.ascontiguousarray(dose, dtype=np.float32)
    nz, ny, nx = d.shape
    xc, yc, zc = (np.ascontiguousarray(b, dtype=np.float32) for b in (xc, yc, zc))
    p = np.ascontiguousarray(pos, dtype=np.float32)
    val, vox = ctypes.c_float(), ctypes.c_float()
    ijk = np.zeros(3, dtype=np.uint32)
    _check(_dose_outputs_lib().gegs_abs_dose(d.ctypes.data, nx, ny, nz, xc.ctypes.data, yc.ctypes.data,
                                             zc.ctypes.data, p.ctypes.data, 1 if fix_q9 else 0, ctypes.byref(val),
                                             ijk.ctypes.data, ctypes.byref(vox)))
    return val.value, tuple(int(v) for v in ijk), vox.value


def reldose_points(dose, xc, yc, zc, start, sizes, dimen):
    L = lib()
    vp, u32 = ctypes.c_void_p, ctypes.c_uint32
    L.gegs_reldose_points.argtypes = [vp, u32, u32, u32, vp, vp, vp, vp, vp, vp, vp]
    d = np.ascontiguousarray(dose, dtype=np.float32)
    nz, ny, nx = d.shape
    xc, yc, zc = (np.ascontiguousarray(b, dtype=np.float32) for b in (xc, yc, zc))
    st, sz = (np.ascontiguousarray(v, dtype=np.float32) for v in (start, sizes))
    dm = np.ascontiguousarray(dimen, dtype=np.uint32)
    out = np.zeros(int(dm.prod()), dtype=np.float32)
    _check(L.gegs_reldose_points(d.ctypes.data, nx, ny, nz, xc.ctypes.data, yc.ctypes.data, zc.ctypes.data,
                                 st.ctypes.data, sz.ctypes.data, dm.ctypes.data, out.ctypes.data))
    return out


def write_relative_dose(root, calc, dimen, measured_path, start, sizes, isocenter, norm_type, array_type, gamma_len,
                        gamma_dose):
    """(file name, summary line) as the C++ writer makes them."""
    L = lib()
    vp, u32 = ctypes.c_void_p, ctypes.c_uint32
    L.gegs_write_relative_dose.argtypes = [ctypes.c_char_p, vp, vp, ctypes.c_char_p, vp, vp, vp, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_char_p,
                                           ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
    c = np.ascontiguousarray(calc, dtype=np.float32)
    dm = np.ascontiguousarray(dimen, dtype=np.uint32)
    st, sz, iso = (np.ascontiguousarray(v, dtype=np.float32) for v in (start, sizes, isocenter))
    fname = ctypes.create_string_buffer(len(root) + 64)
    summary = ctypes.create_string_buffer(512)
    _check(L.gegs_write_relative_dose(root.encode(), c.ctypes.data, dm.ctypes.data,
                                      measured_path.encode() if measured_path else None, st.ctypes.data,
                                      sz.ctypes.data, iso.ctypes.data, int(norm_type), int(array_type),
                                      float(gamma_len), float(gamma_dose), fname, len(fname), summary,
                                      len(summary)))
    return fname.value.decode(), summary.value.decode()


# ------------------------------------------------------------------ CT front end
def _check_ct(rc):
    if rc:
        raise HostError(rc, lib().gegs_ct_last_error().decode())


def read_dicom_series(list_path):
    """gegs_dicom_read_series as a ctimport.CTVolume."""
    from .ctimport import CTVolume
    v = CCtVolume()
    _check_ct(lib().gegs_dicom_read_series(list_path.encode(), ctypes.byref(v)))
    try:
        nx, ny, nz = v.size
        hu = np.ctypeslib.as_array(v.hu, (nz * ny * nx,)).reshape(nz, ny, nx).copy()
        return CTVolume(hu, list(v.voxel), list(v.offset))
    finally:
        lib().gegs_ct_volume_free(ctypes.byref(v))


def _ramp_array(ramp):
    arr = (CCtRamp * MXMED)()
    if not 1 <= len(ramp) <= MXMED:
        raise ValueError("a CT ramp has 1 to %d media" % MXMED)
    for i, (name, up, dlo, dhi) in enumerate(ramp):
        arr[i].name = name.encode()
        arr[i].ct_upper = int(up)
        arr[i].density_lower = float(dlo)
        arr[i].density_upper = float(dhi)
    return arr


def read_ct_ramps(path):
    """gegs_read_ct_ramps: ([(name, CT upper bound, density low, density high)], lowest CT number),
    densities as the binary32 values the library holds."""
    arr = (CCtRamp * MXMED)()
    n = ctypes.c_uint32()
    lower = ctypes.c_int32()
    _check_ct(lib().gegs_read_ct_ramps(path.encode(), arr, ctypes.byref(n), ctypes.byref(lower)))
    return [(arr[i].name.decode(), arr[i].ct_upper, arr[i].density_lower, arr[i].density_upper)
            for i in range(n.value)], lower.value


def ct_to_medium(hu, ramp, lower, fix_q10=False):
    """gegs_ct_to_medium: (1-based media, binary32 densities) of int16 CT numbers."""
    hu = np.ascontiguousarray(hu, dtype=np.int16).reshape(-1)
    med = np.zeros(hu.size, dtype=np.int32)
    dens = np.zeros(hu.size, dtype=np.float32)
    _check_ct(lib().gegs_ct_to_medium(hu.ctypes.data, hu.size, _ramp_array(ramp), len(ramp), int(lower),
                                      int(bool(fix_q10)), med.ctypes.data, dens.ctypes.data))
    return med, dens


def ctcreate(list_path, thickness, box_min, box_max, ramp, lower, fix_q10=False):
    """gegs_ctcreate: the phantom of a DICOM series as an Egsphant (C memory owned)."""
    f3 = lambda v: (ctypes.c_float * 3)(*[float(x) for x in v])
    c = CEgsphant()
    _check_ct(lib().gegs_ctcreate(list_path.encode(), f3(thickness), f3(box_min), f3(box_max), _ramp_array(ramp),
                                  len(ramp), int(lower), int(bool(fix_q10)), ctypes.byref(c)))
    return Egsphant._owning(c)


def write_egsphant_ctcreate(path, phantom):
    """gegs_egsphant_write_ctcreate of any phantom with media_names, bounds,
    media_idx and density (x fastest)."""
    nx, ny, nz = len(phantom.xb) - 1, len(phantom.yb) - 1, len(phantom.zb) - 1
    keep = [np.ascontiguousarray(phantom.xb, dtype=np.float32), np.ascontiguousarray(phantom.yb, dtype=np.float32),
            np.ascontiguousarray(phantom.zb, dtype=np.float32),
            np.ascontiguousarray(np.asarray(phantom.media_idx).reshape(-1), dtype=np.int32),
            np.ascontiguousarray(np.asarray(phantom.density).reshape(-1), dtype=np.float32)]
    names = _names_block(phantom.media_names, max(1, len(phantom.media_names)))
    fp = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    c = CEgsphant(nx, ny, nz, len(phantom.media_names), ctypes.addressof(names), fp(keep[0], ctypes.c_float),
                  fp(keep[1], ctypes.c_float), fp(keep[2], ctypes.c_float), fp(keep[3], ctypes.c_int32),
                  fp(keep[4], ctypes.c_float))
    _check_ct(lib().gegs_egsphant_write_ctcreate(path.encode(), ctypes.byref(c)))
