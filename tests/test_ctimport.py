"""CT front end: DICOM series reader (ReadCT_DICOM.c:316-909), SubsizeCT
(ctcreate.c:40-274) and ResampleCT (ctcreate.c:277-823).

The reader is checked on series written by ``ctimport.write_ct_slice``
(explicit and implicit VR, with and without preamble, slices listed out of
order, a localizer and a repeated z).  Cropping and resampling run in the
library; they are checked against loop-by-loop binary32 restatements of the
reference written here (np.float32 scalars, the reference's loop order), bit
for bit on the integers they return, and against closed forms: a constant CT
resamples to the same constant, a step resamples to the means of whole
output voxels, and a box on voxel faces selects exactly those voxels."""
import os

import numpy as np
import pytest

from gegs_amd import ctimport, phantom

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gegs_amd", "libgegs_hip.so")
needs_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="libgegs_hip.so not built")
f32 = np.float32


def _series(tmp, zs_mm, rows=6, cols=6, spacing=(1.5, 1.5), seed=0, **kw):
    rng = np.random.default_rng(seed)
    raw = {}
    names = []
    for n, z in enumerate(zs_mm):
        pix = rng.integers(0, 3000, (rows, cols)).astype(np.int16)
        nm = "s%02d.dcm" % n
        ctimport.write_ct_slice(os.path.join(tmp, nm), pix, (-4.5, -3.0, z), spacing,
                                implicit=(n % 2 == 1), preamble=(n % 3 != 2), **kw)
        raw[nm] = (z, pix)
        names.append(nm)
    return raw, names


def test_reader_orders_slices_and_drops_localizer_and_repeats(tmp_path):
    d = str(tmp_path)
    raw, names = _series(d, [12.0, 0.0, 6.0, 3.0, 9.0])
    # a localizer and a second slice at z = 6 mm, listed after the first
    loc = np.full((6, 6), 7, dtype=np.int16)
    ctimport.write_ct_slice(os.path.join(d, "loc.dcm"), loc, (0, 0, 1.0), (1.5, 1.5), localizer=True)
    ctimport.write_ct_slice(os.path.join(d, "dup.dcm"), loc, (-4.5, -3.0, 6.0), (1.5, 1.5))
    order = ["loc.dcm", names[2], names[0], "dup.dcm", names[1], names[3], names[4]]
    with open(os.path.join(d, "list.txt"), "w") as f:
        f.write("\n".join(order) + "\n")
    vol = ctimport.read_dicom_series(os.path.join(d, "list.txt"))
    assert tuple(vol.size) == (6, 6, 5)
    by_z = sorted(raw.values(), key=lambda t: t[0])
    for k, (z, pix) in enumerate(by_z):
        # slope 1, intercept -1024: CT number = stored value - 1024
        assert np.array_equal(vol.hu[k], (pix.astype(np.int32) - 1024).astype(np.int16))
    # dz from the first two slices; corner = first pixel centre - half a voxel
    assert vol.voxel[2] == f32(f32(0.3) - f32(0.0))
    assert vol.voxel[0] == f32(1.5 / 10.0) and vol.voxel[1] == f32(1.5 / 10.0)
    assert vol.offset[0] == f32(np.float64(f32(-0.45)) - np.float64(f32(0.15)) / 2.0)
    assert vol.offset[2] == f32(np.float64(f32(0.0)) - np.float64(vol.voxel[2]) / 2.0)


def test_rescale_needs_slope_and_intercept_nonzero(tmp_path):
    d = str(tmp_path)
    raw, names = _series(d, [0.0, 2.0], slope=2.0, intercept=0.0)
    with open(os.path.join(d, "list.txt"), "w") as f:
        f.write("\n".join(names))
    vol = ctimport.read_dicom_series(os.path.join(d, "list.txt"))
    # intercept 0: the reference leaves the stored values as they are
    assert np.array_equal(vol.hu[0], raw[names[0]][1])
    raw, names = _series(d, [0.0, 2.0], slope=0.5, intercept=-1000.0, seed=3)
    with open(os.path.join(d, "list.txt"), "w") as f:
        f.write("\n".join(names))
    vol = ctimport.read_dicom_series(os.path.join(d, "list.txt"))
    p = raw[names[0]][1]
    expect = np.trunc(p.astype(np.float32) * f32(0.5) + f32(-1000.0)).astype(np.int16)
    assert np.array_equal(vol.hu[0], expect)


# ------------------------------------------------------------------ restatements (binary32, reference loops)
def _bounds(lo, w, n):
    b = [f32(lo)]
    for _ in range(n):
        b.append(f32(b[-1] + f32(w)))
    return b


def _subsize_ref(hu, voxel, offset, bmin, bmax):
    nz, ny, nx = hu.shape
    n = [nx, ny, nz]
    first, last = [0, 0, 0], [0, 0, 0]
    whole = all(v == 0 for v in list(bmin) + list(bmax))
    for a in range(3):
        b = _bounds(offset[a], voxel[a], n[a])
        lo, hi = f32(bmin[a]), f32(bmax[a])
        if whole:
            lo, hi = b[0], b[n[a]]
        if (lo < b[0] and hi < b[0]) or (hi > b[n[a]] and lo > b[n[a]]):
            lo, hi = b[0], b[n[a]]
        if hi <= lo:
            lo, hi = b[0], b[n[a]]
        lo = max(lo, b[0]) if lo < b[0] else lo
        hi = b[n[a]] if hi > b[n[a]] else hi
        for i in range(n[a]):
            if b[i] <= lo < b[i + 1]:
                first[a] = i
            if b[i] <= hi <= b[i + 1]:
                last[a] = i
        if a == 2 and b[n[2] - 1] == hi:
            last[2] = n[2] - 1
    sub = hu[first[2]:last[2] + 1, first[1]:last[1] + 1, first[0]:last[0] + 1]
    off = [f32(f32(offset[a]) + f32(f32(voxel[a]) * f32(first[a]))) for a in range(3)]
    return sub, off


def _resample_ref(hu, voxel, offset, thick):
    nz, ny, nx = hu.shape
    n = [nx, ny, nz]
    out = [int(f32(f32(n[a]) * f32(voxel[a])) / f32(thick[a])) for a in range(3)]
    t = [f32(f32(f32(n[a]) * f32(voxel[a])) / f32(out[a])) for a in range(3)]
    xb = [_bounds(offset[a], t[a], out[a]) for a in range(3)]
    acc = np.zeros((out[2], out[1], out[0]), dtype=np.float32)
    w = [np.zeros(out[a], dtype=np.float32) for a in range(3)]
    lo_hi = [[0, 0] for _ in range(3)]

    def weights(a, c):
        b = xb[a]
        ctw = f32(voxel[a])
        lo = f32(b[0] + f32(ctw * f32(c)))
        hi = f32(b[0] + f32(ctw * f32(c + 1)))
        for i in range(out[a]):
            if b[i] <= lo and b[i + 1] >= lo:
                lo_hi[a][0] = i
            if b[i] <= hi and b[i + 1] >= hi:
                lo_hi[a][1] = i
                break
        i0, i1 = lo_hi[a]
        if i0 == i1:
            w[a][i0] = f32(ctw / f32(b[i0 + 1] - b[i0]))
            return
        for i in range(i0, i1 + 1):
            if b[i] >= lo and b[i + 1] <= hi:
                w[a][i] = f32(1.0)
            elif b[i] <= lo and b[i + 1] <= hi:
                w[a][i] = f32(f32(b[i + 1] - lo) / f32(b[i + 1] - b[i]))
            else:
                w[a][i] = f32(f32(hi - b[i]) / f32(b[i + 1] - b[i]))

    for ic in range(nx):
        weights(0, ic)
        for jc in range(ny):
            weights(1, jc)
            for kc in range(nz):
                weights(2, kc)
                v = f32(hu[kc, jc, ic])
                for i in range(lo_hi[0][0], lo_hi[0][1] + 1):
                    for j in range(lo_hi[1][0], lo_hi[1][1] + 1):
                        for k in range(lo_hi[2][0], lo_hi[2][1] + 1):
                            acc[k, j, i] = f32(acc[k, j, i] + f32(f32(f32(v * w[0][i]) * w[1][j]) * w[2][k]))
    return np.trunc(acc).astype(np.int32).astype(np.int16), xb


@needs_lib
@pytest.mark.parametrize("box", [((0, 0, 0), (0, 0, 0)), ((-1.0, -0.3, 0.2), (0.4, 0.9, 0.71)),
                                 ((-0.6, -0.45, 0.0), (0.3, 0.45, 0.6)), ((-9, -9, -9), (9, 9, 9)),
                                 ((2.0, 0.1, 0.1), (3.0, 0.2, 0.2)), ((0.5, 0.5, 0.5), (0.1, 0.1, 0.1))])
def test_subsize_matches_restatement(box):
    rng = np.random.default_rng(1)
    hu = rng.integers(-1000, 2000, (7, 9, 11)).astype(np.int16)
    vol = ctimport.CTVolume(hu, [0.15, 0.15, 0.3], [-0.825, -0.675, -0.15])
    got = ctimport.subsize(vol, *box)
    sub, off = _subsize_ref(hu, vol.voxel, vol.offset, *box)
    assert np.array_equal(got.hu, sub)
    assert [got.offset[a] for a in range(3)] == off


@needs_lib
def test_subsize_box_on_voxel_faces():
    hu = np.arange(5 * 6 * 7, dtype=np.int16).reshape(5, 6, 7)
    vol = ctimport.CTVolume(hu, [0.5, 0.5, 0.5], [0.0, 0.0, 0.0])
    got = ctimport.subsize(vol, (0.5, 1.0, 0.25), (1.75, 2.25, 1.25))
    # from the voxel holding the lower corner to the voxels whose closed extent holds the upper one
    assert np.array_equal(got.hu, hu[0:3, 2:5, 1:4])
    assert list(got.offset) == [0.5, 1.0, 0.0]


@needs_lib
@pytest.mark.parametrize("shape,voxel,thick", [((5, 7, 9), (0.15, 0.15, 0.3), (0.4, 0.5, 0.7)),
                                               ((4, 8, 8), (0.25, 0.25, 0.5), (0.5, 0.5, 1.0)),
                                               ((6, 5, 7), (0.117, 0.131, 0.29), (0.3, 0.29, 0.61))])
def test_resample_matches_restatement(shape, voxel, thick):
    rng = np.random.default_rng(7)
    hu = rng.integers(-1000, 2500, shape).astype(np.int16)
    vol = ctimport.CTVolume(hu, voxel, [-0.3, 0.2, 1.1])
    got, xb, yb, zb = ctimport.resample(vol, thick)
    ref, rb = _resample_ref(hu, vol.voxel, vol.offset, thick)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    assert list(xb) == rb[0] and list(yb) == rb[1] and list(zb) == rb[2]


@needs_lib
def test_resample_closed_forms_and_errors():
    hu = np.full((8, 8, 8), 1000, dtype=np.int16)
    vol = ctimport.CTVolume(hu, [0.25, 0.25, 0.25], [0.0, 0.0, 0.0])
    got, xb, _, _ = ctimport.resample(vol, (1.0, 0.5, 2.0))
    assert got.shape == (1, 4, 2) and np.all(got == 1000)
    assert list(xb) == [0.0, 1.0, 2.0]
    # a step in x averaged over whole output voxels
    hu = np.zeros((2, 2, 8), dtype=np.int16)
    hu[:, :, 4:] = 800
    got, _, _, _ = ctimport.resample(ctimport.CTVolume(hu, [0.25, 0.5, 0.5], [0, 0, 0]), (0.75, 1.0, 1.0))
    # int(2 cm / 0.75) = 2 output voxels of 1 cm: columns 0-3 and 4-7
    assert got.shape == (1, 1, 2)
    assert list(got[0, 0]) == [0, 800]
    for thick, code in (((0.0, 1, 1), 5002), ((3.0, 1, 1), 5003), ((1e-5, 1, 1), 5004)):
        with pytest.raises(ValueError, match=str(code)):
            ctimport.resample(ctimport.CTVolume(hu, [0.25, 0.5, 0.5], [0, 0, 0]), thick)


@needs_lib
def test_ctcreate_end_to_end(tmp_path):
    d = str(tmp_path)
    rng = np.random.default_rng(2)
    names = []
    for k in range(6):
        pix = (rng.integers(0, 20, (8, 8)) + np.where(np.arange(8) < 4, 24, 1024)[None, :]).astype(np.int16)
        nm = "c%d.dcm" % k
        ctimport.write_ct_slice(os.path.join(d, nm), pix, (-6.0, -6.0, 3.0 * k), (1.5, 1.5))
        names.append(nm)
    with open(os.path.join(d, "list.txt"), "w") as f:
        f.write("\n".join(reversed(names)))
    ph = ctimport.ctcreate(os.path.join(d, "list.txt"), (0.3, 0.3, 0.6))
    assert (ph.nx, ph.ny, ph.nz) == (4, 4, 3)
    vol = ctimport.read_dicom_series(os.path.join(d, "list.txt"))
    hu, _, _, _ = ctimport.resample(vol, (0.3, 0.3, 0.6))
    med, dens = phantom.ct_to_medium(hu.reshape(-1))
    assert np.array_equal(ph.media_idx, med) and np.array_equal(ph.density, dens)
    # left half near -1000 HU (air), right half near 0 HU (tissue)
    m = med.reshape(ph.nz, ph.ny, ph.nx)
    xc = 0.5 * (ph.xb[:-1] + ph.xb[1:])
    assert np.all(m[:, :, xc < 0] == 1) and np.all(m[:, :, xc > 0] == 3)


def test_ct_ramps_file(tmp_path):
    p = str(tmp_path / "ramps.inp")
    with open(p, "w") as f:
        f.write("2,-1000\n")
        f.write("AIR700ICRU".ljust(24) + "\n")
        f.write("-900,0.001,0.1,0\n")
        f.write("H2O700ICRU".ljust(24) + "\n")
        f.write("3000,0.9,2.5,0\n")
    ramp, lower = ctimport.read_ct_ramps(p)
    assert lower == -1000
    assert ramp == [("AIR700ICRU", -900, float(np.float32(0.001)), float(np.float32(0.1))),
                    ("H2O700ICRU", 3000, float(np.float32(0.9)), 2.5)]
    med, dens = phantom.ct_to_medium(np.array([-1000, -950, -900, 0, 3000, 3001]), ramp=ramp, lower=lower)
    assert list(med) == [1, 1, 1, 2, 2, 2]
    # below the ramp: vacuum as the reference (Q10); with the fix, the first medium at its lowest density
    med, dens = phantom.ct_to_medium(np.array([-1024, -1001, -1000]), ramp=ramp, lower=lower)
    assert list(med) == [0, 0, 1] and list(dens[:2]) == [0.0, 0.0]
    med, dens = phantom.ct_to_medium(np.array([-1024, -1001, -1000]), ramp=ramp, lower=lower, fix_q10=True)
    assert list(med) == [1, 1, 1] and dens[0] == np.float32(0.001) and dens[2] == np.float32(0.001)
    with open(p, "w") as f:
        f.write("0,0\n")
    assert ctimport.read_ct_ramps(p)[0] == list(phantom.DEFAULT_RAMP)
    with open(p, "w") as f:
        f.write("9,0\n")
    with pytest.raises(ValueError, match="5005"):
        ctimport.read_ct_ramps(p)


@needs_lib
def test_phantom_from_run_input(tmp_path):
    from gegs_amd import runinput
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "dicom"))
    names = []
    for k in range(4):
        pix = np.full((8, 8), 1024, dtype=np.int16)
        nm = "dicom/c%d.dcm" % k
        ctimport.write_ct_slice(os.path.join(d, nm), pix, (-6.0, -6.0, 3.0 * k), (1.5, 1.5))
        names.append(os.path.basename(nm))
    with open(os.path.join(d, "dicom", "list.dlt"), "w") as f:
        f.write("\n".join(names))
    with open(os.path.join(d, "ramps.inp"), "w") as f:
        f.write("0,0\n")
    text = open(os.path.join(os.path.dirname(__file__), "golden", "ini", "TestCalibration.ini")).read()
    text = text.replace("use ctcreate = No", "use ctcreate = Yes")
    text = text.replace('CT ramps file = "CT_ramps.inp"', 'CT ramps file = "ramps.inp"')
    text = text.replace('dicom file list = "..\\data\\Dicom\\PELICAN\\LungAbs.dlt"', 'dicom file list = "dicom\\list.dlt"')
    text = text.replace("xyz boundaries = -17.65 12.85 -26.95 -6.45 -19.65 -7.65", "xyz boundaries = -0.5 0.5 -0.4 0.4 0.0 0.8")
    text = text.replace("3d voxel sizes = 0.5 0.5 0.5\n\n[Phantom]", "3d voxel sizes = 0.3 0.3 0.3\n\n[Phantom]")
    ini = os.path.join(d, "run.ini")
    with open(ini, "w") as f:
        f.write(text)
    ri = runinput.read_run_input(ini)
    ph = ctimport.phantom_from_run_input(ri)
    vol = ctimport.subsize(ctimport.read_dicom_series(os.path.join(d, "dicom", "list.dlt")),
                           (-0.5, -0.4, 0.0), (0.5, 0.4, 0.8))
    hu, xb, _, _ = ctimport.resample(vol, (0.3, 0.3, 0.3))
    assert (ph.nx, ph.ny, ph.nz) == (hu.shape[2], hu.shape[1], hu.shape[0])
    assert np.array_equal(ph.xb, xb) and np.all(ph.media_idx == 3)
