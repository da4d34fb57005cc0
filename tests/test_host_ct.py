"""The C++ CT front end (gegs_amd/csrc/host_ct.cpp) against the package's
Python front end (gegs_amd/ctimport.py, phantom.ct_to_medium) and the
reference's rules: the DICOM series reader (readct_dicom,
ReadCT_DICOM.c:316-909) gives the same volume bit for bit on series with
explicit and implicit VR, with and without preamble, listed out of order, with
a localizer, a repeated z, blank and CRLF list lines and a rescale; both
refuse the same broken series with 5008; the ramps file reader
(ctcreate.c:886-1043) gives the same ramp and refuses with 5005, 5006 and
5007; CTToMedium (ctcreate.c:1052-1104) is computed in binary32 with the
ramp's densities held as binary32, as the reference declares them
(ctcreate.c:862-863), pinned by the ramp's end points and by a case where
the decimal density difference gives another density; ctcreate and the
egsphant text it writes (ctcreate.c:1249-1311) are equal byte for byte.
Host only: nothing here touches a GPU."""
import os

import numpy as np
import pytest

from gegs_amd import ctimport, hostio, phantom

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gegs_amd", "libgegs_hip.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libgegs_hip.so not built")
f32 = np.float32


def _series(d, zs_mm, rows=6, cols=7, spacing=(1.5, 1.25), seed=0, **kw):
    rng = np.random.default_rng(seed)
    names = []
    for n, z in enumerate(zs_mm):
        pix = rng.integers(0, 3000, (rows, cols)).astype(np.int16)
        nm = "s%02d.dcm" % n
        ctimport.write_ct_slice(os.path.join(d, nm), pix, (-4.5, -3.0, z), spacing,
                                implicit=(n % 2 == 1), preamble=(n % 3 != 2), **kw)
        names.append(nm)
    return names


def _same_volume(a, b):
    assert tuple(a.size) == tuple(b.size)
    assert np.array_equal(a.voxel, b.voxel) and np.array_equal(a.offset, b.offset)
    assert np.array_equal(a.hu, b.hu)


def test_series_reader_matches_python(tmp_path):
    d = str(tmp_path)
    names = _series(d, [12.0, 0.0, 6.0, 3.0, 9.0])
    loc = np.full((6, 7), 7, dtype=np.int16)
    ctimport.write_ct_slice(os.path.join(d, "loc.dcm"), loc, (0, 0, 1.0), (1.5, 1.25), localizer=True)
    ctimport.write_ct_slice(os.path.join(d, "dup.dcm"), loc, (-4.5, -3.0, 6.0), (1.5, 1.25))
    order = ["loc.dcm", names[2], "", names[0], "dup.dcm", "  ", names[1], names[3], names[4]]
    lst = os.path.join(d, "list.txt")
    with open(lst, "wb") as f:
        f.write("\r\n".join(order).encode())
    py, cc = ctimport.read_dicom_series(lst), hostio.read_dicom_series(lst)
    _same_volume(py, cc)
    assert tuple(cc.size) == (7, 6, 5)
    # rows along y, columns along x: x spacing is the second pixel spacing
    assert cc.voxel[0] == f32(1.25 / 10.0) and cc.voxel[1] == f32(1.5 / 10.0)
    assert cc.voxel[2] == f32(f32(0.3) - f32(0.0))


@pytest.mark.parametrize("slope,intercept", [(0.5, -1000.0), (2.0, 0.0), (0.0, -1024.0), (1.7, -1024.5),
                                             (30.0, 5.0)])
def test_rescale_matches_python(tmp_path, slope, intercept):
    """Slope and intercept both non-zero rescale in binary32 (ReadCT_DICOM.c:879-882);
    values past 16 bits (slope 30) keep the low half of the 32-bit conversion."""
    d = str(tmp_path)
    names = _series(d, [0.0, 2.0, 4.0], slope=slope, intercept=intercept, seed=3)
    lst = os.path.join(d, "list.txt")
    with open(lst, "w") as f:
        f.write("\n".join(names))
    py, cc = ctimport.read_dicom_series(lst), hostio.read_dicom_series(lst)
    _same_volume(py, cc)


def _broken(d, case):
    names = _series(d, [0.0, 2.0, 4.0])
    lst = os.path.join(d, "list.txt")
    p0 = os.path.join(d, names[0])
    raw = bytearray(open(p0, "rb").read())
    if case == "missing slice":
        names[1] = "nothere.dcm"
    elif case == "one slice":
        names = names[:1]
    elif case == "short pixels":
        open(p0, "wb").write(raw[:-4])
    elif case == "no pixel data":
        k = raw.rfind(b"\xe0\x7f\x10\x00")
        open(p0, "wb").write(raw[:k])
    elif case == "bad position":
        k = raw.find(b"-4.500000")
        raw[k:k + 9] = b"-4.5x0000"
        open(p0, "wb").write(raw)
    elif case == "position not finite":
        k = raw.find(b"-4.500000")
        raw[k:k + 9] = b"nan      "
        open(p0, "wb").write(raw)
    elif case == "no list":
        return os.path.join(d, "absent.txt")
    with open(lst, "w") as f:
        f.write("\n".join(names))
    return lst


@pytest.mark.parametrize("case", ["missing slice", "one slice", "short pixels", "no pixel data", "bad position",
                                  "position not finite", "no list"])
def test_series_errors_are_5008_in_both(tmp_path, case):
    lst = _broken(str(tmp_path), case)
    with pytest.raises(ValueError, match="5008"):
        ctimport.read_dicom_series(lst)
    with pytest.raises(hostio.HostError) as e:
        hostio.read_dicom_series(lst)
    assert e.value.code == 5008


def _ramp_file(path, text):
    with open(path, "w") as f:
        f.write(text)
    return path


def test_ramps_file_matches_python(tmp_path):
    p = _ramp_file(str(tmp_path / "r.inp"),
                   "3, -1000\n" + "AIR700ICRU".ljust(24) + "\n-900,0.001,0.1,0\n"
                   + "  H2O700ICRU   with a long tail\n" + "300 , 0.9 ,1.1,0\n"
                   + "ICRPBONE700ICRU\n" + "3000,1.1,2.5\n")
    py, cc = ctimport.read_ct_ramps(p), hostio.read_ct_ramps(p)
    assert py == cc
    assert cc[1] == -1000 and [r[0] for r in cc[0]] == ["AIR700ICRU", "H2O700ICRU   with a lo", "ICRPBONE700ICRU"]
    p = _ramp_file(str(tmp_path / "d.inp"), "0,0\n")
    ramp, lower = hostio.read_ct_ramps(p)
    assert lower == -1024 and [r[:2] for r in ramp] == [r[:2] for r in phantom.DEFAULT_RAMP]
    assert [(r[2], r[3]) for r in ramp] == [(float(f32(r[2])), float(f32(r[3]))) for r in phantom.DEFAULT_RAMP]


@pytest.mark.parametrize("text,code", [
    ("7,0\n", 5005), ("-1,0\n", 5005),
    ("1,-1000\nAIR\n-1001,0.001,0.1,0\n", 5006),       # upper bound below the lowest CT number
    ("1,-1000\nAIR\n-900,-0.001,0.1,0\n", 5006),       # negative density
    ("1,-1000\nAIR\n-900,0.2,0.1,0\n", 5006),          # densities out of order
    ("", 5007), ("x,0\n", 5007), ("2,-1000\nAIR\n-900,0.001,0.1,0\n", 5007),
    ("1,-1000\nAIR\n-900,0.001\n", 5007)])
def test_ramps_file_errors_keep_the_codes(tmp_path, text, code):
    p = _ramp_file(str(tmp_path / "r.inp"), text)
    with pytest.raises(ValueError, match=str(code)):
        ctimport.read_ct_ramps(p)
    with pytest.raises(hostio.HostError) as e:
        hostio.read_ct_ramps(p)
    assert e.value.code == code
    with pytest.raises(hostio.HostError) as e:
        hostio.read_ct_ramps(str(tmp_path / "absent.inp"))
    assert e.value.code == 5007


def _reference_ct_to_medium(hu, ramp, lower):
    """CTToMedium with the ramp in binary32 arrays, loop by loop (ctcreate.c:1056-1104)."""
    lo_d = [f32(r[2]) for r in ramp]
    hi_d = [f32(r[3]) for r in ramp]
    up = [r[1] for r in ramp]
    med, dens = [], []
    for ct in (int(v) for v in hu):
        if ct < lower:
            med.append(0)
            dens.append(f32(0.0))
        elif ct > up[-1]:
            med.append(len(ramp))
            dens.append(hi_d[-1])
        else:
            for i in range(len(ramp)):
                if ct <= up[i]:
                    base = lower if i == 0 else up[i - 1]
                    dens.append(lo_d[i] + f32(ct - base) * (hi_d[i] - lo_d[i]) / f32(up[i] - base))
                    med.append(i + 1)
                    break
    return np.array(med, dtype=np.int32), np.array(dens, dtype=np.float32)


def test_ct_to_medium_is_the_reference_binary32_arithmetic():
    rng = np.random.default_rng(11)
    hu = np.concatenate([np.arange(-1100, 2100), rng.integers(-2000, 4000, 5000)]).astype(np.int16)
    custom = [("A", -900, float(f32(0.001)), float(f32(0.1))), ("B", 300, float(f32(0.9)), float(f32(1.1))),
              ("C", 3000, float(f32(1.1)), 2.5)]
    for ramp, lower in ((phantom.DEFAULT_RAMP, phantom.DEFAULT_RAMP_LOWER), (custom, -1000)):
        want = _reference_ct_to_medium(hu, ramp, lower)
        for fix in (False, True):
            m_py, d_py = phantom.ct_to_medium(hu, ramp=ramp, lower=lower, fix_q10=fix)
            m_cc, d_cc = hostio.ct_to_medium(hu, ramp, lower, fix_q10=fix)
            assert np.array_equal(m_py, m_cc) and np.array_equal(d_py, d_cc)
            if not fix:
                assert np.array_equal(m_cc, want[0]) and np.array_equal(d_cc, want[1])
            else:
                below = hu < lower
                assert np.all(m_cc[below] == 1) and np.all(d_cc[below] == f32(ramp[0][2]))
                assert np.array_equal(d_cc[~below], want[1][~below])
    # closed forms: each ramp starts at its lower density, values past the top clamp
    lo = [phantom.DEFAULT_RAMP_LOWER] + [r[1] for r in phantom.DEFAULT_RAMP[:-1]]
    m, d = hostio.ct_to_medium(np.array([lo[0]] + [v + 1 for v in lo[1:]] + [1977, 32767]),
                               phantom.DEFAULT_RAMP, phantom.DEFAULT_RAMP_LOWER)
    assert list(m) == [1, 2, 3, 4, 4, 4]
    assert d[0] == f32(0.001) and d[4] == d[5] == f32(2.088)
    # the density difference between the binary32 ends, not the decimal one:
    # on the built-in ramp the two give different densities for tissue and bone
    tissue = np.arange(-723, 102)
    span_dec = f32(1.101 - 0.302)
    alt = f32(0.302) + (tissue + 724).astype(np.float32) * span_dec / f32(825)
    _, d = hostio.ct_to_medium(tissue, phantom.DEFAULT_RAMP, phantom.DEFAULT_RAMP_LOWER)
    assert np.any(alt != d)


def _ct_dataset(d, n=6, rows=10, cols=12):
    rng = np.random.default_rng(2)
    names = []
    for k in range(n):
        pix = (rng.integers(0, 40, (rows, cols)) + np.where(np.arange(cols) < 5, 24, 1024)[None, :]
               + np.where(np.arange(rows) > 6, 900, 0)[:, None]).astype(np.int16)
        nm = "c%d.dcm" % k
        ctimport.write_ct_slice(os.path.join(d, nm), pix, (-9.0, -7.5, 2.5 * k), (1.5, 1.5))
        names.append(nm)
    lst = os.path.join(d, "list.txt")
    with open(lst, "w") as f:
        f.write("\n".join(reversed(names)))
    return lst


@pytest.mark.parametrize("box,thick,fix", [((0, 0, 0, 0, 0, 0), (0.3, 0.3, 0.5), False),
                                           ((-0.5, 0.6, -0.4, 0.5, 0.1, 1.0), (0.2, 0.25, 0.3), True)])
def test_ctcreate_and_its_egsphant_text_match_python(tmp_path, box, thick, fix):
    d = str(tmp_path)
    lst = _ct_dataset(d)
    ramp, lower = phantom.DEFAULT_RAMP, phantom.DEFAULT_RAMP_LOWER
    bmin, bmax = (box[0], box[2], box[4]), (box[1], box[3], box[5])
    py = ctimport.ctcreate(lst, thick, bmin, bmax, ramp=ramp, lower=lower, fix_q10=fix)
    cc = hostio.ctcreate(lst, thick, bmin, bmax, ramp, lower, fix_q10=fix)
    assert (cc.nx, cc.ny, cc.nz) == (py.nx, py.ny, py.nz) and cc.media_names == py.media_names
    for a in ("xb", "yb", "zb"):
        assert np.array_equal(getattr(cc, a), getattr(py, a))
    assert np.array_equal(cc.media_idx, np.asarray(py.media_idx).reshape(-1))
    assert np.array_equal(cc.density, np.asarray(py.density, dtype=np.float32).reshape(-1))
    assert len(set(cc.media_idx.tolist())) >= 2
    p_py, p_cc = str(tmp_path / "py.egsphant"), str(tmp_path / "cc.egsphant")
    ctimport.write_egsphant_ctcreate(p_py, py)
    hostio.write_egsphant_ctcreate(p_cc, cc)
    assert open(p_cc, "rb").read() == open(p_py, "rb").read()
    back = hostio.Egsphant(p_cc)
    ref = phantom.read_egsphant(p_py)
    assert np.array_equal(back.media_idx, ref.media_idx) and np.array_equal(back.density, ref.density)


def test_ctcreate_refusals_pass_through(tmp_path):
    d = str(tmp_path)
    lst = _ct_dataset(d, n=3)
    with pytest.raises(hostio.HostError) as e:
        hostio.ctcreate(lst, (0.0, 0.3, 0.3), (0, 0, 0), (0, 0, 0), phantom.DEFAULT_RAMP, -1024)
    assert e.value.code == 5002
    with pytest.raises(hostio.HostError) as e:
        hostio.ctcreate(lst, (0.3, 0.3, 50.0), (0, 0, 0), (0, 0, 0), phantom.DEFAULT_RAMP, -1024)
    assert e.value.code == 5003
    with pytest.raises(hostio.HostError) as e:
        hostio.ctcreate(str(tmp_path / "absent.txt"), (0.3, 0.3, 0.3), (0, 0, 0), (0, 0, 0), phantom.DEFAULT_RAMP,
                        -1024)
    assert e.value.code == 5008
