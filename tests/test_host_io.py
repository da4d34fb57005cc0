"""The C++ host layer (include/gegs_host.h, gegs_amd/csrc/host_io.cpp) against
the package's Python readers and writers and the reference's rules in closed
form: the media-table and head files round trip byte for byte; the egsphant
reader gives the phantom, voxel masses and region records of
phantom.read_egsphant / voxel_mass / region_table, accepts what the
reference's fscanf reader accepts (media digits running across line breaks)
and refuses with its codes (8006-8013); the phase-space reader gives
the records of source.PhaseSpaceFiles (quirks Q4, Q15 and their fixes) and
refuses where the reference's read returns -1; the batches of a field equal
run.field_schedule (quirk Q16 and its fix); the dose file equals
postproc.voxel_dose.  Host only: nothing here touches a GPU."""
import ctypes
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from gegs_amd import hatch, head, hostio, phantom as ph, postproc, run as grun, source, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gegs_amd", "libgegs_hip.so")
RUNNER = os.path.join(ROOT, "gegs_amd", "gegs_run")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libgegs_hip.so not built")

F32 = np.float32


@pytest.fixture(scope="module")
def tables():
    return hatch.MediaTables(["AIR700ICRU", "LUNG700ICRU", "ICRUTISSUE700ICRU", "ICRPBONE700ICRU", "W700ICRU"],
                             mlc_medium="W700ICRU")


# ------------------------------------------------------------------ media-table file
def test_media_file_round_trip(tmp_path, tables):
    p = str(tmp_path / "t.gegsmed")
    hostio.write_media_file(p, tables)
    c, names = hostio.read_media_file(p)
    assert names == tables.names
    assert bytes(c) == bytes(tables.c)


def test_media_file_refuses_a_corrupt_file(tmp_path, tables):
    p = str(tmp_path / "t.gegsmed")
    hostio.write_media_file(p, tables)
    raw = bytearray(open(p, "rb").read())
    raw[20 + 5 * 64 + 1000] ^= 1                       # one bit of the table block
    open(p, "wb").write(bytes(raw))
    with pytest.raises(hostio.HostError, match="checksum"):
        hostio.read_media_file(p)
    open(p, "wb").write(b"NOTATABLEFILE")
    with pytest.raises(hostio.HostError) as e:
        hostio.read_media_file(p)
    assert e.value.code == 10029


# ------------------------------------------------------------------ egsphant
@pytest.fixture(scope="module")
def ct_file(tmp_path_factory):
    d = tmp_path_factory.mktemp("phan")
    phan = ph.ct_phantom(12, 10, 7, 0.5, seed=4)
    p = str(d / "ct.egsphant")
    ph.write_egsphant(p, phan)
    return p, phan


def test_egsphant_matches_the_python_reader(ct_file, tables):
    p, _ = ct_file
    py = ph.read_egsphant(p)
    cc = hostio.Egsphant(p)
    assert (cc.nx, cc.ny, cc.nz) == (py.nx, py.ny, py.nz)
    assert cc.media_names == py.media_names
    for a, b in ((cc.xb, py.xb), (cc.yb, py.yb), (cc.zb, py.zb), (cc.density, py.density)):
        assert np.array_equal(a, b)
    assert np.array_equal(cc.media_idx, py.media_idx)
    assert np.array_equal(cc.voxel_mass, py.voxel_mass())
    others = ["AIR700ICRU", "W700ICRU", "W700ICRU", "AIR700ICRU", "AIR700ICRU"]
    for flags in (ph.DEFAULT_FLAGS, ph.KN_FLAGS):
        want = ph.region_table(py, tables, others, flags=flags)
        got = cc.region_table(tables.c, tables.names, others, flags=flags)
        assert got.tobytes() == want.tobytes()
    with pytest.raises(hostio.HostError) as e:
        cc.region_table(tables.c, tables.names, ["AIR700ICRU", "NOSUCHMEDIUM", "W700ICRU", "AIR700ICRU",
                                                 "AIR700ICRU"])
    assert e.value.code == 10029
    cc.close()


def test_voxel_mass_closed_form(tmp_path):
    """init.cu:1256-1271: |(dx * dy) * dz| * density, a zero mass becomes 1."""
    xb = np.array([0.0, 0.5, 1.5], dtype=F32)
    yb = np.array([-1.0, 0.0, 0.25], dtype=F32)
    zb = np.array([2.0, 2.5, 4.0], dtype=F32)
    dens = np.array([1.0, 0.0, 2.5, 1.2, 0.3, 0.7, 1.0, 1.9], dtype=F32)
    mi = np.where(dens > 0, 1, 0)
    phan = ph.Phantom(["H2O700ICRU"], xb, yb, zb, mi, dens)
    p = str(tmp_path / "m.egsphant")
    ph.write_egsphant(p, phan)
    cc = hostio.Egsphant(p)
    want = []
    for k in range(2):
        for j in range(2):
            for i in range(2):
                v = abs(F32(F32(F32(xb[i + 1] - xb[i]) * F32(yb[j + 1] - yb[j])) * F32(zb[k + 1] - zb[k])))
                m = F32(v * dens[i + 2 * j + 4 * k])
                want.append(m if m != 0 else F32(1.0))
    assert np.array_equal(cc.voxel_mass, np.array(want, dtype=F32))


def _egsphant_text(rows_per_line):
    """A 3x2x2 phantom with two media; media digits of a z slice either one
    row per line or the whole slice on one line."""
    lines = ["2", "AIR700ICRU", "H2O700ICRU", "0.25 0.25", "    3    2    2",
             "0 1 2 3", "0 1 2", "0 1 2"]
    dig = [["120", "210"], ["111", "022"]]
    for k in range(2):
        lines += dig[k] if rows_per_line else ["".join(dig[k])]
        lines.append("")
    for k in range(2):
        for j in range(2):
            lines.append(" ".join("%.3f" % (0.001 * (1 + i + 3 * j + 6 * k)) for i in range(3)))
        lines.append("")
    return "\n".join(lines) + "\n"


def test_egsphant_digits_may_run_across_lines(tmp_path):
    """The reference reads the media digits with fscanf("%1d"), which skips
    line breaks: a slice written on one line reads as the same phantom."""
    a, b = tmp_path / "a.egsphant", tmp_path / "b.egsphant"
    a.write_text(_egsphant_text(True))
    b.write_text(_egsphant_text(False))
    pa, pb = hostio.Egsphant(str(a)), hostio.Egsphant(str(b))
    assert np.array_equal(pa.media_idx, pb.media_idx)
    assert list(pa.media_idx) == [1, 2, 0, 2, 1, 0, 1, 1, 1, 0, 2, 2]
    assert np.array_equal(pa.media_idx, ph.read_egsphant(str(a)).media_idx)


@pytest.mark.parametrize("edit,code", [
    (lambda t: t.replace("\n210\n", "\n210 7\n"), 8011),           # text after slice 1's last digit
    (lambda t: t.replace("\n\n0.001", "\n9\n0.001", 1), 8013),      # a tenth digit: densities shift by one
    (lambda t: t.replace("210", "2x0"), 8010),                      # not a digit
    (lambda t: t.replace("210", "290"), 8010),                      # medium 9 of 2
    (lambda t: t[:t.rindex("0.012")], 8012),                        # truncated densities
    (lambda t: t.replace("0 1 2 3", "0 1 two 3"), 8007),            # x bound
    (lambda t: t.replace("    3    2    2", "three"), 8006),
])
def test_egsphant_errors_keep_the_reference_codes(tmp_path, edit, code):
    p = tmp_path / "bad.egsphant"
    p.write_text(edit(_egsphant_text(True)))
    with pytest.raises(hostio.HostError) as e:
        hostio.Egsphant(str(p))
    assert e.value.code == code


# ------------------------------------------------------------------ egsphsp
def _records(n, tag):
    rng = np.random.default_rng(tag)
    rec = np.zeros(n, dtype=source.PHSP_DTYPE)
    rec["latch"] = rng.choice([0, 0x20000000, 0x40000000, 0x00000005], n)
    rec["e"] = rng.uniform(0.2, 6.0, n).astype(F32) * rng.choice([1, -1], n)
    rec["x"] = rng.uniform(-5, 5, n)
    rec["y"] = rng.uniform(-5, 5, n)
    u = rng.uniform(-0.6, 0.6, n).astype(F32)
    v = rng.uniform(-0.6, 0.6, n).astype(F32)
    rec["u"], rec["v"] = u, v
    rec["wt"] = rng.uniform(0.5, 1.5, n).astype(F32) * rng.choice([1, -1], n)
    rec["u"][0], rec["v"][0] = F32(0.8), F32(0.6000001)          # 1 - u^2 - v^2 < 0: W = 0
    return rec


@pytest.mark.parametrize("n_files,fix_q4,fix_q15", [(3, False, False), (3, False, True), (1, False, False),
                                                   (1, True, False)])
def test_phase_space_reader_matches_python(tmp_path, n_files, fix_q4, fix_q15):
    base = str(tmp_path / "ps")
    sizes = [700, 500, 300] if n_files > 1 else [900]
    recs = [_records(n, 10 + i) for i, n in enumerate(sizes)]
    if n_files > 1:
        for i, r in enumerate(recs):
            source.write_egsphsp("%s_p%d.egsphsp1" % (base, i), r)
    else:
        source.write_egsphsp(base + ".egsphsp1", recs[0])
        # the reference's single-file read takes its records from _p0 (Q4)
        source.write_egsphsp(base + "_p0.egsphsp1", _records(900, 99))
    fixes = (hostio.HOST_FIX_Q4 if fix_q4 else 0) | (hostio.HOST_FIX_Q15 if fix_q15 else 0)
    cc = hostio.PhaseSpace(base, n_files, fixes)
    py = source.PhaseSpaceFiles(base, n_files, fix_q4=fix_q4, fix_q15=fix_q15)
    assert cc.total == py.total == sum(sizes)
    assert cc.header_paths == py.header_paths
    z = F32(16.44)
    for istart, length in ((0, 10), (650, 100), (699, 2), (1150, 80), (5, sum(sizes) - 5), (sum(sizes) - 1, 1)):
        if length > sum(sizes) - istart:
            continue
        try:
            want = py.read(istart, length)
        except ValueError:
            # a Q15 read that takes more records from file i than it holds:
            # both readers refuse (the reference reads past the file's end)
            with pytest.raises(hostio.HostError, match="fewer than"):
                cc.read(istart, length, z)
            continue
        got = cc.read(istart, length, z)
        assert got.tobytes() == source.decode_phsp(want, z).tobytes(), (istart, length)
    assert cc.read(sum(sizes), 1, z) is None                       # istart past the total
    assert cc.read(sum(sizes) - 3, 4, z) is None                   # fewer records left than asked
    if n_files > 1 or fix_q4:
        assert cc.read(0, 1, z)["w"][0] == 0.0                    # record 0 has 1 - u^2 - v^2 < 0
    cc.close()


def test_phase_space_decoding_closed_form(tmp_path):
    """CUDA_EGS.cu:56-67: LATCH bit 30 electron, bit 29 positron; |E|; W from
    U, V with the sign of WT; |WT|."""
    rec = np.zeros(4, dtype=source.PHSP_DTYPE)
    rec["latch"] = [0x40000000, 0x20000000, 0x60000000, 7]
    rec["e"] = [-1.5, 2.0, 3.0, -0.25]
    rec["u"] = [0.6, 0.0, 0.0, 0.28]
    rec["v"] = [0.0, 0.8, 0.0, 0.96]
    rec["wt"] = [1.0, -2.0, 0.5, -1.0]
    base = str(tmp_path / "d")
    source.write_egsphsp(base + ".egsphsp1", rec)
    s = hostio.PhaseSpace(base, 1, hostio.HOST_FIX_Q4).read(0, 4, 3.0)
    assert list(s["charge"]) == [-1, 1, -1, 0]
    assert list(s["e"]) == [1.5, 2.0, 3.0, 0.25]
    assert list(s["wt"]) == [1.0, 2.0, 0.5, 1.0]
    assert list(s["z"]) == [3.0] * 4
    w = np.sqrt(F32(1) - rec["u"] * rec["u"] - rec["v"] * rec["v"]).astype(F32)
    assert np.array_equal(s["w"], np.array([w[0], -w[1], w[2], -w[3]], dtype=F32))


def test_phase_space_missing_p0_names_the_quirk(tmp_path):
    base = str(tmp_path / "one")
    source.write_egsphsp(base + ".egsphsp1", _records(10, 1))
    cc = hostio.PhaseSpace(base, 1, 0)
    with pytest.raises(hostio.HostError, match="Q4"):
        cc.read(0, 5, 0.0)
    assert hostio.PhaseSpace(base, 1, hostio.HOST_FIX_Q4).read(0, 5, 0.0) is not None


# ------------------------------------------------------------------ batches of a field
@pytest.mark.parametrize("num,total", [(2500, 1000), (999, 1000), (1, 7), (100_000_000, 20_000_000),
                                       (99_999_999, 20_000_000), (7_000_001, 500_000), (12345, 12345),
                                       (9998, 10000 // 2), (14 * 480_000 + 3, 480_000)])
@pytest.mark.parametrize("fix_q16", [False, True])
def test_field_schedule_matches_python(num, total, fix_q16):
    py = grun.field_schedule(num, total, fix_q16=fix_q16)
    s, loads = hostio.field_schedule(num, total, hostio.HOST_FIX_Q16 if fix_q16 else 0)
    assert (s.n_repeat, s.batch_size, s.n_batch, s.last_batch) == (py.n_repeat, py.batch_size, py.n_batch,
                                                                   py.last_batch)
    assert [tuple(map(int, l)) for l in loads] == py.loads
    assert s.n_dropped == len(py.dropped) and s.dropped_records == sum(n for _, n in py.dropped)
    assert list(s.repeat_counts[:s.n_repeat + 1]) == py.repeat_counts
    assert not any(s.repeat_counts[s.n_repeat + 1:])


def test_field_schedule_refuses_what_the_reference_refuses():
    with pytest.raises(hostio.HostError):
        hostio.field_schedule(0, 10)
    with pytest.raises(hostio.HostError):
        hostio.field_schedule(257 * 10, 10)                      # nRepeat 256 > 255
    with pytest.raises(hostio.HostError, match="wraps"):
        hostio.field_schedule(256 * 10, 10)                      # nRepeat 255: unsigned char wraps


# ------------------------------------------------------------------ dose file and head file
def test_dose_file_matches_voxel_dose(tmp_path):
    rng = np.random.default_rng(3)
    edep = rng.exponential(5.0, 4 * 3 * 2) * (rng.random(24) > 0.2)
    mass = rng.uniform(0.001, 2.0, 24).astype(F32)
    unit = F32(F32(F32(1.6e-8) * F32(6.02e13)) / F32(1000))
    root = str(tmp_path / "Run")
    got, fname = hostio.write_dose(root, 4, 3, 2, edep, mass, unit)
    assert fname == postproc.dose_file_name(root, 4, 3, 2)
    want = postproc.voxel_dose(edep, mass, unit)
    assert got.tobytes() == want.tobytes()
    assert np.fromfile(fname, dtype=np.float32).tobytes() == want.tobytes()


def test_head_file_round_trip(tmp_path):
    mlc = head.build_mlc(head.millennium_like())
    jaws = head.build_jaws(**head.CLINAC_LIKE_JAWS)
    bw = head.build_block_wedge(workloads.BLOCK_PYRAMIDS, workloads.WEDGE_PYRAMIDS)
    lp = np.stack([head.random_aperture(i) for i in range(3)]).astype(F32)
    jp = np.array([[-4, 4, -4, 4], [-3, 5, -4, 3], [-1, 1, -2, 2]], dtype=F32)
    p = str(tmp_path / "h.gegshead")
    hostio.write_head_file(p, mlc=mlc, jaws=jaws, block_wedge=bw, leaf_positions=lp, jaw_positions=jp)
    h = hostio.read_head_file(p)
    assert bytes(h["mlc"]) == bytes(mlc) and bytes(h["jaws"]) == bytes(jaws)
    assert np.array_equal(h["leaf_positions"], lp) and np.array_equal(h["jaw_positions"], jp)
    assert bytes(h["block_wedge"]) == bytes(bw)
    p2 = str(tmp_path / "j.gegshead")
    hostio.write_head_file(p2, jaws=jaws, jaw_positions=jp[:1])
    h2 = hostio.read_head_file(p2)
    assert h2["mlc"] is None and h2["leaf_positions"] is None and h2["block_wedge"] is None
    assert np.array_equal(h2["jaw_positions"], jp[:1])


# ------------------------------------------------------------------ prepared run
def test_prepare_writes_what_the_python_driver_loads(tmp_path):
    sys.path.insert(0, os.path.dirname(__file__))
    from rundata import FIELDS, HPM, make_dataset
    ini, data = make_dataset(str(tmp_path))
    rf = str(tmp_path / "plan.gegsrun")
    grun.prepare(ini, rf, base_dir=str(tmp_path), seed=9, n_batches=4, echo=False)
    kv = {}
    fields = []
    for line in open(rf):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        k, v = (s.strip() for s in line.split("=", 1))
        if k == "field":
            fields.append(v.split())
        else:
            kv[k] = v
    assert kv["histories per mu"] == str(HPM) and kv["phsp files"] == "2" and kv["seed"] == "9"
    assert kv["module first"] == str(head.M_SECJAWY) and kv["host fixes"] == "0"
    assert [int(f[0]) for f in fields] == [f[5] for f in FIELDS]
    # the field geometry's binary32 cos and sin, written to round-trip exactly
    a = float(F32(90.0 * 3.1415926 / 180.0))
    assert F32(fields[1][2]) == F32(math.cos(a)) and F32(fields[1][3]) == F32(math.sin(a))
    c, names = hostio.read_media_file(kv["media file"])
    tables = hatch.MediaTables(names, mlc_medium="W700ICRU")
    assert bytes(c) == bytes(tables.c)
    h = hostio.read_head_file(kv["head file"])
    assert h["leaf_positions"].shape == (len(FIELDS), 120) and h["jaw_positions"].shape == (len(FIELDS), 4)
    assert np.array_equal(h["jaw_positions"], np.array([f[1:5] for f in FIELDS], dtype=F32))
    cc = hostio.Egsphant(kv["egsphant file"])
    assert cc.nx == 32 and cc.media_names == ph.read_egsphant(kv["egsphant file"]).media_names


@pytest.mark.skipif(not os.path.exists(RUNNER), reason="gegs_run not built")
def test_runner_refuses_a_bad_run_file(tmp_path):
    """The C++ driver's input checks run before any device call."""
    p = tmp_path / "bad.gegsrun"
    p.write_text("media file = x.gegsmed\nno such key = 1\n")
    r = subprocess.run([RUNNER, str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "unknown key" in r.stderr
    p.write_text("media file = x.gegsmed\n")
    r = subprocess.run([RUNNER, str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "required" in r.stderr


# ------------------------------------------------------------------ head inputs
def test_secjaws_reader_matches_python(tmp_path):
    p = str(tmp_path / "j.egsinp")
    head.write_egsinp_secjaws(p, head.build_jaws(**head.CLINAC_LIKE_JAWS), "W700ICRU")
    py_j, py_med = head.read_egsinp_secjaws(p)
    cc_j, cc_med = hostio.read_secjaws(p)
    assert bytes(cc_j) == bytes(py_j) and cc_med == py_med == "W700ICRU"
    q = str(tmp_path / "none.egsinp")
    open(q, "w").write("no components here\n")
    assert hostio.read_secjaws(q) is None and head.read_egsinp_secjaws(q) is None


def test_dynvmlc_reader_and_builder_match_python(tmp_path):
    """The whole gegs_mlc (cross-sections, boundaries, leaf ends, default
    openings) built by the C++ reader equals head.build_mlc's, byte for byte,
    for the synthetic MLC and for one with tongue-and-groove widths, a shifted
    start and several opening groups."""
    import dataclasses
    p = str(tmp_path / "m.egsinp")
    d = head.millennium_like()
    head.write_dynvmlc(p, d)
    m_py = head.build_mlc(head.read_dynvmlc(p))
    m_cc, med = hostio.read_dynvmlc(p)
    assert bytes(m_cc) == bytes(m_py) and med == d.medium
    lf = head.Leaf(0.474, 0.03, 0.02, 0.05, 0.1, 0.07, 49.1, 50.0, 48.3, 54.0, 51.0, 52.0, 0.0, 50.5, 52.5)
    d2 = dataclasses.replace(d, full=lf, x_start=-14.1, rad_end=8.5, zthick=6.7,
                             openings=[(-1.5, 2.0, 20), (0.5, 0.5, 30), (-3.25, 4.0, 10)])
    p2 = str(tmp_path / "m2.egsinp")
    head.write_dynvmlc(p2, d2)
    assert bytes(hostio.read_dynvmlc(p2)[0]) == bytes(head.build_mlc(head.read_dynvmlc(p2)))
    q = str(tmp_path / "none.egsinp")
    open(q, "w").write("nothing\n")
    assert hostio.read_dynvmlc(q) is None


def test_pyramids_readers_and_builder_match_python(tmp_path):
    p = str(tmp_path / "bw.egsinp")
    head.write_egsinp_pyramids(p, block=workloads.BLOCK_PYRAMIDS, wedge=workloads.WEDGE_PYRAMIDS)
    b_py, bmed_py = head.read_egsinp_block(p)
    r_py, l_py, wmed_py = head.read_egsinp_wedge(p)
    b_cc, bmed_cc = hostio.read_block(p)
    r_cc, l_cc, wmed_cc = hostio.read_wedge(p)
    assert np.array_equal(b_cc, np.asarray(b_py, dtype=F32)) and bmed_cc == bmed_py
    assert F32(r_cc) == F32(r_py) and np.array_equal(l_cc, l_py) and wmed_cc == wmed_py
    for block, wedge in ((b_py, (r_py, l_py)), (b_py, None), (None, (r_py, l_py))):
        want = head.build_block_wedge(block, wedge)
        got = hostio.build_block_wedge(block, wedge)
        assert bytes(got) == bytes(want)
    # every slanted-side case of the wedge tracer
    for layers in ([[0, 1, 2, 2, -3, -1]], [[0, 1, 2, 2, -1, -3], [1, 2, 2, 2, -1.5, -3]],
                   [[0, 1, 1, 3, -2, -2]], [[0, 1, 3, 1, -2, -2], [1, 2, 3, 0.5, -2, -2]]):
        w = (7.5, np.asarray(layers, dtype=F32))
        assert bytes(hostio.build_block_wedge(None, w)) == bytes(head.build_block_wedge(None, w))


def test_mlcinfo_leaf_files_and_output_factors_match_python(tmp_path):
    sys.path.insert(0, os.path.dirname(__file__))
    from rundata import make_dataset
    ini, data = make_dataset(str(tmp_path))
    mi = os.path.join(data, "HEAD", "plan.mlcinfo")
    py = head.read_mlcinfo(mi)
    cc = hostio.read_mlcinfo(mi)
    assert len(cc) == len(py) == 2
    for a, b in zip(cc, py):
        assert (a["angle"], a["x1"], a["x2"], a["y1"], a["y2"], a["mu"]) == (b.angle, b.x1, b.x2, b.y1, b.y2, b.mu)
        assert a["leaf_file"] == b.leaf_file
        lp = os.path.join(data, b.leaf_file.replace("\\", "/"))
        assert np.array_equal(hostio.read_leaf_file(lp), head.read_leaf_file(lp))
    of = os.path.join(data, "output_factor.txt")
    s_py, f_py = postproc.read_output_factor_table(of)
    s_cc, f_cc = hostio.read_output_factors(of)
    assert np.array_equal(s_cc, s_py) and np.array_equal(f_cc, f_py)
    for x1, x2, y1, y2 in ((-4, 4, -4, 4), (-3, 5, -4, 3), (-1.1, 0.7, -2.3, 5.9), (-0.5, 0.5, -0.5, 0.5),
                           (-12, 12, -12, 12)):
        want = postproc.field_output_factor(x1, x2, y1, y2, s_py, f_py)
        got = hostio.field_output_factor(x1, x2, y1, y2, s_cc, f_cc)
        assert (got is None and want is None) or got == want, (x1, x2, y1, y2)


def _run_file_dict(path):
    kv, fields = {}, []
    for line in open(path):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        k, v = (s.strip() for s in line.split("=", 1))
        if k == "field":
            fields.append(v)
        else:
            kv[k] = v
    return kv, fields


@pytest.mark.skipif(not os.path.exists(RUNNER), reason="gegs_run not built")
@pytest.mark.parametrize("variant", ["plain", "fixes", "block_wedge", "interpolation", "ctcreate", "ctcreate_q10"])
def test_cpp_run_input_reader_matches_prepare(tmp_path, variant):
    """gegs_run -i <file.ini> --dry-run reads the run input, the .egsinp
    components, the field list, the leaf files and the output factors in
    C++ and writes the run file and head file that run.prepare writes from
    the package's Python readers: every key and every byte of the head
    geometry equal; with "use ctcreate = Yes" the phantom built from the
    DICOM series (the built-in ramp over the whole series, or a ramp file
    whose lowest CT number leaves the air below it, with and without the Q10
    fix, over a cropped box) is the same egsphant text byte for byte."""
    sys.path.insert(0, os.path.dirname(__file__))
    from rundata import add_ctcreate, make_dataset
    ini, data = make_dataset(str(tmp_path))
    fixes = ()
    if variant == "fixes":
        fixes = ("Q1", "Q7", "Q9", "Q15", "Q16")
    if variant == "ctcreate":
        add_ctcreate(str(tmp_path), ini)
    if variant == "ctcreate_q10":
        add_ctcreate(str(tmp_path), ini, ramp_text="2,-900\nAIR700ICRU\n-500,0.001,0.5,0\nICRUTISSUE700ICRU\n"
                                                   "3000,0.9,2.0,0\n",
                     box="-2.2 2.4 -1.9 2.6 -1.0 1.5", voxel="0.3 0.2 0.25")
        fixes = ("Q10",)
    if variant == "interpolation":
        open(ini, "a").write("[Interpolation]\ndo interpolate = Yes\nstarting point = -2.5 -1.0 -3.0\n"
                             "xyz dimensions = 21 11 9\n3d voxel sizes = 0.25 0.2 0.5\n")
    if variant == "block_wedge":
        head.write_egsinp_pyramids(os.path.join(data, "bw.egsinp"), block=workloads.BLOCK_PYRAMIDS,
                                   wedge=workloads.WEDGE_PYRAMIDS)
        text = open(ini).read() + "[Block]\nmodule Block = On\nblock file = \"bw.egsinp\"\n" \
                                  "[Wedge]\nmodule Wedge = On\nwedge file = \"bw.egsinp\"\n"
        open(ini, "w").write(text)
    rf_py = str(tmp_path / "py.gegsrun")
    grun.prepare(ini, rf_py, base_dir=str(tmp_path), seed=9, n_batches=4, fixes=fixes, echo=False)
    rf_cc = str(tmp_path / "cc.gegsrun")
    cmd = [RUNNER, "-i", ini, "--media", str(tmp_path / "py.gegsmed"), "--base-dir", str(tmp_path), "--seed", "9",
           "--batches", "4", "--dry-run", rf_cc]
    if fixes:
        cmd += ["--fix", ",".join(fixes)]
    ct_py = None
    if variant.startswith("ctcreate"):
        ct_py = open(os.path.join(str(tmp_path), "out", "Syn.egsphant"), "rb").read()
        os.remove(os.path.join(str(tmp_path), "out", "Syn.egsphant"))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    if ct_py is not None:
        ct_cc = open(os.path.join(str(tmp_path), "out", "Syn.egsphant"), "rb").read()
        assert ct_cc == ct_py
        phan = ph.read_egsphant(os.path.join(str(tmp_path), "out", "Syn.egsphant"))
        assert 0 not in phan.media_idx and len(set(phan.media_idx.tolist())) >= 2
        if variant == "ctcreate_q10":                                 # without the fix the air is vacuum
            from gegs_amd import ctimport
            ramp, lower = ctimport.read_ct_ramps(os.path.join(data, "ramps.inp"))
            bare = ctimport.ctcreate(os.path.join(str(tmp_path), "dicom", "list.dlt"), (0.3, 0.2, 0.25),
                                     (-2.2, -1.9, -1.0), (2.4, 2.6, 1.5), ramp=ramp, lower=lower)
            assert 0 in bare.media_idx
    kv_py, f_py = _run_file_dict(rf_py)
    kv_cc, f_cc = _run_file_dict(rf_cc)
    h_py, h_cc = kv_py.pop("head file"), kv_cc.pop("head file")
    assert kv_cc == kv_py
    assert f_cc == f_py
    assert open(h_cc, "rb").read() == open(h_py, "rb").read()
    if variant == "block_wedge":
        assert hostio.read_head_file(h_cc)["block_wedge"] is not None


@pytest.mark.skipif(not os.path.exists(RUNNER), reason="gegs_run not built")
def test_cpp_run_input_errors_keep_the_reference_codes(tmp_path):
    sys.path.insert(0, os.path.dirname(__file__))
    from rundata import make_dataset
    ini, data = make_dataset(str(tmp_path))
    text = open(ini).read()
    cases = [(text.replace("histories = ", "histories = x"), "ERROR (2014)"),
             (text.replace("[Phase space]\nphsp file number = 2\n", "[Phase space]\n"), "ERROR (5013)"),
             (text + "[CTcreate]\nuse ctcreate = Yes\n", "ERROR (2012)"),
             (text + "[CTcreate]\nuse ctcreate = Yes\nCT ramps file = \"r.inp\"\ndicom file list = \"l.dlt\"\n"
                     "xyz boundaries = 0 0 0 0 0\n3d voxel sizes = 0.2 0.2 0.2\n", "ERROR (2011)"),
             (text + "[CTcreate]\nuse ctcreate = Yes\nCT ramps file = \"r.inp\"\ndicom file list = \"l.dlt\"\n"
                     "xyz boundaries = 0 0 0 0 0 0\n3d voxel sizes = 0.2 0.2 0.2\n", "ERROR (5007)"),
             (text.replace("plan.mlcinfo", "nothere.mlcinfo"), "MLC info  file reading error")]
    for k, (t, want) in enumerate(cases):
        p = str(tmp_path / ("bad%d.ini" % k))
        open(p, "w").write(t)
        r = subprocess.run([RUNNER, "-i", p, "--media", "m.gegsmed", "--base-dir", str(tmp_path), "--dry-run",
                            str(tmp_path / "x.gegsrun")], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0 and want in r.stderr, (k, r.stderr)


def test_dose_outputs_match_postproc():
    """Interpolated and absolute dose of the C++ host equal postproc's (the
    reference's binary32 order), on host coordinates from host_bounds, with
    quirk Q9 and its fix; Q9 past the x bounds is refused by both."""
    rng = np.random.default_rng(8)
    nx, ny, nz = 9, 7, 12
    dose = rng.exponential(2.0, (nz, ny, nx)).astype(F32)
    xb = np.cumsum(np.r_[-4.0, rng.uniform(0.4, 1.1, nx)]).astype(F32)
    yb = np.cumsum(np.r_[-3.0, rng.uniform(0.4, 1.1, ny)]).astype(F32)
    zb = np.cumsum(np.r_[-5.0, rng.uniform(0.4, 1.1, nz)]).astype(F32)
    xc, yc, zc = (postproc.host_bounds(b) for b in (xb, yb, zb))
    for b, c in ((xb, xc), (yb, yc), (zb, zc)):
        assert np.array_equal(hostio.host_bounds(b), c)
    start = [F32(F32(-4.3) + F32(0.37) / F32(2)), F32(F32(-3.1) + F32(0.5) / F32(2)), F32(F32(-5.2) + F32(0.61) / F32(2))]
    sizes, dimen = [F32(0.37), F32(0.5), F32(0.61)], [23, 13, 19]
    want = postproc.interp_dose(dose, xc, yc, zc, start, sizes, dimen)
    got = hostio.interp_dose(dose, xc, yc, zc, start, sizes, dimen)
    assert got.tobytes() == want.tobytes() and got.any()
    for pos in ((0.1, 0.2, -1.0), (-3.3, 2.0, 3.9), (1.5, -1.9, -4.1)):
        for fix in (False, True):
            try:
                w = postproc.abs_dose(dose, xc, yc, zc, pos, fix_q9=fix)
            except ValueError:
                with pytest.raises(hostio.HostError, match="Q9"):
                    hostio.abs_dose(dose, xc, yc, zc, pos, fix_q9=fix)
                continue
            g = hostio.abs_dose(dose, xc, yc, zc, pos, fix_q9=fix)
            assert F32(g[0]) == F32(w[0]) and g[1] == w[1] and F32(g[2]) == F32(w[2]), (pos, fix)


@pytest.mark.parametrize("norm_type,array_type", [(1, 2), (2, 2), (0, 2), (2, 1), (1, 1)])
def test_relative_dose_matches_postproc(tmp_path, norm_type, array_type):
    """Points, normalisation, pass marks, the file and its summary line equal
    postproc's (output_RelDos_II's rules) for every normalisation and both
    array types; without the measured file both write the calculated values
    alone."""
    rng = np.random.default_rng(norm_type * 10 + array_type)
    nx, ny, nz = 14, 9, 16
    dose = (rng.exponential(1.0, (nz, ny, nx)) + 2.0).astype(F32)
    xb, yb, zb = (np.linspace(-a, a, n + 1).astype(F32) for a, n in ((3.5, nx), (2.25, ny), (4.0, nz)))
    xc, yc, zc = (postproc.host_bounds(b) for b in (xb, yb, zb))
    iso = [F32(0.1), F32(0.0), F32(-0.2)]
    start = [F32(F32(-2.0) + iso[0]), F32(F32(0.0) + iso[1]), F32(F32(-2.5) + iso[2])]
    sizes, dimen = [F32(0.5), F32(0.0), F32(0.5)], [9, 1, 11]
    want = postproc.reldose_points_nearest(dose, xc, yc, zc, start, sizes, dimen)
    got = hostio.reldose_points(dose, xc, yc, zc, start, sizes, dimen)
    assert got.tobytes() == want.tobytes()
    meas_path = str(tmp_path / "measured.txt")
    meas = (want * F32(1.0) + rng.normal(0, 0.08, len(want)).astype(F32)).astype(F32)
    with open(meas_path, "w") as f:
        f.write("".join("%g\n" % v for v in meas))
    kw = dict(start=start, sizes=sizes, isocenter=iso, norm_type=norm_type, array_type=array_type,
              gamma_len=F32(F32(3) * F32(0.1)), gamma_dose=F32(F32(3) * F32(0.01)))
    fn_py, lines = postproc.write_relative_dose(str(tmp_path / "py"), want, dimen, meas_path, **kw)
    fn_cc, summary = hostio.write_relative_dose(str(tmp_path / "cc"), got, dimen, meas_path, **kw)
    assert summary == lines[-1] and "points compared" in summary
    assert open(fn_cc).read() == open(fn_py).read()
    fn_py, lines = postproc.write_relative_dose(str(tmp_path / "py2"), want, dimen, str(tmp_path / "none.txt"), **kw)
    fn_cc, summary = hostio.write_relative_dose(str(tmp_path / "cc2"), got, dimen, str(tmp_path / "none.txt"), **kw)
    assert summary == lines[-1] == "Error: dose data file reading error!"
    assert open(fn_cc).read() == open(fn_py).read()
