"""The run-input driver's host logic (gegs_amd/run.py) against closed forms:
the field schedule (nRepeat, init_stack's batch size, the last batch and
quirk Q16), the phase-space files of a run (quirks Q4 and Q15), the
SECJAWS reader, the relative-dose outputs and the absolute-dose report.
The transport itself is covered by tests/test_gpu_run_driver.py."""
import os

import numpy as np
import pytest

from gegs_amd import head, postproc, run as grun, source

f32 = np.float32


# ------------------------------------------------------------------ schedule
def test_batch_size_of_init_stack():
    assert grun.batch_size_for(0) == 480000
    assert grun.batch_size_for(12) == 480000                  # 480000 * 13 * 1.5 = 9.36e6 < 1e7
    assert grun.batch_size_for(13) == int(1e7 / (14 * 1.5))   # 1.008e7 >= 1e7: 476190
    with pytest.raises(grun.RunError):
        grun.batch_size_for(255)                              # unsigned char wraps to 0


@pytest.mark.parametrize("num,total", [(2500, 1000), (999, 1000), (1, 7), (100_000_000, 20_000_000),
                                       (99_999_999, 20_000_000), (7_000_001, 500_000), (12345, 12345)])
def test_field_schedule_properties(num, total):
    s = grun.field_schedule(num, total)
    r = s.n_repeat + 1
    assert s.n_repeat == (num - 1) // total and r * total >= num
    loads = s.loads + s.dropped
    assert [a for a, _ in loads] == [b * s.batch_size for b in range(s.n_batch + 1)]
    assert all(n == s.batch_size for _, n in loads[:-1])
    # the loads end one record past num / r: floor(num / r) + 1 records in all
    assert loads[-1][0] + loads[-1][1] == num // r + 1
    # a load past the end of the phase space is refused (Q16)
    assert bool(s.dropped) == (num // r + 1 > total)
    assert s.transported == r * s.records
    # the fix transports exactly num histories from ceil(num / r) records
    x = grun.field_schedule(num, total, fix_q16=True)
    assert x.transported == num and not x.dropped
    assert x.records == -(-num // r)


def test_field_schedule_exact_multiple_drops_the_last_batch():
    s = grun.field_schedule(3000, 1000)                  # nRepeat 2, last batch 1001 records of 1000
    assert s.n_repeat == 2 and s.loads == [] and s.dropped == [(0, 1001)] and s.transported == 0
    s = grun.field_schedule(100_000_000, 20_000_000)    # 41 full batches, then 320001 records past the end
    assert s.n_batch == 41 and s.dropped == [(41 * 480000, 320001)] and s.transported == 5 * 41 * 480000


# ------------------------------------------------------------------ phase-space files
def _records(n, tag):
    rec = np.zeros(n, dtype=source.PHSP_DTYPE)
    rec["e"] = 1.0
    rec["x"] = np.arange(n, dtype=np.float32)
    rec["y"] = tag
    rec["wt"] = 1.0
    return rec


def test_single_file_reads_records_from_p0(tmp_path):
    base = str(tmp_path / "ph")
    source.write_egsphsp(base + ".egsphsp1", _records(50, 1.0))
    source.write_egsphsp(base + "_p0.egsphsp1", _records(60, 2.0))
    q4 = source.PhaseSpaceFiles(base, 1)
    assert q4.total == 50                                   # header of <name>.egsphsp1
    assert np.all(q4.read(10, 40)["y"] == 2.0)              # records of <name>_p0.egsphsp1 (Q4)
    fixed = source.PhaseSpaceFiles(base, 1, fix_q4=True)
    r = fixed.read(10, 40)
    assert np.all(r["y"] == 1.0) and list(r["x"][:3]) == [10, 11, 12]
    assert q4.read(10, 41) is None and q4.read(50, 1) is None


def test_reads_across_files(tmp_path):
    base = str(tmp_path / "ph")
    for i, n in enumerate((30, 20, 25)):
        source.write_egsphsp("%s_p%d.egsphsp1" % (base, i), _records(n, float(i)))
    p = source.PhaseSpaceFiles(base, 3)
    assert p.total == 75
    r = p.read(35, 10)                                      # inside file 1
    assert np.all(r["y"] == 1.0) and list(r["x"]) == list(range(5, 15))
    r = p.read(25, 10)                                      # 5 from file 0, then 5 from file 0 again (Q15)
    assert list(r["y"]) == [0.0] * 10 and list(r["x"]) == [25, 26, 27, 28, 29, 0, 1, 2, 3, 4]
    r = source.PhaseSpaceFiles(base, 3, fix_q15=True).read(25, 10)
    assert list(r["y"]) == [0.0] * 5 + [1.0] * 5 and list(r["x"]) == [25, 26, 27, 28, 29, 0, 1, 2, 3, 4]
    assert p.read(70, 6) is None


def test_missing_p0_file_names_the_quirk(tmp_path):
    base = str(tmp_path / "ph")
    source.write_egsphsp(base + ".egsphsp1", _records(5, 1.0))
    with pytest.raises(ValueError, match="Q4"):
        source.PhaseSpaceFiles(base, 1).read(0, 5)


# ------------------------------------------------------------------ head inputs
def test_secjaws_block_round_trip(tmp_path):
    p = str(tmp_path / "head.egsinp")
    head.write_dynvmlc(p, head.millennium_like())
    jaws = head.build_jaws(**head.CLINAC_LIKE_JAWS)
    head.write_egsinp_secjaws(p, jaws, "PB700ICRU", append=True)
    j, med = head.read_egsinp_secjaws(p)
    assert med == "PB700ICRU"
    assert (j.rmax, j.zminy, j.zmaxy, j.zminx, j.zmaxx) == (f32(20.0), f32(28.0), f32(35.8), f32(36.7), f32(44.5))
    assert head.read_dynvmlc(p).medium == "W700ICRU"         # both components in one file
    open(str(tmp_path / "none.egsinp"), "w").write("nothing\n")
    assert head.read_egsinp_secjaws(str(tmp_path / "none.egsinp")) is None


def test_host_paths():
    assert grun.host_path("/a/b", "..\\data") == "/a/data"
    assert grun.host_path("/a/b", "HEAD\\x.mlcinfo") == "/a/b/HEAD/x.mlcinfo"
    assert grun.host_path("/a/b", "/abs/f") == "/abs/f"


# ------------------------------------------------------------------ dose outputs
def test_host_bounds_are_voxel_centres():
    b = np.array([0.0, 1.0, 3.0, 6.0], dtype=np.float32)
    assert list(postproc.host_bounds(b)) == [0.5, 2.0, 4.5, 6.0]


def test_voxel_dose_order():
    e = np.array([1.0 / 3.0, 2.0], dtype=np.float64)
    m = np.array([0.7, 1.0], dtype=np.float32)
    d = postproc.voxel_dose(e, m, 1.6e-8)
    assert d[0] == f32(f32(f32(1.0 / 3.0) / f32(0.7)) * f32(1.6e-8))


def test_relative_dose_points_take_the_nearest_voxel():
    nz, ny, nx = 4, 3, 5
    dose = np.arange(nz * ny * nx, dtype=np.float32).reshape(nz, ny, nx)
    b = np.arange(6, dtype=np.float32)
    xb, yb, zb = postproc.host_bounds(b), postproc.host_bounds(b[:4]), postproc.host_bounds(b[:5])
    pts = postproc.reldose_points_nearest(dose, xb, yb, zb, (0.5, 1.5, 0.5), (1.0, 0.0, 1.0), (5, 1, 4))
    assert np.array_equal(pts.reshape(4, 5), dose[:, 1, :])
    # 0.4 voxels past a centre still takes that voxel (the +0.1 before truncation)
    pts = postproc.reldose_points_nearest(dose, xb, yb, zb, (1.89, 0.5, 0.5), (0.0, 0.0, 0.0), (1, 1, 1))
    assert pts[0] == dose[0, 0, 1]
    with pytest.raises(ValueError):
        postproc.reldose_points_nearest(dose, xb, yb, zb, (0.5, 0.5, 0.5), (1.0, 0.0, 0.0), (6, 1, 1))


def test_compare_ptw729_by_dose_difference():
    meas = np.array([1.0, 1.0, 0.05, 1.0], dtype=np.float32)          # cGy -> x 100 without normalisation
    calc = np.array([102.0, 104.0, 50.0, 97.5], dtype=np.float32)
    rows, total, count, pct = postproc.compare_measured(calc, meas, (4, 1, 1), (0, 0, 0), (1, 0, 0), (0, 0, 0),
                                                        norm_type=0, array_type=1, gamma_len=0.3, gamma_dose=0.03)
    assert [p for _, _, p in rows] == [1, -1, 0, 1]                   # 10% threshold drops the third point
    assert (total, count) == (3, 2) and pct == f32(f32(2) / f32(3))
    assert rows[0][0] == 100.0


def test_compare_ptw1000_takes_a_neighbour_within_the_distance():
    # 3 x 1 x 1 points; the middle one misses by 5% but its left neighbour
    # is within 1.5% of its measured value: sqrt((0.5)^2 + 0.25^2/0.3^2) < 1
    meas = np.array([1.0, 1.0, 1.0], dtype=np.float32)
    calc = np.array([101.5, 105.0, 108.0], dtype=np.float32)
    rows, total, count, _ = postproc.compare_measured(calc, meas, (3, 1, 1), (0, 0, 0), (1, 0, 0), (0, 0, 0),
                                                      norm_type=0, array_type=2, gamma_len=0.3, gamma_dose=0.03)
    assert [p for _, _, p in rows] == [1, 1, -1]
    # 2% off is too far: (2/3)^2 + 0.694 > 1
    calc2 = np.array([102.0, 105.0, 108.0], dtype=np.float32)
    rows, _, _, _ = postproc.compare_measured(calc2, meas, (3, 1, 1), (0, 0, 0), (1, 0, 0), (0, 0, 0),
                                              norm_type=0, array_type=2, gamma_len=0.3, gamma_dose=0.03)
    assert [p for _, _, p in rows] == [1, -1, -1]


def test_compare_normalisations():
    meas = np.array([1.0, 2.0, 4.0], dtype=np.float32)
    calc = np.array([10.0, 20.0, 40.0], dtype=np.float32)
    for norm in (1, 2):
        rows, total, count, _ = postproc.compare_measured(calc, meas, (3, 1, 1), (-1.0, 0, 0), (1.0, 0, 0),
                                                          (0.0, 0, 0), norm_type=norm, array_type=1,
                                                          gamma_len=0.3, gamma_dose=0.01)
        assert count == total == 3                                   # proportional profiles agree
        if norm == 2:                                                # point 1 sits at the isocentre
            assert rows[1][0] == 1.0 and rows[1][1] == 1.0


def test_relative_dose_file(tmp_path):
    mp = str(tmp_path / "PTW.txt")
    open(mp, "w").write("1.0 first\n1.0\n0.01\n")
    calc = np.array([101.0, 95.0, 1.0], dtype=np.float32)
    fn, log = postproc.write_relative_dose(str(tmp_path / "plan"), calc, (3, 1, 1), mp, start=(0, 0, 0),
                                           sizes=(1, 0, 0), isocenter=(0, 0, 0), norm_type=0, array_type=1,
                                           gamma_len=0.3, gamma_dose=0.03)
    assert os.path.basename(fn) == "plan_RelativeDose_003x001x001.txt"
    lines = open(fn).read().splitlines()
    assert lines[0] == "100.000000\t 101.000000\t 1" and lines[1] == "100.000000\t 95.000000\t -1"
    assert lines[3] == "Totally 2 points compared, 1 passed, percentage = 50 %" and log[-1] == lines[3]
    fn, log = postproc.write_relative_dose(str(tmp_path / "plan"), calc, (3, 1, 1), str(tmp_path / "missing"))
    assert open(fn).read().splitlines() == ["101.000000", "95.000000", "1.000000"]


def test_abs_dose_report_formats():
    dose = np.full((4, 4, 4), 200.0, dtype=np.float32)
    b = postproc.host_bounds(np.arange(5, dtype=np.float32))
    lines = postproc.abs_dose_report(dose, b, b, b, (1.2, 1.2, 1.2), True, 190.0)
    assert lines[1] == "  measured = 190.0 cGy, calculated = 200.0 cGy, difference = 5.3 %"
    lines = postproc.abs_dose_report(dose, b, b, b, (1.2, 1.2, 1.2), False, 0.0)
    assert lines[0] == "Absolute dose at (1.20, 1.20, 1.20) = 2.000E+02 cGy"


# ------------------------------------------------------------------ the driver's composition, transport faked
class FakeContext:
    """Stands in for native.Context: every history slot k of a job deposits
    its record's energy times (1 + h mod 3) into voxel h * 2654435761 mod
    nvox of its plan, h = history_offset + k - first, so that the dose
    depends on which record each history index gets and on the plan."""

    def __init__(self):
        self.calls = []
        self.nvox = 0

    def load_media(self, tables):
        self.calls.append(("media", tuple(tables.names)))

    def load_phantom(self, phan, regions):
        self.nvox = phan.nvox
        self.calls.append(("phantom", phan.nx, phan.ny, phan.nz))

    def load_mlc(self, mlc):
        self.calls.append(("mlc",))

    def load_jaws(self, jaws):
        self.calls.append(("jaws", jaws.zminy))

    def set_leaf_positions(self, pos):
        self.leaves = np.asarray(pos, dtype=np.float32).reshape(-1, 120)

    def set_jaw_positions(self, xy):
        self.jaw_xy = np.asarray(xy, dtype=np.float32).reshape(-1, 4)

    def set_fixes(self, mask):
        self.fixes = mask

    def set_source(self, src):
        self.src = np.array(src)

    def set_field(self, f):
        self.field = f

    def _deposit(self, dose, first, count, offset):
        h = offset + np.arange(count, dtype=np.int64)
        vox = (h * 2654435761) % self.nvox
        np.add.at(dose, vox, self.src["e"][first:first + count].astype(np.float64) * (1 + h % 3))

    def run_batched(self, jobs, n_plans, seed, n_batches=10):
        dose = np.zeros((n_plans, self.nvox))
        for j in jobs:
            self._deposit(dose[j["plan"]], j["first"], j["count"], j["history_offset"])
        return dict(dose=dose, counters=dict(overflow=0))

    def run(self, seed, n_batches=10, history_offset=0, first=0, count=0):
        # the plain kernel's history index is history_offset + record index
        # (gegs_run_plain), added modulo 2^64
        dose = np.zeros(self.nvox)
        self._deposit(dose, first, count, (history_offset + first) % (1 << 64))
        return dict(dose=dose)


def test_driver_composition_with_a_faked_transport(tmp_path):
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from rundata import HPM, by_hand, make_dataset
    ini, data = make_dataset(str(tmp_path))
    fake = FakeContext()
    res = grun.run(ini, base_dir=str(tmp_path), method="batched", seed=3, echo=False, context=fake)
    ref = FakeContext()
    dose, field_edep, hist = by_hand(data, 3, ref)
    assert res.histories == hist and [s.n_repeat for s in res.schedules] == [0, 1]
    assert np.array_equal(fake.src, ref.src)                          # records in the reference's order
    assert np.array_equal(fake.leaves, ref.leaves) and np.array_equal(fake.jaw_xy, ref.jaw_xy)
    assert np.array_equal(res.field_edep, field_edep)
    assert np.array_equal(res.dose, dose)
    assert np.array_equal(np.fromfile(res.files["dose"], dtype=np.float32), dose)
    plain = grun.run(ini, base_dir=str(tmp_path), method="plain", seed=3, echo=False, context=FakeContext())
    assert np.allclose(plain.field_edep, field_edep, rtol=1e-13, atol=0)
    log = open(res.files["log"]).read()
    assert "histories %d" % (2 * HPM) in log and "Absolute dose at (0.30, 1.20, -0.40):" in log
