"""A run input end to end on the GPU (gegs_amd/run.py) against the same run
composed from the library's low-level calls.

The synthetic data set has the shape of the reference's TestHeAbDo_5Fields
inputs: an egsphant phantom, one .egsinp file with the DYNVMLC and SECJAWS
components, an .mlcinfo list of fields with Varian leaf files, a phase space
split over two files (<name>_p0/_p1.egsphsp1), an output-factor table,
calibration, absolute dose and a relative-dose profile against a measured
file.  The composition by hand follows the reference's rules stated in
closed form: field i transports floor(MU_i x hpm / r_i) + 1 records r_i
times, r_i = (MU_i x hpm - 1) // N + 1, the records that run past file 0
come from the start of file 0 again (quirk Q15), the plan's deposit is the
sum of the fields' deposits times their output factors, and the dose file
holds (float) E / m x dose_unit with dose_unit = 1.6e-8 x hpc / hpm."""
import os

import numpy as np
import pytest

from gegs_amd import native, run as grun

from rundata import HPM, by_hand, make_dataset

pytestmark = pytest.mark.gpu


def test_run_input_matches_low_level_composition(tmp_path):
    ini, data = make_dataset(str(tmp_path))
    print("run driver: batched, 2 fields, %d histories per MU" % HPM, flush=True)
    res = grun.run(ini, base_dir=str(tmp_path), method="batched", seed=31, echo=False)
    ctx = native.Context(0)
    try:
        dose, field_edep, hist = by_hand(data, 31, ctx)
    finally:
        ctx.close()
    assert res.histories == hist
    assert [s.n_repeat for s in res.schedules] == [0, 1]
    # the same deposits; FP64 atomics add them in another order in each run
    assert np.allclose(res.field_edep, field_edep, rtol=1e-12, atol=1e-12 * field_edep.max())
    on_disk = np.fromfile(res.files["dose"], dtype=np.float32)
    assert os.path.basename(res.files["dose"]) == "Syn_DoseResults_032x032x024.bin"
    assert np.array_equal(on_disk, res.dose)
    assert np.allclose(res.dose, dose, rtol=3e-7, atol=0)
    assert dose.max() > 0
    log = open(res.files["log"]).read()
    assert "Absolute dose at (0.30, 1.20, -0.40):" in log and "points compared" in log
    assert os.path.exists(res.files["relative"])


def test_run_input_plain_equals_batched(tmp_path):
    ini, _ = make_dataset(str(tmp_path))
    print("run driver: plain and batched", flush=True)
    b = grun.run(ini, base_dir=str(tmp_path), method="batched", seed=5, echo=False)
    p = grun.run(ini, base_dir=str(tmp_path), method="plain", seed=5, echo=False)
    # the same histories: deposits differ only by the order of the FP64 sums
    assert np.allclose(p.field_edep, b.field_edep, rtol=1e-9, atol=1e-9 * b.field_edep.max())


RUNNER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gegs_amd", "gegs_run")


def test_cpp_driver_matches_the_python_driver(tmp_path):
    """The prepared run (run.prepare) transported by the C++ driver
    gegs_amd/gegs_run -- its own egsphant and phase-space readers, batch
    schedule, region records and dose file -- gives the Python driver's dose
    file; the deposits differ only by the order of the FP64 atomic sums."""
    import subprocess
    ini, _ = make_dataset(str(tmp_path))
    doses = {}
    for method in ("batched", "plain"):
        rf = str(tmp_path / ("plan_%s.gegsrun" % method))
        grun.prepare(ini, rf, base_dir=str(tmp_path), method=method, seed=31, n_batches=10, echo=False)
        print("C++ run driver: %s" % method, flush=True)
        r = subprocess.run([RUNNER, rf], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        out = os.path.join(str(tmp_path), "out", "Syn_DoseResults_032x032x024.bin")
        doses[method] = np.fromfile(out, dtype=np.float32)
        log = open(os.path.join(str(tmp_path), "out", "Syn_log.txt")).read()
        assert "repeat times 2" in log and "Output factor" in log
    print("python run driver: batched", flush=True)
    res = grun.run(ini, base_dir=str(tmp_path), method="batched", seed=31, echo=False)
    assert doses["batched"].max() > 0
    assert np.allclose(doses["batched"], res.dose, rtol=3e-7, atol=0)
    assert np.allclose(doses["plain"], res.dose, rtol=3e-7, atol=0)


def test_cpp_driver_reads_the_run_input(tmp_path):
    """gegs_run -i <file.ini> --media <tables>: the C++ driver reads the run
    input, head components, field list, leaf files and output factors itself
    and gives the Python driver's dose file (to the FP64 atomics' order)."""
    import subprocess
    ini, _ = make_dataset(str(tmp_path))
    open(ini, "a").write("[Interpolation]\ndo interpolate = Yes\nstarting point = -2.5 -1.0 -3.0\n"
                         "xyz dimensions = 21 11 9\n3d voxel sizes = 0.25 0.2 0.5\n")
    rf = str(tmp_path / "tables.gegsrun")
    grun.prepare(ini, rf, base_dir=str(tmp_path), seed=31, n_batches=10, echo=False)   # writes tables.gegsmed
    print("C++ run driver: -i run input", flush=True)
    r = subprocess.run([RUNNER, "-i", ini, "--media", str(tmp_path / "tables.gegsmed"), "--base-dir", str(tmp_path),
                        "--seed", "31"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    outdir = os.path.join(str(tmp_path), "out")
    cpp = np.fromfile(os.path.join(outdir, "Syn_DoseResults_032x032x024.bin"), dtype=np.float32)
    cpp_interp = open(os.path.join(outdir, "Syn_Interpolate_021x011x009.bin"), "rb").read()
    cpp_log = open(os.path.join(outdir, "Syn_log.txt")).read()
    cpp_rel = open(os.path.join(outdir, "Syn_RelativeDose_005x001x005.txt")).read().splitlines()
    print("python run driver: batched", flush=True)
    res = grun.run(ini, base_dir=str(tmp_path), method="batched", seed=31, echo=False)
    assert cpp.max() > 0
    assert np.allclose(cpp, res.dose, rtol=3e-7, atol=0)
    py_interp = open(res.files["interpolate"], "rb").read()
    assert cpp_interp[:12] == py_interp[:12]                                # int32 dimensions x, y, z
    a, b = np.frombuffer(cpp_interp[12:], dtype=np.float32), np.frombuffer(py_interp[12:], dtype=np.float32)
    assert a.any() and np.allclose(a, b, rtol=1e-6, atol=1e-6 * b.max())
    assert "Absolute dose at (0.30, 1.20, -0.40):" in cpp_log and "voxel dose of index" in cpp_log
    py_rel = open(res.files["relative"]).read().splitlines()
    assert len(cpp_rel) == len(py_rel) == 26 and cpp_rel[-1].startswith("Totally 25 points compared")
    assert cpp_rel[-1] in cpp_log


def test_cpp_driver_builds_the_ct_phantom(tmp_path):
    """gegs_run -i with "use ctcreate = Yes": the C++ driver builds the
    phantom from the DICOM series and the CT ramps file itself, writes it to
    <output root>.egsphant and transports the plan on it; the dose file equals
    the Python driver's on its own ctcreate phantom (to the FP64 atomics'
    order)."""
    import glob
    import subprocess
    from rundata import add_ctcreate
    ini, _ = make_dataset(str(tmp_path))
    add_ctcreate(str(tmp_path), ini)
    rf = str(tmp_path / "tables.gegsrun")
    grun.prepare(ini, rf, base_dir=str(tmp_path), seed=17, n_batches=10, echo=False)   # writes tables.gegsmed
    print("C++ run driver: -i run input with ctcreate", flush=True)
    r = subprocess.run([RUNNER, "-i", ini, "--media", str(tmp_path / "tables.gegsmed"), "--base-dir", str(tmp_path),
                        "--seed", "17"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    outdir = os.path.join(str(tmp_path), "out")
    (dose_file,) = glob.glob(os.path.join(outdir, "Syn_DoseResults_*.bin"))
    assert dose_file.endswith("_024x024x019.bin")
    cpp = np.fromfile(dose_file, dtype=np.float32)
    ct_cpp = open(os.path.join(outdir, "Syn.egsphant"), "rb").read()
    print("python run driver: batched with ctcreate", flush=True)
    res = grun.run(ini, base_dir=str(tmp_path), method="batched", seed=17, echo=False)
    assert open(os.path.join(outdir, "Syn.egsphant"), "rb").read() == ct_cpp
    assert cpp.max() > 0
    assert np.allclose(cpp, res.dose, rtol=3e-7, atol=0)
