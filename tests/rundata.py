"""A synthetic run-input data set shaped like the reference's
TestHeAbDo_5Fields inputs, shared by the CPU and GPU tests of the run-input
driver (gegs_amd/run.py)."""
import os

import numpy as np

from gegs_amd import ctimport, hatch, head, phantom as ph, postproc, source, workloads

F32 = np.float32
N0, N1 = 6000, 4000
FIELDS = [  # angle, x1, x2, y1, y2, MU
    (0.0, -4.0, 4.0, -4.0, 4.0, 2),
    (90.0, -3.0, 5.0, -4.0, 3.0, 3),
]
HPM = 4999
HPC = 6.02e5


def make_dataset(tmp):
    data = os.path.join(tmp, "data")
    os.makedirs(os.path.join(data, "HEAD"))
    phan = ph.ct_phantom(32, 32, 24, 0.5, seed=5)
    ph.write_egsphant(os.path.join(data, "ct.egsphant"), phan)
    egsinp = os.path.join(data, "VarianMLC.egsinp")
    head.write_dynvmlc(egsinp, head.millennium_like())
    head.write_egsinp_secjaws(egsinp, head.build_jaws(**head.CLINAC_LIKE_JAWS), "W700ICRU", append=True)
    fl = []
    for i, (a, x1, x2, y1, y2, mu) in enumerate(FIELDS):
        leaf = "HEAD/f%d.mlc" % i
        head.write_leaf_file(os.path.join(data, leaf), head.random_aperture(10 + i))
        fl.append(head.MlcField(a, x1, x2, y1, y2, mu, "HEAD\\f%d.mlc" % i))
    head.write_mlcinfo(os.path.join(data, "HEAD", "plan.mlcinfo"), fl)
    rec = source.encode_phsp(source.head_phase_space(N0 + N1, workloads.PHSP_Z, 100.0, 10.0, 3))
    source.write_egsphsp(os.path.join(data, "VarNoJaws_p0.egsphsp1"), rec[:N0])
    source.write_egsphsp(os.path.join(data, "VarNoJaws_p1.egsphsp1"), rec[N0:])
    with open(os.path.join(data, "output_factor.txt"), "w") as f:
        f.write("field size, output factor\n4\n2 0.9\n6 0.96\n10 1.0\n20 1.05\n")
    with open(os.path.join(data, "PTW1000.txt"), "w") as f:
        f.write("".join("%g\n" % (1.0 + 0.01 * k) for k in range(5 * 5)))
    ini = os.path.join(tmp, "run.ini")
    with open(ini, "w") as f:
        f.write("""[General]
output prefix = "Syn"
output directory = "out"
histories = %d
[Data]
data directory = "data"
egsinp file = "VarianMLC.egsinp"
pegs file = "700icru.pegs4dat"
photon xsections = "si"
atomic ff file = "pgs4form.dat"
default medium = AIR700ICRU
[Phase space]
phsp file number = 2
phase space file = "VarNoJaws"
z position = 16.44
[SecJaws]
module SecJaws = On
secjaws file = "VarianMLC.egsinp"
[MLC]
module MLC = On
mlcinfo file = "HEAD\\plan.mlcinfo"
[Phantom]
change coordinate = Yes
egsphant file = "ct.egsphant"
isocenter location = 0 1 0
source to isocenter = 100
[Calibration]
dose calibration = Yes
histories per cGy = %g
[Output factor]
use output factor = Yes
output factor file = "output_factor.txt"
[Dose output]
absolute dose = Yes
dose position = 0.3 1.2 -0.4
dose measured = 1.0
relative dose = Yes
starting point = -2.0 0.0 -2.0
xyz dimensions = 5 1 5
3d voxel sizes = 1.0 0.0 1.0
PTW729/PTW1000 = PTW1000
gamma index (mm,%%) = 3 3
norm to Max/Center = Max
measured data = "PTW1000.txt"
""" % (HPM, HPC))
    return ini, data


def add_ctcreate(tmp, ini, ramp_text="0,0\n", box="0 0 0 0 0 0", voxel="0.25 0.25 0.25"):
    """Switch the run input to a phantom built from DICOM (use ctcreate =
    Yes): a 40 x 40 x 16 CT series of 1.5 mm pixels and 3 mm slices (air
    around an elliptic water body with a lung and a bone insert, stored as
    CT number + 1024 with slope 1 and intercept -1024), listed in reverse
    order in dicom/list.dlt, and the CT ramps file data/ramps.inp."""
    d = os.path.join(tmp, "dicom")
    os.makedirs(d)
    rng = np.random.default_rng(8)
    yy, xx = np.mgrid[0:40, 0:40]
    names = []
    for k in range(16):
        hu = np.full((40, 40), -1000)
        hu[((xx - 19.5) / 18) ** 2 + ((yy - 19.5) / 15) ** 2 < 1] = 0
        hu[((xx - 12) / 5) ** 2 + ((yy - 18) / 7) ** 2 < 1] = -800
        hu[(xx - 28) ** 2 + (yy - 24) ** 2 < 16] = 1000
        pix = (hu + 1024 + rng.integers(-20, 21, hu.shape)).astype(np.int16)
        nm = "ct%02d.dcm" % k
        ctimport.write_ct_slice(os.path.join(d, nm), pix, (-29.25, -29.25, -22.5 + 3.0 * k), (1.5, 1.5))
        names.append(nm)
    with open(os.path.join(d, "list.dlt"), "w") as f:
        f.write("\n".join(names[::-1]) + "\n")
    with open(os.path.join(tmp, "data", "ramps.inp"), "w") as f:
        f.write(ramp_text)
    with open(ini, "a") as f:
        f.write('[CTcreate]\nuse ctcreate = Yes\nCT ramps file = "ramps.inp"\ndicom file list = "dicom\\list.dlt"\n'
                "xyz boundaries = %s\n3d voxel sizes = %s\n" % (box, voxel))


def by_hand(data, seed, ctx):
    """The same run from the library's calls, with the schedule in closed form."""
    phan = ph.read_egsphant(os.path.join(data, "ct.egsphant"))
    others = ["AIR700ICRU", "W700ICRU", "W700ICRU", "AIR700ICRU", "AIR700ICRU"]
    names = list(phan.media_names) + [n for n in ("W700ICRU",) if n not in phan.media_names]
    tables = hatch.MediaTables(names, mlc_medium="W700ICRU")
    regions = ph.region_table(phan, tables, others)
    _, r0 = source.read_egsphsp(os.path.join(data, "VarNoJaws_p0.egsphsp1"))
    srcs, jobs, ofs, hist = [], [], [], 0
    sizes = np.array([2, 6, 10, 20], dtype=F32)
    factors = np.array([0.9, 0.96, 1.0, 1.05], dtype=F32)
    first = 0
    for i, (a, x1, x2, y1, y2, mu) in enumerate(FIELDS):
        num = mu * HPM
        r = (num - 1) // (N0 + N1) + 1
        n = num // r + 1
        assert N0 < n <= N0 + N1                      # the load runs past file 0 ...
        rec = np.concatenate([r0, r0[:n - N0]])       # ... and continues in file 0 again (Q15)
        srcs.append(source.decode_phsp(rec, F32(16.44)))
        fld = workloads.make_field(True, 100.0, float(F32(a)), (0.0, 1.0, 0.0), head.M_SECJAWY)
        for q in range(r):
            jobs.append(dict(field=fld, plan=i, leaf_set=i, first=first, count=n, history_offset=hist))
            hist += n
        first += n
        ofs.append(postproc.field_output_factor(x1, x2, y1, y2, sizes, factors))
    ctx.load_media(tables)
    ctx.load_phantom(phan, regions)
    ctx.load_mlc(head.build_mlc(head.read_dynvmlc(os.path.join(data, "VarianMLC.egsinp"))))
    # the leaf files hold the positions to 4 decimals
    ctx.set_leaf_positions(np.stack([head.read_leaf_file(os.path.join(data, "HEAD", "f%d.mlc" % i))
                                     for i in range(len(FIELDS))]))
    ctx.load_jaws(head.build_jaws(**head.CLINAC_LIKE_JAWS))
    ctx.set_jaw_positions(np.array([f[1:5] for f in FIELDS], dtype=F32))
    ctx.set_source(np.concatenate(srcs))
    out = ctx.run_batched(jobs, len(FIELDS), seed, n_batches=10)
    edep = postproc.plan_dose(list(out["dose"]), ofs)
    dose_unit = F32(F32(F32(1.6e-8) * F32(HPC)) / F32(HPM))
    return postproc.voxel_dose(edep, phan.voxel_mass(), dose_unit), out["dose"], hist
