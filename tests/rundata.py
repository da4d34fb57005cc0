"""A synthetic run-input data set shaped like the reference's
TestHeAbDo_5Fields inputs, shared by the CPU and GPU tests of the run-input
driver (gegs_amd/run.py)."""
import os

import numpy as np

from gegs_amd import hatch, head, phantom as ph, postproc, source, workloads

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
