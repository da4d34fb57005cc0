"""A reference run input end to end: ``python -m gegs_amd.run -i <file.ini>``.

Restates the reference's main() and runSimulation() (CUDA_EGS.cu:367-900,
902-1773) over this package's transport:

1. the run input (``runinput.read_run_input``) and the log file
   ``<output directory>/<prefix>_log.txt`` (logstr, output.c:41-49);
2. the phase-space headers (read_egsphsp_header, read_egsphsp.cpp:60-165);
3. the phantom: ctcreate from the DICOM list, written to
   ``<prefix>.egsphant`` and read back as the reference does
   (CUDA_EGS.cu:1327-1383, ctcreate.c:1249-1311), or the egsphant file;
4. the head: the SECJAWS and DYNVMLC components of the .egsinp files (read
   whether or not the modules are on), block and wedge when on; the media
   list is the phantom's plus the default, jaw, MLC, block and wedge media
   (init_phantom_and_MLC, init.cu:1164-1376);
5. the fields of the .mlcinfo list with their leaf files and output factors
   (read_mlcinp, init.cu:1377-1503);
6. per field MU x histories-per-MU histories: the phase-space records are
   recycled nRepeat + 1 times and loaded in batches (runSimulation,
   CUDA_EGS.cu:367-411; init_stack, init.cu:29-41);
7. the transport of every field (one batched event-queued launch sequence
   with one dose buffer per field, or the plain kernel field after field);
8. the plan's energy deposit, each field times its output factor when on
   (CUDA_EGS.cu:1712-1735), written as ``<prefix>_DoseResults_NNNxNNNxNNN.bin``
   in cGy x dose unit, then the interpolated dose, absolute dose and
   relative dose outputs (write_dose, output.c:873-918).

Kept from the reference, each with a switch (``fixes``) that corrects it:
Q4 and Q15 (phase-space file names and reads across files, see
``source.PhaseSpaceFiles``), Q16 (the last batch of a field holds
floor(rest / (nRepeat + 1)) + 1 records, so a field whose histories are a
multiple of nRepeat + 1 reads one record past the file and the reference's
read refuses the batch: its records are not transported; the fix transports
exactly MU x histories-per-MU histories), Q9 (absolute-dose voxel) and Q10
(CT values below the ramp).  Device-side corrections (Q1, Q7, Q8) are the
``gegs_set_fixes`` mask.

``--prepare <run file>`` stops after step 5 and writes what the C++ driver
gegs_amd/gegs_run needs to do steps 2, 6, 7 and the dose file without Python:
the run file, the media tables and the head geometry (``prepare``).

The reference's RNG is cuRAND; this transport uses Philox streams keyed by
a history index, so histories are independent of the reference's but each
(record, repeat) pair of the schedule is transported once, as there.
"""
from __future__ import annotations

import argparse
import dataclasses
import os
import sys
import time

import numpy as np

from . import ctimport, hatch, head, phantom as ph, postproc, runinput, source as srcmod
from .workloads import M_PHANTOM, make_field

# launch constants of the reference's build (CUDA_EGS.h:111-138)
SIMULATION_NUM_THREADS = 5 * 15 * 4 * 32
PHSP_STACK_SIZE = 50 * SIMULATION_NUM_THREADS
MAX_STACK_SIZE = 10_000_000
PHOTON_STACK_FACTOR = np.float32(1.5)       # relative_stack_size[p_photon_step], init.cu:31
MAXFIELDS = 20                              # CUDA_EGS.h:266
MAX_JOBS = 512                              # jobs per gegs_run_batched call

FIXES = ("Q1", "Q4", "Q7", "Q8", "Q9", "Q10", "Q15", "Q16")


class RunError(RuntimeError):
    pass


# ------------------------------------------------------------------ schedule of a field
@dataclasses.dataclass
class FieldSchedule:
    histories: int          # MU x histories per MU
    n_repeat: int           # nRepeat: extra uses of every record
    batch_size: int         # records per load (init_stack)
    n_batch: int            # full batches before the last one
    last_batch: int         # records of the last load
    loads: list             # (first record, records) of every load the reference's read accepts
    dropped: list           # loads the read refuses (Q16)
    repeat_counts: list     # records transported with repeat q, q = 0 .. n_repeat

    @property
    def records(self):
        return sum(n for _, n in self.loads)

    @property
    def transported(self):
        return int(sum(self.repeat_counts))


def batch_size_for(n_repeat):
    """init_stack's batch size (init.cu:29-41): PHSP_STACK_SIZE, reduced to
    MAX_STACK_SIZE / ((nRepeat + 1) x 1.5) when the photon stack would
    reach MAX_STACK_SIZE; nRepeat + 1 is an unsigned char."""
    nr = (n_repeat + 1) & 0xFF
    if nr == 0:
        raise RunError("nRepeat = 255: the reference's unsigned-char repeat count wraps to 0 and its particle "
                       "stacks are allocated empty (init.cu:33); use fewer histories per record")
    if np.float32(np.float32(PHSP_STACK_SIZE * nr) * PHOTON_STACK_FACTOR) >= np.float32(MAX_STACK_SIZE):
        return int(np.float32(MAX_STACK_SIZE) / np.float32(np.float32(nr) * PHOTON_STACK_FACTOR))
    return PHSP_STACK_SIZE


def field_schedule(histories, total_records, fix_q16=False):
    """The batches of one field (CUDA_EGS.cu:377-411, :462-476, :635-650)."""
    num = int(histories)
    if num <= 0 or total_records <= 0:
        raise RunError("The repeat number is not acceptable, try another number of hitories or phase space file")
    n_repeat = (num - 1) // total_records
    if n_repeat > 255:
        raise RunError("The repeat number is not acceptable, try another number of hitories or phase space file")
    r = n_repeat + 1
    bs = batch_size_for(n_repeat)
    n_batch = (num - 1) // (bs * r)
    last = (num - n_batch * (bs * r)) // r + 1
    loads, dropped = [], []
    for b in range(n_batch + 1):
        start = b * bs
        n = bs if b < n_batch else last
        if fix_q16 and b == n_batch:
            n = -(-(num - n_batch * bs * r) // r)          # ceil: the records the remaining histories need
        (loads if start + n <= total_records else dropped).append((start, n))
    recs = sum(n for _, n in loads)
    if fix_q16 and not dropped:
        # exactly num histories: every record r times but the last, which
        # takes what is left
        full = num // r
        counts = [full + (1 if q < num - full * r else 0) for q in range(r)]
    else:
        counts = [recs] * r
    return FieldSchedule(num, n_repeat, bs, n_batch, last, loads, dropped, counts)


# ------------------------------------------------------------------ paths and log
def host_path(base, name):
    """A file name of the run input on this host: backslashes as separators,
    relative names against `base` (the reference resolves them with
    GetFullPathName against its working directory)."""
    name = name.replace("\\", "/")
    return name if os.path.isabs(name) else os.path.normpath(os.path.join(base, name))


class RunLog:
    """logstr (output.c:41-49): every line to the log file and to stdout."""

    def __init__(self, path, echo=True):
        self.path = path
        self.f = open(path, "w")
        self.echo = echo
        self.lines = []

    def __call__(self, text=""):
        for line in text.split("\n"):
            self.lines.append(line)
            self.f.write(line + "\n")
            if self.echo:
                print(line, flush=True)
        self.f.flush()

    def close(self):
        self.f.close()


# ------------------------------------------------------------------ the run
@dataclasses.dataclass
class RunResult:
    root: str
    phantom: object
    fields: list
    schedules: list
    output_factors: list | None
    field_edep: np.ndarray          # [field, voxel] FP64 energy deposit (MeV)
    edep: np.ndarray                # the plan's deposit, output factors applied
    dose: np.ndarray                # float32 [voxel], as written
    dose_unit: float
    files: dict
    seconds: float
    histories: int


def _reference_data(data_dir, ri):
    files = hatch.reference_data_files(data_dir, ri.pegs_file, ri.photon_xsections, ri.atomic_ff_file)
    need = [files["pegs"], files["ff"], files["msnew"], files["photo_relax"], files["photo_cs"], files["incoh"],
            files["eii"]] + [files["xsections"] + s for s in ("_photo.data", "_rayleigh.data", "_pair.data",
                                                              "_triplet.data")]
    return files if all(os.path.exists(p) for p in need) else None


def _fix_mask(fixes):
    from .native import FIX_Q1, FIX_Q7, FIX_Q8
    return (FIX_Q1 if "Q1" in fixes else 0) | (FIX_Q7 if "Q7" in fixes else 0) | (FIX_Q8 if "Q8" in fixes else 0)


def run(path, base_dir=None, method="batched", seed=1, fixes=(), n_batches=10, device=0, tuning=None,
        echo=True, context=None):
    """Run the input at `path`; file names in it resolve against base_dir
    (default: the working directory, as the reference's GetFullPathName)."""
    fixes = set(fixes)
    unknown = fixes - set(FIXES)
    if unknown:
        raise RunError("unknown fixes %s (known: %s)" % (sorted(unknown), ", ".join(FIXES)))
    t0 = time.perf_counter()
    base = base_dir or os.getcwd()
    ri = runinput.read_run_input(path)
    data_dir = host_path(base, ri.data_directory)
    out_dir = host_path(base, ri.output_directory)
    os.makedirs(out_dir, exist_ok=True)
    root = os.path.join(out_dir, ri.output_prefix)
    log = RunLog(root + "_log.txt", echo)
    try:
        return _run(ri, path, base, data_dir, root, log, method, seed, fixes, n_batches, device, tuning, context, t0)
    finally:
        log.close()


def _setup(ri, base, data_dir, root, log, fixes):
    """Steps 2-5: phase-space headers, phantom, head components, media tables
    and region records, fields with leaf files and output factors."""
    d = lambda name: os.path.join(data_dir, name.replace("\\", "/"))
    # phase space (read_egsphsp_header)
    phsp = srcmod.PhaseSpaceFiles(d(ri.phase_space_file), ri.phsp_file_number, fix_q4="Q4" in fixes,
                                  fix_q15="Q15" in fixes)
    for i, p in enumerate(phsp.header_paths):
        log("  Phase space file #%d . . . . %s" % (i, p) if phsp.n_files > 1 else "  Phase space file  . . . . . %s" % p)
    log("  total particles . . . . . . %d" % phsp.total)
    log("  z position of phsp  . . . . %.2f" % ri.phsp_z)

    # phantom
    if ri.use_ctcreate:
        ramp, lower = ctimport.read_ct_ramps(d(ri.ct_ramps_file))
        b = ri.xyz_boundaries
        ct = ctimport.ctcreate(host_path(base, ri.dicom_file_list), ri.ct_voxel_sizes, (b[0], b[2], b[4]),
                               (b[1], b[3], b[5]), ramp=ramp, lower=lower, fix_q10="Q10" in fixes)
        egsphant = root + ".egsphant"
        ctimport.write_egsphant_ctcreate(egsphant, ct)
    else:
        egsphant = d(ri.egsphant_file)
    phan = ph.read_egsphant(egsphant)
    log("  egsphant file . . . . . . . %s" % egsphant)
    log("  phantom voxels  . . . . . . %d x %d x %d" % (phan.nx, phan.ny, phan.nz))

    # head components (read whether or not the modules are on)
    jaws_rec = head.read_egsinp_secjaws(d(ri.secjaws_file)) if os.path.exists(d(ri.secjaws_file)) else None
    if ri.secjaws and jaws_rec is None:
        raise RunError('no JAWS component SECJAWS in "%s"' % d(ri.secjaws_file))
    jaws, jaw_med = jaws_rec if jaws_rec is not None else (None, ri.default_medium)
    dyn = None
    if os.path.exists(d(ri.egsinp_file)):
        try:
            dyn = head.read_dynvmlc(d(ri.egsinp_file))
        except (ValueError, StopIteration):
            dyn = None
    if ri.mlc and dyn is None:
        raise RunError('no DYNVMLC component in "%s"' % d(ri.egsinp_file))
    mlc_med = dyn.medium if dyn is not None else ri.default_medium
    block = wedge = None
    block_med = wedge_med = ri.default_medium
    if ri.block:
        r = head.read_egsinp_block(d(runinput.get(ri.sections, "Block", "block file")))
        if r is None:
            raise RunError("no PYRAMIDS component BLOCK in the block file")
        block, block_med = r
    if ri.wedge:
        r = head.read_egsinp_wedge(d(runinput.get(ri.sections, "Wedge", "wedge file")))
        if r is None:
            raise RunError("no PYRAMIDS component WEDGE in the wedge file")
        wedge, wedge_med = (r[0], r[1]), r[2]
    others = [ri.default_medium, jaw_med, mlc_med, block_med, wedge_med]
    names = list(phan.media_names)
    for n in others:
        if n not in names:
            names.append(n)
    module_first = (head.M_SECJAWY if ri.secjaws else head.M_MLC if ri.mlc else head.M_BLOCK if ri.block
                    else head.M_WEDGE if ri.wedge else M_PHANTOM)

    # media tables: the run's own EGSnrc data set when present, else pegs-lite
    files = _reference_data(data_dir, ri)
    if files is None:
        log("WARNING: the EGSnrc data files the run input names are not all in %s; the pegs-lite data of this "
            "package are used" % data_dir)
    tables = hatch.MediaTables(names, mlc_medium=mlc_med, files=files)
    regions = ph.region_table(phan, tables, others)
    log("  media . . . . . . . . . . . %s" % ", ".join(names))

    # fields (read_mlcinp)
    mlcinfo = d(ri.mlcinfo_file)
    if not os.path.exists(mlcinfo):
        raise RunError("Error: MLC info  file reading error! No simulation will be done!")
    fields = head.read_mlcinfo(mlcinfo)
    if len(fields) > MAXFIELDS:
        log("Warning! number of MLC fields is larger than the MAXFIELDS=%d" % MAXFIELDS)
        fields = fields[:MAXFIELDS]
    for f in fields:
        f.positions = head.read_leaf_file(d(f.leaf_file))
    ofs = None
    if ri.output_factor:
        sizes, factors = postproc.read_output_factor_table(d(ri.output_factor_file))
        ofs = []
        for i, f in enumerate(fields):
            o = postproc.field_output_factor(f.x1, f.x2, f.y1, f.y2, sizes, factors)
            if o is None:
                raise RunError("field #%d: its equivalent square lies outside the output-factor table (the "
                               "reference leaves its factor unset)" % i)
            ofs.append(o)
    total_mu = sum(int(f.mu) for f in fields)
    log("  Total number of MU  . . . . %d" % total_mu)
    log("  Histories per MU  . . . . . %d" % ri.histories)

    return dict(phsp=phsp, egsphant=egsphant, phan=phan, jaws=jaws, dyn=dyn, block=block, wedge=wedge,
                others=others, names=names, module_first=module_first, tables=tables, regions=regions,
                fields=fields, ofs=ofs)


def _run(ri, path, base, data_dir, root, log, method, seed, fixes, n_batches, device, tuning, context, t0):
    d = lambda name: os.path.join(data_dir, name.replace("\\", "/"))
    log("\nGEGS (MI355X) started at %s" % time.strftime("%Y-%m-%d %H:%M:%S"))
    log("\nSimulation")
    log("  Input file  . . . . . . . . %s" % os.path.abspath(path))
    for w in ri.warnings:
        log("WARNING: " + w)
    log("  Output prefix . . . . . . . %s" % ri.output_prefix)
    log("  Histories per MU  . . . . . %.3E" % ri.histories_input)
    log("  Data directory  . . . . . . %s" % data_dir)

    S = _setup(ri, base, data_dir, root, log, fixes)
    phsp, phan, dyn, jaws, block, wedge = S["phsp"], S["phan"], S["dyn"], S["jaws"], S["block"], S["wedge"]
    module_first, tables, regions, fields, ofs = S["module_first"], S["tables"], S["regions"], S["fields"], S["ofs"]

    # schedules and source records
    scheds, seqs, seq_of = [], {}, []
    for f in fields:
        s = field_schedule(int(f.mu) * ri.histories, phsp.total, fix_q16="Q16" in fixes)
        key = tuple(s.loads)
        if key not in seqs:
            recs = [phsp.read(a, n) for a, n in s.loads]
            seqs[key] = srcmod.decode_phsp(np.concatenate(recs), ri.phsp_z) if recs else \
                np.zeros(0, dtype=srcmod.SOURCE_DTYPE)
        scheds.append(s)
        seq_of.append(key)
    keys = list(seqs)
    offsets = np.cumsum([0] + [len(seqs[k]) for k in keys])
    source = np.concatenate([seqs[k] for k in keys]) if keys else np.zeros(0, dtype=srcmod.SOURCE_DTYPE)
    first_of = [int(offsets[keys.index(k)]) for k in seq_of]

    # transport
    from . import native
    ctx = context or native.Context(device)
    try:
        _load(ctx, tables, phan, regions, dyn, jaws, block, wedge, fields)
        ctx.set_fixes(_fix_mask(fixes))
        if tuning:
            ctx.set_tuning(tuning)
        ctx.set_source(source)
        gfields = [make_field(ri.change_coordinate, ri.source_to_isocenter, float(np.float32(f.angle)),
                              tuple(ri.isocenter), module_first) for f in fields]
        field_edep = np.zeros((len(fields), phan.nvox), dtype=np.float64)
        hist = 0
        jobs = []
        for i, (s, f) in enumerate(zip(scheds, fields)):
            log("Field #%d: MU %d, gantry %.1f, histories %d, repeat times %d, batches %d of %d, last %d%s"
                % (i, int(f.mu), f.angle, s.histories, s.n_repeat + 1, s.n_batch + 1, s.batch_size,
                   s.last_batch, " (%d records not read, quirk Q16)" % sum(n for _, n in s.dropped)
                   if s.dropped else ""))
            if ofs is not None:
                log("  Output factor . . . . . . . %f" % ofs[i])
            for q, cnt in enumerate(s.repeat_counts):
                if cnt:
                    jobs.append(dict(field=gfields[i], plan=i, leaf_set=i, first=first_of[i], count=cnt,
                                     history_offset=hist, fidx=i))
                    hist += cnt
        if method == "batched":
            for c0 in range(0, len(jobs), MAX_JOBS):
                chunk = jobs[c0:c0 + MAX_JOBS]
                fl = sorted({j["fidx"] for j in chunk})
                remap = {fi: k for k, fi in enumerate(fl)}
                cj = [dict(j, plan=remap[j["fidx"]]) for j in chunk]
                out = ctx.run_batched(cj, len(fl), seed, n_batches=n_batches)
                if out["counters"]["overflow"]:
                    raise RunError("particle banks overflowed")
                for fi in fl:
                    field_edep[fi] += out["dose"][remap[fi]]
        elif method == "plain":
            for j in jobs:
                i = j["fidx"]
                ctx.set_leaf_positions(fields[i].positions)
                ctx.set_jaw_positions([fields[i].x1, fields[i].x2, fields[i].y1, fields[i].y2])
                ctx.set_field(j["field"])
                # the plain kernel keys history k of the range by history_offset
                # + record index, the batched loop by the job's history_offset
                # + k: the same histories need the offset less the first record
                # (modulo 2^64, as the library adds it)
                off = (j["history_offset"] - j["first"]) % (1 << 64)
                out = ctx.run(seed, n_batches=n_batches, history_offset=off, first=j["first"], count=j["count"])
                field_edep[i] += out["dose"]
        else:
            raise RunError("method must be batched or plain")
    finally:
        if context is None:
            ctx.close()

    # write_dose
    edep = postproc.plan_dose(list(field_edep), ofs)
    mass = phan.voxel_mass()
    dose = postproc.voxel_dose(edep, mass, ri.dose_unit)
    files_out = {"log": log.path}
    fname = postproc.dose_file_name(root, phan.nx, phan.ny, phan.nz)
    dose.tofile(fname)
    log(fname)
    files_out["dose"] = fname
    xb, yb, zb = (postproc.host_bounds(b) for b in (phan.xb, phan.yb, phan.zb))
    grid = dose.reshape(phan.nz, phan.ny, phan.nx)
    if ri.interpolate:
        g = postproc.interp_dose(grid, xb, yb, zb, ri.interp_start, ri.interp_sizes, ri.interp_dimen)
        files_out["interpolate"] = postproc.write_interp_dose(root, g)
        log(files_out["interpolate"])
    if ri.absolute_dose:
        for line in postproc.abs_dose_report(grid, xb, yb, zb, ri.dose_position, ri.dose_calibration,
                                             ri.dose_measured, fix_q9="Q9" in fixes):
            log(line)
    if ri.relative_dose:
        if list(ri.reldose_dimen) == [1, 1, 1]:
            for line in postproc.abs_dose_report(grid, xb, yb, zb, ri.reldose_start, ri.dose_calibration,
                                                 ri.dose_measured or 0.0, fix_q9="Q9" in fixes):
                log(line)
        else:
            calc = postproc.reldose_points_nearest(grid, xb, yb, zb, ri.reldose_start, ri.reldose_sizes,
                                                   ri.reldose_dimen)
            fn, lines = postproc.write_relative_dose(
                root, calc, ri.reldose_dimen, d(ri.measured_data), start=ri.reldose_start, sizes=ri.reldose_sizes,
                isocenter=ri.isocenter, norm_type=ri.norm_type, array_type=ri.array_type,
                gamma_len=ri.gamma_len, gamma_dose=ri.gamma_dose)
            files_out["relative"] = fn
            for line in lines:
                log(line)
    secs = time.perf_counter() - t0
    log("\nTotal CPU/GPU . . . . . . . . %.2f ms (%.2f min)" % (secs * 1e3, secs / 60.0))
    return RunResult(root, phan, fields, scheds, ofs, field_edep, edep, dose, ri.dose_unit, files_out, secs,
                     sum(s.transported for s in scheds))


def prepare(path, run_file, base_dir=None, method="batched", seed=1, fixes=(), n_batches=10, echo=True):
    """Steps 1-5 of the run input at `path`, written for the C++ driver
    (gegs_amd/gegs_run, gegs_amd/csrc/gegs_run.cpp), which does the rest
    without Python: the run file `run_file` (key = value lines), the media
    tables (<run file stem>.gegsmed) and the head geometry with one aperture
    set per field (<run file stem>.gegshead).  The driver reads the egsphant
    and phase-space files itself.  Returns the run file's path."""
    from . import hostio
    fixes = set(fixes)
    unknown = fixes - set(FIXES)
    if unknown:
        raise RunError("unknown fixes %s (known: %s)" % (sorted(unknown), ", ".join(FIXES)))
    base = base_dir or os.getcwd()
    ri = runinput.read_run_input(path)
    data_dir = host_path(base, ri.data_directory)
    out_dir = host_path(base, ri.output_directory)
    os.makedirs(out_dir, exist_ok=True)
    root = os.path.join(out_dir, ri.output_prefix)
    stem = os.path.splitext(os.path.abspath(run_file))[0]
    log = RunLog(stem + "_prepare_log.txt", echo)
    try:
        log("\nGEGS (MI355X) run preparation at %s" % time.strftime("%Y-%m-%d %H:%M:%S"))
        log("  Input file  . . . . . . . . %s" % os.path.abspath(path))
        S = _setup(ri, base, data_dir, root, log, fixes)
    finally:
        log.close()
    fields = S["fields"]
    media_file = stem + ".gegsmed"
    hostio.write_media_file(media_file, S["tables"])
    head_file = None
    if S["dyn"] is not None or S["jaws"] is not None or S["block"] is not None or S["wedge"] is not None:
        head_file = stem + ".gegshead"
        hostio.write_head_file(
            head_file,
            mlc=head.build_mlc(S["dyn"]) if S["dyn"] is not None else None,
            jaws=S["jaws"],
            block_wedge=head.build_block_wedge(S["block"], S["wedge"])
            if S["block"] is not None or S["wedge"] is not None else None,
            leaf_positions=np.stack([f.positions for f in fields]) if S["dyn"] is not None else None,
            jaw_positions=np.array([[f.x1, f.x2, f.y1, f.y2] for f in fields], dtype=np.float32)
            if S["jaws"] is not None else None)
    g = lambda v: "%.9g" % float(np.float32(v))
    host_fixes = ((1 if "Q4" in fixes else 0) | (2 if "Q15" in fixes else 0) | (4 if "Q16" in fixes else 0)
                  | (8 if "Q9" in fixes else 0))
    lines = ["# gegs run file: %s prepared for gegs_amd/gegs_run" % os.path.abspath(path),
             "media file = %s" % media_file]
    if head_file:
        lines.append("head file = %s" % head_file)
    lines += ["egsphant file = %s" % os.path.abspath(S["egsphant"])]
    lines += ["region medium %d = %s" % (i + 1, n) for i, n in enumerate(S["others"])]
    lines += ["region flags = 0x%04x" % ph.DEFAULT_FLAGS,
              "phsp base = %s" % os.path.abspath(os.path.join(data_dir, ri.phase_space_file.replace("\\", "/"))),
              "phsp files = %d" % ri.phsp_file_number,
              "phsp z = %s" % g(ri.phsp_z),
              "host fixes = %d" % host_fixes,
              "device fixes = %d" % _fix_mask(fixes),
              "change xyz = %d" % int(ri.change_coordinate),
              "module first = %d" % S["module_first"],
              "source to isocenter = %s" % g(ri.source_to_isocenter),
              "isocenter = %s" % " ".join(g(v) for v in ri.isocenter),
              "histories per mu = %d" % ri.histories,
              "dose unit = %s" % g(ri.dose_unit),
              "output root = %s" % os.path.abspath(root),
              "seed = %d" % seed,
              "batches = %d" % n_batches,
              "method = %s" % method]
    if ri.interpolate:
        lines.append("interpolation = %s %s" % (" ".join(g(v) for v in list(ri.interp_start) + list(ri.interp_sizes)),
                                                " ".join("%d" % v for v in ri.interp_dimen)))
    if ri.absolute_dose:
        lines.append("absolute dose = %s %d %s" % (" ".join(g(v) for v in ri.dose_position),
                                                    1 if ri.dose_calibration else 0, g(ri.dose_measured)))
    if ri.relative_dose:
        lines.append("relative dose = %s %s %s %d %d %s %s %d %s %s" % (
            " ".join(g(v) for v in list(ri.reldose_start) + list(ri.reldose_sizes)),
            " ".join("%d" % v for v in ri.reldose_dimen), " ".join(g(v) for v in ri.isocenter),
            ri.norm_type, ri.array_type, g(ri.gamma_len), g(ri.gamma_dose), 1 if ri.dose_calibration else 0,
            g(ri.dose_measured or 0.0), os.path.abspath(os.path.join(data_dir, ri.measured_data.replace("\\", "/")))))
    for i, f in enumerate(fields):
        fl = make_field(ri.change_coordinate, ri.source_to_isocenter, float(np.float32(f.angle)),
                        tuple(ri.isocenter), S["module_first"])
        of = g(S["ofs"][i]) if S["ofs"] is not None else "-"
        lines.append("field = %d %s %s %s %s" % (int(f.mu), g(f.angle), g(fl.cos_angle), g(fl.sin_angle), of))
    with open(run_file, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return run_file


def _load(ctx, tables, phan, regions, dyn, jaws, block, wedge, fields):
    """Media, phantom and head geometry of the run on the context; leaf set
    and jaw set i belong to field i."""
    ctx.load_media(tables)
    ctx.load_phantom(phan, regions)
    if dyn is not None:
        ctx.load_mlc(head.build_mlc(dyn))
        ctx.set_leaf_positions(np.stack([f.positions for f in fields]))
    if jaws is not None:
        ctx.load_jaws(jaws)
        ctx.set_jaw_positions(np.array([[f.x1, f.x2, f.y1, f.y2] for f in fields], dtype=np.float32))
    if block is not None or wedge is not None:
        ctx.load_block_wedge(head.build_block_wedge(block, wedge))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m gegs_amd.run", description=__doc__.split("\n")[0])
    ap.add_argument("-i", dest="input", required=True, help="run input (.ini)")
    ap.add_argument("--method", default="batched", choices=("batched", "plain"))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--batches", type=int, default=10)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--base-dir", default=None, help="folder the input's relative paths start from (default: cwd)")
    ap.add_argument("--fix", default="", help="comma list of reference quirks to correct: " + ", ".join(FIXES))
    ap.add_argument("--prepare", default=None, metavar="RUN_FILE",
                    help="write the run file, media tables and head geometry for the C++ driver gegs_amd/gegs_run "
                         "instead of running")
    a = ap.parse_args(argv)
    fixes = [f.strip().upper() for f in a.fix.split(",") if f.strip()]
    if a.prepare:
        try:
            rf = prepare(a.input, a.prepare, base_dir=a.base_dir, method=a.method, seed=a.seed, fixes=fixes,
                         n_batches=a.batches)
        except (runinput.RunInputError, RunError, ValueError) as e:
            print("ERROR: %s" % e, file=sys.stderr)
            return getattr(e, "code", 1) if isinstance(getattr(e, "code", None), int) else 1
        print("run file %s: gegs_amd/gegs_run %s" % (rf, rf))
        return 0
    try:
        r = run(a.input, base_dir=a.base_dir, method=a.method, seed=a.seed, fixes=fixes, n_batches=a.batches,
                device=a.device)
    except (runinput.RunInputError, RunError, ValueError) as e:
        print("ERROR: %s" % e, file=sys.stderr)
        return getattr(e, "code", 1) if isinstance(getattr(e, "code", None), int) else 1
    print("%d histories in %.2f s -> %s" % (r.histories, r.seconds, r.files["dose"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
