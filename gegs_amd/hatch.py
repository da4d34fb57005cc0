"""HATCH: build the device physics tables from PEGS4-format inputs.

A restatement of the reference's host-side table builders:

* ``read_pegs_file``      media.cu:332-830  (PEGS4 reader, dunit scaling, zbrang)
* ``init_user_photon``    media.cu:832-1058 (ge/gmfp/gbr1/gbr2/cohe from element data)
* ``kn_sigma0``           media.cu:316-330
* ``get_data``            media.cu:260-311
* ``init_rayleigh_data``  media.cu:1060-1292
* ``egs_set_defaults``    PhysicsInit.cu:187-315
* ``init_ms_sr``          PhysicsInit.cu:860-906
* ``mscati``              PhysicsInit.cu:908-1182 (skin depth, sig_e, range_ep, tmxs)
* ``fix_brems``           PhysicsInit.cu:1552-1641

The builders run in float64 and round to float32 when stored, where the
reference mixes float and double host arithmetic; the result is the same
table to within float32 rounding of the inputs.  The output is a
:class:`MediaTables` holding numpy arrays named and laid out exactly as the
``gegs_media_tables`` block of ``include/gegs_media.h``.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

MXMED = 6
MXGE = 200
MXEKE = 150
MXEL = 10
MXRAYFF = 100
RAYCDFSIZE = 100
MAXL_MS, MAXQ_MS, MAXU_MS = 63, 7, 31
MS_SIZE = (MAXU_MS + 1) * (MAXQ_MS + 1) * (MAXL_MS + 1)
LAMBMIN_MS, LAMBMAX_MS, QMAX_MS = 1.0, 1e5, 0.5
RM_F = float(np.float32(0.5110034))       # ELECTRON_REST_MASS_FLOAT promoted to double
E_MAX_RR = 0.8
MAX_SMAX = 1e10
MXELEMENT, MXSHELL, MXTRANS, MXEDGE = 100, 6, 39, 16
MXTOTSH, MXMDSH = 1538, 50
MAX_EII_SHELLS, N_EII_BINS = 40, 250

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


class F2(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float)]


class Ray(ctypes.Structure):
    _fields_ = [("xgrid", ctypes.c_float), ("fcum", ctypes.c_float),
                ("b_array", ctypes.c_float), ("c_array", ctypes.c_float)]


def _arr(t, n):
    return t * n


class CMediaTables(ctypes.Structure):
    """ctypes mirror of ``gegs_media_tables`` (include/gegs_media.h)."""
    _fields_ = [
        ("nmed", ctypes.c_int32), ("eii_flag", ctypes.c_int32), ("nbr_split", ctypes.c_int32),
        ("ibrdst", ctypes.c_int32), ("iprdst", ctypes.c_int32), ("exact_bca", ctypes.c_int32),
        ("spin_effects", ctypes.c_int32), ("transport_algorithm", ctypes.c_int32),
        ("skindepth_for_bca", ctypes.c_float), ("dllambi", ctypes.c_float),
        ("dqmsi", ctypes.c_float), ("rm", ctypes.c_float),
        ("rho", _arr(ctypes.c_float, MXMED)), ("te", _arr(ctypes.c_float, MXMED)),
        ("thmoll", _arr(ctypes.c_float, MXMED)), ("ae", _arr(ctypes.c_float, MXMED)),
        ("ap", _arr(ctypes.c_float, MXMED)), ("log_ap", _arr(ctypes.c_float, MXMED)),
        ("smaxir", _arr(ctypes.c_float, MXMED)), ("e_max_rr", _arr(ctypes.c_float, MXMED)),
        ("zbrang", _arr(ctypes.c_float, MXMED)), ("lzbrang", _arr(ctypes.c_float, MXMED)),
        ("delcm", _arr(ctypes.c_float, MXMED)), ("blcc", _arr(ctypes.c_float, MXMED)),
        ("xcc", _arr(ctypes.c_float, MXMED)),
        ("i_do_rr", _arr(ctypes.c_int32, MXMED)), ("meke", _arr(ctypes.c_int32, MXMED)),
        ("nne", _arr(ctypes.c_int32, MXMED)),
        ("zelem", _arr(ctypes.c_float, MXEL * MXMED)), ("pz", _arr(ctypes.c_float, MXEL * MXMED)),
        ("dl1", _arr(ctypes.c_float, 8 * MXMED)), ("dl2", _arr(ctypes.c_float, 8 * MXMED)),
        ("dl3", _arr(ctypes.c_float, 8 * MXMED)), ("dl4", _arr(ctypes.c_float, 8 * MXMED)),
        ("dl5", _arr(ctypes.c_float, 8 * MXMED)), ("dl6", _arr(ctypes.c_float, 8 * MXMED)),
        ("bpar", _arr(ctypes.c_float, 2 * MXMED)), ("sig_e", _arr(ctypes.c_float, 2 * MXMED)),
        ("sig_ismonotone", _arr(ctypes.c_int32, 2 * MXMED)),
        ("eke01", _arr(F2, MXMED)),
        ("sig", _arr(F2, 2 * MXMED * MXEKE)), ("dedx", _arr(F2, 2 * MXMED * MXEKE)),
        ("ebr1", _arr(F2, MXMED * MXEKE)), ("pbr1", _arr(F2, MXMED * MXEKE)),
        ("pbr2", _arr(F2, MXMED * MXEKE)), ("tmxs", _arr(F2, MXMED * MXEKE)),
        ("range_ep", _arr(ctypes.c_float, 2 * MXMED * MXEKE)),
        ("e_array", _arr(ctypes.c_float, MXMED * MXEKE)),
        ("ge", _arr(F2, MXMED)), ("gmfp", _arr(F2, MXMED * MXGE)), ("gbr1", _arr(F2, MXMED * MXGE)),
        ("gbr2", _arr(F2, MXMED * MXGE)), ("cohe", _arr(F2, MXMED * MXGE)),
        ("pmax", _arr(F2, MXMED * MXGE)),
        ("rayleigh_data", _arr(Ray, MXMED * MXRAYFF)),
        ("i_array", _arr(ctypes.c_uint32, MXMED * RAYCDFSIZE)),
        ("ums", _arr(ctypes.c_float, MS_SIZE)), ("fms", _arr(ctypes.c_float, MS_SIZE)),
        ("wms", _arr(ctypes.c_float, MS_SIZE)), ("ims", _arr(ctypes.c_int16, MS_SIZE)),
        ("binding_energies", _arr(ctypes.c_float, MXSHELL * MXELEMENT)),
        ("interaction_prob", _arr(ctypes.c_float, MXSHELL * MXELEMENT)),
        ("relaxation_prob", _arr(ctypes.c_float, MXTRANS * MXELEMENT)),
        ("edge_energies", _arr(ctypes.c_float, MXEDGE * MXELEMENT)),
        ("edge_a", _arr(ctypes.c_float, MXEDGE * MXELEMENT)), ("edge_b", _arr(ctypes.c_float, MXEDGE * MXELEMENT)),
        ("edge_c", _arr(ctypes.c_float, MXEDGE * MXELEMENT)), ("edge_d", _arr(ctypes.c_float, MXEDGE * MXELEMENT)),
        ("edge_number", _arr(ctypes.c_int32, MXELEMENT)),
        ("n_shell", _arr(ctypes.c_int32, MXMED)),
        ("shell_array", _arr(ctypes.c_int32, MXMDSH * MXMED)),
        ("eno_array", _arr(ctypes.c_float, MXMDSH * MXMED)),
        ("eno_atbin_array", _arr(ctypes.c_int32, MXMDSH * MXMED)),
        ("iz_array", _arr(ctypes.c_int32, MXTOTSH)), ("shn_array", _arr(ctypes.c_int32, MXTOTSH)),
        ("be_array", _arr(ctypes.c_float, MXTOTSH)), ("jo_array", _arr(ctypes.c_float, MXTOTSH)),
        ("erfjo_array", _arr(ctypes.c_float, MXTOTSH)),
        ("eii_nsh", _arr(ctypes.c_int32, MXMED)),
        ("eii_no", _arr(ctypes.c_int32, MXEL * MXMED)), ("eii_first", _arr(ctypes.c_int32, MXEL * MXMED)),
        ("eii_cons", _arr(ctypes.c_float, MXMED)),
        ("eii", _arr(F2, MAX_EII_SHELLS)),
        ("eii_xsection", _arr(F2, N_EII_BINS * MAX_EII_SHELLS)),
    ]


# ------------------------------------------------------------------ PEGS4 reader
class PegsMedium:
    pass


def _kv_line(line):
    out = {}
    for entry in line.strip().split(","):
        if "=" not in entry:
            continue
        k, v = entry.split("=", 1)
        out[k.replace(" ", "")] = v.strip()
    return out


def read_pegs_file(path, names):
    """PEGS4 reader following media.cu:332-608: media are matched by name,
    in the order requested; a missing medium raises (reference code 10029)."""
    with open(path) as f:
        lines = f.read().split("\n")
    found = {}
    i = 0
    while i < len(lines) and len(found) < len(names):
        line = lines[i]
        i += 1
        if not line.startswith(" MEDIUM="):
            continue
        name = line[8:32].split(" ")[0].split(",")[0]
        if name not in names:
            continue
        m = PegsMedium()
        m.name = name
        kv = _kv_line(lines[i]); i += 1
        m.rho = float(kv["RHO"])
        m.ne = int(kv["NE"])
        m.z, m.wa, m.pz, m.rhoz = [], [], [], []
        for _ in range(m.ne):
            kv = _kv_line(lines[i]); i += 1
            m.z.append(float(kv["Z"])); m.wa.append(float(kv["A"]))
            m.pz.append(float(kv["PZ"])); m.rhoz.append(float(kv["RHOZ"]))
        # the rest is whitespace-separated numbers in a fixed order
        tok = []
        j = i

        def take(n):
            nonlocal j
            while len(tok) < n:
                tok.extend(lines[j].split())
                j += 1
            vals = tok[:n]
            del tok[:n]
            return vals
        m.rlc, m.ae, m.ap, m.ue, m.up = map(float, take(5))
        m.te = m.ae - RM_F
        m.thmoll = m.te * 2 + RM_F
        ints = list(map(int, take(9)))
        m.msge, m.mge, m.mseke, m.meke, m.mleke, m.mcmfp, m.mrange, m.irayl = ints[:8]
        take(36)                                   # dl1..dl6 (fix_brems recomputes)
        m.delcm = float(take(1)[0])
        take(6)                                    # alphi, bpar, delpos
        m.xr0, m.teff0, m.blcc, m.xcc = map(float, take(4))
        m.eke0, m.eke1 = map(float, take(2))
        rows = np.array(list(map(float, take(16 * m.meke)))).reshape(m.meke, 16)
        m.esig, m.psig = rows[:, 0:2].copy(), rows[:, 2:4].copy()
        m.ededx, m.pdedx = rows[:, 4:6].copy(), rows[:, 6:8].copy()
        m.ebr1, m.pbr1, m.pbr2 = rows[:, 8:10].copy(), rows[:, 10:12].copy(), rows[:, 12:14].copy()
        m.tmxs = rows[:, 14:16].copy()
        take(3)                                    # ebinda, ge0, ge1
        take(6 * m.mge)
        if m.irayl == 1:
            ngr = int(take(1)[0])
            take(2 + 2 * ngr + 2 * m.mge)
        i = j
        found[name] = m
    missing = [n for n in names if n not in found]
    if missing:
        raise RuntimeError("ERROR (10029): media not found in %s: %s" % (path, " ".join(missing)))
    media = [found[n] for n in names]
    # distance-unit scaling, dunit = 1 cm (media.cu:697-770)
    for m in media:
        dfacti = 1.0 / m.rlc
        for arr in (m.esig, m.psig, m.ededx, m.pdedx, m.tmxs):
            arr *= dfacti
        m.blcc *= dfacti
        m.xcc *= math.sqrt(dfacti)
        m.smaxir = MAX_SMAX
        zz = sum(p * z * (z + 1.0) for p, z in zip(m.pz, m.z))
        pznorm = sum(m.pz)
        m.zbrang = 8.116224e-05 * (zz / pznorm) ** (1.0 / 3.0)
        m.lzbrang = -math.log(m.zbrang)
    return media


# ------------------------------------------------------------------ photons
def read_xsec_data(path):
    """Element cross sections (media.cu:114-142): per Z, n pairs (lnE, ln sigma)."""
    with open(path) as f:
        tok = f.read().split()
    out = []
    p = 0
    for _ in range(100):
        n = int(tok[p]); p += 1
        d = np.array(tok[p:p + 2 * n], dtype=np.float64).reshape(n, 2)
        p += 2 * n
        out.append(d)
    return out


def kn_sigma0(e):
    """Klein-Nishina total cross section per electron in barns (media.cu:316-330)."""
    con = float(np.float32(0.1274783851))
    ko = e / RM_F
    if ko < 0.01:
        return 8.0 * con / 3.0 * (1.0 - ko * (2.0 - ko * (5.2 - 13.3 * ko))) / RM_F
    c1 = 1.0 / (ko * ko)
    c2 = 1.0 - 2.0 * (1.0 + ko) * c1
    c3 = (1.0 + 2.0 * ko) * c1
    eps2 = 1.0
    eps1 = 1.0 / (1.0 + 2.0 * ko)
    return (c1 * (1.0 / eps1 - 1.0 / eps2) + c2 * math.log(eps2 / eps1)
            + eps2 * (c3 + 0.5 * eps2) - eps1 * (c3 + 0.5 * eps1)) / e * con


def get_data(flag, zs, pzs, data, ge):
    """Mixture cross section on the MXGE grid (media.cu:260-311)."""
    res = np.zeros(MXGE)
    for z, pz in zip(zs, pzs):
        d = data[int(z + 0.5) - 1].copy()
        if flag != 0:
            eth = 2.0 * RM_F if flag == 1 else 4.0 * RM_F
            d[:, 1] -= 3.0 * np.log(1.0 - eth / np.exp(d[:, 0]))
            d = np.vstack([[math.log(eth), d[0, 1]], d])
        n = len(d)
        for j in range(MXGE):
            gle = (j - ge[0]) / ge[1]
            e = math.exp(gle)
            if gle < d[0, 0] or gle >= d[n - 1, 0]:
                if flag == 0:
                    raise RuntimeError("ERROR (10003): energy %g outside data range" % e)
                sig = 0.0 if gle < d[0, 0] else math.exp(d[n - 1, 1])
            else:
                k = int(np.searchsorted(d[:, 0], gle, side="right") - 1)
                p = (gle - d[k, 0]) / (d[k + 1, 0] - d[k, 0])
                sig = math.exp(p * d[k + 1, 1] + (1.0 - p) * d[k, 1])
            if flag != 0 and e > eth:
                sig *= (1.0 - eth / e) ** 3
            res[j] += pz * sig
    return res


def init_user_photon(media, prefix):
    """ge, gmfp, gbr1, gbr2, cohe (media.cu:832-1058)."""
    photo = read_xsec_data(prefix + "_photo.data")
    rayl = read_xsec_data(prefix + "_rayleigh.data")
    pair = read_xsec_data(prefix + "_pair.data")
    trip = read_xsec_data(prefix + "_triplet.data")
    nmed = len(media)
    ge = np.zeros((nmed, 2))
    tabs = {k: np.zeros((nmed, MXGE, 2)) for k in ("gmfp", "gbr1", "gbr2", "cohe")}
    for i, m in enumerate(media):
        ge[i, 1] = (MXGE - 1) / math.log(m.up / m.ap)
        ge[i, 0] = -ge[i, 1] * math.log(m.ap)
        order = np.argsort(np.array(m.z), kind="stable")
        zs = [m.z[k] for k in order]
        pzs = [m.pz[k] for k in order]
        sum_a = sum(p * a for p, a in zip(m.pz, m.wa))
        sum_z = sum(p * z for p, z in zip(m.pz, m.z))
        con2 = m.rho / (sum_a * float(np.float32(1.6605655)))
        s_ph = get_data(0, zs, pzs, photo, ge[i])
        s_ra = get_data(0, zs, pzs, rayl, ge[i])
        s_pa = get_data(1, zs, pzs, pair, ge[i])
        s_tr = get_data(2, zs, pzs, trip, ge[i])
        vals = {k: np.zeros(MXGE) for k in tabs}
        gles = np.zeros(MXGE)
        for j in range(MXGE):
            gle = (j - ge[i, 0]) / ge[i, 1]
            gles[j] = gle
            e = math.exp(gle)
            sig_kn = sum_z * kn_sigma0(e)
            sig_p = s_pa[j] + s_tr[j]
            sigma = sig_kn + sig_p + s_ph[j]
            vals["gmfp"][j] = 1.0 / (sigma * con2)
            vals["gbr1"][j] = sig_p / sigma
            vals["gbr2"][j] = vals["gbr1"][j] + sig_kn / sigma
            vals["cohe"][j] = sigma / (s_ra[j] + sigma)
        for k, v in vals.items():
            t = tabs[k][i]
            for j in range(1, MXGE):
                t[j - 1, 1] = (v[j] - v[j - 1]) * ge[i, 1]
                t[j - 1, 0] = v[j] - t[j - 1, 1] * gles[j]
            t[MXGE - 1, 1] = t[MXGE - 2, 1]
            t[MXGE - 1, 0] = v[MXGE - 1] - t[MXGE - 1, 1] * gles[MXGE - 1]
    return ge, tabs


def read_ff_data(path):
    with open(path) as f:
        tok = np.array(f.read().split(), dtype=np.float64)
    xval = tok[:MXRAYFF]
    aff = tok[MXRAYFF:MXRAYFF + 100 * MXRAYFF].reshape(100, MXRAYFF)
    return xval, aff


def init_rayleigh_data(media, ff_path):
    """Rayleigh sampling tables (media.cu:1060-1292)."""
    xval, aff = read_ff_data(ff_path)
    nmed = len(media)
    ray = np.zeros((nmed, MXRAYFF, 4))
    i_arr = np.zeros((nmed, RAYCDFSIZE), dtype=np.uint32)
    pmax = np.zeros((nmed, MXGE, 2))
    for i, m in enumerate(media):
        ff = np.zeros(MXRAYFF)
        xgrid = xval.copy()
        for j in range(MXRAYFF):
            v = 0.0
            for z, pz in zip(m.z, m.pz):
                v += pz * aff[int(z) - 1, j] ** 2
            ff[j] = float(np.sqrt(np.float32(v)))   # sqrtf in the reference
        if xgrid[0] < 1e-6:
            xgrid[0] = float(np.float32(0.0001))
        ge1 = (MXGE - 1) / math.log(m.up / m.ap)
        ge0 = -ge1 * math.log(m.ap)
        emin = math.exp(-ge0 / ge1)
        emax = math.exp((MXGE - 1.0 - ge0) / ge1)
        ff[ff == 0.0] = 5e-324
        fcum = np.zeros(MXRAYFF)
        b_arr = np.zeros(MXRAYFF)
        s0 = 0.0
        for j in range(MXRAYFF - 1):
            b = math.log(ff[j + 1] / ff[j]) / math.log(xgrid[j + 1] / xgrid[j])
            b_arr[j] = b
            x1, x2 = xgrid[j], xgrid[j + 1]
            p1, p2 = x1 ** (2 * b), x2 ** (2 * b)
            s0 += ff[j] ** 2 * (x2 * x2 * p2 - x1 * x1 * p1) / ((1.0 + b) * p1)
            fcum[j + 1] = s0
        b_arr[MXRAYFF - 1] = 0.0
        pe = np.zeros(MXGE)
        dle = math.log(emax / emin) / (MXGE - 1.0)
        idx = 1
        for j in range(1, MXGE + 1):
            e = emin * math.exp(dle * (j - 1.0))
            xmax = 20.607544 * 2.0 * e / 0.5110034
            k = 1
            while k <= MXRAYFF - 1:
                if xgrid[k - 1] <= xmax < xgrid[k]:
                    break
                k += 1
            idx = k
            b = b_arr[idx - 1]
            x1, x2 = xgrid[idx - 1], xmax
            p1, p2 = x1 ** (2 * b), x2 ** (2 * b)
            pe[j - 1] = fcum[idx - 1] + ff[idx - 1] ** 2 * (x2 * x2 * p2 - x1 * x1 * p1) / ((1.0 + b) * p1)
        i_arr[i, RAYCDFSIZE - 1] = idx
        anorm = 1.0 / float(np.sqrt(np.float32(pe[MXGE - 1])))
        anorm1 = 1.005 / pe[MXGE - 1]
        anorm2 = 1.0 / pe[MXGE - 1]
        pe = np.minimum(pe * anorm1, 1.0)
        ff = ff * anorm
        fcum = fcum * anorm2
        c_arr = (1.0 + b_arr) / ((xgrid * ff) ** 2)
        # uniform cumulative bins
        dw = 1.0 / (RAYCDFSIZE - 1.0)
        xold = xgrid[0]
        ibin = 1
        b = b_arr[0]
        p1 = xgrid[0] ** (2 * b)
        i_arr[i, 0] = 1
        for j in range(2, RAYCDFSIZE):
            w = dw
            while True:
                x1 = xold
                x2 = xgrid[ibin]
                t = x1 * x1 * x1 ** (2 * b)
                p2 = x2 ** (2 * b)
                aux = ff[ibin - 1] ** 2 * (x2 * x2 * p2 - t) / ((1.0 + b) * p1)
                if aux > w:
                    xold = math.exp(math.log(t + w * (1.0 + b) * p1 / ff[ibin - 1] ** 2) / (2.0 + 2.0 * b))
                    i_arr[i, j - 1] = ibin
                    break
                w -= aux
                xold = x2
                ibin += 1
                b = b_arr[ibin - 1]
                p1 = xold ** (2 * b)
        b_arr = 0.5 / (1.0 + b_arr)
        for j in range(MXGE - 1):
            gle = (j - ge0) / ge1
            pmax[i, j, 1] = (pe[j + 1] - pe[j]) * ge1
            pmax[i, j, 0] = pe[j] - pmax[i, j, 1] * gle
        pmax[i, MXGE - 1] = pmax[i, MXGE - 2]
        ray[i, :, 0] = xgrid
        ray[i, :, 1] = fcum
        ray[i, :, 2] = b_arr
        ray[i, :, 3] = c_arr
    return ray, i_arr, pmax


# ------------------------------------------------------------------ electrons
def init_ms_sr(path):
    """q2+ surface reader (PhysicsInit.cu:860-906)."""
    with open(path) as f:
        tok = f.read().split()
    ums = np.zeros(MS_SIZE)
    fms = np.zeros(MS_SIZE)
    wms = np.zeros(MS_SIZE)
    ims = np.zeros(MS_SIZE, dtype=np.int64)
    p = 0
    L1, Q1 = MAXL_MS + 1, MAXQ_MS + 1

    def ix(k, j, i):
        return k * L1 * Q1 + j * L1 + i
    for i in range(MAXL_MS + 1):
        for j in range(MAXQ_MS + 1):
            for k in range(MAXU_MS + 1):
                ums[ix(k, j, i)] = float(tok[p]); p += 1
            for k in range(MAXU_MS + 1):
                fms[ix(k, j, i)] = float(tok[p]); p += 1
            for k in range(MAXU_MS):
                wms[ix(k, j, i)] = float(tok[p]); p += 1
            for k in range(MAXU_MS):
                ims[ix(k, j, i)] = int(tok[p]); p += 1
            for k in range(MAXU_MS):
                fms[ix(k, j, i)] = fms[ix(k + 1, j, i)] / fms[ix(k, j, i)] - 1.0
                ims[ix(k, j, i)] -= 1
            fms[ix(MAXU_MS, j, i)] = fms[ix(MAXU_MS - 1, j, i)]
    llammin = math.log(LAMBMIN_MS)
    llammax = math.log(LAMBMAX_MS)
    dllamb = (llammax - llammin) / MAXL_MS
    dqms = QMAX_MS / MAXQ_MS
    return ums, fms, wms, ims, 1.0 / dllamb, 1.0 / dqms


def mscati(media, region_ecuts, ximax=0.5, estepe=0.25):
    """PhysicsInit.cu:908-1182 for PRESTA-I with inexact BCA (the reference default)."""
    # skin depth for the BCA from the minimum ECUT over regions
    ecutmn = min(region_ecuts) if region_ecuts else 1e30
    tstbmn = 1e30
    for m in media:
        tstbm = (ecutmn - 0.5110034) * (ecutmn + 0.5110034) / (ecutmn * ecutmn)
        r = ecutmn / m.xcc
        tstbm = m.blcc * tstbm * r * r
        aux = math.log(tstbm)
        tstbm = math.log(tstbm / aux)
        tstbmn = min(tstbmn, tstbm)
    skindepth = math.exp(tstbmn)
    for m in media:
        m.blcc *= 1.16699413758864573
        m.xcc = m.xcc * m.xcc
    rm = RM_F
    for m in media:
        neke = m.meke
        e0, e1 = m.eke0, m.eke1
        sige, sigp = 1e-15, 1e-15
        mono_e = mono_p = True
        old_e = old_p = -1.0
        for i in range(1, neke + 1):
            ei = math.exp((i - e0) / e1)
            eil = math.log(ei)
            s = (m.esig[i - 1, 1] * eil + m.esig[i - 1, 0]) / (m.ededx[i - 1, 1] * eil + m.ededx[i - 1, 0])
            sige = max(sige, s)
            if s < old_e:
                mono_e = False
            old_e = s
            s = (m.psig[i - 1, 1] * eil + m.psig[i - 1, 0]) / (m.pdedx[i - 1, 1] * eil + m.pdedx[i - 1, 0])
            sigp = max(sigp, s)
            if s < old_p:
                mono_p = False
            old_p = s
        m.sig_e = (sige, sigp)
        m.sig_ismonotone = (mono_e, mono_p)
        # e_array, range_ep
        e_array = np.zeros(MXEKE)
        rng = np.zeros((MXEKE, 2))
        ei = math.exp((1 - e0) / e1)
        e_array[0] = ei
        for i in range(1, neke):
            eip1 = math.exp((i + 1 - e0) / e1)
            e_array[i] = eip1
            eke = (eip1 + ei) * 0.5
            elke = math.log(eke)
            lelke = int(e1 * elke + e0)
            for q, tab in ((1, m.pdedx), (0, m.ededx)):
                ededx = tab[lelke - 1, 1] * elke + tab[lelke - 1, 0]
                aux = tab[i - 1, 1] / ededx
                r1 = (eip1 - ei) / eke
                rng[i, q] = rng[i - 1, q] + (eip1 - ei) / ededx * (1 + aux * (aux * 2 + 1) * r1 * r1 / 24)
            ei = eip1
        m.e_array = e_array
        m.range_ep = rng
        # tmxs
        tm = np.zeros((MXEKE, 2))
        eil = (1 - e0) / e1
        ei = math.exp(eil)
        p2 = ei * (ei + rm * 2)
        beta2 = p2 / (p2 + rm * rm)
        chi_a2 = m.xcc / (p2 * 4 * m.blcc)
        dedx0 = m.ededx[0, 1] * eil + m.ededx[0, 0]
        estepx = p2 * 2 * beta2 * dedx0 / ei / m.xcc / (math.log(1.0 / chi_a2 + 1) * (chi_a2 + 1) - 1)
        estepx = min(estepx * ximax, estepe)
        si = estepx * ei / dedx0
        for i in range(1, neke):
            elke = (i + 1 - e0) / e1
            eke = math.exp(elke)
            lelke = i + 1
            p2 = eke * (eke + rm * 2)
            beta2 = p2 / (p2 + rm * rm)
            chi_a2 = m.xcc / (p2 * 4 * m.blcc)
            ededx = m.ededx[lelke - 1, 1] * elke + m.ededx[lelke - 1, 0]
            estepx = p2 * 2 * beta2 * ededx / eke / m.xcc / (math.log(1.0 / chi_a2 + 1) * (chi_a2 + 1) - 1)
            estepx = min(estepx * ximax, estepe)
            ekef = (1 - estepx) * eke
            if ekef <= e_array[0]:
                sip1 = (e_array[0] - ekef) / dedx0
                ekef = e_array[0]
                elkef = (1 - e0) / e1
                lelkef = 1
            else:
                elkef = math.log(ekef)
                lelkef = int(e1 * elkef + e0)
                leip1l = lelkef + 1
                eip1l = (leip1l - e0) / e1
                eip1 = e_array[leip1l - 1]
                aux = (eip1 - ekef) / eip1
                elktmp = 0.5 * (elkef + eip1l + 0.25 * aux * aux * (1 + aux * (1 + 0.875 * aux)))
                ektmp = (ekef + eip1) * 0.5
                ededx = m.ededx[lelkef - 1, 1] * elktmp + m.ededx[lelkef - 1, 0]
                aux = m.ededx[lelkef - 1, 1] / ededx
                r1 = (eip1 - ekef) / ektmp
                sip1 = (eip1 - ekef) / ededx * (1 + aux * (aux * 2 + 1) * r1 * r1 / 24)
            sip1 = sip1 + rng[lelke - 1, 0] - rng[lelkef, 0]
            tm[i - 1, 1] = (sip1 - si) * e1
            tm[i - 1, 0] = sip1 - tm[i - 1, 1] * elke
            si = sip1
        tm[neke - 1] = tm[neke - 2]
        m.tmxs_ms = tm
    return skindepth


def fix_brems(m):
    """Screening-function coefficients dl1..dl6 and bpar (PhysicsInit.cu:1552-1641)."""
    from .pegs_lite import fcoulc, xsif
    m.log_ap = math.log(m.ap)
    zt = zb = zf = 0.0
    for z, p in zip(m.z, m.pz):
        fc = fcoulc(z)
        xi = xsif(z)
        aux = p * z * (z + xi)
        zt += aux
        zb -= aux * math.log(z) / 3
        zf += aux * fc
    zv = (zb - zf) / zt
    zg = zb / zt
    f1 = 2 * (20.863 + 4 * zg) - 2 * (20.029 + 4 * zg) / 3
    f2 = 2 * (20.863 + 4 * zv) - 2 * (20.029 + 4 * zv) / 3
    dl = np.zeros((6, 8))
    dl[:, 0] = [(20.863 + 4 * zg) / f1, -3.242 / f1, 0.625 / f1, (21.12 + 4 * zg) / f1, -4.184 / f1, 0.952]
    dl[:, 1] = [(20.029 + 4 * zg) / f1, -1.93 / f1, -0.086 / f1, (21.12 + 4 * zg) / f1, -4.184 / f1, 0.952]
    dl[:, 2] = [(20.863 + 4 * zv) / f2, -3.242 / f2, 0.625 / f2, (21.12 + 4 * zv) / f2, -4.184 / f2, 0.952]
    dl[:, 3] = [(20.029 + 4 * zv) / f2, -1.93 / f2, -0.086 / f2, (21.12 + 4 * zv) / f2, -4.184 / f2, 0.952]
    dl[:, 4] = [3 * (20.863 + 4 * zg) - (20.029 + 4 * zg), 3 * (-3.242) - (-1.930), 3 * 0.625 - (-0.086),
                2 * 21.12 + 8 * zg, 2 * (-4.184), 0.952]
    dl[:, 5] = [3 * (20.863 + 4 * zg) + (20.029 + 4 * zg), 3 * (-3.242) + (-1.930), 3 * 0.625 + (-0.086),
                4 * 21.12 + 16 * zg, 4 * (-4.184), 0.952]
    dl[:, 6] = [3 * (20.863 + 4 * zv) - (20.029 + 4 * zv), 3 * (-3.242) - (-1.930), 3 * 0.625 - (-0.086),
                2 * 21.12 + 8 * zv, 2 * (-4.184), 0.952]
    dl[:, 7] = [3 * (20.863 + 4 * zv) + (20.029 + 4 * zv), 3 * (-3.242) + (-1.930), 3 * 0.625 + (-0.086),
                4 * 21.12 + 16 * zv, 4 * (-4.184), 0.952]
    m.dl = dl
    m.bpar = (12 * dl[0, 7] / (3 * dl[0, 7] + dl[0, 6]), dl[0, 6] / (3 * dl[0, 7] + dl[0, 6]))


# ------------------------------------------------------------------ assembly
# ------------------------------------------------------------------ atomic data
def _numbers(path):
    with open(path) as f:
        return f.read().split()


def edgset(relax_path, cs_path):
    """PhysicsInit.cu:1184-1334: binding energies (eV -> MeV), shell
    interaction probabilities (the sixth set to 1.01), the 38 relaxation
    probabilities per element, and the photo-absorption edges and fits."""
    tok = _numbers(relax_path)
    k = 0

    def take(n):
        nonlocal k
        out = tok[k:k + n]
        k += n
        return out
    be = np.zeros((MXELEMENT, MXSHELL))
    ip = np.zeros((MXELEMENT, MXSHELL))
    rp = np.zeros((MXELEMENT, MXTRANS))
    for i in range(MXELEMENT):
        take(1)
        be[i] = [float(v) * 1e-6 for v in take(MXSHELL)]
    for i in range(MXELEMENT):
        take(1)
        ip[i, :MXSHELL - 1] = [float(v) for v in take(MXSHELL - 1)]
        ip[i, MXSHELL - 1] = 1.01
    for lo, hi in ((0, 19), (19, 26), (26, 32), (32, 37), (37, 38)):
        for i in range(MXELEMENT):
            take(1)
            rp[i, lo:hi] = [float(v) for v in take(hi - lo)]
    tok = _numbers(cs_path)
    k = 0
    ne = np.zeros(MXELEMENT, dtype=np.int32)
    ed = np.zeros((5, MXELEMENT, MXEDGE))        # a, b, c, d, energy
    for i in range(MXELEMENT):
        _, n = take(2)
        ne[i] = int(n)
        for j in range(ne[i]):
            ed[:, i, j] = [float(v) for v in take(5)]
    return be, ip, rp, ne, ed


def _erf_jo(jo):
    """0.82436063535 erf(0.70710678119 (1 + 0.3 jo)) as init_compton stores it
    (PhysicsInit.cu:1397: the reference drops the -1 of EGSnrc's erf1(x) - 1)."""
    return 0.82436063535 * math.erf(0.70710678119 * (1.0 + 0.3 * jo))


def init_compton(media, path):
    """PhysicsInit.cu:1336-1519: the sub-shell data (J0 x 137, binding in
    units of m) and per medium Walker's alias table over the sub-shells of
    its elements weighted by pz x occupation, built as the reference does."""
    with open(path) as f:
        lines = f.read().split("\n")[18:]
    tok = " ".join(lines).split()
    iz = np.zeros(MXTOTSH, dtype=np.int32)
    shn = np.zeros(MXTOTSH, dtype=np.int32)
    ne = np.zeros(MXTOTSH, dtype=np.int32)
    jo = np.zeros(MXTOTSH, dtype=np.float32)
    be = np.zeros(MXTOTSH, dtype=np.float32)
    erfjo = np.zeros(MXTOTSH, dtype=np.float32)
    f32 = np.float32
    for j in range(MXTOTSH):
        a = tok[5 * j:5 * j + 5]
        iz[j], shn[j], ne[j] = int(a[0]), int(a[1]), int(a[2])
        jo[j] = f32(f32(a[3]) * f32(137.0))
        be[j] = f32(float(f32(a[4])) * 1e-6 / float(f32(0.5110034)))
        erfjo[j] = _erf_jo(float(jo[j]))
    nshell = np.zeros(MXMED, dtype=np.int32)
    shell = np.zeros((MXMED, MXMDSH), dtype=np.int32)
    eno = np.zeros((MXMED, MXMDSH), dtype=np.float32)
    atbin = np.zeros((MXMED, MXMDSH), dtype=np.int32)
    for im, m in enumerate(media):
        pztot = f32(0.0)
        nsh = 0
        for iele in range(m.ne):
            z = int(m.z[iele])
            for j in range(MXTOTSH):
                if iz[j] == z:
                    nsh += 1
                    if nsh > MXMDSH:
                        raise RuntimeError("medium %s has more than MXMDSH Compton sub-shells" % m.name)
                    shell[im, nsh - 1] = j + 1
                    aux = f32(f32(m.pz[iele]) * f32(ne[j]))
                    eno[im, nsh - 1] = aux
                    pztot = f32(pztot + aux)
        if nsh == 0:
            raise RuntimeError("medium %s has no Compton sub-shells" % m.name)
        nshell[im] = nsh
        e = eno[im]
        for i in range(nsh):
            e[i] = f32(-f32(e[i] / pztot))
            atbin[im, i] = i + 1
        atav = f32(f32(1.0) / f32(nsh))
        for i in range(nsh - 1):
            jh = nsh - 1
            for t in range(nsh - 1):
                if e[t] < 0 and abs(e[t]) > atav:
                    jh = t
                    break
            jl = nsh - 1
            for t in range(nsh - 1):
                if e[t] < 0 and abs(e[t]) < atav:
                    jl = t
                    break
            aux = f32(atav - abs(e[jl]))
            e[jh] = f32(e[jh] + aux)
            e[jl] = f32(f32(-e[jl] / atav) + f32(jl + 1))
            atbin[im, jl] = jh + 1
            if i == nsh - 2:
                e[jh] = f32(jh + 1 + 1)
        for i in range(nsh):
            if e[i] < 0:
                e[i] = f32(1 + i + 1)
    return dict(iz=iz, shn=shn, ne=ne, jo=jo, be=be, erfjo=erfjo, n_shell=nshell, shell=shell, eno=eno,
                atbin=atbin)


def eii_init(media, c, binding, path, eii_flag):
    """PhysicsInit.cu:1643-2047: EII shells (K..L3 with binding energy above
    the lowest electron/photon cut of the media), their cross-section fits,
    and for media with such shells the electron tables corrected for EII
    (dedx, sig, ebr1, sig_e, sig_ismonotone; after mscati, as the reference
    calls it).  Returns the EII flag in force."""
    rm = RM_F
    nshells = np.zeros(MXELEMENT, dtype=np.int32)
    if not eii_flag:
        return 0
    e_min = min(min(m.ae - rm for m in media), min(m.ap for m in media))
    for m in media:
        for iele in range(m.ne):
            z = int(m.z[iele] + 0.5)
            if nshells[z - 1] == 0:
                nshells[z - 1] = sum(1 for ish in range(4) if binding[z - 1, ish] > e_min)
    nsh_tot = int(nshells.sum())
    if nsh_tot == 0:
        return 0
    if nsh_tot > MAX_EII_SHELLS:
        raise RuntimeError("more than MAX_EII_SHELLS EII shells")
    first = np.concatenate([[0], np.cumsum(nshells)[:-1]])
    for im, m in enumerate(media):
        nsh = 0
        for iele in range(m.ne):
            z = int(m.z[iele] + 0.5)
            c.eii_no[iele * MXMED + im] = int(nshells[z - 1])
            nsh += int(nshells[z - 1])
            c.eii_first[MXMED * iele + im] = int(first[z - 1]) + 1 if nshells[z - 1] > 0 else 0
        c.eii_nsh[im] = nsh
    tok = _numbers(path)
    nskip = int(tok[0])
    with open(path) as f:
        lines = f.read().split("\n")
    tok = " ".join(lines[1 + nskip:]).split()
    emax, nbin = float(tok[0]), int(tok[1])
    if nbin != N_EII_BINS:
        raise RuntimeError("inconsistent EII data file")
    k = 2
    ii = 0
    eii_z, eii_sh = [], []
    for _ in range(MXELEMENT):
        z, nsh = int(tok[k]), int(tok[k + 1])
        k += 2
        if nsh < nshells[z - 1]:
            raise RuntimeError("EII data lack shells of element %d" % z)
        for ish in range(1, nsh + 1):
            fmax = float(np.float32(tok[k]))
            aux = np.array(tok[k + 1:k + 1 + nbin], dtype=np.float32).astype(np.float64)
            k += 1 + nbin
            if ish <= nshells[z - 1]:
                ii += 1
                eii_z.append(z)
                eii_sh.append(ish)
                u = float(binding[z - 1, ish - 1])
                ex = nbin / math.log(emax / u)
                ey = 1.0 - ex * math.log(u)
                c.eii[ii - 1].x, c.eii[ii - 1].y = ex, ey
                exf, eyf = float(c.eii[ii - 1].x), float(c.eii[ii - 1].y)
                for kk in range(1, nbin + 1):
                    sigo = fmax * aux[kk - 2] if kk > 1 else 0.0
                    loge = (kk - eyf) / exf
                    xs = (fmax * aux[kk - 1] - sigo) * exf
                    q = nbin * (ii - 1) + kk - 1
                    c.eii_xsection[q].x = xs
                    c.eii_xsection[q].y = sigo - float(np.float32(xs)) * loge
        if ii == nsh_tot:
            break
    # electron tables with EII (the reference's shell lookup below reads
    # eii[] at the shell's index within its element, jjj, where the shell's
    # own entry is jj: the same when the first EII element is the only one)
    occn = (2, 2, 2, 4)
    for im, m in enumerate(media):
        ec = m.ae - rm
        ecc = min(ec, m.ap)
        sum_a = sum(m.pz[i] * m.wa[i] for i in range(m.ne))
        con_med = float(np.float32(m.rho / 1.6605655 / sum_a))
        c.eii_cons[im] = con_med
        if c.eii_nsh[im] <= 0:
            continue
        e0, e1 = float(c.eke01[im].x), float(c.eke01[im].y)
        is_monotone = True
        sigma_max = 0.0
        dedx_old = sigma_old = wbrem_old = 0.0
        base = im * MXEKE
        for j in range(m.meke):
            loge = (j + 1 - e0) / e1
            e = math.exp(loge)
            tau = e / rm
            beta2 = tau * (tau + 2) / ((tau + 1) * (tau + 1))
            p2 = 2 * rm * tau * (tau + 2)
            dd, sg, eb = c.dedx[base + j], c.sig[base + j], c.ebr1[base + j]
            dedx = float(dd.y) * loge + float(dd.x)
            sig = float(sg.y) * loge + float(sg.x)             # the reference's condition is always true
            if e > 2 * ec:
                wbrem = float(eb.y) * loge + float(eb.x)
                sigm = sig * (1 - wbrem)
            else:
                sigm = 0.0
                wbrem = 1.0
            sum_occn = sum_sigma = sum_dedx = 0.0
            for iele in range(m.ne):
                z = int(m.z[iele] + 0.5)
                sum_sh = 0.0
                for ish in range(c.eii_no[MXMED * iele + im]):
                    jj = c.eii_first[MXMED * iele + im] + ish
                    jjj = eii_sh[jj - 1]
                    u = float(binding[z - 1, jjj - 1])
                    wmax = (e + u) / 2
                    uwm = u / wmax
                    if u < e and u > ecc:
                        sum_sh += occn[jjj - 1]
                        ss_0 = 2 * (math.log(p2 / u) - uwm ** 3 * math.log(p2 / wmax)
                                    - (beta2 + 0.833333) * (1 - uwm ** 3)) / 3 / u
                        sh_0 = ((1 - uwm) * (1 + uwm / (2 - uwm)) + u * (wmax - u) / (e + rm) ** 2
                                - (2 * tau + 1) / (tau + 1) ** 2 * uwm / 2 * math.log((2 - uwm) / uwm)) / u
                        ss_1 = math.log(p2 / u) - uwm * uwm * math.log(p2 / wmax) - (beta2 + 1) * (1 - uwm * uwm)
                        sh_1 = (math.log(wmax / u / (2 - uwm)) + 2 * (wmax - u) / (2 * wmax - u)
                                + (wmax * wmax - u * u) / (e + rm) ** 2 / 2
                                - (2 * tau + 1) / (tau + 1) ** 2 * math.log((2 * wmax - u) / wmax))
                        av_e = (ss_1 + sh_1) / (ss_0 + sh_0)
                        ex, ey = float(c.eii[jjj - 1].x), float(c.eii[jjj - 1].y)
                        i = int(ex * loge + ey)
                        i = (jj - 1) * N_EII_BINS + i
                        xs = c.eii_xsection[i - 1]
                        sig_j = (float(xs.x) * loge + float(xs.y)) * m.pz[iele] * con_med
                        sum_sigma += sig_j
                        sum_dedx += sig_j * av_e
                sum_occn += sum_sh * m.pz[iele]
            sigm += sum_sigma
            dedx -= sum_dedx
            aux = ec / e
            if e > 2 * ec:
                r1 = tau / (tau + 1)
                r2 = tau + 1
                k0 = float(np.float32(np.float32(0.153536) * np.float32(m.rho)))
                sigo = (sum_occn * k0 / (beta2 * ec) * ((1 - aux * 2) * (1 + aux / (1 - aux) + r1 * r1 * aux / 2)
                        - (tau * 2 + 1) / (r2 * r2) * aux * math.log((1 - aux) / aux)) / sum_a)
                de = (sum_occn * k0 / beta2 * (math.log(0.25 / aux / (1 - aux)) + (1 - aux * 2) / (1 - aux)
                      + r1 * r1 * (1 - aux * 4 * aux) / 8 - (tau * 2 + 1) / (r2 * r2) * math.log((1 - aux) * 2))
                      / sum_a)
                sigm -= sigo
                dedx += de
            sigma = sigm + wbrem * sig
            sigma_max = max(sigma_max, sigma / dedx)
            wbrem = wbrem * sig / sigma if sigma > 0 else 1.0
            if j > 0:
                d = c.dedx[base + j - 1]
                d.y = (dedx - dedx_old) * e1
                d.x = dedx - float(np.float32(d.y)) * loge
                q = c.sig[base + j - 1]
                q.y = (sigma - sigma_old) * e1
                q.x = sigma - float(np.float32(q.y)) * loge
                b = c.ebr1[base + j - 1]
                b.y = (wbrem - wbrem_old) * e1
                b.x = wbrem - float(np.float32(b.y)) * loge
                if sigma / dedx < sigma_old / dedx_old:
                    is_monotone = False
            dedx_old, sigma_old, wbrem_old = dedx, sigma, wbrem
        last, prev = base + m.meke - 1, base + m.meke - 2
        for arr in (c.dedx, c.sig, c.ebr1):
            arr[last].x, arr[last].y = arr[prev].x, arr[prev].y
        c.sig_ismonotone[2 * im] = int(is_monotone)
        c.sig_e[im] = sigma_max
    return 1


def data_files(data_dir=DATA_DIR, pegs="pegslite700.pegs4dat", prefix="pegslite"):
    """Paths of the data set: the PEGS4 file, the photon cross-section prefix
    (``<prefix>_photo.data`` ...), the atomic form factors, the
    multiple-scattering surface, the relaxation and photo-absorption edge
    data, the bound-Compton shells and the EII cross sections, in the
    pegs-lite naming of this package."""
    j = lambda n: os.path.join(data_dir, n)
    return dict(pegs=j(pegs), xsections=j(prefix), ff=j(prefix + "_ff.data"), msnew=j(prefix + "_msnew.data"),
                photo_relax=j(prefix + "_photo_relax.data"), photo_cs=j(prefix + "_photo_cs.data"),
                incoh=j(prefix + "_incoh.data"), eii=j(prefix + "_eii_ik.data"))


def reference_data_files(data_dir, pegs_file, photon_xsections, atomic_ff_file):
    """The same set as a run input names it, in the reference's naming
    (media.cu:151, :844-848; PhysicsInit.cu:865, :1209, :1311, :1377,
    :1800): <data dir>/<pegs file>, <data dir>/<photon xsections>_*.data,
    <data dir>/<atomic ff file>, msnew.data, photo_relax.data,
    photo_cs.data, incoh.data and eii_ik.data in the data directory."""
    j = lambda n: os.path.join(data_dir, n)
    return dict(pegs=j(pegs_file), xsections=j(photon_xsections), ff=j(atomic_ff_file), msnew=j("msnew.data"),
                photo_relax=j("photo_relax.data"), photo_cs=j("photo_cs.data"), incoh=j("incoh.data"),
                eii=j("eii_ik.data"))


class MediaTables:
    """Device physics tables for an ordered list of media."""

    def __init__(self, names, data_dir=DATA_DIR, pegs="pegslite700.pegs4dat", prefix="pegslite",
                 mlc_medium=None, eii_flag=1, edit=None, files=None, transport="presta-I", bca="inexact"):
        """edit: optional callable applied to the list of PegsMedium records
        read from the PEGS file before any table is derived from them (test
        media with synthetic stopping powers or cross sections).  files:
        optional paths that replace the pegs-lite names, with the keys of
        ``data_files`` (a run input's EGSnrc data set).  transport
        ("presta-I", the reference's default, or "presta-II") and bca
        ("inexact", the default, or "exact") are the electron-step options
        the reference selects at compile time (ESA_PII and BCA_EXACT,
        CUDA_EGS.h:35-38; PhysicsInit.cu:263-282, :923-996).  The inexact
        crossing takes the skin depth mscati computes; the exact one EGSnrc's
        3 elastic mean free paths (the reference's SKIN_DEPTH_FOR_BCA = 0
        would hold an electron on a boundary with zero-length steps)."""
        if transport not in ("presta-I", "presta-II") or bca not in ("inexact", "exact"):
            raise ValueError("transport is presta-I or presta-II, bca inexact or exact")
        if len(names) > MXMED:
            raise ValueError("at most %d media (MXMED)" % MXMED)
        self.names = list(names)
        fp = data_files(data_dir, pegs, prefix)
        fp.update(files or {})
        self.files = fp
        media = read_pegs_file(fp["pegs"], self.names)
        if edit is not None:
            edit(media)
        ge, ptabs = init_user_photon(media, fp["xsections"])
        ray, i_arr, pmax = init_rayleigh_data(media, fp["ff"])
        ums, fms, wms, ims, dllambi, dqmsi = init_ms_sr(fp["msnew"])
        # every region cut equals its medium's AE (init.cu:1005-1006)
        skin = mscati(media, [m.ae for m in media])
        be, ip, rp, ne_edges, ed = edgset(fp["photo_relax"], fp["photo_cs"])
        comp = init_compton(media, fp["incoh"])
        for m in media:
            fix_brems(m)
        c = CMediaTables()
        c.nmed = len(media)
        c.eii_flag = eii_flag
        c.nbr_split = 1
        c.ibrdst = 1
        c.iprdst = 1
        c.exact_bca = 1 if bca == "exact" else 0
        c.spin_effects = 0
        c.transport_algorithm = 0 if transport == "presta-II" else 1
        # the exact crossing takes EGSnrc's default of 3 elastic mean free
        # paths: with the reference's SKIN_DEPTH_FOR_BCA = 0 an electron on a
        # boundary (tperp = 0) gets tustep = 0 and never leaves it
        c.skindepth_for_bca = 3.0 if bca == "exact" else skin
        c.dllambi = dllambi
        c.dqmsi = dqmsi
        c.rm = RM_F
        for im, m in enumerate(media):
            c.rho[im] = m.rho
            c.te[im] = m.te
            c.thmoll[im] = m.thmoll
            c.ae[im] = m.ae
            c.ap[im] = m.ap
            c.log_ap[im] = m.log_ap
            c.smaxir[im] = m.smaxir
            c.e_max_rr[im] = E_MAX_RR
            c.zbrang[im] = m.zbrang
            c.lzbrang[im] = m.lzbrang
            c.delcm[im] = m.delcm
            c.blcc[im] = m.blcc
            c.xcc[im] = m.xcc
            c.i_do_rr[im] = 1 if (mlc_medium is not None and m.name == mlc_medium) else 0
            c.meke[im] = m.meke
            c.nne[im] = m.ne
            for ie in range(m.ne):
                c.zelem[MXMED * ie + im] = m.z[ie]
                c.pz[MXMED * ie + im] = m.pz[ie]
            for l in range(8):
                c.dl1[8 * im + l] = m.dl[0, l]
                c.dl2[8 * im + l] = m.dl[1, l]
                c.dl3[8 * im + l] = m.dl[2, l]
                c.dl4[8 * im + l] = m.dl[3, l]
                c.dl5[8 * im + l] = m.dl[4, l]
                c.dl6[8 * im + l] = m.dl[5, l]
            c.bpar[2 * im] = m.bpar[0]
            c.bpar[2 * im + 1] = m.bpar[1]
            c.sig_e[im] = m.sig_e[0]
            c.sig_e[MXMED + im] = m.sig_e[1]
            c.sig_ismonotone[2 * im] = int(m.sig_ismonotone[0])
            c.sig_ismonotone[2 * im + 1] = int(m.sig_ismonotone[1])
            c.eke01[im].x = m.eke0
            c.eke01[im].y = m.eke1
            for i in range(m.meke):
                for q, (st, dt) in enumerate(((m.esig, m.ededx), (m.psig, m.pdedx))):
                    k = q * MXEKE * MXMED + im * MXEKE + i
                    c.sig[k].x, c.sig[k].y = st[i, 0], st[i, 1]
                    c.dedx[k].x, c.dedx[k].y = dt[i, 0], dt[i, 1]
                k = im * MXEKE + i
                c.ebr1[k].x, c.ebr1[k].y = m.ebr1[i]
                c.pbr1[k].x, c.pbr1[k].y = m.pbr1[i]
                c.pbr2[k].x, c.pbr2[k].y = m.pbr2[i]
                c.tmxs[k].x, c.tmxs[k].y = m.tmxs_ms[i]
                c.range_ep[(im * MXEKE + i) * 2] = m.range_ep[i, 0]
                c.range_ep[(im * MXEKE + i) * 2 + 1] = m.range_ep[i, 1]
                c.e_array[k] = m.e_array[i]
            c.ge[im].x, c.ge[im].y = ge[im]
            for j in range(MXGE):
                k = im * MXGE + j
                for name in ("gmfp", "gbr1", "gbr2", "cohe"):
                    getattr(c, name)[k].x = ptabs[name][im, j, 0]
                    getattr(c, name)[k].y = ptabs[name][im, j, 1]
                c.pmax[k].x, c.pmax[k].y = pmax[im, j]
            for j in range(MXRAYFF):
                r = c.rayleigh_data[im * MXRAYFF + j]
                r.xgrid, r.fcum, r.b_array, r.c_array = ray[im, j]
            for j in range(RAYCDFSIZE):
                c.i_array[im * RAYCDFSIZE + j] = int(i_arr[im, j])
        for name, arr, dt in (("ums", ums, np.float32), ("fms", fms, np.float32),
                              ("wms", wms, np.float32), ("ims", ims, np.int16),
                              ("binding_energies", be, np.float32), ("interaction_prob", ip, np.float32),
                              ("relaxation_prob", rp, np.float32), ("edge_number", ne_edges, np.int32),
                              ("edge_a", ed[0], np.float32), ("edge_b", ed[1], np.float32),
                              ("edge_c", ed[2], np.float32), ("edge_d", ed[3], np.float32),
                              ("edge_energies", ed[4], np.float32),
                              ("n_shell", comp["n_shell"], np.int32), ("shell_array", comp["shell"], np.int32),
                              ("eno_array", comp["eno"], np.float32), ("eno_atbin_array", comp["atbin"], np.int32),
                              ("iz_array", comp["iz"], np.int32), ("shn_array", comp["shn"], np.int32),
                              ("be_array", comp["be"], np.float32), ("jo_array", comp["jo"], np.float32),
                              ("erfjo_array", comp["erfjo"], np.float32)):
            buf = np.ascontiguousarray(arr, dtype=dt)     # keep alive across the copy
            assert buf.nbytes == ctypes.sizeof(getattr(c, name)), name
            ctypes.memmove(getattr(c, name), buf.ctypes.data, buf.nbytes)
        # EII last: it corrects the electron tables written above (init.cu:1691-1695 order)
        c.eii_flag = eii_init(media, c, be, fp["eii"], eii_flag)
        self.compton = comp
        self.c = c
        self.media = media

    def raw(self):
        """Bytes of the ``gegs_media_tables`` block."""
        return ctypes.string_at(ctypes.addressof(self.c), ctypes.sizeof(self.c))

    def ptr(self):
        return ctypes.addressof(self.c)

    def index(self, name):
        return self.names.index(name)


def main(argv=None):
    """``python -m gegs_amd.hatch -o tables.gegsmed [--mlc-medium M] MEDIUM ...``:
    the media tables of the named media written as a media-table file, the
    physics input of the C++ run driver (gegs_amd/gegs_run -i <file.ini>
    --media tables.gegsmed).  With --data-dir and the run input's --pegs,
    --xsections and --ff names the EGSnrc data set there is used (as
    ``python -m gegs_amd.run`` does when it finds it), else pegs-lite."""
    import argparse
    from . import hostio
    ap = argparse.ArgumentParser(prog="python -m gegs_amd.hatch", description=main.__doc__.split("\n")[0])
    ap.add_argument("media", nargs="+", help="medium names, in table order")
    ap.add_argument("-o", dest="out", required=True, help="media-table file to write")
    ap.add_argument("--mlc-medium", default=None, help="the MLC medium (range rejection on)")
    ap.add_argument("--data-dir", default=None)
    ap.add_argument("--pegs", default="700icru.pegs4dat")
    ap.add_argument("--xsections", default="si")
    ap.add_argument("--ff", default="pgs4form.dat")
    a = ap.parse_args(argv)
    files = reference_data_files(a.data_dir, a.pegs, a.xsections, a.ff) if a.data_dir else None
    tables = MediaTables(a.media, mlc_medium=a.mlc_medium, files=files)
    hostio.write_media_file(a.out, tables)
    print("%s: %s" % (a.out, ", ".join(tables.names)))
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
