"""pegs-lite: PEGS4-format media data generated from published formulas.

The reference reads EGSnrc's data sets (a PEGS4 file ``700icru.pegs4dat``,
``si_{photo,rayleigh,pair,triplet}.data``, ``pgs4form.dat`` and
``msnew.data``: media.cu:332-830, media.cu:832-1058, media.cu:1060-1292,
PhysicsInit.cu:860-906).  None of them ships with the reference, so this
module writes files in the *same formats* from textbook physics:

* electron tables (PEGS4 ``ELECIN``): Berger-Seltzer restricted collision
  stopping powers with the Sternheimer-Peierls density effect; restricted
  radiative stopping power and bremsstrahlung cross sections from the
  Bethe-Heitler law with the screening functions the samplers use
  (PhysicsInit.cu:1552-1641); Moller and Bhabha cross sections above
  TE = AE - m; Heitler in-flight annihilation; Moliere constants BLCC/XCC.
* photon cross sections per element: pair/triplet from the same
  Bethe-Heitler integrand, a smooth photo-electric fit calibrated on oxygen
  at 100 keV, Rayleigh from a screened hydrogenic form factor; Compton is
  Klein-Nishina and is computed by the table builder itself
  (media.cu:316-330), as in the reference.
* the multiple-scattering q^(2+) surface from the Legendre expansion of the
  screened Rutherford law (``tablegen/msgen.c``).

These are approximations chosen to be physically sensible; they are *inputs*
shared by the GPU path and the oracle, not a claim of EGSnrc data parity.
Every value is a plain function of the element list below, so regenerating
the files reproduces them bit for bit.
"""
from __future__ import annotations

import math
import os
import subprocess
import tempfile

import numpy as np

RM = 0.5110034            # MeV, the reference's electron rest mass (CUDA_EGS.h:152)
RE = 2.8179403e-13        # cm
FINE = 137.03604
NA = 6.0221367e23
HC = 0.012398418          # MeV * Angstrom
C2PI = 2.0 * math.pi * RE * RE * RM * NA     # 0.1535 MeV cm^2 / mol
C4AR = 4.0 / FINE * RE * RE * NA             # 4 alpha r_e^2 N_A, cm^2 / mol

# symbol, Z, A, mean excitation energy (eV) -- standard values (ICRU 37)
ELEMENTS = {
    "H": (1, 1.00794, 19.2), "C": (6, 12.011, 78.0), "N": (7, 14.0067, 82.0),
    "O": (8, 15.9994, 95.0), "NA": (11, 22.98977, 149.0),
    "MG": (12, 24.305, 156.0), "P": (15, 30.97376, 173.0),
    "S": (16, 32.066, 180.0), "CL": (17, 35.453, 174.0),
    "AR": (18, 39.948, 188.0), "K": (19, 39.0983, 190.0),
    "CA": (20, 40.078, 191.0), "FE": (26, 55.847, 286.0),
    "ZN": (30, 65.39, 330.0), "W": (74, 183.85, 727.0),
}

# name, density (g/cm^3), I (eV), gas, mass fractions (at most MXEL = 10
# elements per medium, EGSnrc_Parameters.h:20; trace Ca/Fe/Zn of ICRU lung are
# folded into the renormalisation)
MEDIA = {
    "AIR700ICRU": (0.00120479, 85.7, True,
                   {"C": 0.000124, "N": 0.755268, "O": 0.231781, "AR": 0.012827}),
    "LUNG700ICRU": (1.05, 75.3, False,
                    {"H": 0.101278, "C": 0.10231, "N": 0.02865, "O": 0.757072,
                     "NA": 0.00184, "MG": 0.00073, "P": 0.0008, "S": 0.00225,
                     "CL": 0.00266, "K": 0.00194}),
    "ICRUTISSUE700ICRU": (1.0, 74.9, False,
                          {"H": 0.101172, "C": 0.111, "N": 0.026, "O": 0.761828}),
    "ICRPBONE700ICRU": (1.85, 91.9, False,
                        {"H": 0.063984, "C": 0.278, "N": 0.027, "O": 0.410016,
                         "MG": 0.002, "P": 0.07, "S": 0.002, "CA": 0.147}),
    "H2O700ICRU": (1.0, 75.0, False, {"H": 0.111894, "O": 0.888106}),
    "W700ICRU": (19.3, 727.0, False, {"W": 1.0}),
}

AE = 0.700      # total-energy electron cut of the "700" data sets
AP = 0.010
UE = 50.511     # upper electron total energy
UP = 50.0
MEKE = 150      # electron mapped intervals (EGSnrc_Parameters.h:19)
MGE = 200       # photon mapped intervals (EGSnrc_Parameters.h:17)


# ---------------------------------------------------------------- screening
def fcoulc(z):
    """Coulomb correction, the closed form of PhysicsInit.cu:1521-1532."""
    asq = (z / FINE) ** 2
    return asq * (1.0 / (asq + 1.0) + 0.20206 + asq * (asq * (asq * (-0.002) + 0.0083) - 0.0369))


def xsif(z):
    """Electron-field ratio xi(Z), PhysicsInit.cu:1533-1550."""
    alrad = [5.31, 4.79, 4.74, 4.71]
    alradp = [6.144, 5.621, 5.805, 5.924]
    if z <= 4:
        iz = int(z)
        return alradp[iz - 1] / (alrad[iz - 1] - fcoulc(z))
    return math.log(1194.0 * z ** (-0.666667)) / (math.log(184.15 * z ** (-0.33333)) - fcoulc(z))


def phi12(delta):
    """Screening functions phi1, phi2 of the brems/pair samplers (brems.cu:163-172)."""
    delta = np.asarray(delta, dtype=np.float64)
    small = delta < 1.0
    p1 = np.where(small, 20.863 - 3.242 * delta + 0.625 * delta ** 2,
                  21.12 - 4.184 * np.log(np.maximum(delta, 1.0) + 0.952))
    p2 = np.where(small, 20.029 - 1.930 * delta - 0.086 * delta ** 2, p1)
    return p1, p2


# ---------------------------------------------------------------- composition
class Composition:
    def __init__(self, name):
        rho, ival, gas, frac = MEDIA[name]
        self.name = name
        self.rho = rho
        self.ival = ival * 1e-6          # MeV
        self.gas = gas
        self.sym = list(frac.keys())
        self.z = np.array([ELEMENTS[s][0] for s in self.sym], dtype=np.float64)
        self.a = np.array([ELEMENTS[s][1] for s in self.sym], dtype=np.float64)
        self.rhoz = np.array([frac[s] for s in self.sym], dtype=np.float64)
        self.rhoz /= self.rhoz.sum()
        self.pz = self.rhoz / self.a     # atoms per "molecule" of 1 g/mol
        self.wm = float(np.sum(self.pz * self.a))   # == 1
        self.za = float(np.sum(self.pz * self.z))   # electrons per (g/mol)
        xi = np.array([xsif(z) for z in self.z])
        fc = np.array([fcoulc(z) for z in self.z])
        aux = self.pz * self.z * (self.z + xi)
        self.zt = float(aux.sum())
        self.zg = float(np.sum(-aux * np.log(self.z) / 3.0) / self.zt)
        # radiation length (g/cm^2): Z^2 (Lrad - f) + Z Lrad'
        lr = np.array([
            [5.31, 4.79, 4.74, 4.71][int(z) - 1] if z <= 4 else math.log(184.15 * z ** (-1.0 / 3.0))
            for z in self.z])
        lrp = np.array([
            [6.144, 5.621, 5.805, 5.924][int(z) - 1] if z <= 4 else math.log(1194.0 * z ** (-2.0 / 3.0))
            for z in self.z])
        inv_x0 = C4AR / self.wm * np.sum(self.pz * (self.z ** 2 * (lr - fc) + self.z * lrp))
        self.x0 = 1.0 / inv_x0                   # g/cm^2
        self.rlc = self.x0 / self.rho            # cm
        self.delcm = 136.0 * RM * math.exp(self.zg)
        # Moliere constants (EGS4 / PEGS4 definitions, per g/cm^2)
        zz = self.pz * self.z * (self.z + 1.0)
        zs = float(zz.sum())
        ze = float(np.sum(zz * np.log(self.z ** (-2.0 / 3.0))))
        zx = float(np.sum(zz * np.log(1.0 + 3.34 * (self.z / FINE) ** 2)))
        self.bc = 6702.33 * zs * math.exp(ze / zs) / (self.wm * math.exp(zx / zs))
        self.xc = 0.39612 * math.sqrt(zs / self.wm)
        self._density_params()

    def _density_params(self):
        """Sternheimer-Peierls general density-effect parameters."""
        if self.name == "H2O700ICRU":
            self.dcbar, self.dx0, self.dx1, self.da, self.dm = 3.5017, 0.2400, 2.8004, 0.09116, 3.4773
            return
        hw = 28.816e-6 * math.sqrt(self.rho * self.za)    # plasma energy, MeV
        cbar = 1.0 + 2.0 * math.log(self.ival / hw)
        if self.gas:
            table = [(10.0, 1.6, 4.0), (10.5, 1.7, 4.0), (11.0, 1.8, 4.0), (11.5, 1.9, 4.0),
                     (12.25, 2.0, 4.0), (13.804, 2.0, 5.0)]
            x0, x1 = 0.326 * cbar - 2.5, 5.0
            for lim, a0, a1 in table:
                if cbar < lim:
                    x0, x1 = a0, a1
                    break
        else:
            if self.ival < 100e-6:
                x1 = 2.0
                x0 = 0.2 if cbar < 3.681 else 0.326 * cbar - 1.0
            else:
                x1 = 3.0
                x0 = 0.2 if cbar < 5.215 else 0.326 * cbar - 1.5
        m = 3.0
        a = (cbar - 4.606 * x0) / (x1 - x0) ** m
        self.dcbar, self.dx0, self.dx1, self.da, self.dm = cbar, x0, x1, a, m

    def density_effect(self, tau):
        x = np.log10(np.sqrt(tau * (tau + 2.0)))
        d = np.where(x < self.dx0, 0.0,
                     np.where(x < self.dx1,
                              4.606 * x - self.dcbar + self.da * np.abs(self.dx1 - x) ** self.dm,
                              4.606 * x - self.dcbar))
        return d


# ---------------------------------------------------------------- electrons
def collision_stopping(comp, t, tcut, positron):
    """Restricted collision stopping power (MeV cm^2/g), Berger-Seltzer form."""
    tau = t / RM
    g = tau + 1.0
    beta2 = tau * (tau + 2.0) / (g * g)
    if positron:
        d = np.minimum(tcut / RM, tau)
        y = 1.0 / (tau + 2.0)
        f = np.log(tau * d) - beta2 / tau * (
            tau + 2.0 * d - 1.5 * d * d * y - (d - d ** 3 / 3.0) * y * y
            - (0.5 * d * d - tau * d ** 3 / 3.0 + 0.25 * d ** 4) * y ** 3)
    else:
        d = np.minimum(tcut / RM, 0.5 * tau)
        f = (-1.0 - beta2 + np.log((tau - d) * d) + tau / (tau - d)
             + (0.5 * d * d + (2.0 * tau + 1.0) * np.log(1.0 - d / tau)) / (g * g))
    lg = np.log(2.0 * (tau + 2.0) / (comp.ival / RM) ** 2)
    return C2PI * comp.za / beta2 * (lg + f - comp.density_effect(tau))


def _brems_s(comp, e_tot, k):
    """Bethe-Heitler brems integrand S(k) (dsigma/dk = C4AR*ZT/WM * S/k)."""
    ep = e_tot - k
    a = ep / e_tot
    delta = comp.delcm * k / (e_tot * ep)
    p1, p2 = phi12(delta)
    s = (1.0 + a * a) * (p1 / 4.0 + comp.zg) - (2.0 / 3.0) * a * (p2 / 4.0 + comp.zg)
    return np.maximum(s, 0.0)


def brems_terms(comp, t):
    """(sigma_brem above AP [cm^2/g], radiative stopping below AP [MeV cm^2/g])."""
    e_tot = t + RM
    const = C4AR * comp.zt / comp.wm
    sig = 0.0
    if t > AP:
        lk = np.linspace(math.log(AP), math.log(t), 400)
        k = np.exp(lk)
        k[-1] = t * (1.0 - 1e-9)
        s = _brems_s(comp, e_tot, k)
        sig = const * np.trapezoid(s, lk)
    kc = min(AP, t)
    kk = np.linspace(0.0, kc * (1.0 - 1e-9), 200)
    s = _brems_s(comp, e_tot, np.maximum(kk, 1e-12))
    srad = const * np.trapezoid(s, kk)
    return sig, srad


def moller_sigma(comp, t):
    te = AE - RM
    if t <= 2.0 * te:
        return 0.0
    tau = t / RM
    g = tau + 1.0
    beta2 = tau * (tau + 2.0) / (g * g)
    e0 = te / t
    g2 = tau * tau / (g * g)
    g3 = (2.0 * tau + 1.0) / (g * g)
    br = 1.0 / e0 - 1.0 / (1.0 - e0) + g2 * (0.5 - e0) - g3 * math.log((1.0 - e0) / e0)
    return C2PI * comp.za / (beta2 * t) * br


def bhabha_sigma(comp, t):
    te = AE - RM
    if t <= te:
        return 0.0
    tau = t / RM
    g = tau + 1.0
    beta2 = tau * (tau + 2.0) / (g * g)
    e0 = te / t
    y = 1.0 / (tau + 2.0)
    yp = 1.0 - 2.0 * y
    b4 = yp ** 3
    b3 = b4 + yp * yp
    b2 = yp * (3.0 + y * y)
    b1 = 2.0 - y * y
    br = (1.0 / e0 - 1.0) - beta2 * (b1 * math.log(1.0 / e0) - b2 * (1.0 - e0)
                                     + b3 * (1.0 - e0 ** 2) / 2.0 - b4 * (1.0 - e0 ** 3) / 3.0)
    return C2PI * comp.za / (beta2 * t) * br


def annih_sigma(comp, t):
    gam = (t + RM) / RM
    s = math.sqrt(gam * gam - 1.0)
    per_e = math.pi * RE * RE / (gam + 1.0) * (
        (gam * gam + 4.0 * gam + 1.0) / (gam * gam - 1.0) * math.log(gam + s) - (gam + 3.0) / s)
    return per_e * NA * comp.za


def electron_tables(comp):
    """PEGS4 ELECIN block: node values on the ln-uniform kinetic grid."""
    te = AE - RM
    tmax = UE - RM
    eke1 = (MEKE - 1) / math.log(tmax / te)
    eke0 = 1.0 - eke1 * math.log(te)
    nodes = np.exp((np.arange(1, MEKE + 1) - eke0) / eke1)
    vals = {k: np.zeros(MEKE) for k in ("esig", "psig", "ededx", "pdedx", "ebr1", "pbr1", "pbr2")}
    for i, t in enumerate(nodes):
        sbr, srad = brems_terms(comp, t)
        sm = moller_sigma(comp, t)
        sb = bhabha_sigma(comp, t)
        sa = annih_sigma(comp, t)
        ecol = collision_stopping(comp, np.array([t]), te, False)[0]
        pcol = collision_stopping(comp, np.array([t]), te, True)[0]
        esig = sm + sbr
        psig = sb + sbr + sa
        vals["esig"][i] = esig * comp.x0
        vals["psig"][i] = psig * comp.x0
        vals["ededx"][i] = (ecol + srad) * comp.x0
        vals["pdedx"][i] = (pcol + srad) * comp.x0
        vals["ebr1"][i] = sbr / esig if esig > 0 else 0.0
        vals["pbr1"][i] = sbr / psig if psig > 0 else 0.0
        vals["pbr2"][i] = (sbr + sb) / psig if psig > 0 else 0.0
    return eke0, eke1, nodes, vals


def node_fit(nodes_ln, values, slope_scale):
    """Piecewise-linear-in-ln(E) coefficients from node values, as PEGS4 stores them."""
    n = len(values)
    a = np.zeros(n)
    b = np.zeros(n)
    for j in range(n - 1):
        b[j] = (values[j + 1] - values[j]) * slope_scale
        a[j] = values[j] - b[j] * nodes_ln[j]
    b[n - 1] = b[n - 2]
    a[n - 1] = values[n - 1] - b[n - 1] * nodes_ln[n - 1]
    return a, b


# ---------------------------------------------------------------- photons
def bh_pair_integral(z, k):
    """Integral over the e- energy fraction of the pair integrand (per Z(Z+xi))."""
    if k <= 2.0 * RM * 1.0001:
        return 0.0
    zg = -math.log(z) / 3.0
    delcm = 136.0 * RM * math.exp(zg)
    emin = RM / k
    eps = np.linspace(emin, 1.0 - emin, 401)
    delta = delcm / (k * eps * (1.0 - eps))
    p1, p2 = phi12(delta)
    f = ((eps ** 2 + (1.0 - eps) ** 2) * (p1 / 4.0 + zg)
         + (2.0 / 3.0) * eps * (1.0 - eps) * (p2 / 4.0 + zg))
    return float(np.trapezoid(np.maximum(f, 0.0), eps))


def pair_xs(z, k):
    """Nuclear-field pair cross section in barns/atom."""
    return 4.0 / FINE * RE * RE * 1e24 * z * z * bh_pair_integral(z, k)


def triplet_xs(z, k):
    """Electron-field pair production (barns/atom), threshold 4 m."""
    if k <= 4.0 * RM * 1.0001:
        return 0.0
    thr = (1.0 - 4.0 * RM / k) ** 3
    return 4.0 / FINE * RE * RE * 1e24 * z * bh_pair_integral(1.0, k) * thr


K_EDGE = {74: (0.0695, 4.4), 26: (0.00711, 7.6), 20: (0.00404, 8.0)}
L3_EDGE = {74: (0.0102, 2.5)}
PE_K = 2.83e-9


def photo_xs(z, e):
    """Smooth photo-effect fit in barns/atom: K Z^4.6 E^-3.3 (1 + (E/0.5)^2.3)."""
    s = PE_K * z ** 4.6 * e ** (-3.3) * (1.0 + (e / 0.5) ** 2.3)
    if z in K_EDGE and e < K_EDGE[z][0]:
        s /= K_EDGE[z][1]
        if z in L3_EDGE and e < L3_EDGE[z][0]:
            s /= L3_EDGE[z][1]
    return s


def form_factor(z, x):
    """Screened hydrogenic atomic form factor F(x, Z), x = sin(theta/2)/lambda in 1/A."""
    x0 = 0.30 * z ** (1.0 / 3.0)
    return z / (1.0 + (x / x0) ** 2) ** 2


def rayleigh_xs(z, e):
    """Coherent cross section (barns/atom) from the form factor above."""
    s = np.concatenate([[0.0], np.logspace(-14, math.log10(2.0), 3000)])
    x = e / HC * np.sqrt(0.5 * s)
    f = form_factor(z, x)
    mu = 1.0 - s
    return math.pi * RE * RE * 1e24 * float(np.trapezoid((1.0 + mu * mu) * f * f, s))


# ---------------------------------------------------------------- writers
def _fmt(v):
    return "%15.7E" % v


def write_pegs4(path, names):
    """Write media in the PEGS4 layout consumed by read_pegs_file (media.cu:332-830)."""
    lines = []
    for name in names:
        c = Composition(name)
        lines.append(" MEDIUM=%-24s,STERNCID=%-24s" % (name, name))
        lines.append(" COMPOUND,RHO=%12.7E,NE=%2d,IUNRST=0,EPSTFL=0,IAPRIM=0" % (c.rho, len(c.sym)))
        for s, z, a, pz, rz in zip(c.sym, c.z, c.a, c.pz, c.rhoz):
            lines.append(" ASYM=%-2s,Z=%3d.,A=%10.3f,PZ=%12.5E,RHOZ=%12.5E" % (s, int(z), a, pz, rz))
        lines.append(" ".join(_fmt(v) for v in (c.rlc, AE, AP, UE, UP)))
        lines.append(" %d %d %d %d %d %d %d %d %d" % (400, MGE, 300, MEKE, 100, 100, 100, 1, 0))
        # dl1..dl6 are recomputed by fix_brems; PEGS values are written for completeness
        for _ in range(6):
            lines.append(" ".join(_fmt(v) for v in (0.0,) * 6))
        lines.append(_fmt(c.delcm))
        for _ in range(2):
            lines.append(" ".join(_fmt(v) for v in (0.0, 0.0, 0.0)))
        lines.append(" ".join(_fmt(v) for v in (c.x0, 0.0, c.bc * c.x0, c.xc * math.sqrt(c.x0))))
        eke0, eke1, nodes, vals = electron_tables(c)
        lines.append(" ".join(_fmt(v) for v in (eke0, eke1)))
        lnn = np.log(nodes)
        fits = {k: node_fit(lnn, v, eke1) for k, v in vals.items()}
        for i in range(MEKE):
            row = []
            for k in ("esig", "psig", "ededx", "pdedx", "ebr1", "pbr1", "pbr2"):
                row += [fits[k][0][i], fits[k][1][i]]
            row += [0.0, 0.0]   # TMXS: recomputed by mscati
            lines.append(" ".join(_fmt(v) for v in row))
        # PHOTIN: EBINDA, GE0, GE1, then mge rows of gmfp/gbr1/gbr2 -- the
        # builder replaces them from the element cross sections
        ge1 = (MGE - 1) / math.log(UP / AP)
        ge0 = -ge1 * math.log(AP)
        lines.append(" ".join(_fmt(v) for v in (0.0, ge0, ge1)))
        for i in range(MGE):
            lines.append(" ".join(_fmt(v) for v in (0.0,) * 6))
        # Rayleigh block (irayl = 1): ngr, rco0, rco1, rsct, cohe -- unused by the builder
        lines.append(" %d" % 2)
        lines.append(" ".join(_fmt(v) for v in (0.0, 0.0)))
        for i in range(2):
            lines.append(" ".join(_fmt(v) for v in (0.0, 0.0)))
        for i in range(MGE):
            lines.append(" ".join(_fmt(v) for v in (1.0, 0.0)))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_xsec(path, fn, emin=1e-3, emax=1e5, npts=240, threshold=0.0):
    """Element cross sections in the read_xsec_data layout (media.cu:114-142):
    for Z = 1..100: n, then n lines 'ln(E) ln(sigma)'."""
    out = []
    for z in range(1, 101):
        lo = max(emin, threshold * 1.02) if threshold > 0 else emin
        le = np.linspace(math.log(lo), math.log(emax), npts)
        rows = []
        for l in le:
            s = fn(float(z), math.exp(l))
            rows.append((l, math.log(max(s, 1e-30))))
        out.append("%d" % len(rows))
        out += ["%.10e %.10e" % r for r in rows]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_ff(path):
    """Form factors in the read_ff_data layout (media.cu:144-171)."""
    x = np.concatenate([[0.0], np.logspace(-3, 4, 99)])
    out = ["%.8e" % v for v in x]
    for z in range(1, 101):
        out += ["%.8e" % form_factor(float(z), v) for v in x]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_msnew(path):
    """Run the q^(2+) generator and store its output as msnew-format text."""
    here = os.path.dirname(os.path.abspath(__file__))
    src = os.path.join(here, "tablegen", "msgen.c")
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "msgen")
        subprocess.check_call(["gcc", "-O2", "-o", exe, src, "-lm"])
        with open(path, "w") as f:
            subprocess.check_call([exe], stdout=f)


def generate(outdir, with_ms=True):
    """Write the complete pegs-lite data set into ``outdir``."""
    os.makedirs(outdir, exist_ok=True)
    write_pegs4(os.path.join(outdir, "pegslite700.pegs4dat"), list(MEDIA.keys()))
    write_xsec(os.path.join(outdir, "pegslite_photo.data"), photo_xs)
    write_xsec(os.path.join(outdir, "pegslite_rayleigh.data"), rayleigh_xs, npts=120)
    write_xsec(os.path.join(outdir, "pegslite_pair.data"), pair_xs, threshold=2 * RM, npts=120)
    write_xsec(os.path.join(outdir, "pegslite_triplet.data"), triplet_xs, threshold=4 * RM, npts=120)
    write_ff(os.path.join(outdir, "pegslite_ff.data"))
    if with_ms:
        write_msnew(os.path.join(outdir, "pegslite_msnew.data"))


if __name__ == "__main__":
    import sys
    generate(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "data"))
