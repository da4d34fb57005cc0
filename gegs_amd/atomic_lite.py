"""Atomic data for the default physics of the reference -- bound Compton
scattering with Doppler broadening, photo-electric shell selection, atomic
relaxation and electron-impact ionisation -- written in the EGSnrc file
formats the reference reads (PhysicsInit.cu edgset:1184-1334,
init_compton:1336-1519, eii_init:1643-2047):

  pegslite_photo_relax.data   binding energies (eV) of K, L1, L2, L3, <M>, <N>;
                              shell interaction probabilities; K/L1/L2/L3/M
                              transition probabilities (fluorescence, Auger,
                              Coster-Kronig) in the order of final_state
  pegslite_photo_cs.data      per element: edges and the four-parameter fits
                              of the photo-absorption cross section per edge
                              interval
  pegslite_incoh.data         per sub-shell: Z, relaxation shell, occupation,
                              Compton-profile value J(0) (atomic units),
                              binding energy (eV)
  pegslite_eii_ik.data        per element and shell K..L3: electron-impact
                              ionisation cross sections on a log grid

The EGSnrc data sets (Biggs et al. Compton profiles, EADL relaxation data,
the 'ik' EII cross sections) are not available here; like pegs-lite these
files are generated from published physics:
  * binding energies: K, L1, L2, L3, M and N edges of the elements of the
    media (H, C, N, O, Na, Mg, P, S, Cl, Ar, K, Ca, W; X-ray data booklet
    values, M and N as occupation-weighted sub-shell means), other elements
    interpolated in Z along Moseley's law sqrt(E) ~ Z between anchor edges;
  * Compton profiles: J(0) = <1/p>/2 of Slater-type orbitals with Slater's
    screening rules (momentum density by Hankel transform);
  * fluorescence yields: omega_K from Bambynek et al. (1972),
    (omega/(1-omega))^(1/4) = 0.015 + 0.0327 Z - 0.64e-6 Z^3; L-shell yields
    and Coster-Kronig fractions after Krause (1979) for tungsten, scaled as
    Z^4 / (Z^4 + a) for other elements; K-alpha2/K-alpha1 = 1/2 ... 0.58
    and K-beta/K-alpha rising with Z;
  * shell interaction probabilities 1 - 1/J from edge jump ratios J (K: the
    pegs-lite photo law's jumps for W, Fe, Ca, 125/Z + 3.5 otherwise);
  * photo-absorption fits: least-squares fits of the pegs-lite photo law per
    edge interval (so element selection in a compound is consistent with the
    photon tables);
  * EII cross sections: the relativistic binary-encounter-Bethe model (Kim,
    Santos and Parente 2000) with the orbital kinetic energy taken equal to
    the binding energy (virial theorem, hydrogenic shells).
Parity with the reference's EGSnrc-derived data is unpinned: these data do
not exist here.
"""
from __future__ import annotations

import math
import os

import numpy as np
from scipy import integrate, special

from . import pegs_lite

HERE = os.path.dirname(os.path.abspath(__file__))
DATA_DIR = os.path.join(HERE, "data")
RM = 0.5110034
MXELEMENT = 100
MXTOTSH = 1538          # EGSnrc_Parameters.h:55
N_EII_BINS = 250        # EGSnrc_Parameters.h:67
EII_EMAX = 1000.0       # MeV, top of the EII grid
RE2_BARN = 0.7940775    # pi r_e^2 in barns
FOUR_PI_RE2 = 4.0 * RE2_BARN

# ------------------------------------------------------------------ binding energies
# edges in eV: K, L1, L2, L3 and occupation-weighted means of the M and N
# sub-shells (0: shell empty)
EDGES = {
    1: (13.6, 0.0, 0.0, 0.0, 0.0, 0.0),
    6: (284.2, 18.0, 11.3, 11.3, 0.0, 0.0),
    7: (409.9, 37.3, 14.5, 14.5, 0.0, 0.0),
    8: (543.1, 41.6, 13.6, 13.6, 0.0, 0.0),
    11: (1070.8, 63.5, 30.81, 30.65, 5.1, 0.0),
    12: (1303.0, 88.7, 49.78, 49.50, 7.6, 0.0),
    15: (2145.5, 189.0, 136.0, 135.0, 10.5, 0.0),
    16: (2472.0, 230.9, 163.6, 162.5, 11.9, 0.0),
    17: (2822.4, 270.0, 202.0, 200.0, 13.7, 0.0),
    18: (3205.9, 326.3, 250.6, 248.4, 19.1, 0.0),
    19: (3608.4, 378.6, 297.3, 294.6, 23.8, 4.3),
    20: (4038.5, 438.4, 349.7, 346.2, 31.7, 6.1),
    74: (69525.0, 12100.0, 11544.0, 10207.0, 2125.0, 213.0),
}
# anchor edges (eV) for other elements
K_ANCHORS = [(1, 13.6), (2, 24.6), (3, 54.7), (6, 284.2), (8, 543.1), (13, 1559.6), (20, 4038.5),
             (26, 7112.0), (29, 8979.0), (32, 11103.0), (40, 17998.0), (47, 25514.0), (50, 29200.0),
             (53, 33169.0), (56, 37441.0), (64, 50239.0), (74, 69525.0), (79, 80725.0), (82, 88005.0),
             (92, 115606.0), (100, 141900.0)]
L_ANCHORS = [  # Z, L1, L2, L3
    (3, 5.0, 0.0, 0.0), (6, 18.0, 11.3, 11.3), (10, 48.5, 21.7, 21.6), (13, 117.8, 72.95, 72.55),
    (20, 438.4, 349.7, 346.2), (26, 844.6, 719.9, 706.8), (29, 1096.7, 952.3, 932.7),
    (32, 1414.6, 1248.1, 1217.0), (40, 2532.0, 2307.0, 2223.0), (47, 3806.0, 3524.0, 3351.0),
    (50, 4465.0, 4156.0, 3929.0), (53, 5188.0, 4852.0, 4557.0), (56, 5989.0, 5624.0, 5247.0),
    (64, 8376.0, 7930.0, 7243.0), (74, 12100.0, 11544.0, 10207.0), (79, 14353.0, 13734.0, 11919.0),
    (82, 15861.0, 15200.0, 13035.0), (92, 21757.0, 20948.0, 17166.0), (100, 27573.0, 26660.0, 20826.0)]
M_ANCHORS = [(11, 5.1), (18, 19.1), (20, 31.7), (26, 70.0), (30, 110.0), (40, 260.0), (50, 560.0),
             (60, 1030.0), (74, 2125.0), (82, 2900.0), (92, 4100.0), (100, 5300.0)]
N_ANCHORS = [(19, 4.3), (30, 9.4), (40, 35.0), (50, 90.0), (60, 130.0), (74, 213.0), (82, 330.0),
             (92, 500.0), (100, 650.0)]


def _moseley(anchors, z):
    zs = np.array([a[0] for a in anchors], dtype=float)
    vs = np.sqrt(np.array([a[1] for a in anchors], dtype=float))
    return float(np.interp(z, zs, vs)) ** 2


def edges(z):
    """Binding energies (eV) of K, L1, L2, L3, <M>, <N> of element z."""
    if z in EDGES:
        return EDGES[z]
    k = _moseley(K_ANCHORS, z)
    if z < 3:
        return (k, 0.0, 0.0, 0.0, 0.0, 0.0)
    l1 = _moseley([(a[0], a[1]) for a in L_ANCHORS], z)
    if z < 5:
        return (k, l1, 0.0, 0.0, 0.0, 0.0)
    l2 = _moseley([(a[0], a[2]) for a in L_ANCHORS if a[2] > 0], z)
    l3 = _moseley([(a[0], a[3]) for a in L_ANCHORS if a[3] > 0], z)
    m = _moseley(M_ANCHORS, z) if z >= 11 else 0.0
    n = _moseley(N_ANCHORS, z) if z >= 19 else 0.0
    return (k, l1, l2, l3, m, n)


# ------------------------------------------------------------------ ground-state sub-shells
MADELUNG = ["1s", "2s", "2p", "3s", "3p", "4s", "3d", "4p", "5s", "4d", "5p", "6s", "4f", "5d", "6p", "7s",
            "5f", "6d", "7p"]
CAP = {"s": 2, "p": 6, "d": 10, "f": 14}
LVAL = {"s": 0, "p": 1, "d": 2, "f": 3}


def configuration(z):
    """Occupations of the nl sub-shells of element z by Madelung's rule."""
    left = z
    occ = []
    for nl in MADELUNG:
        if left <= 0:
            break
        k = min(CAP[nl[1]], left)
        occ.append((int(nl[0]), nl[1], k))
        left -= k
    return occ


def split_j(n, l, k):
    """Occupations of the j = l - 1/2 and j = l + 1/2 levels of the 2p shell
    (L2 and L3, in proportion to 2j + 1); other shells are kept whole."""
    if not (n == 2 and l == "p"):
        return [(0, k)]
    lo = int(round(k / 3.0))
    return [(-1, lo), (1, k - lo)]


def relax_shell(n, l, j):
    """Shell index of the relaxation data: 1 K, 2 L1, 3 L2, 4 L3, 5 M, 6 N, 7 beyond."""
    if n == 1:
        return 1
    if n == 2:
        return 2 if l == "s" else (3 if j < 0 else 4)
    return 5 if n == 3 else (6 if n == 4 else 7)


def slater_zeta(z, occ, n, l):
    """Slater (1930) orbital exponent of the nl electrons: (Z - s) / n*."""
    grp = lambda nn, ll: (nn, 0) if ll in "sp" else (nn, LVAL[ll])
    g = grp(n, l)
    s = 0.0
    for (nn, ll, k) in occ:
        gg = grp(nn, ll)
        same = gg == g
        cnt = k - 1 if (nn == n and ll == l) else k
        if same or (nn == n and ll in "sp" and l in "sp"):
            s += cnt * (0.30 if n == 1 else 0.35)
        elif l in "sp":
            if nn == n - 1:
                s += cnt * 0.85
            elif nn < n - 1:
                s += cnt * 1.0
            elif nn == n and ll in "df":
                s += 0.0
        else:                           # d and f: everything in groups to the left screens fully
            if nn < n or (nn == n and LVAL[ll] < LVAL[l]):
                s += cnt * 1.0
    nstar = {1: 1.0, 2: 2.0, 3: 3.0, 4: 3.7, 5: 4.0, 6: 4.2, 7: 4.4}[n]
    return max(z - s, 0.5) / nstar, nstar


_J0_CACHE = {}


def sto_j0(nstar, lval, zeta):
    """J(0) = <1/p>/2 (atomic units) of the Slater-type orbital
    r^(n*-1) exp(-zeta r) Y_lm, via the Hankel transform to momentum space;
    J(0) scales as 1/zeta."""
    key = (round(nstar, 3), lval)
    if key not in _J0_CACHE:
        r = np.linspace(1e-6, 60.0, 6000)
        rad = r ** (nstar - 1.0) * np.exp(-r)
        rad /= math.sqrt(np.trapezoid(rad * rad * r * r, r))
        p = np.concatenate([np.linspace(1e-5, 4.0, 800), np.linspace(4.0, 60.0, 700)[1:]])
        phi = np.array([np.trapezoid(r * r * rad * special.spherical_jn(lval, pp * r), r) for pp in p])
        rho = (2.0 / math.pi) * phi * phi
        norm = np.trapezoid(rho * p * p, p)
        _J0_CACHE[key] = 0.5 * np.trapezoid(rho * p, p) / norm
    return _J0_CACHE[key] / zeta


# binding energies (eV) of whole nl sub-shells beyond L3 where the
# occupation-weighted mean <M> or <N> would misplace them (X-ray data booklet;
# the 5d and 6s of tungsten at its valence scale)
SUBSHELL_EV = {
    74: {(3, "s"): 2820.0, (3, "p"): 2379.0, (3, "d"): 1834.0, (4, "s"): 594.1, (4, "p"): 445.9,
         (4, "d"): 248.4, (4, "f"): 32.3, (5, "s"): 75.6, (5, "p"): 39.6, (5, "d"): 6.0, (6, "s"): 5.0},
    20: {(3, "s"): 44.3, (3, "p"): 25.4, (4, "s"): 6.1},
    19: {(3, "s"): 34.8, (3, "p"): 18.3, (4, "s"): 4.3},
    18: {(3, "s"): 29.3, (3, "p"): 15.9},
}


def valence_binding(z, n, l):
    """Binding energy (eV) of an outer sub-shell from its Slater exponent
    (hydrogenic 13.6 zeta^2, at least the first ionisation scale)."""
    zeta, nstar = slater_zeta(z, configuration(z), n, l)
    return max(13.6 * zeta * zeta, 5.0)


def subshells(z):
    """(relaxation shell, occupation, J0 a.u., binding eV) per sub-shell of element z."""
    occ = configuration(z)
    ed = edges(z)
    table = SUBSHELL_EV.get(z, {})
    out = []
    for (n, l, k) in occ:
        zeta, nstar = slater_zeta(z, occ, n, l)
        j0 = sto_j0(nstar, LVAL[l], zeta)
        for j, kk in split_j(n, l, k):
            if kk <= 0:
                continue
            sh = relax_shell(n, l, j)
            if (n, l) in table:
                b = table[(n, l)]
            elif sh <= 4 and ed[sh - 1] > 0:
                b = ed[sh - 1]
            elif sh == 5 and ed[4] > 0:
                b = ed[4]
            elif sh == 6 and ed[5] > 0:
                b = ed[5]
            else:
                b = valence_binding(z, n, l)
            out.append((sh, kk, j0, b))
    return out


# ------------------------------------------------------------------ relaxation probabilities
def omega_k(z):
    """K fluorescence yield, Bambynek et al. (1972)."""
    x = (0.015 + 0.0327 * z - 0.64e-6 * z ** 3) ** 4
    return x / (1.0 + x)


def transitions(z):
    """Transition probabilities of a K, L1, L2, L3 and M vacancy in the order
    of final_state (commonBlocks dev.h:260-262):
      K:  L3 L2 M N (fluorescence), Auger L1L1 L2L1 L3L1 L3L3 L3L2 L2L2 ML1 ML2
          ML3 NL1 NL2 NL3 MM NM NN
      L1: CK L2, CK L3, fluorescence M N, Auger MM NM NN
      L2: CK L3, fluorescence M N, Auger MM NM NN
      L3: fluorescence M N, Auger MM NM NN
      M:  fluorescence N (the Auger NN remainder is implied)
    Transitions into an empty shell have probability 0; each list sums to 1."""
    ed = edges(z)
    has = [e > 0.0 for e in ed]
    # K
    wk = omega_k(z) if has[1] else 0.0
    ratio_a2 = 0.5 + 0.08 * min(1.0, max(0.0, (z - 20) / 54.0))       # K-alpha2 / K-alpha1
    kb = 0.0 if not has[4] else min(0.23, 0.13 * min(1.0, (z - 10) / 15.0) + 0.10 * max(0.0, (z - 30) / 44.0))
    kn = 0.0 if not has[5] else min(0.05, 0.05 * max(0.0, (z - 30) / 44.0))
    rad = np.array([1.0, ratio_a2, kb * (1 + ratio_a2), kn * (1 + ratio_a2)])
    rad[0] *= has[3]
    rad[1] *= has[2]
    rad[2] *= has[4]
    rad[3] *= has[5]
    aug = np.array([0.08, 0.09, 0.12, 0.13, 0.20, 0.03, 0.05, 0.06, 0.10, 0.01, 0.02, 0.03, 0.05, 0.025, 0.005])
    a_sh = [(2, 2), (3, 2), (4, 2), (4, 4), (4, 3), (3, 3), (5, 2), (5, 3), (5, 4), (6, 2), (6, 3), (6, 4),
            (5, 5), (6, 5), (6, 6)]
    for i, (p, q) in enumerate(a_sh):
        if not (has[p - 1] and has[q - 1]):
            aug[i] = 0.0
    k = np.zeros(19)
    if rad.sum() > 0:
        k[:4] = wk * rad / rad.sum()
    if aug.sum() > 0:
        k[4:] = (1.0 - k[:4].sum()) * aug / aug.sum()
    elif rad.sum() > 0:
        k[:4] *= 1.0 / k[:4].sum()
    # L shells: yields after Krause (1979) at W, scaled as Z^4 / (Z^4 + a)
    z4 = float(z) ** 4
    w1 = 0.147 * (z4 / (z4 + 8.8e7)) / (74.0 ** 4 / (74.0 ** 4 + 8.8e7))
    w23 = z4 / (z4 + 8.8e7)
    w2, w3 = min(0.9, 1.06 * w23), min(0.9, w23)
    def shell_list(ck, w, mw_frac):
        """[CK..., fluorescence M, N, Auger MM, NM, NN]"""
        fl = np.array([1.0 - mw_frac, mw_frac]) * w
        fl[0] *= has[4]
        fl[1] *= has[5]
        au = np.array([0.70, 0.23, 0.07])
        au[0] *= has[4]
        au[1] *= has[4] and has[5]
        au[2] *= has[5]
        rest = 1.0 - sum(ck) - fl.sum()
        au = rest * au / au.sum() if au.sum() > 0 else np.zeros(3)
        if au.sum() == 0 and fl.sum() == 0:
            ck = [c / sum(ck) for c in ck] if sum(ck) > 0 else ck
        return np.concatenate([ck, fl, au])
    l1 = shell_list([0.12 * has[2], 0.30 * has[3]], w1, 0.25) if has[1] and has[4] else np.zeros(7)
    l2 = shell_list([0.13 * has[3]], w2, 0.22) if has[2] and has[4] else np.zeros(6)
    l3 = shell_list([], w3, 0.18) if has[3] and has[4] else np.zeros(5)
    m = np.array([0.025 * (z4 / (z4 + 2.0e8)) / (74.0 ** 4 / (74.0 ** 4 + 2.0e8)) if has[5] else 0.0])
    return k, l1, l2, l3, m


def interaction_probs(z):
    """Probability that the photo-absorption is on shell k (K, L1, L2, L3, M)
    given that the photon is above that shell's edge and was not absorbed on
    an earlier shell: 1 - 1/J with J the jump ratio of the edge."""
    jk = {74: 4.4, 26: 7.6, 20: 8.0}.get(z, 125.0 / z + 3.5) if z >= 3 else 1e9
    p = [1.0 - 1.0 / jk, 1.0 - 1.0 / 1.16, 1.0 - 1.0 / 1.4, 1.0 - 1.0 / 2.5, 0.7]
    # the outermost occupied shell of K..M takes the rest, so that the
    # selection never ends on a shell the element does not have
    ed = edges(z)
    last = max(k for k in range(5) if ed[k] > 0)
    if ed[5] <= 0:
        p[last] = 1.0
    return tuple(p)


# ------------------------------------------------------------------ photo-absorption fits
def photo_fit(z):
    """Edges (MeV, descending) of element z at or above 1 keV (at least the K
    edge) and per interval the coefficients a, b, c, d: above the K edge
    sigma = (a + b/E + c/E^2.5 + d/E^3)/E, below it sigma = exp(a + b lnE +
    c ln^2 E + d ln^3 E) (photo.cu:100-153), least-squares fits of the
    pegs-lite photo law."""
    ed = edges(z)
    es = [e * 1e-6 for e in ed[:4] if e > 0]
    es = sorted(set(es), reverse=True)
    keep = [es[0]] + [e for e in es[1:] if e >= 1e-3]
    if z in EDGES and ed[4] >= 1000.0:
        keep.append(ed[4] * 1e-6)
    keep = sorted(set(keep), reverse=True)
    out = []
    hi = 100.0
    for i, e_lo in enumerate(keep):
        # sample inside the interval, clear of the edges (the photo law's own
        # edges lie within 0.1% of these binding energies)
        lo = max(e_lo, 1e-3) * 1.005
        top = hi if i == 0 else keep[i - 1] * 0.995
        e = np.exp(np.linspace(math.log(lo), math.log(top), 120))
        s = np.array([pegs_lite.photo_xs(float(z), float(x)) for x in e])
        if i == 0:
            # sigma E = a + b/E + c/E^2.5 + d/E^3, relative least squares in
            # x = E / lo (scaled so the columns are of one order)
            x = e / lo
            A = np.stack([np.ones_like(x), 1 / x, x ** -2.5, x ** -3.0], axis=1) / (s * e)[:, None]
            c = np.linalg.lstsq(A, np.ones_like(x), rcond=None)[0]
            coef = np.array([c[0], c[1] * lo, c[2] * lo ** 2.5, c[3] * lo ** 3])
        else:
            le = np.log(e)
            A = np.stack([np.ones_like(le), le, le ** 2, le ** 3], axis=1)
            coef = np.linalg.lstsq(A, np.log(s), rcond=None)[0]
        out.append((coef, e_lo))
        hi = e_lo
    return out


# ------------------------------------------------------------------ EII (RBEB)
def rbeb(t, b, n):
    """Relativistic binary-encounter-Bethe ionisation cross section (barns) of
    a shell with binding energy b and n electrons by an electron of kinetic
    energy t (MeV); orbital kinetic energy U = B (Kim, Santos, Parente 2000)."""
    if t <= b:
        return 0.0
    tt = t / b
    tp, bp = t / RM, b / RM
    bt2 = 1.0 - 1.0 / (1.0 + tp) ** 2
    bb2 = 1.0 - 1.0 / (1.0 + bp) ** 2
    bu2 = bb2
    pre = FOUR_PI_RE2 * n / ((bt2 + bu2 + bb2) * 2.0 * bp)
    br = (0.5 * (math.log(bt2 / (1.0 - bt2)) - bt2 - math.log(2.0 * bp)) * (1.0 - 1.0 / tt ** 2)
          + 1.0 - 1.0 / tt - math.log(tt) / (tt + 1.0) * (1.0 + 2.0 * tp) / (1.0 + tp / 2.0) ** 2
          + bp * bp / (1.0 + tp / 2.0) ** 2 * (tt - 1.0) / 2.0)
    return max(pre * br, 0.0)


OCCN = (2, 2, 2, 4)          # electrons of K, L1, L2, L3 (eii_init's occn_numbers)


# ------------------------------------------------------------------ writers
def write_photo_relax(path):
    lines = []
    for z in range(1, MXELEMENT + 1):
        lines.append("%d " % z + " ".join("%.6g" % e for e in edges(z)))
    for z in range(1, MXELEMENT + 1):
        lines.append("%d " % z + " ".join("%.6f" % p for p in interaction_probs(z)))
    tr = [transitions(z) for z in range(1, MXELEMENT + 1)]
    for part in range(5):
        for z in range(1, MXELEMENT + 1):
            lines.append("%d " % z + " ".join("%.6f" % p for p in tr[z - 1][part]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_photo_cs(path):
    lines = []
    for z in range(1, MXELEMENT + 1):
        fit = photo_fit(z)
        lines.append("%d %d" % (z, len(fit)))
        for coef, e in fit:
            lines.append(" ".join("%.9e" % v for v in coef) + " %.9e" % e)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_incoh(path):
    head = ["Sub-shell Compton data generated by gegs_amd.atomic_lite:",
            "Z, relaxation shell (1 K, 2 L1, 3 L2, 4 L3, 5 M, 6 N, 7 beyond),",
            "occupation, J(0) in atomic units (Slater-type orbitals),",
            "binding energy in eV."] + [""] * 14
    rows = []
    for z in range(1, MXELEMENT + 1):
        for sh, ne, j0, b in subshells(z):
            rows.append("%d %d %d %.6f %.4f" % (z, sh, ne, j0, b))
    if len(rows) > MXTOTSH:
        raise RuntimeError("%d sub-shells exceed MXTOTSH" % len(rows))
    # records beyond the last element carry Z = 0, which no medium matches
    rows += ["0 7 0 1.000000 0.0000"] * (MXTOTSH - len(rows))
    with open(path, "w") as f:
        f.write("\n".join(head + rows) + "\n")


def write_eii(path):
    head = ["EII cross sections (barns) generated by gegs_amd.atomic_lite (RBEB, U = B);",
            "per element: Z, shells; per shell: scale, then 250 values x scale at",
            "E_k = B (Emax / B)^(k / 250), k = 1 .. 250."]
    lines = ["%d" % len(head)] + head + ["%.6e %d" % (EII_EMAX, N_EII_BINS)]
    for z in range(1, MXELEMENT + 1):
        ed = edges(z)
        shells = [k for k in range(4) if ed[k] > 0]
        lines.append("%d %d" % (z, len(shells)))
        for k in shells:
            b = ed[k] * 1e-6
            lg = math.log(EII_EMAX / b)
            sig = np.array([rbeb(b * math.exp(lg * i / N_EII_BINS), b, OCCN[k]) for i in range(1, N_EII_BINS + 1)])
            fmax = float(sig.max()) if sig.max() > 0 else 1.0
            lines.append("%.6e" % fmax)
            vals = sig / fmax
            for i in range(0, N_EII_BINS, 10):
                lines.append(" ".join("%.6e" % v for v in vals[i:i + 10]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def generate(outdir=DATA_DIR, prefix="pegslite"):
    os.makedirs(outdir, exist_ok=True)
    write_photo_relax(os.path.join(outdir, prefix + "_photo_relax.data"))
    write_photo_cs(os.path.join(outdir, prefix + "_photo_cs.data"))
    write_incoh(os.path.join(outdir, prefix + "_incoh.data"))
    write_eii(os.path.join(outdir, prefix + "_eii_ik.data"))


if __name__ == "__main__":
    import sys
    generate(sys.argv[1] if len(sys.argv) > 1 else DATA_DIR)
