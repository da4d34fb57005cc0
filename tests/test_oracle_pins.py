"""Closed-form pins of the oracle samplers that energy conservation alone
left open: the photo-electron angle, in-flight annihilation, the brems
photon angle, the electron step's energy loss and range, and PRESTA-I's
lateral and longitudinal displacement.

Each test compares the oracle with a formula derived independently of the
sampler (a published cross section or an exact moment), and shows its own
power: the same samples are rejected by the formula with a plausible mistake
planted in it (a sign, a power or a dropped term)."""
import math
import types

import numpy as np
import pytest
from scipy import integrate, special, stats

from gegs_amd import hatch, native, workloads
from gegs_amd import phantom as ph
from oracle import oracle

RM = 0.5110034
P_PHOTO = 2
E_STEP, E_BREMS, E_ANNIH = 5, 7, 9


@pytest.fixture(scope="module")
def wl():
    return workloads.water(n=20, voxel=1.5)


def with_regions(wl, ecut=None, flags=None):
    """The workload with every voxel's electron cut (total energy) and/or
    region flags replaced."""
    reg = wl.regions.copy()
    vox = reg["med"] != 0xFFFF
    if ecut is not None:
        reg["ecut"][vox] = np.float32(ecut)
    if flags is not None:
        reg["flags"][vox] = flags
    return types.SimpleNamespace(phantom=wl.phantom, tables=wl.tables, field=wl.field, regions=reg)


def particle(status, e, charge=0, lin=1):
    p = np.zeros(1, dtype=native.PARTICLE_DTYPE)[0]
    p["e"] = e
    p["x"], p["y"], p["z"] = 0.1, 0.2, 100.0 + 15.0      # centre of the 30 cm water cube, head frame
    p["u"], p["v"], p["w"] = 0.0, 0.0, 1.0
    p["wt"] = 1.0
    p["dnear"] = 1e5
    p["charge"] = charge
    p["status"] = status
    p["region"] = 6 + 10 + 10 * 20 + 10 * 400
    p["lin"] = lin
    return p


def chi2_equal_bins(samples, cdf, grid, nbins=25):
    """Chi-square of samples against the distribution whose CDF is tabulated
    on `grid`, with bins of equal expected content (robust against the
    float32 granularity of the sampled variable)."""
    c = cdf(grid)
    edges = np.interp(np.linspace(0.0, 1.0, nbins + 1), c, grid)
    edges[0], edges[-1] = grid[0] - 1e-12, grid[-1] + 1e-12
    counts, _ = np.histogram(samples, bins=edges)
    exp = np.diff(np.interp(edges, grid, c)) * len(samples)
    keep = exp > 5
    c2 = np.sum((counts[keep] - exp[keep]) ** 2 / exp[keep])
    return stats.chi2.sf(c2, keep.sum() - 1)


def tabulated_cdf(density, lo, hi, n=200001):
    g = np.linspace(lo, hi, n)
    f = density(g)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(g))])
    cum /= cum[-1]
    return g, (lambda x: np.interp(x, g, cum))


# ----------------------------------------------------------------- Sauter
def sauter(c, gamma, mistake=None):
    """Sauter (1931) K-shell photo-electron distribution in cos(theta) for an
    electron of total energy gamma (units of m):
    sin^2 / (1 - beta c)^4 [1 + gamma (gamma - 1)(gamma - 2)(1 - beta c) / 2]."""
    beta = math.sqrt(1.0 - 1.0 / gamma ** 2)
    x = 1.0 + beta * c if mistake == "mirror" else 1.0 - beta * c
    k = 0.5 * gamma * (gamma - 1.0) * (gamma - 2.0)
    if mistake == "sign":
        k = -k
    power = 3 if mistake == "power" else 4
    return (1.0 - c * c) / x ** power * (1.0 + k * x)


@pytest.mark.parametrize("t,mistake", [(0.005, "mirror"), (0.3, "power"), (3.0, "sign"), (80.0, "power")])
def test_sauter_photoelectron_angle(wl, t, mistake):
    """photo.cu:197-244: the sampled cos(theta) of the photo-electron against
    the Sauter distribution.  5 keV and 80 MeV take the small beta/alpha
    branch (an approximation of Sauter good to O((beta/alpha)^2)), 80 MeV
    also the gamma > 100 form of cos(theta); 0.3 and 3 MeV the exact branch."""
    # without atomic relaxation the photo-electron takes the whole photon energy
    w = with_regions(wl, RM + 0.0009 if t < 0.1 else None, ph.KN_FLAGS)
    n = 6000
    cs, gam = [], None
    for i in range(n):
        o, _ = oracle.interaction(w, particle(P_PHOTO, t, lin=20000 + i), 5)
        el = o[o["charge"] == -1]
        assert len(el) == 1
        g = float(el["e"][0]) / RM
        gam = g if gam is None else gam
        assert abs(g - gam) < 1e-6 * g            # one electron energy per photon energy
        # the photon flies along +z: the electron's polar cosine is its w
        cs.append(float(el["w"][0]))
    cs = np.array(cs)
    lo = -1.0
    if t > 10:
        lo = 0.999                                # all of the distribution is within a few 1/gamma
    grid, cdf = tabulated_cdf(lambda c: sauter(c, gam), lo, 1.0)
    assert chi2_equal_bins(cs, cdf, grid) > 1e-4
    grid_m, cdf_m = tabulated_cdf(lambda c: sauter(c, gam, mistake), lo, 1.0)
    assert chi2_equal_bins(cs, cdf_m, grid_m) < 1e-6


# ----------------------------------------------------------------- Heitler
def heitler(eps, a, mistake=None):
    """Two-photon annihilation of a positron of total energy (a - 1) m on an
    electron at rest, in the energy fraction eps = k / (a m) of one photon.
    From the spin-averaged matrix element of e+ e- -> gamma gamma (the
    crossed Compton amplitude; Heitler 1954, section 29) the density is
    x2/x1 + x1/x2 + 2 (1/x1 + 1/x2) - (1/x1 + 1/x2)^2 with x1 = a eps,
    x2 = a (1 - eps) the photon energies in units of m."""
    x1, x2 = a * eps, a * (1.0 - eps)
    s = 1.0 / x1 + 1.0 / x2
    if mistake == "sign":
        return x2 / x1 + x1 / x2 - 2.0 * s - s * s
    if mistake == "noterm":
        return x2 / x1 + x1 / x2
    return x2 / x1 + x1 / x2 + 2.0 * s - s * s


@pytest.mark.parametrize("t,mistake,n", [(0.3, "noterm", 5000), (2.0, "sign", 5000), (10.0, "noterm", 20000)])
def test_heitler_inflight_annihilation(wl, t, mistake, n):
    """annih.cu:77-184: photon energies against the Heitler cross section,
    each photon's angle against two-body kinematics, and momentum balance."""
    e_pos = t + RM
    a = (e_pos + RM) / RM
    gam = a - 1.0
    p_pos = math.sqrt(t * (t + 2 * RM))
    eps = []
    for i in range(n):
        o, ed = oracle.interaction(wl, particle(E_ANNIH, e_pos, charge=1, lin=30000 + i), 9)
        assert len(o) == 2 and np.all(o["charge"] == 0) and ed == 0.0
        k = o["e"].astype(np.float64)
        d = np.stack([o["u"], o["v"], o["w"]], axis=1).astype(np.float64)
        assert abs(k.sum() - (e_pos + RM)) < 2e-6 * (e_pos + RM)
        # momentum: k1 d1 + k2 d2 = (0, 0, p_pos), up to the float32 granularity of
        # sin(theta) = sqrt(1 - cos^2) near cos = 1 (at most sqrt(2 * 6e-8) per photon)
        mom = (k[:, None] * d).sum(axis=0)
        tol = 4e-4 * k.sum() + 2e-5 * (e_pos + RM)
        assert np.allclose(mom, [0.0, 0.0, p_pos], atol=tol)
        # two-body kinematics: cos(theta) = a (x - 1) / (x p) with x = k / m
        x = k / RM
        pm = math.sqrt(gam * gam - 1.0)
        assert np.allclose(d[:, 2], np.minimum(1.0, a * (x - 1.0) / (x * pm)), atol=2e-5)
        eps.extend(k / (e_pos + RM))
    eps = np.array(eps)
    e0 = 1.0 / (a + math.sqrt(gam * gam - 1.0))
    grid, cdf = tabulated_cdf(lambda e: heitler(e, a), e0, 1.0 - e0)
    assert chi2_equal_bins(eps, cdf, grid) > 1e-4
    grid_m, cdf_m = tabulated_cdf(lambda e: heitler(e, a, mistake), e0, 1.0 - e0)
    assert chi2_equal_bins(eps, cdf_m, grid_m) < 1e-6


# ----------------------------------------------------------------- Koch-Motz 2BS
def km2bs(y2, e0, e, ztarg, mistake=None):
    """Koch and Motz (1959) formula 2BS, per unit y^2 (y = E0 theta):
    {16 y^2 E / ((y^2+1)^4 E0) - (E0 + E)^2 / ((y^2+1)^2 E0^2)
     + [(E0^2 + E^2) / ((y^2+1)^2 E0^2) - 4 y^2 E / ((y^2+1)^4 E0)] ln M(y)},
    1/M = (k / (2 E0 E))^2 + (Z^(1/3) / (111 (y^2+1)))^2, energies in units
    of m, with ztarg = (Z^(1/3) / 111)^2."""
    k = e0 - e
    q = 1.0 + y2
    lnm = -np.log((k / (2.0 * e0 * e)) ** 2 + ztarg / q ** 2)
    if mistake == "screening":
        lnm = -np.log((k / (2.0 * e0 * e)) ** 2)
    first = 16.0 * y2 * e / (q ** 4 * e0)
    if mistake == "sign":
        first = -first
    return (first - (e0 + e) ** 2 / (q ** 2 * e0 ** 2)
            + ((e0 ** 2 + e ** 2) / (q ** 2 * e0 ** 2) - 4.0 * y2 * e / (q ** 4 * e0)) * lnm)


def pit_km(y2, e0, e, y2max, ztarg, mistake=None, ngrid=3001):
    """Probability integral transform of each sample under its own 2BS density
    on [0, y2max] (integrated over s = (1 + y^2)^(-1/2), where it is smooth)."""
    s = np.linspace(1.0 / math.sqrt(1.0 + y2max), 1.0, ngrid)
    yg = 1.0 / s ** 2 - 1.0                                    # decreasing
    out = np.empty(len(y2))
    for i in range(len(y2)):
        f = km2bs(yg, e0, e[i], ztarg, mistake) * 2.0 / s ** 3
        cum = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(s))])
        # CDF in y^2 = mass at s' > s(y^2)
        cy = (cum[-1] - cum) / cum[-1]
        out[i] = np.interp(y2[i], yg[::-1], cy[::-1])
    return out


@pytest.mark.parametrize("t", [2.0, 6.0])
def test_koch_motz_brems_angle(wl, t):
    """brems.cu:191-270: the photon's polar angle against formula 2BS, in the
    variable y^2 = (1 - cos theta) beta (1 + beta) E0^2 that the reference
    uses to extend y = E0 theta to the whole sphere (it tends to (E0 theta)^2
    as theta -> 0 and beta -> 1).  Every photon is tested under the density of
    its own energy (probability integral transform, KS against uniform)."""
    e_tot = t + RM
    e0 = e_tot / RM
    beta = math.sqrt((e0 - 1.0) * (e0 + 1.0)) / e0
    y2max = 2.0 * beta * (1.0 + beta) * e0 * e0
    ztarg = float(wl.tables.c.zbrang[0])
    y2, ee = [], []
    for i in range(4000):
        o, _ = oracle.interaction(wl, particle(E_BREMS, e_tot, charge=-1, lin=40000 + i), 13)
        ph = o[o["charge"] == 0]
        if len(ph) != 1:
            continue                    # photon below PCUT: scored, not pushed
        k = float(ph["e"][0])
        c = float(ph["w"][0])
        y2.append((1.0 - c) * 0.5 * y2max)
        ee.append((e_tot - k) / RM)
    y2, ee = np.array(y2), np.array(ee)
    assert len(y2) > 3000
    u = pit_km(y2, e0, ee, y2max, ztarg)
    assert stats.kstest(u, "uniform").pvalue > 1e-4
    for mistake in ("sign", "screening"):
        um = pit_km(y2, e0, ee, y2max, ztarg, mistake)
        assert stats.kstest(um, "uniform").pvalue < 1e-6, mistake


# ----------------------------------------------------------------- CSDA range and energy loss
A_S, B_S = 1.9, 0.08           # synthetic stopping power S(T) = A + B ln T, MeV/cm, T in MeV


def csda_medium(media):
    """Water with S(T) = A_S + B_S ln T for electrons and positrons in every
    bin (the tables' log-linear form, exact), no discrete events, and elastic
    scattering scaled down 1e9-fold (multiple scattering off)."""
    m = media[0]
    for arr in (m.ededx, m.pdedx):
        arr[:, 0] = A_S
        arr[:, 1] = B_S
    m.esig[:] = 0.0
    m.psig[:] = 0.0
    m.blcc *= 1e-9
    m.xcc *= math.sqrt(1e-9)


def csda_range(t1, t0):
    """Path from kinetic energy t0 down to t1: int dT / (A + B ln T)
    = (e^(-A/B) / B) [Ei((A + B ln t0)/B) - Ei((A + B ln t1)/B)] (density 1)."""
    f = lambda t: math.exp(-A_S / B_S) / B_S * special.expi((A_S + B_S * math.log(t)) / B_S)
    return f(t0) - f(t1)


@pytest.fixture(scope="module")
def csda_tables():
    return hatch.MediaTables(["H2O700ICRU", "AIR700ICRU"], edit=csda_medium)


def test_range_table_is_the_csda_integral(csda_tables):
    """range_ep and the step's range (electron.cu:1414-1432, mscati's range_ep)
    equal int dT/S(T) from the bottom of the table (e_array[0]) to T."""
    tb = csda_tables
    t_lo = float(tb.c.e_array[0])
    for q in (0, 1):
        for t in (0.25, 0.6, 1.0, 2.5, 7.0, 15.0):
            r = oracle.electron_range(tb, q, 0, 1.0, t)
            exact = csda_range(t_lo, t)
            assert abs(r - exact) < 2e-5 * exact, (q, t, r, exact)
            # a range built from S at the bin's lower edge alone misses by far more
            assert abs(t / (A_S + B_S * math.log(t)) - exact) > 50 * abs(r - exact)


@pytest.mark.parametrize("t,step", [(2.0, 0.01), (2.0, 0.15), (0.5, 0.05), (8.0, 0.9), (1.0, 0.3)])
def test_step_energy_loss_is_the_csda_inverse(csda_tables, t, step):
    """compute_eloss (electron.cu:1649-1701): the energy lost over a path s
    is T - T1 with int_{T1}^{T} dT/S = s, within a bin and across bins."""
    tb = csda_tables
    rng = oracle.electron_range(tb, 0, 0, 1.0, t)
    de = oracle.compute_eloss(tb, 0, 0, rng, 1.0, step, t)
    from scipy.optimize import brentq
    t1 = brentq(lambda x: csda_range(x, t) - step, 1e-3, t)
    exact = t - t1
    assert abs(de - exact) < 2e-5 * t, (de, exact)
    # the first-order loss S(T) s misses by far more (the correction terms matter)
    first = (A_S + B_S * math.log(t)) * step
    assert abs(first - exact) > 20 * abs(de - exact) or abs(first - exact) < 1e-6 * t


@pytest.mark.parametrize("fix", [0, 1])
def test_csda_straight_ahead_path_to_ecut(csda_tables, fix):
    """A whole electron track with discrete events and multiple scattering
    off: the summed step lengths from T0 down to the cut equal the CSDA range
    int dT/S between T0 and the table's lowest energy (the last step spends
    the remaining range), with the reference's hownear offset (Q1) kept and
    fixed."""
    wl = workloads.water(n=20, voxel=1.5)
    w = types.SimpleNamespace(phantom=wl.phantom, tables=csda_tables, field=wl.field,
                              regions=hatch_regions(wl, csda_tables))
    t_lo = float(csda_tables.c.e_array[0])
    oracle.lib().oracle_set_fixes(1 if fix else 0)
    try:
        for t0 in (0.5, 2.0, 6.0):
            o, ed = oracle.interaction(w, particle(E_STEP, t0 + RM, charge=-1, lin=7), 3)
            assert len(o) == 0                      # no discrete event, nothing created
            assert abs(ed - t0) < 1e-5 * t0        # every MeV deposited along the track
            path = oracle.last_path()
            exact = csda_range(t_lo, t0)
            assert abs(path - exact) < 1e-4 * exact, (t0, path, exact)
    finally:
        oracle.lib().oracle_set_fixes(0)


def hatch_regions(wl, tables):
    from gegs_amd import phantom as ph
    return ph.region_table(wl.phantom, tables, wl.other_media)


# ----------------------------------------------------------------- PRESTA-I msdist
def screened_rutherford_g(lam, chia2, l):
    """G_l = lam (1 - <P_l>) for single scattering density 1 / (1 - mu + 2 chia2)^2."""
    f = lambda mu: 1.0 / (1.0 - mu + 2.0 * chia2) ** 2
    norm = integrate.quad(f, -1.0, 1.0, points=[1.0 - 1e-3, 1.0 - 1e-6], limit=400)[0]
    pl = integrate.quad(lambda mu: special.eval_legendre(l, mu) * f(mu), -1.0, 1.0,
                        points=[1.0 - 1e-3, 1.0 - 1e-6], limit=400)[0] / norm
    return lam * (1.0 - pl)


@pytest.mark.parametrize("t_kin,tustep,direction", [
    (2.0, 0.05, (0.0, 0.0, 1.0)),
    (1.0, 0.2, (0.48, 0.6, 0.64)),
    (0.4, 0.05, (0.0, -0.6, 0.8)),
])
def test_presta1_displacement_moments(wl, t_kin, tustep, direction):
    """msdist_pI (electron.cu:618-882) for a step of path t without energy
    loss, against what PRESTA-I fixes:
      * the displacement along the initial direction is t (1 - e^-G1)/G1,
        Lewis's exact mean longitudinal displacement (every sample: PRESTA-I
        sets it to the mean);
      * the final direction relative to the initial one has <cos theta> =
        e^-G1 and <sin^2 theta> = 2 (1 - e^-G2)/3 (Goudsmit-Saunderson);
      * the lateral displacement is t sin(theta)/2 along the final
        direction's azimuth, shortened to t sqrt(1 - z^2) where that would
        make the chord longer than t; the returned straight-line distance is
        the chord's length.
    G1 and G2 come from screened Rutherford scattering with the tables'
    Moliere constants: lam = t rho blcc / beta^2 / (1 + chia2), chia2 =
    xcc / (4 p^2 blcc)."""
    tb = wl.tables
    c = tb.c
    blcc, xcc = float(c.blcc[0]), float(c.xcc[0])
    p2 = t_kin * (t_kin + 2 * RM)
    beta2 = p2 / (p2 + RM * RM)
    chia2 = xcc / p2 / (4.0 * blcc)
    lam = tustep * blcc / beta2 / (1.0 + chia2)
    g1 = screened_rutherford_g(lam, chia2, 1)
    g2 = screened_rutherford_g(lam, chia2, 2)
    assert abs(g1 - 2 * lam * chia2 * ((1 + chia2) * math.log(1 + 1 / chia2) - 1)) < 1e-6 * g1
    n = 40000
    d0 = np.array(direction, dtype=np.float64)
    out = oracle.msdist(tb, 0, 1.0, t_kin + RM, 0.0, tustep, d0, n, seed=17).astype(np.float64)
    disp, om, ustep = out[:, :3], out[:, 3:6], out[:, 6]
    assert np.allclose(np.linalg.norm(disp, axis=1), ustep, rtol=2e-6, atol=1e-9)
    # longitudinal: Lewis's mean, for every sample
    along = disp @ d0
    z_lewis = (1.0 - math.exp(-g1)) / g1
    assert np.all(np.abs(along / tustep - z_lewis) < 2e-5)
    # angular moments of the final direction about the initial one
    cost = om @ d0
    sin2 = 1.0 - cost * cost
    err = cost.std() / math.sqrt(n)
    assert abs(cost.mean() - math.exp(-g1)) < 4 * err + 2e-3 * (1 - math.exp(-g1))
    s2 = 2.0 * (1.0 - math.exp(-g2)) / 3.0
    es = sin2.std() / math.sqrt(n)
    assert abs(sin2.mean() - s2) < 4 * es + 5e-3 * s2, (sin2.mean(), s2)
    # lateral: t sin(theta)/2 along the final azimuth, capped at the chord t
    lat = disp - along[:, None] * d0
    om_perp = om - cost[:, None] * d0
    rho = np.linalg.norm(lat, axis=1) / tustep
    cap = math.sqrt(max(0.0, 1.0 - z_lewis * z_lewis))
    want = np.minimum(0.5 * np.sqrt(sin2), cap)
    assert np.all(np.abs(rho - want) < 2e-4 + 2e-3 * want)
    par = (lat * om_perp).sum(axis=1)
    norm = np.linalg.norm(lat, axis=1) * np.linalg.norm(om_perp, axis=1)
    moved = norm > 1e-12 * tustep
    assert np.all(par[moved] / norm[moved] > 1.0 - 1e-3)     # same azimuth, not the opposite
    assert np.mean(moved) > 0.5


# ----------------------------------------------------------------- PRESTA-II msdist
def single_scattering_walk(lam, chia2, n, rng):
    """Brute-force transport over a path of unit length: elastic scatterings
    at mean rate lam along the path, each with the screened Rutherford angle
    1 - cos = 2 chia2 r / (1 - r + chia2) and a uniform azimuth, starting
    along +z.  Returns the final displacement and direction (n x 3)."""
    pos = np.zeros((n, 3))
    d = np.tile([0.0, 0.0, 1.0], (n, 1))
    left = np.ones(n)
    while True:
        act = left > 0.0
        if not act.any():
            return pos, d
        step = rng.exponential(1.0 / lam, n)
        go = np.where(act, np.minimum(step, left), 0.0)
        pos += d * go[:, None]
        scat = act & (step < left)
        left = left - go
        k = np.flatnonzero(scat)
        if len(k) == 0:
            continue
        r = rng.random(len(k))
        cost = 1.0 - 2.0 * chia2 * r / (1.0 - r + chia2)
        sint = np.sqrt(np.maximum(0.0, 1.0 - cost * cost))
        phi = rng.random(len(k)) * 2.0 * math.pi
        u, v, w = d[k, 0], d[k, 1], d[k, 2]
        st = np.sqrt(u * u + v * v)
        cp, sp = np.cos(phi), np.sin(phi)
        # rotate (theta, phi) about the current direction (along +-z: directly)
        along = st < 1e-10
        sts = np.where(along, 1.0, st)
        nu = np.where(along, sint * cp, u * cost + sint * (u * w * cp - v * sp) / sts)
        nv = np.where(along, sint * sp, v * cost + sint * (v * w * cp + u * sp) / sts)
        nw = np.where(along, w * cost, w * cost - sint * cp * st)
        d[k] = np.stack([nu, nv, nw], axis=1)
        d[k] /= np.linalg.norm(d[k], axis=1)[:, None]


@pytest.mark.parametrize("t_kin,tustep", [(0.1, 0.005), (0.3, 0.02), (0.3, 0.05), (1.0, 0.1)])
def test_presta2_displacement_moments(t_kin, tustep):
    """msdist_pII (electron.cu:282-614) for a step of path t without energy
    loss, against brute-force single-scattering transport of the same
    screened Rutherford scattering (the process whose moments PRESTA-II is
    built to reproduce, Kawrakow & Bielajew 1998): the means of the
    longitudinal displacement z and of z^2 within 0.5%, of the lateral
    displacement x^2 + y^2 within 2.5% (the algorithm's own bias here is
    +1% to +2%; exchanging the substep lengths b and c moves it by -4% to
    -12%, a sign error in delta by +3% to +7%), and the final direction's
    <cos theta> = e^-G1.  The energy-loss corrections of the algorithm are
    not pinned by this test."""
    tb = hatch.MediaTables(["H2O700ICRU"], transport="presta-II")
    c = tb.c
    assert c.transport_algorithm == 0
    blcc, xcc = float(c.blcc[0]), float(c.xcc[0])
    p2 = t_kin * (t_kin + 2 * RM)
    beta2 = p2 / (p2 + RM * RM)
    chia2 = xcc / p2 / (4.0 * blcc)
    lam = tustep * blcc / beta2 / (1.0 + chia2)
    g1 = 2 * lam * chia2 * ((1 + chia2) * math.log(1 + 1 / chia2) - 1)
    n = 200000
    out = oracle.msdist(tb, 0, 1.0, t_kin + RM, 0.0, tustep, np.array([0.0, 0.0, 1.0]), n, seed=23).astype(np.float64)
    disp, om, ustep = out[:, :3] / tustep, out[:, 3:6], out[:, 6] / tustep
    assert np.allclose(np.linalg.norm(disp, axis=1), ustep, rtol=2e-6, atol=1e-9)
    ss_pos, ss_dir = single_scattering_walk(lam, chia2, 100000, np.random.default_rng(5))

    def rel_agree(a, b, syst):
        err = math.sqrt(a.var() / len(a) / a.mean() ** 2 + b.var() / len(b) / b.mean() ** 2)
        return abs(a.mean() / b.mean() - 1.0) <= 3 * err + syst, (a.mean() / b.mean() - 1.0, err)

    ez = math.exp(-g1)
    # the final direction: both angles come from mscat, whose q^(2+) tables
    # reproduce 1 - <cos> of screened Rutherford scattering to about 1-3% here
    # (pinned on their own in test_samplers.py); the brute-force walk exactly
    assert abs(om[:, 2].mean() - ez) < 4 * om[:, 2].std() / math.sqrt(n) + 0.03 * (1 - ez)
    assert abs(ss_dir[:, 2].mean() - ez) < 4 * ss_dir[:, 2].std() / math.sqrt(len(ss_dir)) + 2e-3 * (1 - ez)
    ok, r = rel_agree(disp[:, 2], ss_pos[:, 2], 0.005)
    assert ok, r
    ok, r = rel_agree(disp[:, 2] ** 2, ss_pos[:, 2] ** 2, 0.005)
    assert ok, r
    ok, r = rel_agree((disp[:, :2] ** 2).sum(axis=1), (ss_pos[:, :2] ** 2).sum(axis=1), 0.025)
    assert ok, r
    # the longitudinal mean of the brute-force walk is Lewis's (1 - e^-G1) / G1
    assert abs(ss_pos[:, 2].mean() - (1 - ez) / g1) < 4 * ss_pos[:, 2].std() / math.sqrt(len(ss_pos)) + 1e-4
