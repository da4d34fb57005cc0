"""Block and wedge head modules (two BEAMnrc PYRAMIDS): input reading,
cross-section construction and the oracle's geometry (restatement of
geom.cu:1037-1149 and :1376-1412 with ROTATE_Block_XY / ROTATE_Wedge_XY).

The geometry is checked against a closed-form model of the solid: a point is
in the module when (y, z) lies in the first cross-section and (x, z) in the
second.  From inside, howfar stops at the first point where the ray leaves
either prism (found by bisection on the model); a ray along z through the
interior enters at the module's top face; a ray that misses a prism hands the
particle to the next module; hownear is the distance to the nearer prism
boundary (point-segment distances in float64)."""
import numpy as np
import pytest

from gegs_amd import head, workloads
from gegs_amd.native import PARTICLE_DTYPE
from oracle import oracle

EPSTEP = 5e-4                       # step past a boundary (geom.cu, EPSTEP)


@pytest.fixture(scope="module")
def bw():
    return head.build_block_wedge(workloads.BLOCK_PYRAMIDS, workloads.WEDGE_PYRAMIDS)


def _poly(arr, n):
    return np.array([[arr[i].x, arr[i].y] for i in range(n)], dtype=np.float64)


def _inside(p, poly):
    odd = False
    n = len(poly)
    j = n - 1
    for i in range(n):
        (ax, ay), (bx, by) = poly[i], poly[j]
        if (ay < p[1] <= by) or (by < p[1] <= ay):
            if ax + (p[1] - ay) / (by - ay) * (bx - ax) < p[0]:
                odd = not odd
        j = i
    return odd


def _seg_dist(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    return float(np.linalg.norm(p - (a + t * ab)))


def _near(p, poly):
    n = len(poly)
    return min(_seg_dist(p, poly[i], poly[(i + 1) % n]) for i in range(n))


def _particle(x, y, z, u, v, w):
    p = np.zeros(1, dtype=PARTICLE_DTYPE)[0]
    p["x"], p["y"], p["z"], p["u"], p["v"], p["w"] = x, y, z, u, v, w
    p["region"] = 1
    p["e"] = 1.0
    return p


def test_pyramids_file_roundtrip(tmp_path):
    path = str(tmp_path / "wb.egsinp")
    head.write_egsinp_pyramids(path, workloads.BLOCK_PYRAMIDS, workloads.WEDGE_PYRAMIDS, "W700ICRU", "PB700ICRU")
    coord, med = head.read_egsinp_block(path)
    assert np.allclose(coord, workloads.BLOCK_PYRAMIDS) and med == "W700ICRU"
    rmax, layers, med = head.read_egsinp_wedge(path)
    assert rmax == 10.0 and med == "PB700ICRU"
    assert np.allclose(layers, workloads.WEDGE_PYRAMIDS[1])
    only_block = str(tmp_path / "b.egsinp")
    head.write_egsinp_pyramids(only_block, block=workloads.BLOCK_PYRAMIDS)
    assert head.read_egsinp_wedge(only_block) is None


def test_block_polygons_follow_the_pyramids_values():
    zmin, zmax, xfp, xbp, xfn, xbn, yfp, ybp, yfn, ybn = workloads.BLOCK_PYRAMIDS
    xz, yz = head.block_polygons(workloads.BLOCK_PYRAMIDS)
    assert np.array_equal(xz, np.float32([[xfp, zmin], [xbp, zmax], [xbn, zmax], [xfn, zmin]]))
    assert np.array_equal(yz, np.float32([[yfp, zmin], [ybp, zmax], [ybn, zmax], [yfn, zmin]]))


def test_wedge_polygon_branches():
    # the + side vertical, - side front > back: traced upwards along the - side fronts
    lay = [[0, 1, 5, 5, 4, 3], [1, 2, 5, 5, 3, 2]]
    g, b = head.wedge_polygon(7.0, lay)
    assert np.array_equal(g, np.float32([[4, 2], [4, 0], [3, 1], [2, 2]]))
    assert np.array_equal(b, np.float32([[-7, 0], [7, 0], [7, 2], [-7, 2]]))
    # the + side vertical, - side front <= back: traced downwards along the - side backs
    lay = [[0, 1, 5, 5, 1, 2], [1, 2, 5, 5, 2, 3]]
    g, _ = head.wedge_polygon(7.0, lay)
    assert np.array_equal(g, np.float32([[3, 0], [3, 2], [2, 1], [1, 0]]))
    # the - side vertical, + side front < back: upwards along the + side fronts
    lay = [[0, 1, 1, 2, -5, -5], [1, 2, 2, 3, -5, -5]]
    g, _ = head.wedge_polygon(7.0, lay)
    assert np.array_equal(g, np.float32([[1, 2], [1, 0], [2, 1], [3, 2]]))
    # the - side vertical, + side front >= back: downwards along the + side backs
    lay = [[0, 1, 4, 3, -5, -5], [1, 2, 3, 2, -5, -5]]
    g, _ = head.wedge_polygon(7.0, lay)
    assert np.array_equal(g, np.float32([[2, 0], [2, 2], [3, 1], [4, 0]]))
    # the workload's wedge: a triangle with a 1.5 cm heel at 2.4 cm
    g, _ = head.wedge_polygon(*workloads.WEDGE_PYRAMIDS)
    assert len(g) == 5 and g[0].tolist() == [2.4000000953674316, 61.5]
    with pytest.raises(ValueError):
        head.wedge_polygon(7.0, [[0, 1, 1, 2, 3, 4]] * 5)
    with pytest.raises(ValueError):
        head.wedge_polygon(7.0, [[0, 1, 1, 2, 3, 4]])


@pytest.mark.parametrize("is_wedge", [0, 1])
def test_geometry_against_the_solid(bw, is_wedge):
    if is_wedge:
        pa, pb = _poly(bw.wedge, bw.wedge_nver), _poly(bw.wedge_bound, 4)
        reg, nxt = 5, 6                      # the wedge hands over to the phantom
    else:
        pa, pb = _poly(bw.block_xz, 4), _poly(bw.block_yz, 4)
        reg, nxt = 4, head.M_WEDGE
    zlo, zhi = pb[:, 1].min(), pb[:, 1].max()

    def inside(x, y, z):
        return _inside((y, z), pa) and _inside((x, z), pb)

    rng = np.random.default_rng(5 + is_wedge)
    n_in = n_hit = 0
    for _ in range(400):
        # points inside: howfar stops where the ray first leaves either prism
        x, y, z = rng.uniform(-3, 3), rng.uniform(-2, 3), rng.uniform(zlo, zhi)
        if not inside(x, y, z):
            continue
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r, t, mod, tperp = oracle.howfar_block_wedge(bw, is_wedge, _particle(x, y, z, *d), 1e9)
        p = np.array([x, y, z])
        lo, hi = 0.0, 1.0
        while inside(*(p + hi * d)):
            hi *= 2
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if inside(*(p + mid * d)) else (lo, mid)
        assert r == 1 and mod == nxt
        assert abs((t - EPSTEP) - lo) < 6e-5 * max(1.0, lo)        # binary32 at z ~ 60 cm: ulp 3.8e-6
        # a step shorter than the exit stays in the module
        r2, t2, mod2, _ = oracle.howfar_block_wedge(bw, is_wedge, _particle(x, y, z, *d), 0.5 * lo)
        assert r2 == reg and mod2 == (head.M_WEDGE if is_wedge else head.M_BLOCK) and t2 == np.float32(0.5 * lo)
        near = min(_near(np.array([y, z]), pa), _near(np.array([x, z]), pb))
        assert abs(tperp - near) < 6e-5
        n_in += 1
    zs = np.linspace(zlo - 0.5, zhi + 0.5, 4001)
    for _ in range(200):
        # rays along z from above: the particle steps to the nearer of the two
        # prisms' entry points and counts as inside the module from there, even
        # where it has not reached the other prism yet (the reference's rule,
        # kept: for the wedge the bounding slab's top comes before the slanted
        # face)
        x, y = rng.uniform(-3, 3), rng.uniform(-2, 3)
        in_a = np.array([_inside((y, z), pa) for z in zs])
        in_b = np.array([_inside((x, z), pb) for z in zs])
        if not (in_a.any() and in_b.any()):
            continue
        entry = min(zs[np.argmax(in_a)], zs[np.argmax(in_b)])
        z0 = zlo - 3.0
        r, t, mod, _ = oracle.howfar_block_wedge(bw, is_wedge, _particle(x, y, z0, 0.0, 0.0, 1.0), 1e9)
        assert r == reg and mod == (head.M_WEDGE if is_wedge else head.M_BLOCK)
        assert abs(t - (entry - z0 + EPSTEP)) < 1.5 * (zs[1] - zs[0])     # the scan's resolution
        n_hit += 1
    assert n_in > 30 and n_hit > 20
    # a ray that runs beside the module misses a prism: handed to the next module at once
    r, t, mod, _ = oracle.howfar_block_wedge(bw, is_wedge, _particle(0.0, 9.0, zlo - 3.0, 0.0, 0.0, 1.0), 1e9)
    assert r == 1 and mod == nxt and t == np.float32(1e9)


def test_module_order_block_then_wedge(bw):
    """With both loaded, the block hands over to the wedge and the wedge to
    the phantom; with the block alone, the block hands over to the phantom."""
    zlo = workloads.BLOCK_PYRAMIDS[0]
    p = _particle(0.0, 9.0, zlo - 3.0, 0.0, 0.0, 1.0)
    assert oracle.howfar_block_wedge(bw, 0, p, 1e9)[2] == head.M_WEDGE
    only = head.build_block_wedge(workloads.BLOCK_PYRAMIDS, None)
    assert oracle.howfar_block_wedge(only, 0, p, 1e9)[2] == 6


def test_histories_beside_the_block_leave_with_their_full_step():
    """Transport through the block module as the reference does it: a ray
    that misses either prism passes to the next module with its step
    unchanged, so a photon beside the block moves on by its whole distance to
    the next interaction (metres in air) and leaves without reaching the
    phantom (quirk of geom.cu:1118-1123, kept); photons through the block are
    attenuated by it and reach the phantom from its exit face."""
    wl = workloads.ct_block_wedge(block=True, wedge=False)
    src = wl.source(6000, 3)
    o = oracle.run(wl, src, seed=9, per_history=True, threads=4)
    z0 = workloads.BLOCK_PYRAMIDS[0]
    s = (z0 - src["z"]) / src["w"]
    x, y = src["x"] + s * src["u"], src["y"] + s * src["v"]
    photon = src["charge"] == 0
    # the block's top face spans y in [0.55, 2.2] and x in [-1.65, 1.65] (z = 55 cm)
    through = (y > 0.6) & (y < 2.1) & (np.abs(x) < 1.6) & photon
    beside = ((y < 0.45) | (y > 2.35)) & photon
    assert through.sum() > 200 and beside.sum() > 1000
    assert np.all(o["edep"][beside] == 0.0)
    assert 0.0 < np.mean(o["edep"][through] > 0) < 0.5


def test_block_wedge_from_run_input(tmp_path):
    """[Block] / [Wedge] of a run input: the PYRAMIDS file is the data
    directory followed by the file property (CUDA_EGS.cu:1231-1275)."""
    from gegs_amd import runinput
    head.write_egsinp_pyramids(str(tmp_path / "WedgeBlock.egsinp"), workloads.BLOCK_PYRAMIDS,
                               workloads.WEDGE_PYRAMIDS, "W700ICRU", "PB700ICRU")
    ini = """[General]
output prefix = "x"
histories = 1e6
[Data]
data directory = "%s/"
egsinp file = "a.egsinp"
pegs file = "p"
photon xsections = "si"
atomic ff file = "ff"
default medium = AIR700ICRU
[Phase space]
phsp file number = 1
phase space file = "ph"
z position = 10
[SecJaws]
secjaws file = "a.egsinp"
[MLC]
mlcinfo file = "f.mlcinfo"
[Block]
module Block = On
block file = "WedgeBlock.egsinp"
[Wedge]
module Wedge = %s
wedge file = "WedgeBlock.egsinp"
[Phantom]
EGSPHANT FILE = "w.egsphant"
isocenter location = 0 0 0
source to isocenter = 100
"""
    for wedge_on in ("On", "Off"):
        p = tmp_path / ("run_%s.ini" % wedge_on)
        p.write_text(ini % (tmp_path, wedge_on))
        ri = runinput.read_run_input(str(p))
        bw, bmed, wmed = head.block_wedge_from_run_input(ri)
        ref = head.build_block_wedge(workloads.BLOCK_PYRAMIDS, workloads.WEDGE_PYRAMIDS if wedge_on == "On" else None)
        assert bytes(bw) == bytes(ref)
        assert bmed == "W700ICRU" and wmed == ("PB700ICRU" if wedge_on == "On" else None)
