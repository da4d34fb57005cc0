"""Pins of the oracle's building blocks: Philox known-answer vectors,
elementary-function accuracy, geometry against brute force."""
import ctypes

import numpy as np
import pytest

from oracle import oracle
from gegs_amd import workloads


def _philox(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    oracle.lib().oracle_philox(ctypes.c_void_p(c.ctypes.data), key[0], key[1])
    return [int(x) for x in c]


# Random123 known-answer vectors for Philox-4x32-10 (Salmon et al., SC'11)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert _philox(ctr, key) == list(expect)


def _ulp_err(got, ref):
    got = np.asarray(got, dtype=np.float32)
    ref32 = np.asarray(ref, dtype=np.float64)
    spacing = np.spacing(np.abs(ref32).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref32) / spacing


def test_log_accuracy():
    L = oracle.lib()
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(1e-7, 1.0, 20000), np.exp(rng.uniform(-20, 12, 20000)),
                        [1.0, 2.0, 0.5, 5.96046448e-08, 1e-30, 3.0e-39]]).astype(np.float32)
    got = np.array([L.oracle_logf(float(v)) for v in x], dtype=np.float32)
    ref = np.log(x.astype(np.float64))
    err = _ulp_err(got, ref)
    # near x = 1 the result is tiny; compare absolutely there
    far = np.abs(ref) > 1e-3
    assert err[far].max() <= 4.0
    assert np.max(np.abs(got[~far] - ref[~far])) < 1e-9
    assert L.oracle_logf(1.0) == 0.0


def test_exp_accuracy():
    L = oracle.lib()
    rng = np.random.default_rng(1)
    x = rng.uniform(-80.0, 80.0, 40000).astype(np.float32)
    got = np.array([L.oracle_expf(float(v)) for v in x], dtype=np.float32)
    err = _ulp_err(got, np.exp(x.astype(np.float64)))
    assert err.max() <= 4.0
    assert L.oracle_expf(0.0) == 1.0
    assert L.oracle_expf(-200.0) == 0.0


def test_sincos_accuracy():
    L = oracle.lib()
    rng = np.random.default_rng(2)
    t = np.concatenate([rng.uniform(0.0, 1.0, 20000), [0.0, 0.125, 0.25, 0.5, 0.75, 1.0]]).astype(np.float32)
    s = ctypes.c_float()
    c = ctypes.c_float()
    got_s, got_c = [], []
    for v in t:
        L.oracle_sincos2pi(float(v), ctypes.byref(s), ctypes.byref(c))
        got_s.append(s.value)
        got_c.append(c.value)
    a = 2.0 * np.pi * t.astype(np.float64)
    assert np.max(np.abs(np.array(got_s) - np.sin(a))) < 3e-7
    assert np.max(np.abs(np.array(got_c) - np.cos(a))) < 3e-7
    # identity to the precision of binary32
    assert np.max(np.abs(np.array(got_s) ** 2 + np.array(got_c) ** 2 - 1.0)) < 1e-6


def test_primary_lineages_distinct():
    L = oracle.lib()
    lins = {L.oracle_primary_lineage(i) for i in range(20000)}
    assert len(lins) == 20000


# ------------------------------------------------------------------ geometry
def _brute_force_exit(ph, pos, d, t_max):
    """Region entered and distance to it by exhaustive plane intersection in
    float64 (truth for howfar_phantom from inside a voxel)."""
    best_t, axis = t_max, -1
    bounds = (ph.xb.astype(np.float64), ph.yb.astype(np.float64), ph.zb.astype(np.float64))
    for k in range(3):
        if d[k] == 0:
            continue
        planes = bounds[k]
        ts = (planes - pos[k]) / d[k]
        ts = ts[ts > 1e-9]
        if ts.size and ts.min() <= best_t:
            best_t, axis = ts.min(), k
    return best_t, axis


def test_howfar_phantom_matches_brute_force():
    """Voxel index arithmetic of howfar_phantom (geom.cu:129-217): the region
    entered equals the neighbour along the first face crossed, found by
    exhaustive plane intersection in double precision."""
    L = oracle.lib()
    w = workloads.ct(nx=8, ny=6, nz=5, voxel=0.7, seed=3)
    ph = w.phantom
    xb, yb, zb = ph.padded_bounds()
    fld = workloads.make_field(False, 100.0)
    rng = np.random.default_rng(5)
    nx, ny, nz = ph.nx, ph.ny, ph.nz
    checked = 0
    for _ in range(3000):
        ix, iy, iz = rng.integers(nx), rng.integers(ny), rng.integers(nz)
        f = rng.uniform(0.05, 0.95, 3)
        pos = np.array([ph.xb[ix] + f[0] * (ph.xb[ix + 1] - ph.xb[ix]),
                        ph.yb[iy] + f[1] * (ph.yb[iy + 1] - ph.yb[iy]),
                        ph.zb[iz] + f[2] * (ph.zb[iz + 1] - ph.zb[iz])], dtype=np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        t_max = np.float32(rng.uniform(0.01, 3.0))
        region = 6 + ix + iy * nx + iz * nx * ny
        # head frame with SID 100: z_head = z + 100
        pd = np.array([pos[0], pos[1], pos[2] + np.float32(100.0), d[0], d[1], d[2]], dtype=np.float32)
        t = ctypes.c_float(float(t_max))
        new = L.oracle_howfar_phantom(nx, ny, nz, xb.ctypes.data, yb.ctypes.data, zb.ctypes.data,
                                      ctypes.addressof(fld), pd.ctypes.data, region, ctypes.byref(t))
        tb, axis = _brute_force_exit(ph, pos.astype(np.float64), d.astype(np.float64), float(t_max))
        # skip rays whose two nearest faces are within float rounding of each other
        if axis >= 0 and abs(t.value - tb) > 1e-5 * max(1.0, tb):
            continue
        if axis < 0:
            assert new == region and t.value == t_max
        else:
            idx = [ix, iy, iz]
            idx[axis] += 1 if d[axis] > 0 else -1
            n = (nx, ny, nz)
            inside = 0 <= idx[axis] < n[axis]
            expect = (6 + idx[0] + idx[1] * nx + idx[2] * nx * ny) if inside else 0
            assert new == expect
        checked += 1
    assert checked > 2500


def test_howfar_entry_from_outside():
    """A ray aimed at the phantom from outside enters the voxel its entry
    point lies in (isWhere bisection, geom.cu:20-36, vs linear search)."""
    L = oracle.lib()
    w = workloads.ct(nx=8, ny=6, nz=5, voxel=0.7, seed=3)
    ph = w.phantom
    xb, yb, zb = ph.padded_bounds()
    fld = workloads.make_field(False, 100.0)
    rng = np.random.default_rng(9)
    for _ in range(500):
        target = np.array([rng.uniform(ph.xb[0], ph.xb[-1]), rng.uniform(ph.yb[0], ph.yb[-1]), ph.zb[0]])
        start = target + np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), -10.0])
        d = target - start
        d = d / np.linalg.norm(d)
        pd = np.array([start[0], start[1], start[2] + 100.0, d[0], d[1], d[2]], dtype=np.float32)
        t = ctypes.c_float(1e8)
        new = L.oracle_howfar_phantom(ph.nx, ph.ny, ph.nz, xb.ctypes.data, yb.ctypes.data, zb.ctypes.data,
                                      ctypes.addressof(fld), pd.ctypes.data, 0, ctypes.byref(t))
        hit = np.float32(pd[0]) + np.float32(pd[3]) * np.float32(t.value)
        hy = np.float32(pd[1]) + np.float32(pd[4]) * np.float32(t.value)
        ix = int(np.searchsorted(ph.xb, hit, side="left")) - 1
        iy = int(np.searchsorted(ph.yb, hy, side="left")) - 1
        ix, iy = max(ix, 0), max(iy, 0)
        assert new == 6 + ix + iy * ph.nx
        assert abs(t.value - 10.0 / d[2]) < 1e-3


def test_hownear_keeps_reference_offset():
    """Quirk Q1 (geom.cu:1487-1490): hownear decodes region-5, i.e. the next
    voxel in x; inside a voxel the distance is then <= 0."""
    L = oracle.lib()
    w = workloads.ct(nx=8, ny=6, nz=5, voxel=0.7, seed=3)
    ph = w.phantom
    xb, yb, zb = ph.padded_bounds()
    fld = workloads.make_field(False, 100.0)
    rng = np.random.default_rng(11)
    for _ in range(500):
        ix, iy, iz = rng.integers(ph.nx - 1), rng.integers(ph.ny), rng.integers(ph.nz)
        f = rng.uniform(0.1, 0.9, 3)
        pos = np.array([ph.xb[ix] + f[0] * 0.7, ph.yb[iy] + f[1] * 0.7, ph.zb[iz] + f[2] * 0.7 + 100.0],
                       dtype=np.float32)
        region = 6 + ix + iy * ph.nx + iz * ph.nx * ph.ny
        t = L.oracle_hownear_phantom(ph.nx, ph.ny, ph.nz, xb.ctypes.data, yb.ctypes.data, zb.ctypes.data,
                                     ctypes.addressof(fld), pos.ctypes.data, region)
        # neighbour voxel ix+1: distance to its lower face is x - xb[ix+1] < 0
        expect = min(pos[0] - ph.xb[ix + 1], ph.xb[ix + 2] - pos[0])
        assert t <= 0.0
        assert t <= expect + 1e-6


def test_hownear_with_q1_fix_measures_the_own_voxel():
    """With GEGS_FIX_Q1 hownear decodes region-6, the particle's own voxel:
    the distance is the smallest of the six distances to its faces (> 0 inside)."""
    L = oracle.lib()
    w = workloads.ct(nx=8, ny=6, nz=5, voxel=0.7, seed=3)
    ph = w.phantom
    xb, yb, zb = ph.padded_bounds()
    fld = workloads.make_field(False, 100.0)
    rng = np.random.default_rng(12)
    oracle.set_fixes(1)
    try:
        for _ in range(500):
            ix, iy, iz = rng.integers(ph.nx), rng.integers(ph.ny), rng.integers(ph.nz)
            f = rng.uniform(0.05, 0.95, 3)
            pos = np.array([ph.xb[ix] + f[0] * 0.7, ph.yb[iy] + f[1] * 0.7, ph.zb[iz] + f[2] * 0.7 + 100.0],
                           dtype=np.float32)
            region = 6 + ix + iy * ph.nx + iz * ph.nx * ph.ny
            t = L.oracle_hownear_phantom(ph.nx, ph.ny, ph.nz, xb.ctypes.data, yb.ctypes.data, zb.ctypes.data,
                                         ctypes.addressof(fld), pos.ctypes.data, region)
            q = pos.astype(np.float64) - np.array([0.0, 0.0, 100.0])
            expect = min(q[0] - ph.xb[ix], ph.xb[ix + 1] - q[0], q[1] - ph.yb[iy], ph.yb[iy + 1] - q[1],
                         q[2] - ph.zb[iz], ph.zb[iz + 1] - q[2])
            assert t > 0.0
            assert abs(t - expect) < 2e-5
    finally:
        oracle.set_fixes(0)


@pytest.mark.parametrize("change_xyz", [0, 1])
def test_phantom_frame_fp64_option_is_correctly_rounded(change_xyz):
    """GEGS_FIX_FP64_FRAME: the transform from the unit-normalised cos/sin
    pair is its binary64 evaluation (recomputed here with numpy in the same
    operation order) rounded once to binary32 per coordinate, where the
    binary32 path of the reference is within two ULP."""
    from gegs_amd import native, workloads
    rng = np.random.default_rng(12)
    oracle.set_fixes(native.FIX_FP64_FRAME)
    try:
        for gantry in (0.0, 33.0, 90.0, 217.5):
            f = workloads.make_field(change_xyz, 100.0, gantry, iso=(1.25, -0.5, 2.0))
            pd = np.zeros((500, 6), dtype=np.float32)
            pd[:, :3] = rng.uniform(-30, 130, (500, 3))
            d = rng.normal(size=(500, 3))
            pd[:, 3:] = d / np.linalg.norm(d, axis=1)[:, None]
            got = oracle.to_phantom(f, pd)
            c, s = np.float64(np.float32(f.cos_angle)), np.float64(np.float32(f.sin_angle))
            r = np.sqrt(c * c + s * s)
            c, s = c / r, s / r
            p = pd.astype(np.float64)
            sid = np.float64(np.float32(f.source_to_isocenter))
            if change_xyz:
                x, y, z, u, v, w = p[:, 1], p[:, 2] - sid, p[:, 0], p[:, 4], p[:, 5], p[:, 3]
            else:
                x, y, z, u, v, w = p[:, 0], p[:, 1], p[:, 2] - sid, p[:, 3], p[:, 4], p[:, 5]
            iso = [np.float64(np.float32(t)) for t in (f.iso_x, f.iso_y, f.iso_z)]
            exact = np.stack([x * c - y * s + iso[0], x * s + y * c + iso[1], z + iso[2],
                              u * c - v * s, u * s + v * c, w], axis=1)
            assert np.array_equal(got, exact.astype(np.float32)), gantry
    finally:
        oracle.set_fixes(0)


@pytest.mark.parametrize("change_xyz", [0, 1])
def test_phantom_frame_within_2ulp_of_fp64(change_xyz):
    """The head -> phantom transform (geom.cu:59-109) runs in binary32 as in
    the reference, so that voxel indexing matches it bit for bit; against the
    same transform in float64 (with the same binary32 cos/sin of the gantry
    angle) every coordinate is within 2 ULP of the terms' scale."""
    from gegs_amd import workloads
    rng = np.random.default_rng(11)
    for gantry in (0.0, 33.0, 90.0, 217.5):
        f = workloads.make_field(change_xyz, 100.0, gantry, iso=(1.25, -0.5, 2.0))
        pd = np.zeros((500, 6), dtype=np.float32)
        pd[:, :3] = rng.uniform(-30, 130, (500, 3))
        d = rng.normal(size=(500, 3))
        pd[:, 3:] = d / np.linalg.norm(d, axis=1)[:, None]
        got = oracle.to_phantom(f, pd).astype(np.float64)
        c, s = np.float64(np.float32(f.cos_angle)), np.float64(np.float32(f.sin_angle))
        p = pd.astype(np.float64)
        if change_xyz:
            x, y, z = p[:, 1], p[:, 2] - np.float64(np.float32(f.source_to_isocenter)), p[:, 0]
            u, v, w = p[:, 4], p[:, 5], p[:, 3]
        else:
            x, y, z = p[:, 0], p[:, 1], p[:, 2] - np.float64(np.float32(f.source_to_isocenter))
            u, v, w = p[:, 3], p[:, 4], p[:, 5]
        iso = [np.float64(np.float32(f.iso_x)), np.float64(np.float32(f.iso_y)), np.float64(np.float32(f.iso_z))]
        exact = np.stack([x * c - y * s + iso[0], x * s + y * c + iso[1], z + iso[2],
                          u * c - v * s, u * s + v * c, w], axis=1)
        scale = np.stack([np.abs(x * c) + np.abs(y * s) + abs(iso[0]), np.abs(x * s) + np.abs(y * c) + abs(iso[1]),
                          np.abs(z) + abs(iso[2]), np.ones_like(u), np.ones_like(u), np.ones_like(u)], axis=1)
        ulp = np.spacing(scale.astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(got - exact) <= 2 * ulp), (gantry, np.max(np.abs(got - exact) / ulp))
    # a quarter turn maps the frame's x onto y
    f = workloads.make_field(0, 100.0, 90.0)
    q = oracle.to_phantom(f, np.array([[3.0, 0.0, 100.0, 1.0, 0.0, 0.0]], dtype=np.float32))[0]
    assert abs(q[0]) < 1e-5 and abs(q[1] - 3.0) < 1e-5 and abs(q[4] - 1.0) < 1e-6
