"""Input/output surface: egsphsp MODE0, egsphant, ctcreate ramp, region table,
dose .bin, and the exported C ABI of the built library."""
import ctypes
import os
import re

import numpy as np
import pytest

from gegs_amd import hatch, phantom as ph, source as src, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_egsphsp_roundtrip_and_decode(tmp_path):
    n = 100
    rng = np.random.default_rng(0)
    rec = np.zeros(n, dtype=src.PHSP_DTYPE)
    q = rng.choice([-1, 0, 1], n)
    rec["latch"] = np.where(q == -1, 0x40000000, np.where(q == 1, 0x20000000, 0)).astype(np.uint32)
    rec["e"] = rng.uniform(0.1, 6, n)
    rec["x"] = rng.uniform(-5, 5, n)
    rec["y"] = rng.uniform(-5, 5, n)
    u = rng.uniform(-0.3, 0.3, n)
    v = rng.uniform(-0.3, 0.3, n)
    rec["u"], rec["v"] = u, v
    sign = rng.choice([-1.0, 1.0], n)
    rec["wt"] = sign * rng.uniform(0.5, 1.5, n)
    path = tmp_path / "t.egsphsp1"
    src.write_egsphsp(str(path), rec)
    assert os.path.getsize(path) == 28 + 28 * n
    hdr, back = src.read_egsphsp(str(path))
    assert hdr["npphsp"] == n and hdr["nphotphsp"] == int(np.sum(q == 0))
    assert np.array_equal(back, rec)
    s = src.decode_phsp(back, 80.0)
    assert np.array_equal(s["charge"], q)
    assert np.all(np.sign(s["w"]) == sign) and np.all(s["wt"] > 0)
    assert np.allclose(s["u"] ** 2 + s["v"] ** 2 + s["w"] ** 2, 1.0, atol=1e-6)
    assert np.all(s["z"] == np.float32(80.0))
    with pytest.raises(ValueError, match="2019"):
        src.read_egsphsp(str(path), start=90, count=20)


def test_ct_ramp_values():
    """Default ctcreate ramp (ctcreate.c:904-1104)."""
    hu = np.array([-1100, -1024, -974, -849, -724, 0, 101, 1976, 3000])
    med, dens = ph.ct_to_medium(hu)
    assert list(med) == [0, 1, 1, 2, 2, 3, 3, 4, 4]
    assert dens[0] == 0.0                         # below the ramp: vacuum (Q10)
    assert abs(dens[1] - 0.001) < 1e-7
    assert abs(dens[2] - 0.044) < 1e-6
    assert abs(dens[3] - (0.044 + 125 * (0.302 - 0.044) / 250)) < 1e-6
    assert abs(dens[5] - (0.302 + 724 * (1.101 - 0.302) / 825)) < 1e-6
    assert abs(dens[8] - 2.088) < 1e-6            # clamped to the top of the ramp


def test_egsphant_roundtrip(tmp_path):
    p = ph.ct_phantom(6, 5, 4, 1.0, seed=1)
    path = tmp_path / "x.egsphant"
    ph.write_egsphant(str(path), p)
    q = ph.read_egsphant(str(path))
    assert q.media_names == p.media_names
    assert np.array_equal(q.media_idx, p.media_idx)
    assert np.allclose(q.density, p.density, atol=1e-6)
    assert np.allclose(q.xb, p.xb) and np.allclose(q.zb, p.zb)


def test_region_table_layout():
    w = workloads.ct(nx=8, ny=6, nz=5, voxel=0.7, seed=3)
    reg = w.regions
    assert len(reg) == 6 + 8 * 6 * 5
    assert reg[0]["med"] == 0xFFFF
    vox = reg[6:]
    mi = w.phantom.media_idx
    assert np.all((vox["med"] == 0xFFFF) == (mi == 0))
    sel = mi > 0
    rho_med = np.array([w.tables.media[m].rho for m in vox["med"][sel]])
    assert np.allclose(vox["rhof"][sel], w.phantom.density[sel] / rho_med, rtol=1e-6)
    assert np.all(vox["ecut"][sel] == np.float32(0.7)) and np.all(vox["pcut"][sel] == np.float32(0.01))


def test_dose_bin_writer(tmp_path):
    w = workloads.water(n=4, voxel=1.0)
    e = np.arange(64, dtype=np.float64)
    out, fname = ph.dose_to_gy_bin(e, w.phantom, str(tmp_path / "run"))
    assert fname.endswith("_DoseResults_004x004x004.bin")
    back = np.fromfile(fname, dtype=np.float32)
    assert np.allclose(back, e.astype(np.float32) / 1.0 * np.float32(1.6e-8))


def _declared_functions():
    """Every function declared in include/*.h (gegs.h, gegs_host.h)."""
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(gegs_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = os.path.join(ROOT, "gegs_amd", "libgegs_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libgegs_hip.so not built")
    L = ctypes.CDLL(lib)
    names = _declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    L.gegs_sizeof_media_tables.restype = ctypes.c_size_t
    assert L.gegs_sizeof_media_tables() == ctypes.sizeof(hatch.CMediaTables)


def test_validate_regions_rejects_unknown_flags():
    """A region asking for physics the kernels do not act on is an error, not a
    silent fallback; the reference's default flags (Rayleigh, relaxation,
    photo-electron angles, bound Compton; init.cu:935-1000) and the bound
    Compton variants pass."""
    lib = os.path.join(ROOT, "gegs_amd", "libgegs_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libgegs_hip.so not built")
    L = ctypes.CDLL(lib)
    L.gegs_validate_regions.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    reg = np.zeros(10, dtype=ph.REGION_DTYPE)
    reg["med"] = 0
    reg["flags"] = ph.DEFAULT_FLAGS
    assert L.gegs_validate_regions(reg.ctypes.data, len(reg)) == 0
    supported = 0x0001 | 0x000E | 0x0010 | 0x0020 | 0x0040
    for bit in range(16):
        r = reg.copy()
        r["flags"][3] = 1 << bit
        want = 0 if (1 << bit) & supported else 2001
        assert L.gegs_validate_regions(r.ctypes.data, len(r)) == want, hex(1 << bit)
    for f in (0x0033, 0x0037, 0x003B, 0x0002):
        r = reg.copy()
        r["flags"][:] = f
        assert L.gegs_validate_regions(r.ctypes.data, len(r)) == 0
    r = reg.copy()
    r["med"][4] = 6
    assert L.gegs_validate_regions(r.ctypes.data, len(r)) == 8001
    r["med"][4] = ph.VACUUM
    assert L.gegs_validate_regions(r.ctypes.data, len(r)) == 0
