"""Dose post-processing (gegs_amd/postproc.py) against closed forms: linear
interpolation reproduces a dose that is linear in the boundary coordinates
it is placed at, points outside the interpolation range get zero, the
output-factor table is interpolated linearly between the bracketing sizes,
and the nearest-voxel choice keeps the reference's z midpoint (quirk Q9)."""
import os

import numpy as np
import pytest

from gegs_amd import postproc


def _linear_phantom():
    xb = np.linspace(-2.0, 2.0, 9).astype(np.float32)
    yb = np.linspace(-1.5, 1.5, 7).astype(np.float32)
    zb = np.linspace(0.0, 5.0, 11).astype(np.float32)
    c0, cx, cy, cz = 3.0, 0.5, -0.25, 0.75
    Z, Y, X = np.meshgrid(zb[:-1], yb[:-1], xb[:-1], indexing="ij")
    dose = (c0 + cx * X + cy * Y + cz * Z).astype(np.float32)
    return dose, xb, yb, zb, (c0, cx, cy, cz)


def test_output_factor_table():
    sizes = np.array([4.0, 6.0, 10.0, 20.0], dtype=np.float32)
    factors = np.array([0.9, 0.95, 1.0, 1.05], dtype=np.float32)
    assert postproc.field_output_factor(-4, 4, -4, 4, sizes, factors) == pytest.approx(0.975, abs=1e-6)
    assert postproc.field_output_factor(-5, 5, -5, 5, sizes, factors) == pytest.approx(1.0, abs=1e-6)
    # sqrt(2 x 8) = 4: the lower size is exclusive, so no bracket
    assert postproc.field_output_factor(-1, 1, -4, 4, sizes, factors) is None
    assert postproc.field_output_factor(-20, 20, -20, 20, sizes, factors) is None


def test_output_factor_file(tmp_path):
    p = str(tmp_path / "of.txt")
    with open(p, "w") as f:
        f.write("field size, output factor\n3\n4 0.9\n10 1.0\n20 1.05\n")
    s, o = postproc.read_output_factor_table(p)
    assert list(s) == [4, 10, 20] and list(o) == [np.float32(0.9), 1.0, np.float32(1.05)]


def test_plan_dose_weights_fields():
    a = np.arange(6.0)
    b = np.ones(6)
    assert np.array_equal(postproc.plan_dose([a, b]), a + b)
    assert np.allclose(postproc.plan_dose([a, b], [np.float32(0.5), np.float32(2.0)]), 0.5 * a + 2.0 * b)
    assert postproc.histories_for_field(250, 1.0e6 / 3) == 83333333


def test_interp_reproduces_linear_dose_and_zeroes_outside(tmp_path):
    dose, xb, yb, zb, (c0, cx, cy, cz) = _linear_phantom()
    start, sizes, dimen = (-2.4, -1.6, -0.3), (0.2, 0.2, 0.25), (25, 17, 23)
    g = postproc.interp_dose(dose, xb, yb, zb, start, sizes, dimen)
    assert g.shape == (23, 17, 25)
    for iz in range(dimen[2]):
        z = np.float32(np.float32(start[2]) + np.float32(sizes[2]) * np.float32(iz))
        for iy in range(dimen[1]):
            y = np.float32(np.float32(start[1]) + np.float32(sizes[1]) * np.float32(iy))
            for ix in range(dimen[0]):
                x = np.float32(np.float32(start[0]) + np.float32(sizes[0]) * np.float32(ix))
                inside = xb[0] < x < xb[-2] and yb[0] < y < yb[-2] and zb[0] < z < zb[-2]
                if inside:
                    assert g[iz, iy, ix] == pytest.approx(c0 + cx * x + cy * y + cz * z, abs=2e-5)
                else:
                    assert g[iz, iy, ix] == 0.0
    fname = postproc.write_interp_dose(str(tmp_path / "plan"), g)
    assert os.path.basename(fname) == "plan_Interpolate_025x017x023.bin"
    raw = np.fromfile(fname, dtype=np.int32, count=3)
    assert list(raw) == [25, 17, 23]
    back = np.fromfile(fname, dtype=np.float32)[3:].reshape(g.shape)
    assert np.array_equal(back, g)


def test_abs_dose_trilinear_and_nearest_voxel():
    dose, xb, yb, zb, (c0, cx, cy, cz) = _linear_phantom()
    for p in ((0.3, -0.2, 2.6), (-1.7, 1.1, 0.4), (1.49, 0.05, 3.9)):
        val, (i, j, k), vox = postproc.abs_dose(dose, xb, yb, zb, p)
        assert val == pytest.approx(c0 + cx * p[0] + cy * p[1] + cz * p[2], abs=2e-5)
        assert vox == dose[k, j, i]
    # z = 1.6 is in z voxel 3 ([1.5, 2.0]), below its midpoint 1.75, so the nearest voxel is 3; the
    # reference compares with (zb[3] + xb[4]) / 2 = (1.5 + 0.0) / 2 = 0.75 and takes voxel 4 (Q9)
    _, (i, j, k), vox = postproc.abs_dose(dose, xb, yb, zb, (0.1, 0.1, 1.6))
    assert k == 4 and vox == dose[4, j, i]
    _, (i, j, k), _ = postproc.abs_dose(dose, xb, yb, zb, (0.1, 0.1, 0.6))
    # z = 0.6 in voxel 1 ([0.5, 1.0]); quirk midpoint (0.5 + xb[2] = -1.0) / 2 = -0.25 < 0.6: voxel 2
    assert k == 2


def test_gamma_pass_rate_closed_forms():
    rng = np.random.default_rng(0)
    ref = rng.uniform(50.0, 100.0, (6, 7, 8))
    spacing = (0.5, 0.5, 0.5)
    assert postproc.gamma_pass_rate(ref, ref, spacing) == 1.0
    # a uniform 0.5% of Dmax shift passes, a 2% shift fails everywhere (neighbours are 5 DTA away)
    dmax = ref.max()
    assert postproc.gamma_pass_rate(ref, ref + 0.005 * dmax, spacing) == 1.0
    assert postproc.gamma_pass_rate(ref, ref + 0.02 * dmax, spacing) == 0.0
    # on a fine grid the distance criterion lets a shifted gradient pass: a ramp of 1% of Dmax per
    # 0.5 mm, shifted by one 0.5 mm voxel, is 1% off in dose but 0.5 mm away
    z = np.arange(40) * 0.05
    ramp = (1.0 + 0.2 * z)[:, None, None] * np.ones((40, 3, 3))
    shifted = np.concatenate([ramp[1:], ramp[-1:]], axis=0)
    fine = (0.05, 0.05, 0.05)
    # every voxel but the first, whose matching neighbour lies outside the grid
    assert postproc.gamma_pass_rate(ramp, shifted, fine, dose_crit=0.005) == 39 / 40
    assert postproc.gamma_pass_rate(ramp, shifted, fine, dose_crit=0.005, dta=0.01) < 0.5


def test_abs_dose_q9_fix_takes_the_z_midpoint():
    dose, xb, yb, zb, _ = _linear_phantom()
    # z = 1.6 in voxel 3 ([1.5, 2.0]), below its midpoint 1.75: the voxel itself with the fix
    _, (i, j, k), vox = postproc.abs_dose(dose, xb, yb, zb, (0.1, 0.1, 1.6), fix_q9=True)
    assert k == 3 and vox == dose[3, j, i]
    _, (i, j, k), _ = postproc.abs_dose(dose, xb, yb, zb, (0.1, 0.1, 1.9), fix_q9=True)
    assert k == 4


def test_abs_dose_q9_past_the_x_bounds_is_a_clear_error():
    # a 4 x 4 x 10 phantom: the point's z cell (8) needs x_bounds[9] on the
    # reference's path (output.c:835), which has only 5 x bounds
    xb = np.linspace(-1.0, 1.0, 5).astype(np.float32)
    yb = np.linspace(-1.0, 1.0, 5).astype(np.float32)
    zb = np.linspace(0.0, 5.0, 11).astype(np.float32)
    dose = np.ones((10, 4, 4), dtype=np.float32)
    with pytest.raises(ValueError, match="Q9"):
        postproc.abs_dose(dose, xb, yb, zb, (0.1, 0.1, 4.2))
    val, (i, j, k), _ = postproc.abs_dose(dose, xb, yb, zb, (0.1, 0.1, 4.2), fix_q9=True)
    assert k == 8 and abs(val - 1.0) < 1e-6
