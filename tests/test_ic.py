"""The seeded workload generator: deterministic transcendental helpers and patch-group cutting."""
import numpy as np

from gamer_amd import ic


def test_det_exp_sin_accuracy():
    x = np.linspace(-6, 6, 20001)
    assert np.max(np.abs(ic.det_exp(x) / np.exp(x) - 1)) < 4e-16
    assert np.max(np.abs(ic.det_sin(3 * x) - np.sin(3 * x))) < 2e-15


def test_cut_and_assemble_roundtrip():
    cons, dt, dh = ic.random_sod_grid((2, 1, 3), seed=5)
    pg = ic.cut_patch_groups(cons, 3)
    assert pg.shape == (6, 5, 22, 22, 22)
    np.testing.assert_array_equal(ic.assemble(ic.pg_interior(pg, 3), (2, 1, 3)), cons)
    # ghost zones hold the periodic neighbours: PG 0's low-x ghost = PG 2's high-x interior
    np.testing.assert_array_equal(pg[0, :, 3:19, 3:19, 0:3], pg[2, :, 3:19, 3:19, 16:19])


def test_workload_shapes_and_physics():
    w = ic.random_sod((1, 2, 2), seed=1, sedov=True)
    u = w.u
    assert u.shape == (4, 5, 22, 22, 22) and w.dt > 0 and w.dh == 1 / 32
    p = (ic.GAMMA - 1) * (u[:, 4] - 0.5 * (u[:, 1] ** 2 + u[:, 2] ** 2 + u[:, 3] ** 2) / u[:, 0])
    assert np.all(u[:, 0] > 0) and np.all(p > 0)
    kh = ic.kelvin_helmholtz((1, 2, 2), seed=2)
    assert set(np.unique(kh.u[:, 0])) == {1.0, 2.0}
