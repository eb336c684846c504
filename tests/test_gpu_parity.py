"""GPU parity: the gfx950 solver (through the C ABI) against the oracle restatement on identical seeded inputs.

FP64 bar: per-cell relative error <= 1e-12 (density/energy relative to themselves, momenta relative to
sqrt(rho E)), fluxes <= 1e-12 of the patch group's flux scale.  FP32-bulk bar: per-field relative L1 <= 1e-5
against the FP64 oracle.
"""
import numpy as np
import pytest

from gamer_amd import ic
from parity import rel_err_cells, rel_l1, flux_rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_L1_TOL = 1e-5


def _gpu(u, w, **kw):
    from gamer_amd import hydro
    dev = torch.device("cuda:0")
    ut = torch.from_numpy(np.ascontiguousarray(u)).to(dev)
    out, flux = hydro.fluid_solver(ut, w.dt, w.dh, **kw)
    torch.cuda.synchronize()
    return out.cpu().numpy(), (flux.cpu().numpy() if flux is not None else None)


def _oracle(u, w, **kw):
    import oracle
    return oracle.fluid_solver(u, w.dt, w.dh, **kw)


@pytest.fixture(scope="module")
def sod():
    return ic.random_sod((2, 2, 4), seed=2024, sedov=True)


@pytest.mark.parametrize("rsolver", ["HLLC", "HLLE", "ROE"])
def test_fp64_matches_oracle(sod, rsolver):
    print(f"[gpu] fp64 parity rsolver={rsolver} npg={sod.npg}", flush=True)
    o, f = _gpu(sod.u, sod, rsolver=rsolver)
    ro, rf = _oracle(sod.u, sod, rsolver=rsolver)
    err = rel_err_cells(o, ro)
    ferr = flux_rel_err(f, rf, ro)
    print(f"  per-field max rel err {err}, flux {ferr:.3e}")
    assert np.all(err <= FP64_TOL), err
    assert ferr <= FP64_TOL


@pytest.mark.parametrize("limiter", ["VANLEER", "GMINMOD", "ALBADA", "VL_GMINMOD", "CENTRAL"])
def test_fp64_limiters(limiter):
    w = ic.random_sod((2, 2, 2), seed=77, sedov=False)
    print(f"[gpu] fp64 limiter={limiter}", flush=True)
    o, f = _gpu(w.u, w, limiter=limiter)
    ro, rf = _oracle(w.u, w, limiter=limiter)
    err = rel_err_cells(o, ro)
    assert np.all(err <= FP64_TOL), err
    assert flux_rel_err(f, rf, ro) <= FP64_TOL


@pytest.mark.parametrize("rsolver", ["HLLC", "HLLE", "ROE"])
def test_fp32_bulk_vs_fp64_oracle(sod, rsolver):
    print(f"[gpu] fp32-bulk rsolver={rsolver}", flush=True)
    o, f = _gpu(sod.u.astype(np.float32), sod, rsolver=rsolver)
    ro, rf = _oracle(sod.u, sod, rsolver=rsolver)
    l1 = rel_l1(o, ro)
    print(f"  per-field rel L1 {l1}")
    assert np.all(l1 <= FP32_L1_TOL), l1


def test_uniform_state_is_fixed_point():
    """All slopes and flux differences vanish: output == input interior bit for bit."""
    n = 22
    u = np.empty((3, 5, n, n, n))
    u[:, 0], u[:, 1], u[:, 2], u[:, 3], u[:, 4] = 1.3, 0.2, -0.1, 0.05, 2.0
    w = ic.Workload(u, 1e-3, 1.0 / 64, (1, 1, 3), 3)
    for dt in (np.float64, np.float32):
        o, f = _gpu(u.astype(dt), w)
        assert np.array_equal(o, ic.pg_interior(u.astype(dt), 3)), dt


def test_conservation_and_determinism():
    """Periodic domain: domain sums of the conserved fields are unchanged to round-off, and two runs agree bitwise."""
    w = ic.random_sod((2, 2, 2), seed=5, sedov=True)
    o1, _ = _gpu(w.u, w)
    o2, _ = _gpu(w.u, w)
    assert np.array_equal(o1, o2)
    s_in = ic.pg_interior(w.u, 3).sum(axis=(0, 2, 3, 4))
    s_out = o1.sum(axis=(0, 2, 3, 4))
    tot = np.abs(ic.pg_interior(w.u, 3)).sum(axis=(0, 2, 3, 4))
    assert np.all(np.abs(s_out - s_in) <= 1e-13 * tot), (s_out - s_in) / tot


def test_empty_batch():
    from gamer_amd import hydro
    u = torch.empty((0, 5, 22, 22, 22), dtype=torch.float64, device="cuda:0")
    out, flux = hydro.fluid_solver(u, 1e-3, 1e-2)
    torch.cuda.synchronize()
    assert out.shape[0] == 0 and flux.shape[0] == 0
