"""GPU: CFL time step (standalone and fused into the solver), ghost-zone preparation / scatter-back for a periodic
uniform level, and multi-step on-device rollouts, all through the C ABI and checked against the oracle."""
import numpy as np
import pytest

from gamer_amd import ic
from parity import rel_err_cells

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _patches(out_pg):
    """(npg, 5, 16,16,16) -> (npg*8, 5, 8,8,8), patch index px + 2 py + 4 pz within each patch group."""
    npg = out_pg.shape[0]
    o = out_pg.reshape(npg, 5, 2, 8, 2, 8, 2, 8)              # pg, v, pz, k, py, j, px, i
    return o.transpose(0, 2, 4, 6, 1, 3, 5, 7).reshape(npg * 8, 5, 8, 8, 8)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_dt_cfl_matches_oracle(dtype):
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] dt_cfl {dtype}", flush=True)
    w = ic.random_sod((2, 2, 2), seed=19, sedov=True)
    patches = _patches(np.ascontiguousarray(ic.pg_interior(w.u, 3)))
    ref = oracle.dt_cfl(patches, w.dh, 0.5)
    got = hydro.dt_cfl(torch.from_numpy(patches).to(DEV, dtype=dtype), w.dh, 0.5)
    torch.cuda.synchronize()
    tol = 1e-14 if dtype == torch.float64 else 1e-6
    np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=tol)


@pytest.mark.parametrize("scheme,ghost", [("MHM_RP", 3), ("MHM", 2)])
def test_fused_dt_matches_cfl_of_output(scheme, ghost):
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] fused dt {scheme}", flush=True)
    w = ic.random_sod((2, 2, 2), seed=23, sedov=True, scheme_ghost=ghost)
    u = torch.from_numpy(w.u).to(DEV)
    dtp = torch.empty((w.npg, 8), dtype=torch.float64, device=DEV)
    out, _ = hydro.fluid_solver(u, w.dt, w.dh, scheme=scheme, dt_patch=dtp, safety=0.4)
    torch.cuda.synchronize()
    ref = oracle.dt_cfl(_patches(out.cpu().numpy()), w.dh, 0.4).reshape(w.npg, 8)
    np.testing.assert_allclose(dtp.cpu().numpy(), ref, rtol=1e-6)   # the patch maxima are reduced in FP32


@pytest.mark.parametrize("ghost", [3, 2])
def test_prepare_and_close_match_numpy(ghost):
    from gamer_amd import hydro
    print(f"[gpu] prepare/close ghost={ghost}", flush=True)
    cons, dt, dh = ic.random_sod_grid((2, 1, 3), seed=7)
    g = torch.from_numpy(cons).to(DEV)
    pg = hydro.prepare_uniform(g, ghost)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(pg.cpu().numpy(), ic.cut_patch_groups(cons, ghost))
    level = torch.zeros_like(g)
    hydro.close_uniform(pg[:, :, ghost:ghost + 16, ghost:ghost + 16, ghost:ghost + 16].contiguous(), level)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(level.cpu().numpy(), cons)


@pytest.mark.parametrize("scheme", ["MHM_RP", "MHM"])
def test_uniform_rollout_matches_oracle(scheme):
    """Three prepare -> solve -> close steps on the device equal the oracle with host-side ghost refills."""
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] uniform rollout {scheme}", flush=True)
    ghost = 3 if scheme == "MHM_RP" else 2
    cons, dt, dh = ic.random_sod_grid((2, 2, 2), seed=41, sedov=True)
    lvl = hydro.UniformLevel(torch.from_numpy(cons).to(DEV), dh, scheme=scheme, safety=0.5)
    ref = cons.copy()
    dts = []
    for _ in range(3):
        dts.append(dt)
        new_dt = float(lvl.step(dt))
        out, _ = oracle.fluid_solver(ic.cut_patch_groups(ref, ghost), dt, dh, scheme=scheme)
        ref = ic.assemble(out, (2, 2, 2))
        dt = new_dt
    torch.cuda.synchronize()
    got = lvl.level.cpu().numpy()
    err = rel_err_cells(ic.cut_patch_groups(got, 0), ic.cut_patch_groups(ref, 0))
    assert np.all(err <= 1e-12), err
    ref_dt = ic.cfl_dt(ref, dh, 0.5)
    assert dt == pytest.approx(ref_dt, rel=1e-6)
