"""GPU: the HLLE rescue of a NaN HLLC flux (RSOLVER_RESCUE; src/Model_Hydro/CPU_Hydro/CPU_FluidSolver_MHM.cpp:737-786
for the half step, CPU_Shared_ComputeFlux.cpp:313-362 for the full step).

A pressureless uniform flow triggers it on every face.  With P = 0 the sound speed is 0, so both HLLC wave-speed
factors temp1_L and temp1_R are 0, the star-state denominator 1/(temp1_L - temp1_R) is infinite and the star
velocity is inf * 0 = NaN.  HLLE then returns the upwind physical flux.  rho = 1, v = (1, 0, 0), E = 1/2 and
P = 0 are exact in FP32 and FP64 and so is every step of the rescued update.  The output equals the input
bitwise, the x-face fluxes are (1, 1, 0, 0, 1/2) and the y- and z-face fluxes vanish.  Without the rescue the
output is NaN.  A mixed batch (cold-flow patch groups between random Sod + Sedov ones) is then compared with the
oracle, which carries the same rescue.
"""
import numpy as np
import pytest

from gamer_amd import ic
from parity import rel_err_cells, flux_rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _cold_flow(npg, dtype):
    u = np.zeros((npg, 5, 22, 22, 22), dtype=dtype)
    u[:, 0] = 1.0
    u[:, 1] = 1.0
    u[:, 4] = 0.5
    return u


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("scheme", ["MHM_RP", "MHM"])
def test_cold_uniform_flow_is_rescued_exactly(prec, scheme):
    from gamer_amd import hydro
    print(f"[gpu] HLLE rescue, cold uniform flow, {scheme} {prec}", flush=True)
    dtype = np.float64 if prec == "fp64" else np.float32
    n = 22 if scheme == "MHM_RP" else 20
    g = (n - 16) // 2
    u = _cold_flow(3, dtype)[:, :, :n, :n, :n].copy()
    out, flux = hydro.fluid_solver(torch.from_numpy(u).to(DEV), 0.01, 1.0 / 64, scheme=scheme, rsolver="HLLC")
    torch.cuda.synchronize()
    o, f = out.cpu().numpy(), flux.cpu().numpy()
    assert not np.isnan(o).any() and not np.isnan(f).any()
    np.testing.assert_array_equal(o, u[:, :, g:g + 16, g:g + 16, g:g + 16])
    want_x = np.array([1.0, 1.0, 0.0, 0.0, 0.5], dtype=dtype)
    np.testing.assert_array_equal(f[:, 0:3], np.broadcast_to(want_x[None, None, :, None, None], f[:, 0:3].shape))
    np.testing.assert_array_equal(f[:, 3:9], 0.0)


def test_mixed_batch_matches_oracle():
    import oracle
    from gamer_amd import hydro
    print("[gpu] HLLE rescue, mixed batch vs oracle (fp64)", flush=True)
    w = ic.random_sod((2, 2, 1), seed=77, sedov=True)
    u = np.concatenate([w.u[:2], _cold_flow(2, np.float64), w.u[2:]], axis=0)
    out, flux = hydro.fluid_solver(torch.from_numpy(u).to(DEV), w.dt, w.dh, rsolver="HLLC")
    torch.cuda.synchronize()
    ro, rf = oracle.fluid_solver(u, w.dt, w.dh, rsolver="HLLC")
    o, f = out.cpu().numpy(), flux.cpu().numpy()
    assert not np.isnan(ro).any()
    assert np.all(rel_err_cells(o, ro) <= 1e-12)
    assert flux_rel_err(f, rf, ro) <= 1e-12
    np.testing.assert_array_equal(o[2:4], u[2:4, :, 3:19, 3:19, 3:19])
