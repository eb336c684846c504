"""GPU: the host-array drop-in for CUAPI_Asyn_FluidSolver (hydro_asyn_fluid_solver through hydro.GpuFluidApi) —
pinned double buffers [2], the NPG / GPU_NStream split with the remainder on the last stream
(src/GPU_API/CUAPI_Asyn_FluidSolver.cu:433-445), GAMER's prepare / solve / close overlap of the two buffers — against
the oracle on the same inputs."""
import numpy as np
import pytest

from gamer_amd import ic
from parity import rel_err_cells, rel_l1, flux_rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
GHOST = {"MHM_RP": 3, "MHM": 2}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("scheme", ["MHM_RP", "MHM"])
def test_host_array_drop_in_matches_oracle(scheme, precision):
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] host-array drop-in {scheme} {precision}", flush=True)
    w = ic.random_sod((4, 4, 4), seed=77, sedov=True, scheme_ghost=GHOST[scheme])
    dt_np = np.float64 if precision == "fp64" else np.float32
    npgs = (61, 37)                                   # not multiples of the 3 streams; two different batches
    inputs = (w.u[:npgs[0]], w.u[64 - npgs[1]:])
    api = hydro.GpuFluidApi(64, gpu_nstream=3, scheme=scheme, precision=precision)

    def fill(aid):
        h_in, h_out, _ = api.host_arrays(aid)
        h_in[:npgs[aid]] = inputs[aid].astype(dt_np)
        h_out[...] = np.nan

    try:
        # GAMER's double-buffered loop: solve buffer 0 while buffer 1 is prepared, synchronise, solve buffer 1
        # while buffer 0 is read back
        fill(0)
        api.asyn_fluid_solver(0, npgs[0], w.dt, w.dh, rsolver="HLLC")
        fill(1)
        api.synchronize()
        api.asyn_fluid_solver(1, npgs[1], w.dt, w.dh, rsolver="HLLC")
        results = [tuple(np.array(a[:npgs[0]], dtype=np.float64) for a in api.host_arrays(0)[1:])]
        api.synchronize()
        results.append(tuple(np.array(a[:npgs[1]], dtype=np.float64) for a in api.host_arrays(1)[1:]))
        for aid, u in enumerate(inputs):
            o, f = results[aid]
            ro, rf = oracle.fluid_solver(u, w.dt, w.dh, scheme=scheme, rsolver="HLLC")
            assert np.isfinite(o).all()
            if precision == "fp64":
                assert np.all(rel_err_cells(o, ro) <= 1e-12)
                assert flux_rel_err(f, rf, ro) <= 1e-12
            else:
                assert np.all(rel_l1(o, ro) <= 1e-5)
    finally:
        api.free()
