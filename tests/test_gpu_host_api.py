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


def test_allocation_beyond_hbm_returns_nomem_and_library_stays_usable():
    """An allocation larger than the device memory returns HYDRO_E_NOMEM (GAMER's CUDA_CHECK_MALLOC -> GAMER_FAILED,
    src/GPU_API/CUAPI_MemAllocate.cu:41-96, on which CUAPI_MemAllocate halves GPU_NSTREAM and retries), frees what
    it had allocated, and leaves the library usable: a normal allocation and solve afterwards match the oracle."""
    import oracle
    from gamer_amd import hydro
    L = hydro.lib()
    assert L.hydro_init(0) == 0
    # 10^6 MHM_RP FP64 patch groups: the input buffer alone is 426 GB, more than the 288 GB of HBM
    print("[gpu] allocation beyond HBM -> HYDRO_E_NOMEM", flush=True)
    rc = L.hydro_mem_allocate_fluid(1_000_000, 4, hydro.SCHEME["MHM_RP"], hydro.PRECISION["fp64"])
    assert rc == hydro.HYDRO_E_NOMEM, (rc, L.hydro_last_error())
    assert b"failed" in L.hydro_last_error()
    args = hydro.make_args(1, 1e-3, 1e-2)
    assert L.hydro_asyn_fluid_solver(None, None, None, args, 1) == hydro.HYDRO_E_STATE
    w = ic.random_sod((1, 2, 2), seed=3, sedov=True)
    api = hydro.GpuFluidApi(4, gpu_nstream=2)
    try:
        h_in, h_out, h_flux = api.host_arrays(0)
        h_in[...] = w.u
        api.asyn_fluid_solver(0, 4, w.dt, w.dh)
        api.synchronize()
        ro, _ = oracle.fluid_solver(w.u, w.dt, w.dh)
        assert np.all(rel_err_cells(np.array(h_out), ro) <= 1e-12)
    finally:
        api.free()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_host_array_dt_solver_matches_oracle(precision):
    """The host-array CFL dt (hydro_asyn_dt_solver, CUAPI_Asyn_dtSolver's DT_FLU_SOLVER path) on the pinned dt buffers,
    8*npg patches split over 3 streams, against the oracle's dt (bit-exact to GAMER's CPU_dtSolver_HydroCFL)."""
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] host-array dt solver {precision}", flush=True)
    w = ic.random_sod((2, 2, 2), seed=12, sedov=True)
    out, _ = oracle.fluid_solver(w.u, w.dt, w.dh)
    npg = out.shape[0] - 1                                 # 56 patches over 3 streams: a remainder on the last
    p = out.reshape(-1, 5, 2, 8, 2, 8, 2, 8).transpose(0, 2, 4, 6, 1, 3, 5, 7).reshape(-1, 5, 8, 8, 8)[:8 * npg]
    ref = oracle.dt_cfl(np.ascontiguousarray(p), w.dh, 0.4)
    api = hydro.GpuFluidApi(8, gpu_nstream=3, precision=precision)
    try:
        h_flu, h_dt = api.host_dt_arrays(1)
        h_flu[:8 * npg] = p
        h_dt[...] = np.nan
        api.asyn_dt_solver(1, npg, w.dh, 0.4)
        api.synchronize()
        got = np.array(h_dt[:8 * npg], dtype=np.float64)
        assert np.isnan(h_dt[8 * npg:]).all()
        rel = np.abs(got - ref) / ref
        assert rel.max() <= (1e-14 if precision == "fp64" else 1e-6), rel.max()
    finally:
        api.free()
