"""GPU: GAMER's own fluid-solver GPU API, end to end.

tests/_build/libgamer_shim_<prec>.so holds the GAMER-side adapter (gamer_amd/gamer_shim/CUAPI_FluidSolver_gfx950.cpp)
compiled against GAMER's headers with a GPU build's flags (MHM_RP, HLLC, PLM; FLOAT8 for fp64) plus
tests/gamer_shim_harness.cpp, which calls CUAPI_SetDevice, CUAPI_MemAllocate_Fluid, CUAPI_Asyn_FluidSolver on the
pinned buffer 0 with GPU_NSTREAM streams, CUAPI_Synchronize and CUAPI_MemFree_Fluid in GAMER's order.  The
outputs are compared with the oracle: FP64 to 1e-12 per cell, FP32 to 1e-5 relative L1 against the FP64
oracle.  The harness is built where the reference headers are (`make -C tests shim`, run by the build step);
without it the tests skip.
"""
import ctypes
import os

import numpy as np
import pytest

from gamer_amd import ic
from parity import flux_rel_err, rel_err_cells, rel_l1

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _harness(prec):
    path = os.path.join(HERE, "_build", f"libgamer_shim_{prec}.so")
    if not os.path.isfile(path):
        pytest.skip("adapter harness not built (it needs the reference headers at build time)")
    from gamer_amd import hydro
    hydro.lib()
    L = ctypes.CDLL(path)
    L.shim_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                           ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                           ctypes.c_int]
    return L


def _run(L, u, dt, dh, nstream=3, limiter=4):
    npg = u.shape[0]
    out = np.empty((npg, 5, 16, 16, 16), dtype=u.dtype)
    flux = np.empty((npg, 9, 5, 16, 16), dtype=u.dtype)
    msg = ctypes.create_string_buffer(1024)
    rc = L.shim_run(0, npg, nstream, dt, dh, limiter, 1.5, u.ctypes.data, out.ctypes.data, flux.ctypes.data, msg, 1024)
    return rc, out, flux, msg.value.decode()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_gamer_cuapi_sequence_matches_oracle(prec):
    import oracle
    print(f"[gpu] GAMER CUAPI_* sequence through the adapter, {prec}", flush=True)
    L = _harness(prec)
    w = ic.random_sod((2, 2, 4), seed=31, sedov=True)
    u = np.ascontiguousarray(w.u if prec == "f64" else w.u.astype(np.float32))
    rc, o, f, msg = _run(L, u, w.dt, w.dh)
    assert rc == 0, msg
    ro, rf = oracle.fluid_solver(w.u, w.dt, w.dh)
    if prec == "f64":
        assert np.all(rel_err_cells(o, ro) <= 1e-12)
        assert flux_rel_err(f, rf, ro) <= 1e-12
    else:
        assert max(rel_l1(o.astype(np.float64), ro)) <= 1e-5


def test_gamer_cuapi_error_reaches_aux_error():
    """A call the library rejects (an unknown limiter) surfaces through Aux_Error with the library's message,
    as CUDA_CHECK_ERROR does in the CUDA build.  The library rejects it on the host before any launch."""
    print("[gpu] GAMER CUAPI_* error path", flush=True)
    L = _harness("f64")
    w = ic.random_sod((1, 1, 2), seed=5)
    rc, _, _, msg = _run(L, np.ascontiguousarray(w.u), w.dt, w.dh, limiter=9)
    assert rc == -1 and "gfx950 fluid solver" in msg, msg


def _bind_extra(L):
    L.shim_dt.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    L.shim_set_eos_gamma.argtypes = [ctypes.c_double, ctypes.c_char_p, ctypes.c_int]
    L.shim_set_mem_size.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p,
                                    ctypes.c_int]
    L.shim_diagnose.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    return L


def _patches(out):
    npg = out.shape[0]
    return np.ascontiguousarray(
        out.reshape(npg, 5, 2, 8, 2, 8, 2, 8).transpose(0, 2, 4, 6, 1, 3, 5, 7).reshape(npg * 8, 5, 8, 8, 8))


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("min_pres", [0.0, 0.3])
def test_gamer_dt_solver_matches_reference_dt(prec, min_pres):
    """CUAPI_MemAllocate -> CUAPI_Asyn_dtSolver(DT_FLU_SOLVER) -> CUAPI_Synchronize through the adapter, against the
    oracle's CFL dt (pinned bit for bit to GAMER's CPU_dtSolver_HydroCFL by tests/test_oracle_ref.py)."""
    import oracle
    print(f"[gpu] GAMER CUAPI_Asyn_dtSolver through the adapter, {prec}, min_pres={min_pres}", flush=True)
    L = _bind_extra(_harness(prec))
    w = ic.random_sod((2, 2, 3), seed=17, sedov=True)
    out, _ = oracle.fluid_solver(w.u, w.dt, w.dh)
    p64 = _patches(out)
    ref = oracle.dt_cfl(p64, w.dh, 0.5, min_pres=min_pres)
    p = p64 if prec == "f64" else p64.astype(np.float32)
    dt = np.empty(p.shape[0], dtype=p.dtype)
    msg = ctypes.create_string_buffer(1024)
    rc = L.shim_dt(0, out.shape[0], 3, w.dh, 0.5, min_pres, p.ctypes.data, dt.ctypes.data, msg, 1024)
    assert rc == 0, msg.value.decode()
    rel = np.abs(dt - ref) / ref
    print(f"  max rel err {rel.max():.3e}", flush=True)
    assert rel.max() <= (1e-14 if prec == "f64" else 1e-6)


def test_gamer_eos_gamma_reaches_the_solver():
    """EoS_Init's CUAPI_SetConstMemory_EoS hands the EOS_GAMMA adiabatic index to the solver calls: a gamma = 1.4 run
    through the adapter matches the oracle at gamma = 1.4 (and not at 5/3)."""
    import oracle
    print("[gpu] GAMER EoS gamma through the adapter", flush=True)
    L = _bind_extra(_harness("f64"))
    msg = ctypes.create_string_buffer(1024)
    assert L.shim_set_eos_gamma(1.4, msg, 1024) == 0, msg.value.decode()
    try:
        w = ic.random_sod((1, 2, 2), seed=23, sedov=True, gamma=1.4)
        rc, o, f, m = _run(L, np.ascontiguousarray(w.u), w.dt, w.dh)
        assert rc == 0, m
        ro, rf = oracle.fluid_solver(w.u, w.dt, w.dh, gamma=1.4)
        assert np.all(rel_err_cells(o, ro) <= 1e-12)
        r53, _ = oracle.fluid_solver(w.u, w.dt, w.dh)
        assert rel_err_cells(o, r53).max() > 1e-6
        assert L.shim_set_eos_gamma(1.0, msg, 1024) == -1 and "gamma" in msg.value.decode()
    finally:
        assert L.shim_set_eos_gamma(5.0 / 3.0, msg, 1024) == 0


def test_gamer_set_mem_size_and_diagnose(tmp_path):
    """CUAPI_SetMemSize's automatic sizes (GPU_NSTREAM 4, FLU_GPU_NPGROUP = 4 x compute units) and
    CUAPI_DiagnoseDevice's Record__Note entry."""
    from gamer_amd import hydro
    print("[gpu] GAMER CUAPI_SetMemSize / CUAPI_DiagnoseDevice through the adapter", flush=True)
    L = _bind_extra(_harness("f64"))
    info = hydro.device_info()
    print(f"  device {info['name']} {info['arch']} CUs={info['compute_units']}", flush=True)
    assert info["wavefront_size"] == 64 and info["compute_units"] > 0 and "gfx950" in info["arch"]
    ns, npg = ctypes.c_int(), ctypes.c_int()
    msg = ctypes.create_string_buffer(1024)
    assert L.shim_set_mem_size(ctypes.byref(ns), ctypes.byref(npg), msg, 1024) == 0, msg.value.decode()
    assert ns.value == 4 and npg.value == 4 * info["compute_units"]
    assert L.shim_diagnose(str(tmp_path).encode(), msg, 1024) == 0, msg.value.decode()
    note = (tmp_path / "Record__Note").read_text()
    assert "Device Diagnosis" in note and info["name"] in note and "Wavefront Size" in note


def test_gamer_mem_allocate_halves_streams_then_reports():
    """CUAPI_MemAllocate's out-of-memory path: FLU_GPU_NPGROUP = 10^6 patch groups (426 GB of input buffer, more than
    the HBM) fails with GPU_NSTREAM 4, 2 and 1 (halving after each failure, src/GPU_API/CUAPI_MemAllocate.cu:73-105)
    and then stops in Aux_Error; the library stays usable for the next sequence."""
    print("[gpu] GAMER CUAPI_MemAllocate out-of-memory path", flush=True)
    L = _bind_extra(_harness("f64"))
    msg = ctypes.create_string_buffer(1024)
    dummy = np.zeros(1)
    rc = L.shim_dt(0, 1_000_000, 4, 0.01, 0.5, 0.0, dummy.ctypes.data, dummy.ctypes.data, msg, 1024)
    assert rc == -1 and "out of memory even with GPU_NSTREAM = 1" in msg.value.decode(), msg.value.decode()
    w = ic.random_sod((1, 1, 2), seed=5)
    rc, o, f, m = _run(L, np.ascontiguousarray(w.u), w.dt, w.dh)
    assert rc == 0, m
