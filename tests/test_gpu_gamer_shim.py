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
