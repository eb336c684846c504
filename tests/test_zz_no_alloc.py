"""GPU: hydro_mhm_solve_batch_noalloc, the persistent batched solve whose work-queue counter lives in the library's
own device array instead of a caller allocation.  Through the C ABI it must give the same bits as
hydro_mhm_solve_batch_async with a caller-allocated counter, on two calls in a row (the second smaller, so a
stale counter or stale outputs would show), on the first and last queue slots, and reject slots outside the
array before any device work."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _args(hydro, u, pg_dt_dh, out, flux, prec):
    return hydro.make_args(u.shape[0], 1.0, 1.0, precision=prec, in_=u.data_ptr(), out=out.data_ptr(),
                           flux=flux.data_ptr())


def _solve(L, hydro, b, n, prec, noalloc, slot=0, g=2, k=1):
    """Run the first n patch groups of the batch through one of the two entry points; returns (out, flux)."""
    u = b.u[:n].contiguous()
    dtdh = b.pg_dt_dh[:n].contiguous()
    out = torch.full((n, 5, 16, 16, 16), float("nan"), dtype=u.dtype, device=DEV)
    flux = torch.full((n, 9, 5, 16, 16), float("nan"), dtype=u.dtype, device=DEV)
    a = _args(hydro, u, dtdh, out, flux, prec)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = L.hydro_mhm_solve_batch_noalloc(ctypes.byref(a), dtdh.data_ptr(), slot, g, k, s)
    else:
        cnt = torch.zeros(1, dtype=torch.int32, device=DEV)
        rc = L.hydro_mhm_solve_batch_async(ctypes.byref(a), dtdh.data_ptr(), cnt.data_ptr(), g, k, s)
    assert rc == 0, L.hydro_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy(), flux.cpu().numpy()


@pytest.mark.parametrize("dtype,prec", [(torch.float64, "fp64"), (torch.float32, "fp32")])
@pytest.mark.parametrize("slot", [0, 63])
def test_noalloc_matches_batch_entry_point(dtype, prec, slot):
    from gamer_amd import amr, hydro
    print(f"[gpu] batch_noalloc {prec} slot={slot}: 48 then 20 patch groups", flush=True)
    L = hydro.lib()
    b = amr.make_level_batch(torch, counts=(8, 16, 24), seed=91, dtype=dtype, scheme_ghost=3, interleave=True)
    for n, (g, k) in ((48, (2, 1)), (20, (1, 2))):            # two calls in a row, the second smaller
        o1, f1 = _solve(L, hydro, b, n, prec, noalloc=True, slot=slot, g=g, k=k)
        o2, f2 = _solve(L, hydro, b, n, prec, noalloc=False, g=g, k=k)
        assert np.isfinite(o1).all() and np.isfinite(f1).all()
        assert np.array_equal(o1.view(np.uint8), o2.view(np.uint8))
        assert np.array_equal(f1.view(np.uint8), f2.view(np.uint8))


def test_noalloc_rejects_slots_outside_the_array():
    from gamer_amd import hydro
    L = hydro.lib()
    a = hydro.make_args(4, 1e-3, 1e-2, in_=16, out=16, flux=16)
    for slot in (-1, 64):
        assert L.hydro_mhm_solve_batch_noalloc(ctypes.byref(a), None, slot, 2, 1, None) == hydro.HYDRO_E_ARG
        assert b"queue_slot" in L.hydro_last_error()
