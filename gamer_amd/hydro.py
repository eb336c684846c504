"""Host-side mirror of GAMER's fluid-solver GPU API over the gfx950 C-ABI library.

``fluid_solver`` is the device-array entry (one stream-ordered launch of ``hydro_mhm_solve_async`` on torch's
current stream); ``GpuFluidApi`` mirrors the host-array path of ``CUAPI_MemAllocate_Fluid`` /
``CUAPI_Asyn_FluidSolver`` / ``CUAPI_Synchronize`` (src/GPU_API/CUAPI_Asyn_FluidSolver.cu:266-749): pinned
double buffers, N streams, asynchronous H2D + kernel + D2H.  Argument names and meanings follow
``CUAPI_Asyn_FluidSolver`` (include/Prototype.h:646); failures raise ``HydroError`` with the library's message
(the reference aborts through Aux_Error instead).

The library must have been built (``python gamer_amd/build.py`` or ``__graft_entry__.build()``); there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GAMER_AMD_HYDRO_LIB: another build of the same library (the sanitizer test loads one whose host code is built
# with ASan/UBSan); there is no other library and no CPU fallback
LIB_PATH = os.environ.get("GAMER_AMD_HYDRO_LIB") or os.path.join(HERE, "libgamer_amd_hydro.so")

HYDRO_OK = 0
HYDRO_E_ARG, HYDRO_E_STATE, HYDRO_E_DEVICE, HYDRO_E_NOMEM = -1, -2, -3, -4
SCHEME = {"MHM": 3, "MHM_RP": 4}
RSOLVER = {"ROE": 2, "HLLE": 3, "HLLC": 4}
LIMITER = {"VANLEER": 1, "GMINMOD": 2, "ALBADA": 3, "VL_GMINMOD": 4, "CENTRAL": 6}
PRECISION = {"fp64": 0, "fp32": 1}
NX = {"MHM_RP": 22, "MHM": 20}

EXPORTS = ("hydro_init", "hydro_mhm_solve_async", "hydro_synchronize", "hydro_last_error",
           "hydro_mem_allocate_fluid", "hydro_mem_free_fluid", "hydro_host_buffers", "hydro_asyn_fluid_solver",
           "hydro_dt_cfl_async", "hydro_prepare_uniform_async", "hydro_close_uniform_async",
           "hydro_mhm_solve_batch_async", "hydro_mhm_adjoint_workspace_bytes", "hydro_mhm_adjoint_async",
           "hydro_mhm_tangent_async", "hydro_prepare_uniform_adjoint_async", "hydro_correct_unphysical_async",
           "hydro_store_flux_async", "hydro_correct_flux_async", "hydro_fixup_flux_async",
           "hydro_fixup_restrict_async", "hydro_close_patches_async", "hydro_interpolate_async",
           "hydro_prepare_patches_async", "hydro_gather_boxes_async", "hydro_assemble_patch_groups_async",
           "hydro_host_dt_buffers", "hydro_asyn_dt_solver", "hydro_device_info", "hydro_mhm_solve_batch_noalloc")


class HydroError(RuntimeError):
    pass


class HydroSolveArgs(ctypes.Structure):
    _fields_ = [
        ("in_", ctypes.c_void_p), ("out", ctypes.c_void_p), ("flux", ctypes.c_void_p),
        ("npg", ctypes.c_int), ("dt", ctypes.c_double), ("dh", ctypes.c_double), ("gamma", ctypes.c_double),
        ("scheme", ctypes.c_int), ("rsolver", ctypes.c_int), ("limiter", ctypes.c_int),
        ("minmod_coeff", ctypes.c_double), ("minmod_maxiter", ctypes.c_int),
        ("min_dens", ctypes.c_double), ("min_pres", ctypes.c_double), ("min_eint", ctypes.c_double),
        ("store_flux", ctypes.c_int), ("precision", ctypes.c_int),
        ("dt_patch", ctypes.c_void_p), ("safety", ctypes.c_double), ("waves_per_group", ctypes.c_int),
    ]


class HydroDeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64), ("device", ctypes.c_int),
        ("device_count", ctypes.c_int), ("compute_units", ctypes.c_int), ("clock_khz", ctypes.c_int),
        ("global_mem_bytes", ctypes.c_size_t), ("lds_per_block_bytes", ctypes.c_size_t),
        ("wavefront_size", ctypes.c_int), ("max_threads_per_block", ctypes.c_int), ("driver_version", ctypes.c_int),
        ("runtime_version", ctypes.c_int), ("ecc", ctypes.c_int), ("async_engines", ctypes.c_int),
        ("concurrent_kernels", ctypes.c_int),
    ]


class HydroCorrectArgs(ctypes.Structure):
    _fields_ = [
        ("in_", ctypes.c_void_p), ("out", ctypes.c_void_p), ("flux", ctypes.c_void_p),
        ("npg", ctypes.c_int), ("dt", ctypes.c_double), ("dh", ctypes.c_double), ("gamma", ctypes.c_double),
        ("scheme", ctypes.c_int), ("mode", ctypes.c_int), ("rsolver", ctypes.c_int),
        ("last_resort_floor", ctypes.c_int), ("auto_reduce_dt", ctypes.c_int), ("fixup_flux", ctypes.c_int),
        ("check_pres", ctypes.c_int), ("min_dens", ctypes.c_double), ("min_pres", ctypes.c_double),
        ("min_eint", ctypes.c_double), ("precision", ctypes.c_int), ("counters", ctypes.c_void_p),
        ("main_rsolver", ctypes.c_int),
    ]


FIRST_FLUX_CORR = {"NONE": 0, "3D": 1, "3D1D": 2}
INT_SCHEME = {"MINMOD3D": 1, "MINMOD1D": 2, "VANLEER": 3, "CQUAD": 4, "QUAD": 5, "CQUAR": 6, "QUAR": 7}


class HydroInterpArgs(ctypes.Structure):
    _fields_ = [
        ("cdata", ctypes.c_void_p), ("fdata", ctypes.c_void_p), ("nbox", ctypes.c_int),
        ("csize", ctypes.c_int * 3), ("cstart", ctypes.c_int * 3), ("crange", ctypes.c_int * 3),
        ("fsize", ctypes.c_int * 3), ("fstart", ctypes.c_int * 3), ("scheme", ctypes.c_int),
        ("monotonic", ctypes.c_int), ("mono_coeff", ctypes.c_double), ("opp_sign_0th", ctypes.c_int),
        ("int_prim", ctypes.c_int), ("reduce_mono", ctypes.c_int), ("mono_max_iter", ctypes.c_int),
        ("gamma", ctypes.c_double), ("min_pres", ctypes.c_double), ("precision", ctypes.c_int),
        ("nfail", ctypes.c_void_p), ("need", ctypes.c_void_p),
    ]

_lib = None


def lib():
    """Load the gfx950 solver library; raises if it was never built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HydroError(f"{LIB_PATH} not built: run `python gamer_amd/build.py`")
        # torch ships its own HIP runtime (soname libamdhip64.so.7, requested by torch as libamdhip64.so).  With
        # torch loaded first, this library binds to that copy; loaded the other way round, the process would map
        # two HIP runtimes and torch's device memory would be foreign to this library's launches.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.hydro_last_error.restype = ctypes.c_char_p
        L.hydro_init.argtypes = [ctypes.c_int]
        L.hydro_mhm_solve_async.argtypes = [ctypes.POINTER(HydroSolveArgs), ctypes.c_void_p]
        L.hydro_mem_allocate_fluid.argtypes = [ctypes.c_int] * 4
        L.hydro_host_buffers.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_void_p)] * 3
        L.hydro_asyn_fluid_solver.argtypes = [ctypes.c_void_p] * 3 + [ctypes.POINTER(HydroSolveArgs), ctypes.c_int]
        L.hydro_host_dt_buffers.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_void_p)] * 2
        L.hydro_asyn_dt_solver.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 4 + \
            [ctypes.c_int]
        L.hydro_device_info.argtypes = [ctypes.POINTER(HydroDeviceInfo)]
        L.hydro_dt_cfl_async.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_void_p]
        L.hydro_prepare_uniform_async.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 5 + [ctypes.c_void_p]
        L.hydro_close_uniform_async.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 4 + [ctypes.c_void_p]
        L.hydro_mhm_solve_batch_async.argtypes = [ctypes.POINTER(HydroSolveArgs), ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.hydro_mhm_solve_batch_noalloc.argtypes = [ctypes.POINTER(HydroSolveArgs), ctypes.c_void_p, ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.hydro_mhm_adjoint_workspace_bytes.argtypes = [ctypes.c_int]
        L.hydro_mhm_adjoint_workspace_bytes.restype = ctypes.c_size_t
        L.hydro_mhm_adjoint_async.argtypes = [ctypes.POINTER(HydroSolveArgs)] + [ctypes.c_void_p] * 4
        L.hydro_mhm_tangent_async.argtypes = [ctypes.POINTER(HydroSolveArgs)] + [ctypes.c_void_p] * 4
        L.hydro_prepare_uniform_adjoint_async.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 5 + [ctypes.c_void_p]
        L.hydro_correct_unphysical_async.argtypes = [ctypes.POINTER(HydroCorrectArgs), ctypes.c_void_p]
        for f in (L.hydro_store_flux_async, L.hydro_correct_flux_async):
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                          ctypes.c_int, ctypes.c_void_p]
        L.hydro_fixup_flux_async.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int] + \
            [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_void_p]
        L.hydro_fixup_restrict_async.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
        L.hydro_close_patches_async.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_int, ctypes.c_void_p]
        L.hydro_interpolate_async.argtypes = [ctypes.POINTER(HydroInterpArgs), ctypes.c_void_p]
        L.hydro_prepare_patches_async.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                  ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        I3p = ctypes.POINTER(ctypes.c_int * 3)
        L.hydro_gather_boxes_async.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, I3p, I3p,
                                               ctypes.c_void_p, I3p, ctypes.c_int, ctypes.c_void_p]
        L.hydro_assemble_patch_groups_async.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p,
                                                        ctypes.c_int, I3p, I3p, ctypes.c_int, ctypes.c_void_p]
        _lib = L
    return _lib


def _check(rc: int):
    if rc != HYDRO_OK:
        raise HydroError(f"hydro error {rc}: {lib().hydro_last_error().decode()}")


def make_args(npg, dt, dh, gamma=5.0 / 3.0, scheme="MHM_RP", rsolver="HLLC", limiter="VL_GMINMOD",
              minmod_coeff=1.5, minmod_maxiter=0, min_dens=0.0, min_pres=0.0, min_eint=0.0, store_flux=True,
              precision="fp64", in_=None, out=None, flux=None, dt_patch=None, safety=0.0,
              waves_per_group=0) -> HydroSolveArgs:
    return HydroSolveArgs(in_, out, flux, int(npg), float(dt), float(dh), float(gamma), SCHEME[scheme],
                          RSOLVER[rsolver], LIMITER[limiter], float(minmod_coeff), int(minmod_maxiter),
                          float(min_dens), float(min_pres), float(min_eint), int(bool(store_flux)),
                          PRECISION[precision], dt_patch, float(safety), int(waves_per_group))


def _stream(dev, stream, *tensors):
    """The launch stream: ``stream`` or torch's current stream.  When the launch goes to another stream than the
    current one, every tensor it touches is recorded on that stream, so the caching allocator does not hand their
    memory out again (e.g. a temporary from ``.contiguous()``) while the launch may still use it."""
    import torch
    cur = torch.cuda.current_stream(dev)
    s = stream if stream is not None else cur
    if s != cur:
        for t in tensors:
            if t is not None:
                t.record_stream(s)
    return ctypes.c_void_p(s.cuda_stream)


def _check_out(t, shape, dtype, device, name):
    """A caller-supplied output: a contiguous device tensor of exactly this shape and dtype on the input's device
    (the kernels write through its raw pointer, sized from the shapes they assume)."""
    if t is None:
        return
    if not getattr(t, "is_cuda", False) or t.device != device:
        raise HydroError(f"{name} must be a device tensor on {device}")
    if t.dtype != dtype:
        raise HydroError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise HydroError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise HydroError(f"{name} must be contiguous")


def _prec(t):
    import torch
    if t.dtype == torch.float64:
        return PRECISION["fp64"]
    if t.dtype == torch.float32:
        return PRECISION["fp32"]
    raise HydroError(f"unsupported dtype {t.dtype}")


def fluid_solver(u, dt, dh, gamma=5.0 / 3.0, scheme="MHM_RP", rsolver="HLLC", limiter="VL_GMINMOD",
                 minmod_coeff=1.5, minmod_maxiter=0, min_dens=0.0, min_pres=0.0, min_eint=0.0, store_flux=True,
                 out=None, flux=None, stream=None, dt_patch=None, safety=0.5, waves_per_group=0):
    """One patch-group solver call on device tensors.

    u: torch tensor on a ROCm device, float64 (FP64 path) or float32 (FP32-bulk path), shape (npg, 5, N, N, N).
    Returns (out (npg,5,16,16,16), flux (npg,9,5,16,16) or None), enqueued on ``stream`` (default: torch's
    current stream) — synchronise before reading on the host.  With ``dt_patch`` (a (npg, 8) tensor) the CFL dt
    of the 8 updated patches of every patch group is computed in the same kernel (index px + 2 py + 4 pz).
    """
    import torch
    if not u.is_cuda:
        raise HydroError("fluid_solver needs a device tensor")
    if u.dtype not in (torch.float64, torch.float32):
        raise HydroError(f"unsupported dtype {u.dtype}")
    n = NX[scheme]
    if u.dim() != 5 or tuple(u.shape[1:]) != (5, n, n, n):
        raise HydroError(f"input shape {tuple(u.shape)} != (npg, 5, {n}, {n}, {n})")
    u = u.contiguous()
    npg = u.shape[0]
    _check_out(out, (npg, 5, 16, 16, 16), u.dtype, u.device, "out")
    _check_out(flux if store_flux else None, (npg, 9, 5, 16, 16), u.dtype, u.device, "flux")
    _check_out(dt_patch, (npg, 8), u.dtype, u.device, "dt_patch")
    if out is None:
        out = torch.empty((npg, 5, 16, 16, 16), dtype=u.dtype, device=u.device)
    if store_flux and flux is None:
        flux = torch.empty((npg, 9, 5, 16, 16), dtype=u.dtype, device=u.device)
    prec = "fp64" if u.dtype == torch.float64 else "fp32"
    args = make_args(npg, dt, dh, gamma, scheme, rsolver, limiter, minmod_coeff, minmod_maxiter, min_dens, min_pres,
                     min_eint, store_flux, prec, u.data_ptr(), out.data_ptr(),
                     flux.data_ptr() if store_flux else None,
                     dt_patch.data_ptr() if dt_patch is not None else None, safety, waves_per_group)
    _check(lib().hydro_mhm_solve_async(ctypes.byref(args),
                                       _stream(u.device, stream, u, out, flux if store_flux else None, dt_patch)))
    return out, (flux if store_flux else None)


def solve_batch(u, pg_dt_dh=None, out=None, flux=None, queue_counter=None, groups_per_cu=2, pgs_per_ticket=1,
                stream=None, dt_patch=None, safety=0.5, dt=0.0, dh=1.0, **kw):
    """One launch over patch groups with per-PG time steps (e.g. several AMR levels at once).

    pg_dt_dh: float64 device tensor (npg, 2) of (dt, dh) per patch group (None: scalar dt, dh for all).
    queue_counter: a 1-element int32 device tensor enables the persistent work-queue mode with ``groups_per_cu``
    workgroups per compute unit, each taking ``pgs_per_ticket`` consecutive patch groups per queue ticket;
    None = one workgroup per PG.
    """
    import torch
    scheme = kw.pop("scheme", "MHM_RP")
    n = NX[scheme]
    if not u.is_cuda or u.dtype not in (torch.float64, torch.float32):
        raise HydroError("solve_batch needs a float64 or float32 device tensor")
    u = u.contiguous()
    if u.dim() != 5 or tuple(u.shape[1:]) != (5, n, n, n):
        raise HydroError(f"input shape {tuple(u.shape)} != (npg, 5, {n}, {n}, {n})")
    npg = u.shape[0]
    if pg_dt_dh is not None and (pg_dt_dh.dtype != torch.float64 or tuple(pg_dt_dh.shape) != (npg, 2)
                                 or not pg_dt_dh.is_cuda or not pg_dt_dh.is_contiguous()):
        raise HydroError("pg_dt_dh must be a contiguous float64 device tensor of shape (npg, 2)")
    if queue_counter is not None and (queue_counter.dtype != torch.int32 or queue_counter.numel() < 1
                                      or not queue_counter.is_cuda):
        raise HydroError("queue_counter must be an int32 device tensor with at least one element")
    store_flux = kw.pop("store_flux", True)
    _check_out(out, (npg, 5, 16, 16, 16), u.dtype, u.device, "out")
    _check_out(flux if store_flux else None, (npg, 9, 5, 16, 16), u.dtype, u.device, "flux")
    _check_out(dt_patch, (npg, 8), u.dtype, u.device, "dt_patch")
    if pg_dt_dh is not None and pg_dt_dh.device != u.device:
        raise HydroError("pg_dt_dh must be on the input's device")
    if out is None:
        out = torch.empty((npg, 5, 16, 16, 16), dtype=u.dtype, device=u.device)
    if store_flux and flux is None:
        flux = torch.empty((npg, 9, 5, 16, 16), dtype=u.dtype, device=u.device)
    args = make_args(npg, dt, dh, scheme=scheme, store_flux=store_flux,
                     precision="fp64" if u.dtype == torch.float64 else "fp32",
                     in_=u.data_ptr(), out=out.data_ptr(), flux=flux.data_ptr() if store_flux else None,
                     dt_patch=dt_patch.data_ptr() if dt_patch is not None else None, safety=safety, **kw)
    _check(lib().hydro_mhm_solve_batch_async(ctypes.byref(args),
                                             pg_dt_dh.data_ptr() if pg_dt_dh is not None else None,
                                             queue_counter.data_ptr() if queue_counter is not None else None,
                                             int(groups_per_cu), int(pgs_per_ticket),
                                             _stream(u.device, stream, u, out, flux if store_flux else None,
                                                     dt_patch, pg_dt_dh, queue_counter)))
    return out, (flux if store_flux else None)


def dt_cfl(u_patch, dh, safety, gamma=5.0 / 3.0, min_pres=0.0, out=None, stream=None):
    """Per-patch CFL dt of (npatch, 5, 8, 8, 8) device data (CUAPI_Asyn_dtSolver for the hydro CFL)."""
    import torch
    if not u_patch.is_cuda:
        raise HydroError("dt_cfl needs a device tensor")
    u_patch = u_patch.contiguous()
    if u_patch.dim() != 5 or tuple(u_patch.shape[1:]) != (5, 8, 8, 8):
        raise HydroError(f"patch shape {tuple(u_patch.shape)} != (npatch, 5, 8, 8, 8)")
    prec = _prec(u_patch)
    _check_out(out, (u_patch.shape[0],), u_patch.dtype, u_patch.device, "out")
    if out is None:
        out = torch.empty(u_patch.shape[0], dtype=u_patch.dtype, device=u_patch.device)
    _check(lib().hydro_dt_cfl_async(u_patch.data_ptr(), out.data_ptr(), u_patch.shape[0], dh, safety, gamma, min_pres,
                                    prec, _stream(u_patch.device, stream, u_patch, out)))
    return out


def prepare_uniform(level, ghost, out=None, stream=None):
    """Patch-group inputs (npg, 5, 16+2g, ...) with periodic ghost zones from a device level (5, Z, Y, X)."""
    import torch
    level = level.contiguous()
    _, nz, ny, nx = level.shape
    n = 16 + 2 * ghost
    npg = (nx // 16) * (ny // 16) * (nz // 16)
    _check_out(out, (npg, 5, n, n, n), level.dtype, level.device, "out")
    if out is None:
        out = torch.empty((npg, 5, n, n, n), dtype=level.dtype, device=level.device)
    _check(lib().hydro_prepare_uniform_async(level.data_ptr(), out.data_ptr(), nx, ny, nz, ghost, _prec(level),
                                             _stream(level.device, stream, level, out)))
    return out


def close_uniform(pg_out, level, stream=None):
    """Scatter (npg, 5, 16, 16, 16) patch-group outputs back into the device level (5, Z, Y, X) in place."""
    _, nz, ny, nx = level.shape
    pg_out = pg_out.contiguous()
    _check(lib().hydro_close_uniform_async(pg_out.data_ptr(), level.data_ptr(), nx, ny, nz,
                                           _prec(level), _stream(level.device, stream, pg_out, level)))
    return level


def adjoint_workspace(npg, device="cuda:0"):
    """Device workspace for the adjoint / tangent entry points (one call at a time)."""
    import torch
    nbytes = int(lib().hydro_mhm_adjoint_workspace_bytes(int(npg)))
    return torch.empty((max(nbytes, 8) + 7) // 8, dtype=torch.float64, device=device)


def _check_fp64_pg(u, name):
    import torch
    if not u.is_cuda or u.dtype != torch.float64:
        raise HydroError(f"{name} must be a float64 device tensor")
    if u.dim() != 5 or tuple(u.shape[1:]) != (5, 22, 22, 22):
        raise HydroError(f"{name} shape {tuple(u.shape)} != (npg, 5, 22, 22, 22)")


def mhm_adjoint(u, lam_out, dt, dh, workspace=None, stream=None, **kw):
    """Vector-Jacobian product of one MHM_RP FP64 step at input ``u`` (npg, 5, 22^3):
    returns lam_in = (d out / d u)^T lam_out, lam_out of shape (npg, 5, 16, 16, 16)."""
    import torch
    _check_fp64_pg(u, "u")
    u = u.contiguous()
    npg = u.shape[0]
    lam_out = lam_out.contiguous()
    if tuple(lam_out.shape) != (npg, 5, 16, 16, 16) or lam_out.dtype != torch.float64:
        raise HydroError("lam_out must be float64 of shape (npg, 5, 16, 16, 16)")
    ws = workspace if workspace is not None else adjoint_workspace(npg, u.device)
    lam_in = torch.empty_like(u)
    args = make_args(npg, dt, dh, store_flux=False, precision="fp64", in_=u.data_ptr(), **kw)
    _check(lib().hydro_mhm_adjoint_async(ctypes.byref(args), lam_out.data_ptr(), lam_in.data_ptr(), ws.data_ptr(),
                                         _stream(u.device, stream, u, lam_out, lam_in, ws)))
    return lam_in


def mhm_tangent(u, v, dt, dh, workspace=None, stream=None, **kw):
    """Jacobian-vector product of one MHM_RP FP64 step: returns (out, J v); v=None gives (out, None) from the
    phase-wise evaluation used by the derivative kernels."""
    import torch
    _check_fp64_pg(u, "u")
    u = u.contiguous()
    npg = u.shape[0]
    if v is not None:
        _check_fp64_pg(v, "v")
        if v.shape != u.shape:
            raise HydroError("v must have the shape of u")
        v = v.contiguous()
    ws = workspace if workspace is not None else adjoint_workspace(npg, u.device)
    out = torch.empty((npg, 5, 16, 16, 16), dtype=torch.float64, device=u.device)
    jv = torch.empty_like(out) if v is not None else None
    args = make_args(npg, dt, dh, store_flux=False, precision="fp64", in_=u.data_ptr(), out=out.data_ptr(), **kw)
    _check(lib().hydro_mhm_tangent_async(ctypes.byref(args), v.data_ptr() if v is not None else None,
                                         jv.data_ptr() if jv is not None else None, ws.data_ptr(),
                                         _stream(u.device, stream, u, v, out, jv, ws)))
    return out, jv


def prepare_uniform_adjoint(pg_lam, level_shape, ghost, stream=None):
    """Transpose of prepare_uniform: sum the (npg, 5, N^3) patch-group copies back onto a (5, Z, Y, X) level."""
    import torch
    pg_lam = pg_lam.contiguous()
    _, nz, ny, nx = level_shape
    lam = torch.empty(tuple(level_shape), dtype=pg_lam.dtype, device=pg_lam.device)
    _check(lib().hydro_prepare_uniform_adjoint_async(pg_lam.data_ptr(), lam.data_ptr(), nx, ny, nz, ghost,
                                                     _prec(pg_lam), _stream(pg_lam.device, stream, pg_lam, lam)))
    return lam


# ---------------- Flu_Close and level synchronisation (device tensors) ----------------
def _int_table(t, shape, name):
    import torch
    if not t.is_cuda or t.dtype != torch.int32 or tuple(t.shape) != tuple(shape) or not t.is_contiguous():
        raise HydroError(f"{name} must be a contiguous int32 device tensor of shape {tuple(shape)}")


def correct_unphysical(u_in, out, flux, dt, dh, mode="3D1D", rsolver="HLLE", scheme="MHM_RP", gamma=5.0 / 3.0,
                       last_resort_floor=True, auto_reduce_dt=False, fixup_flux=True, check_pres=False,
                       min_dens=0.0, min_pres=0.0, min_eint=0.0, stream=None, counters=None, main_rsolver=None):
    """CorrectUnphysical on device: corrects ``out`` (and the coarse-fine planes of ``flux``) in place.
    Returns a 2-element int32 device tensor: [corrected cells, cells still unphysical].  With ``counters`` (an
    int32 device tensor of 2 elements) the kernel adds its counts to it and returns it, with no extra launches.
    ``main_rsolver`` is the fluid solver's Riemann solver (GAMER's compile-time RSOLVER; None: ``rsolver``): a
    first-order Roe correction checks its intermediate states only when it is "ROE"."""
    import torch
    n = NX[scheme]
    if tuple(u_in.shape[1:]) != (5, n, n, n) or tuple(out.shape) != (u_in.shape[0], 5, 16, 16, 16):
        raise HydroError("correct_unphysical: shapes must be (npg, 5, N^3) and (npg, 5, 16^3)")
    if not u_in.is_cuda or u_in.device != out.device:
        raise HydroError("correct_unphysical: u_in and out must be device tensors on one device")
    if u_in.dtype != out.dtype:
        raise HydroError("correct_unphysical: dtypes differ")
    _check_out(out, (u_in.shape[0], 5, 16, 16, 16), u_in.dtype, u_in.device, "out")
    _check_out(flux, (u_in.shape[0], 9, 5, 16, 16), u_in.dtype, u_in.device, "flux")
    if counters is None:
        cnt = torch.zeros(2, dtype=torch.int32, device=out.device)
    elif (counters.dtype != torch.int32 or counters.numel() != 2 or not counters.is_contiguous()
          or counters.device != out.device):
        raise HydroError("correct_unphysical: counters must be a contiguous int32 tensor of 2 elements on out's device")
    else:
        cnt = counters
    u_in = u_in.contiguous()
    a = HydroCorrectArgs(u_in.data_ptr(), out.data_ptr(), flux.data_ptr() if flux is not None else None,
                         u_in.shape[0], float(dt), float(dh), float(gamma), SCHEME[scheme], FIRST_FLUX_CORR[mode],
                         RSOLVER[rsolver], int(last_resort_floor), int(auto_reduce_dt), int(fixup_flux),
                         int(check_pres), float(min_dens), float(min_pres), float(min_eint), _prec(out),
                         cnt.data_ptr(), 0 if main_rsolver is None else RSOLVER[main_rsolver])
    _check(lib().hydro_correct_unphysical_async(ctypes.byref(a), _stream(out.device, stream, u_in, out, flux, cnt)))
    return cnt


def store_flux(flux, slot, dt, reg, stream=None):
    """StoreFlux: reg[slot[pg, LocalID, s]] = dt * patch-face fluxes (reg: float64 (nreg, 5, 8, 8))."""
    _int_table(slot, (flux.shape[0], 8, 6), "slot")
    flux = flux.contiguous()
    _check(lib().hydro_store_flux_async(flux.data_ptr(), slot.data_ptr(), flux.shape[0], float(dt),
                                        reg.data_ptr(), _prec(flux), _stream(flux.device, stream, flux, slot, reg)))
    return reg


def correct_flux(flux, slot, dt, reg, stream=None):
    """CorrectFlux: reg[slot[pg, s]] -= dt/4 * the fine fluxes on patch-group face s."""
    _int_table(slot, (flux.shape[0], 6), "slot")
    flux = flux.contiguous()
    _check(lib().hydro_correct_flux_async(flux.data_ptr(), slot.data_ptr(), flux.shape[0], float(dt),
                                          reg.data_ptr(), _prec(flux), _stream(flux.device, stream, flux, slot, reg)))
    return reg


def fixup_flux(patch, slot, reg, dh, gamma=5.0 / 3.0, min_dens=0.0, min_pres=0.0, min_eint=0.0, stream=None):
    """Flu_FixUp_Flux on patches (npatch, 5, 8, 8, 8) in place, registers reg[slot[patch, s]]."""
    _int_table(slot, (patch.shape[0], 6), "slot")
    _check(lib().hydro_fixup_flux_async(patch.data_ptr(), slot.data_ptr(), reg.data_ptr(), patch.shape[0],
                                        float(dh), float(gamma), float(min_dens), float(min_pres), float(min_eint),
                                        _prec(patch), _stream(patch.device, stream, patch, slot, reg)))
    return patch


def fixup_restrict(son, fa, son0, fa_idx, min_eint=0.0, stream=None):
    """Flu_FixUp_Restrict: fa[fa_idx[q]] = mean of son[son0[q] + LocalID], then the Eint floor (in place)."""
    _int_table(son0, (son0.shape[0],), "son0")
    _int_table(fa_idx, (son0.shape[0],), "fa_idx")
    son = son.contiguous()
    _check(lib().hydro_fixup_restrict_async(son.data_ptr(), fa.data_ptr(), son0.shape[0],
                                            son0.data_ptr(), fa_idx.data_ptr(), float(min_eint), _prec(fa),
                                            _stream(fa.device, stream, son, fa, son0, fa_idx)))
    return fa


def close_patches(pg_out, patch, pid0, stream=None):
    """Copy (npg, 5, 16^3) patch-group outputs into patches pid0[pg] + LocalID of (npatch, 5, 8, 8, 8)."""
    _int_table(pid0, (pg_out.shape[0],), "pid0")
    pg_out = pg_out.contiguous()
    _check(lib().hydro_close_patches_async(pg_out.data_ptr(), patch.data_ptr(), pg_out.shape[0],
                                           pid0.data_ptr(), _prec(patch), _stream(patch.device, stream, pg_out, patch, pid0)))
    return patch


def interpolate(cdata, fdata, cstart, crange, fstart=(0, 0, 0), scheme="CQUAD", monotonic=0x1F, mono_coeff=2.0,
                opp_sign_0th=False, int_prim=True, reduce_mono=True, mono_max_iter=10, gamma=5.0 / 3.0, min_pres=0.0,
                need=None, stream=None, nfail=None):
    """Coarse -> fine interpolation of boxes cdata (nbox, 5, cz, cy, cx) into fdata (nbox, 5, fz, fy, fx), in place.
    Index tuples are (x, y, z).  ``need``: optional uint8 device mask (nbox, crz, cry, crx) of the coarse cells whose
    fine cells are used (the others are skipped and their fine cells left as they are).
    Returns an int32 device tensor [number of boxes still unphysical].  With ``nfail`` (an int32 device tensor of
    1 element) the kernel adds its count to it and returns it, with no extra launch."""
    import torch
    nbox = cdata.shape[0]
    if cdata.dim() != 5 or fdata.dim() != 5 or fdata.shape[0] != nbox or cdata.shape[1] != 5 or fdata.shape[1] != 5:
        raise HydroError("interpolate: cdata/fdata must be (nbox, 5, z, y, x)")
    if cdata.dtype != fdata.dtype or not cdata.is_contiguous() or not fdata.is_contiguous():
        raise HydroError("interpolate: contiguous tensors of one dtype required")
    if nfail is None:
        nf = torch.zeros(1, dtype=torch.int32, device=fdata.device)
    elif (nfail.dtype != torch.int32 or nfail.numel() != 1 or nfail.device != fdata.device):
        raise HydroError("interpolate: nfail must be an int32 tensor of 1 element on fdata's device")
    else:
        nf = nfail
    I3 = ctypes.c_int * 3
    a = HydroInterpArgs(cdata.data_ptr(), fdata.data_ptr(), nbox, I3(*reversed(cdata.shape[2:])), I3(*cstart),
                        I3(*crange), I3(*reversed(fdata.shape[2:])), I3(*fstart), INT_SCHEME[scheme], int(monotonic),
                        float(mono_coeff), int(opp_sign_0th), int(int_prim), int(reduce_mono), int(mono_max_iter),
                        float(gamma), float(min_pres), _prec(fdata), nf.data_ptr(), None)
    if need is not None:
        if (need.dtype != torch.uint8 or not need.is_cuda or not need.is_contiguous()
                or tuple(need.shape) != (nbox,) + tuple(reversed(crange))):
            raise HydroError("interpolate: need must be a contiguous uint8 device tensor (nbox, crz, cry, crx)")
        a.need = need.data_ptr()
    _check(lib().hydro_interpolate_async(ctypes.byref(a), _stream(fdata.device, stream, cdata, fdata, nf, need)))
    return nf


def prepare_patches(patch, pg_in, slot, ghost, stream=None):
    """Same-level patch-group inputs from patches (npatch, 5, 8, 8, 8); slot (npg, 4, 4, 4) int32, -1 = keep."""
    _int_table(slot, (pg_in.shape[0], 4, 4, 4), "slot")
    patch = patch.contiguous()
    _check(lib().hydro_prepare_patches_async(patch.data_ptr(), pg_in.data_ptr(), pg_in.shape[0],
                                             slot.data_ptr(), int(ghost), _prec(pg_in),
                                             _stream(pg_in.device, stream, patch, pg_in, slot)))
    return pg_in


def gather_boxes(patch, box, slot, origin, stream=None):
    """Coarse boxes (nbox, 5, sz, sy, sx) from patches (npatch, 5, 8, 8, 8), in place: box cell (x,y,z) comes from the
    patch covering cell origin + (x,y,z) of the box's slot table slot (nbox, tz, ty, tx) (-1 = keep).  (x,y,z) tuples."""
    nbox = box.shape[0]
    tdim = tuple(reversed(slot.shape[1:]))
    _int_table(slot, (nbox,) + tuple(slot.shape[1:]), "slot")
    if slot.dim() != 4:
        raise HydroError("gather_boxes: slot must be (nbox, tz, ty, tx)")
    I3 = ctypes.c_int * 3
    patch = patch.contiguous()
    _check(lib().hydro_gather_boxes_async(patch.data_ptr(), box.data_ptr(), nbox,
                                          ctypes.byref(I3(*reversed(box.shape[2:]))), ctypes.byref(I3(*origin)),
                                          slot.data_ptr(), ctypes.byref(I3(*tdim)), _prec(box),
                                          _stream(box.device, stream, patch, box, slot)))
    return box


def assemble_patch_groups(patch, fbox, pg_in, slot, ghost, fb_origin, stream=None):
    """Patch-group inputs: same-level patches where slot (npg, 4, 4, 4) >= 0, else the interpolated fine boxes
    fbox (npg, 5, fz, fy, fx) at window cell + fb_origin (x, y, z)."""
    npg = pg_in.shape[0]
    _int_table(slot, (npg, 4, 4, 4), "slot")
    if fbox.shape[0] != npg or fbox.dtype != pg_in.dtype or not fbox.is_contiguous():
        raise HydroError("assemble_patch_groups: fbox must be a contiguous (npg, 5, fz, fy, fx) tensor of pg_in's dtype")
    I3 = ctypes.c_int * 3
    patch = patch.contiguous()
    _check(lib().hydro_assemble_patch_groups_async(patch.data_ptr(), fbox.data_ptr(), pg_in.data_ptr(),
                                                   npg, slot.data_ptr(), int(ghost),
                                                   ctypes.byref(I3(*reversed(fbox.shape[2:]))),
                                                   ctypes.byref(I3(*fb_origin)), _prec(pg_in),
                                                   _stream(pg_in.device, stream, patch, fbox, pg_in, slot)))
    return pg_in


def prepare_refined_patch_groups(coarse_patch, fine_patch, coarse_slot, fine_slot, pg_in, ghost=3, scheme="CQUAD",
                                 need=None, stream=None, nfail=None, **interp_kw):
    """Prepare_PatchData for fine patch groups at a refinement boundary, on the device: the coarse box of each patch
    group (father patch + 26 siblings: coarse_slot (npg, 3, 3, 3)) -> Interpolate() over coarse cells -2..9 around the
    father (the 22^3 window of ghost 3 plus one fine cell) -> same-level fine patches (fine_slot (npg, 4, 4, 4)) and
    the interpolated cells into pg_in (npg, 5, 16+2g, ...).  ``need`` (npg, 12, 12, 12) uint8, optional: the coarse
    cells whose fine cells have no same-level patch (gamer_amd.amr.refined_interp_need); only those are
    interpolated.  Returns the interpolation failure counter (``nfail``, when given, is added to and returned)."""
    import torch
    if ghost != 3:
        raise HydroError("prepare_refined_patch_groups: ghost 3 (the MHM_RP window) only")
    g = 2 if scheme in ("CQUAR", "QUAR") else 1
    npg = pg_in.shape[0]
    n = 12 + 2 * g
    cbox = torch.empty((npg, 5, n, n, n), dtype=pg_in.dtype, device=pg_in.device)
    gather_boxes(coarse_patch, cbox, coarse_slot, (6 - g,) * 3, stream=stream)
    fbox = torch.empty((npg, 5, 24, 24, 24), dtype=pg_in.dtype, device=pg_in.device)
    nfail = interpolate(cbox, fbox, (g,) * 3, (12, 12, 12), scheme=scheme, need=need, stream=stream, nfail=nfail,
                        **interp_kw)
    assemble_patch_groups(fine_patch, fbox, pg_in, fine_slot, ghost, (1, 1, 1), stream=stream)
    return nfail


class UniformLevel:
    """A periodic uniform level advanced entirely on the device: ghost-zone preparation -> patch-group solver
    (with the fused CFL dt) -> scatter back, one pass of the Flu_AdvanceDt pipeline (src/Fluid/Flu_AdvanceDt.cpp:38-100)
    for a single level without refinement."""

    def __init__(self, level, dh, scheme="MHM_RP", safety=0.5, **solver_kw):
        import torch
        self.level = level.contiguous()
        self.dh, self.scheme, self.safety, self.kw = dh, scheme, safety, solver_kw
        self.ghost = 3 if scheme == "MHM_RP" else 2
        _, nz, ny, nx = level.shape
        self.npg = (nx // 16) * (ny // 16) * (nz // 16)
        n = 16 + 2 * self.ghost
        self.pg_in = torch.empty((self.npg, 5, n, n, n), dtype=level.dtype, device=level.device)
        self.pg_out = torch.empty((self.npg, 5, 16, 16, 16), dtype=level.dtype, device=level.device)
        self.flux = torch.empty((self.npg, 9, 5, 16, 16), dtype=level.dtype, device=level.device)
        self.dt_patch = torch.empty((self.npg, 8), dtype=level.dtype, device=level.device)

    def step(self, dt):
        """Advance by dt; returns the CFL-limited dt of the new state (device scalar tensor)."""
        prepare_uniform(self.level, self.ghost, out=self.pg_in)
        fluid_solver(self.pg_in, dt, self.dh, scheme=self.scheme, out=self.pg_out, flux=self.flux,
                     dt_patch=self.dt_patch, safety=self.safety, **self.kw)
        close_uniform(self.pg_out, self.level)
        return self.dt_patch.min()


def device_info():
    """Properties of the current device (HydroDeviceInfo as a dict)."""
    info = HydroDeviceInfo()
    _check(lib().hydro_device_info(ctypes.byref(info)))
    d = {k: getattr(info, k) for k, _ in HydroDeviceInfo._fields_}
    d["name"], d["arch"] = info.name.decode(), info.arch.decode()
    return d


class GpuFluidApi:
    """The GPU_API host-array path: pinned double buffers [2] + GPU_NStream streams (CUAPI_MemAllocate_Fluid)."""

    def __init__(self, flu_npg: int, gpu_nstream: int = 4, scheme: str = "MHM_RP", precision: str = "fp64",
                 device: int = 0):
        L = lib()
        _check(L.hydro_init(device))
        _check(L.hydro_mem_allocate_fluid(flu_npg, gpu_nstream, SCHEME[scheme], PRECISION[precision]))
        self.flu_npg, self.nstream, self.scheme, self.precision = flu_npg, gpu_nstream, scheme, precision
        self.dtype = np.float64 if precision == "fp64" else np.float32
        n = NX[scheme]
        self._views = []
        for aid in (0, 1):
            pi, po, pf = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
            _check(L.hydro_host_buffers(aid, ctypes.byref(pi), ctypes.byref(po), ctypes.byref(pf)))
            self._views.append((self._np(pi, (flu_npg, 5, n, n, n)), self._np(po, (flu_npg, 5, 16, 16, 16)),
                                self._np(pf, (flu_npg, 9, 5, 16, 16)), pi, po, pf))

    def _np(self, ptr, shape):
        count = int(np.prod(shape))
        buf = (ctypes.c_char * (count * np.dtype(self.dtype).itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=self.dtype).reshape(shape)

    def host_arrays(self, array_id: int):
        """(h_Flu_Array_In, h_Flu_Array_Out, h_Flux_Array) pinned numpy views of double buffer ``array_id``."""
        return self._views[array_id][:3]

    def asyn_fluid_solver(self, array_id: int, npg: int, dt, dh, gpu_nstream=None, **kw):
        """CUAPI_Asyn_FluidSolver on double buffer ``array_id``; returns before the GPU finishes."""
        v = self._views[array_id]
        args = make_args(npg, dt, dh, scheme=self.scheme, precision=self.precision, **kw)
        _check(lib().hydro_asyn_fluid_solver(v[3], v[4], v[5], ctypes.byref(args),
                                             gpu_nstream if gpu_nstream is not None else self.nstream))

    def host_dt_arrays(self, array_id: int):
        """(h_Flu_Array_T (8 npg, 5, 8, 8, 8), h_dt_Array_T (8 npg,)) pinned numpy views of double buffer ``array_id``."""
        pf, pd = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().hydro_host_dt_buffers(array_id, ctypes.byref(pf), ctypes.byref(pd)))
        return self._np(pf, (8 * self.flu_npg, 5, 8, 8, 8)), self._np(pd, (8 * self.flu_npg,))

    def asyn_dt_solver(self, array_id: int, npg: int, dh, safety, gamma=5.0 / 3.0, min_pres=0.0, gpu_nstream=None):
        """CUAPI_Asyn_dtSolver (DT_FLU_SOLVER) on double buffer ``array_id``; returns before the GPU finishes."""
        pf, pd = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().hydro_host_dt_buffers(array_id, ctypes.byref(pf), ctypes.byref(pd)))
        _check(lib().hydro_asyn_dt_solver(pf, pd, int(npg), float(dh), float(safety), float(gamma), float(min_pres),
                                          gpu_nstream if gpu_nstream is not None else self.nstream))

    @staticmethod
    def synchronize():
        _check(lib().hydro_synchronize())

    def free(self):
        self._views = []
        _check(lib().hydro_mem_free_fluid())
