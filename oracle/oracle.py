"""ORACLE bindings — test infrastructure only.

ctypes access to
  * ``liboracle.so``: the from-scratch CPU restatement of GAMER's MHM/MHM_RP solver (hydro_oracle.cpp), and
  * ``_ref/libgamer_ref_<scheme>_<rsolver>_<prec>.so``: GAMER's own CPU_FluidSolver_MHM, compiled from the
    reference sources by oracle/Makefile (absent where the reference tree was never available).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, and only as
the checker or the CPU baseline — never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SCHEME = {"MHM": 3, "MHM_RP": 4}
RSOLVER = {"ROE": 2, "HLLE": 3, "HLLC": 4}
LIMITER = {"VANLEER": 1, "GMINMOD": 2, "ALBADA": 3, "VL_GMINMOD": 4, "CENTRAL": 6}

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SOLVER_ARGS = [_P, _P, _P, _I, _D, _D, _D, _I, _I, _I, _I, _D, _I, _D, _D, _D, _I]
_REF_ARGS = [_P, _P, _P, _I, _D, _D, _D, _I, _I, _D, _I, _D, _D, _D, _I]

_lib = None


def lib():
    """The restatement library (built by ``make -C oracle oracle``)."""
    global _lib
    if _lib is None:
        # ORACLE_LIB: another build of the same restatement (the sanitizer test loads an ASan/UBSan build)
        path = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle oracle`")
        L = ctypes.CDLL(path)
        for name in ("oracle_fluid_solver_f64", "oracle_fluid_solver_f32"):
            getattr(L, name).argtypes = _SOLVER_ARGS
            getattr(L, name).restype = _I
        L.oracle_riemann_f64.argtypes = [_I, _I, _P, _P, _D, _D, _P]
        L.oracle_limit_slope_f64.argtypes = [_P, _P, _P, _I, _D, _P]
        L.oracle_dt_cfl_f64.argtypes = [_P, _P, _I, _D, _D, _D, _D, _I]
        for suf in ("f64", "f32"):
            getattr(L, "oracle_correct_unphysical_" + suf).argtypes = [_P, _P, _P, _I, _I, _D, _D, _D, _I, _I, _I, _I,
                                                                         _I, _I, _D, _D, _D, _I, _P]
            getattr(L, "oracle_store_flux_" + suf).argtypes = [_P, _I, _P, _D, _P]
            getattr(L, "oracle_correct_flux_" + suf).argtypes = [_P, _I, _P, _D, _P]
            getattr(L, "oracle_fixup_flux_" + suf).argtypes = [_P, _I, _P, _P, _D, _D, _D, _D, _D]
            getattr(L, "oracle_fixup_restrict_" + suf).argtypes = [_P, _P, _I, _P, _P, _D]
            getattr(L, "oracle_close_patches_" + suf).argtypes = [_P, _P, _I, _P]
            getattr(L, "oracle_interpolate_" + suf).argtypes = [_P, _P, _I, _P, _P, _P, _P, _P, _I, _I, _D, _I, _I,
                                                                  _I, _I, _D, _D, _P]
            getattr(L, "oracle_prepare_patches_" + suf).argtypes = [_P, _P, _I, _P, _I]
            getattr(L, "oracle_gather_boxes_" + suf).argtypes = [_P, _P, _I, _P, _P, _P, _P]
            getattr(L, "oracle_assemble_patch_groups_" + suf).argtypes = [_P, _P, _P, _I, _P, _I, _P, _P]
        _lib = L
    return _lib


def ref_path(scheme: str, rsolver: str, prec: str) -> str:
    return os.path.join(HERE, "_ref", f"libgamer_ref_{scheme.lower()}_{rsolver.lower()}_{prec}.so")


def ref_available(scheme="MHM_RP", rsolver="HLLC", prec="f64") -> bool:
    return os.path.exists(ref_path(scheme, rsolver, prec))


_refs = {}


def ref_lib(scheme: str, rsolver: str, prec: str):
    key = (scheme, rsolver, prec)
    if key not in _refs:
        L = ctypes.CDLL(ref_path(scheme, rsolver, prec))
        L.gamer_ref_fluid_solver.argtypes = _REF_ARGS
        L.gamer_ref_fluid_solver.restype = _I
        L.gamer_ref_dt_cfl.argtypes = [_P, _P, _I, _D, _D, _D, _D, _I]
        L.gamer_ref_dt_cfl.restype = _I
        L.gamer_ref_interpolate.argtypes = [_P, _P, _I, _P, _P, _P, _P, _P, _I, _I, _D, _I, _I, _I, _I, _D, _D]
        L.gamer_ref_interpolate.restype = _I
        L.gamer_ref_correct_unphysical.argtypes = [_P, _P, _P, _I, _D, _D, _D, _I, _I, _I, _I, _I, _I, _D, _D, _D,
                                                   _P, _I]
        L.gamer_ref_correct_unphysical.restype = _I
        L.gamer_ref_store_flux.argtypes = [_P, _I, _P, _D, _P]
        L.gamer_ref_correct_flux.argtypes = [_P, _I, _P, _D, _P]
        L.gamer_ref_fixup_flux.argtypes = [_P, _I, _P, _P, _D, _D, _D, _D, _D]
        L.gamer_ref_fixup_restrict.argtypes = [_P, _P, _I, _P, _P, _I, _D]
        L.gamer_ref_close_patches.argtypes = [_P, _P, _I, _P, _I]
        for f in ("store_flux", "correct_flux", "fixup_flux", "fixup_restrict", "close_patches"):
            getattr(L, "gamer_ref_" + f).restype = _I
        _refs[key] = L
    return _refs[key]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _alloc_out(npg, dtype, store_flux):
    out = np.empty((npg, 5, 16, 16, 16), dtype=dtype)
    flux = np.empty((npg, 9, 5, 16, 16), dtype=dtype) if store_flux else None
    return out, flux


def fluid_solver(u, dt, dh, gamma=5.0 / 3.0, scheme="MHM_RP", rsolver="HLLC", store_flux=True,
                 limiter="VL_GMINMOD", minmod_coeff=1.5, minmod_maxiter=0, min_dens=0.0, min_pres=0.0,
                 min_eint=0.0, nthreads=0):
    """Restatement: u (npg, 5, N, N, N) float64/float32 -> (out (npg,5,16,16,16), flux (npg,9,5,16,16))."""
    u = np.ascontiguousarray(u)
    npg = u.shape[0]
    out, flux = _alloc_out(npg, u.dtype, store_flux)
    fn = lib().oracle_fluid_solver_f64 if u.dtype == np.float64 else lib().oracle_fluid_solver_f32
    rc = fn(_ptr(u), _ptr(out), _ptr(flux), npg, dt, dh, gamma, SCHEME[scheme], RSOLVER[rsolver],
            int(store_flux), LIMITER[limiter], minmod_coeff, minmod_maxiter, min_dens, min_pres, min_eint, nthreads)
    if rc != 0:
        raise ValueError(f"oracle_fluid_solver returned {rc}")
    return out, flux


def ref_fluid_solver(u, dt, dh, gamma=5.0 / 3.0, scheme="MHM_RP", rsolver="HLLC", store_flux=True,
                     limiter="VL_GMINMOD", minmod_coeff=1.5, minmod_maxiter=0, min_dens=0.0, min_pres=0.0,
                     min_eint=0.0, nthreads=0):
    """GAMER's own CPU_FluidSolver_MHM (compiled reference) on the same arguments."""
    u = np.ascontiguousarray(u)
    prec = "f64" if u.dtype == np.float64 else "f32"
    L = ref_lib(scheme, rsolver, prec)
    npg = u.shape[0]
    out, flux = _alloc_out(npg, u.dtype, store_flux)
    rc = L.gamer_ref_fluid_solver(_ptr(u), _ptr(out), _ptr(flux), npg, dt, dh, gamma, int(store_flux),
                                  LIMITER[limiter], minmod_coeff, minmod_maxiter, min_dens, min_pres, min_eint,
                                  nthreads)
    if rc != 0:
        raise ValueError(f"gamer_ref_fluid_solver returned {rc}")
    return out, flux


def riemann(rsolver, d, L, R, gamma=5.0 / 3.0, min_pres=0.0):
    L = np.ascontiguousarray(L, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64)
    f = np.empty(5)
    rc = lib().oracle_riemann_f64(RSOLVER[rsolver], d, _ptr(L), _ptr(R), gamma, min_pres, _ptr(f))
    if rc != 0:
        raise ValueError(rc)
    return f


def limit_slope(L, C, R, limiter, coeff=1.5):
    L, C, R = (np.ascontiguousarray(a, dtype=np.float64) for a in (L, C, R))
    s = np.empty(5)
    rc = lib().oracle_limit_slope_f64(_ptr(L), _ptr(C), _ptr(R), LIMITER[limiter], coeff, _ptr(s))
    if rc != 0:
        raise ValueError(rc)
    return s


def ref_dt_cfl(u_patch, dh, safety, gamma=5.0 / 3.0, min_pres=0.0, nthreads=0):
    """GAMER's own CPU_dtSolver_HydroCFL (compiled reference, FP64 build) on (npatch, 5, 8, 8, 8) float64,
    npatch a multiple of 8 (the reference loops over the 8 patches of each patch group)."""
    u = np.ascontiguousarray(u_patch, dtype=np.float64)
    dt = np.empty(u.shape[0])
    rc = ref_lib("MHM_RP", "HLLC", "f64").gamer_ref_dt_cfl(_ptr(u), _ptr(dt), u.shape[0], dh, safety, gamma,
                                                             min_pres, nthreads)
    if rc != 0:
        raise ValueError(f"gamer_ref_dt_cfl returned {rc}")
    return dt


def ref_interpolate(cdata, fdata, cstart, crange, fstart=(0, 0, 0), scheme="CQUAD", monotonic=0x1F, mono_coeff=2.0,
                    opp_sign_0th=False, int_prim=True, reduce_mono=True, mono_max_iter=10, gamma=5.0 / 3.0,
                    min_pres=0.0):
    """GAMER's own Interpolate() (compiled reference, the build of cdata's precision) box by box, with the
    arguments of ``interpolate``: returns (cdata after the call, fdata).  The reference aborts the process if a
    box still fails after its last iteration, so only inputs it accepts may be passed."""
    c, f = np.array(cdata, copy=True), np.array(fdata, copy=True)
    i3 = lambda t: np.ascontiguousarray(t, dtype=np.int32)
    L = ref_lib("MHM_RP", "HLLC", "f64" if c.dtype == np.float64 else "f32")
    rc = L.gamer_ref_interpolate(_ptr(c), _ptr(f), c.shape[0], _ptr(i3(c.shape[:1:-1])), _ptr(i3(cstart)),
                                 _ptr(i3(crange)), _ptr(i3(f.shape[:1:-1])), _ptr(i3(fstart)), _INT[scheme],
                                 int(monotonic), mono_coeff, int(opp_sign_0th), int(int_prim), int(reduce_mono),
                                 int(mono_max_iter), gamma, min_pres)
    if rc != 0:
        raise ValueError(f"gamer_ref_interpolate returned {rc}")
    return c, f


def dt_cfl(u_patch, dh, safety, gamma=5.0 / 3.0, min_pres=0.0, nthreads=0):
    """Per-patch CFL dt; u_patch (npatch, 5, 8, 8, 8) float64."""
    u = np.ascontiguousarray(u_patch, dtype=np.float64)
    dt = np.empty(u.shape[0])
    rc = lib().oracle_dt_cfl_f64(_ptr(u), _ptr(dt), u.shape[0], dh, safety, gamma, min_pres, nthreads)
    if rc != 0:
        raise ValueError(rc)
    return dt


# ---------------- Flu_Close and the AMR level synchronisation ----------------
_FFC = {"NONE": 0, "3D": 1, "3D1D": 2}


def _suffix(a):
    return "f64" if a.dtype == np.float64 else "f32"


def correct_unphysical(u_in, out, flux, dt, dh, mode="3D1D", rsolver="HLLE", scheme="MHM_RP", gamma=5.0 / 3.0,
                       last_resort_floor=True, auto_reduce_dt=False, fixup_flux=True, check_pres=False,
                       min_dens=0.0, min_pres=0.0, min_eint=0.0, main_rsolver=None):
    """Restatement of CorrectUnphysical; returns (out, flux, counters) as new arrays.  main_rsolver is the
    fluid solver's Riemann solver (GAMER's compile-time RSOLVER; None: the same as rsolver): a first-order Roe
    correction checks its intermediate states only when it is "ROE"."""
    u_in = np.ascontiguousarray(u_in)
    out = np.array(out, copy=True)
    flux = None if flux is None else np.array(flux, copy=True)
    cnt = np.zeros(2, dtype=np.int64)
    fn = getattr(lib(), "oracle_correct_unphysical_" + _suffix(out))
    rc = fn(_ptr(u_in), _ptr(out), _ptr(flux) if flux is not None else None, out.shape[0], SCHEME[scheme],
            ctypes.c_double(dt), ctypes.c_double(dh), ctypes.c_double(gamma), _FFC[mode], RSOLVER[rsolver],
            int(last_resort_floor), int(auto_reduce_dt), int(fixup_flux), int(check_pres),
            ctypes.c_double(min_dens), ctypes.c_double(min_pres), ctypes.c_double(min_eint),
            0 if main_rsolver is None else RSOLVER[main_rsolver], _ptr(cnt))
    if rc != 0:
        raise ValueError(f"oracle_correct_unphysical returned {rc}")
    return out, flux, cnt


_RSOLVER_1ST = {"ROE": 1, "HLLC": 2, "HLLE": 3}     # OptRSolver1st_t (include/Typedef.h:469-474)


def ref_correct_unphysical(u_in, out, flux, dt, dh, mode="3D1D", rsolver="HLLE", scheme="MHM_RP", gamma=5.0 / 3.0,
                           last_resort_floor=True, auto_reduce_dt=False, fixup_flux=True, check_pres=False,
                           min_dens=0.0, min_pres=0.0, min_eint=0.0, main_rsolver="HLLC", nthreads=0):
    """GAMER's own CorrectUnphysical (src/Fluid/Flu_Close.cpp, compiled reference of out's precision, the given
    scheme and RSOLVER = main_rsolver) on the same arguments as correct_unphysical.  Returns (out, flux, (ncorr, failed)): ncorr is what
    the call adds to NCorrUnphy (0 when a cell failed), failed is 1 when FluStatus_ThisRank was set to
    GAMER_FAILED.  A cell that stays unphysical with auto_reduce_dt False makes the reference abort the process,
    so callers check such inputs with the restatement first."""
    u_in = np.ascontiguousarray(u_in)
    out = np.array(out, copy=True)
    flux = None if flux is None else np.array(flux, copy=True)
    cnt = np.zeros(2, dtype=np.int64)
    # CorrectUnphysical reads the first-order solver from OPT__1ST_FLUX_CORR_SCHEME at run time; the build's
    # RSOLVER decides whether its Roe solver checks intermediate states (include/CUFLU.h:178-186)
    L = ref_lib(scheme, main_rsolver, _suffix(out))
    rc = L.gamer_ref_correct_unphysical(_ptr(u_in), _ptr(out), _ptr(flux), out.shape[0], dt, dh, gamma, _FFC[mode],
                                        _RSOLVER_1ST[rsolver], int(last_resort_floor), int(auto_reduce_dt),
                                        int(fixup_flux), int(check_pres), min_dens, min_pres, min_eint,
                                        _ptr(cnt), nthreads)
    if rc != 0:
        raise ValueError(f"gamer_ref_correct_unphysical returned {rc}")
    return out, flux, cnt


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def store_flux(flux, slot, dt, reg):
    flux, slot, reg = np.ascontiguousarray(flux), _i32(slot), np.array(reg, dtype=np.float64, copy=True)
    getattr(lib(), "oracle_store_flux_" + _suffix(flux))(_ptr(flux), flux.shape[0], _ptr(slot), ctypes.c_double(dt),
                                                         _ptr(reg))
    return reg


def correct_flux(flux, slot, dt, reg):
    flux, slot, reg = np.ascontiguousarray(flux), _i32(slot), np.array(reg, dtype=np.float64, copy=True)
    getattr(lib(), "oracle_correct_flux_" + _suffix(flux))(_ptr(flux), flux.shape[0], _ptr(slot),
                                                           ctypes.c_double(dt), _ptr(reg))
    return reg


def fixup_flux(patch, slot, reg, dh, gamma=5.0 / 3.0, min_dens=0.0, min_pres=0.0, min_eint=0.0):
    patch, slot, reg = np.array(patch, copy=True), _i32(slot), np.ascontiguousarray(reg, dtype=np.float64)
    getattr(lib(), "oracle_fixup_flux_" + _suffix(patch))(_ptr(patch), patch.shape[0], _ptr(slot), _ptr(reg),
                                                          ctypes.c_double(dh), ctypes.c_double(gamma),
                                                          ctypes.c_double(min_dens), ctypes.c_double(min_pres),
                                                          ctypes.c_double(min_eint))
    return patch


def fixup_restrict(son, fa, son0, fa_idx, min_eint=0.0):
    son, fa = np.ascontiguousarray(son), np.array(fa, copy=True)
    son0, fa_idx = _i32(son0), _i32(fa_idx)
    getattr(lib(), "oracle_fixup_restrict_" + _suffix(fa))(_ptr(son), _ptr(fa), son0.shape[0], _ptr(son0),
                                                           _ptr(fa_idx), ctypes.c_double(min_eint))
    return fa


def close_patches(pg_out, patch, pid0):
    pg_out, patch, pid0 = np.ascontiguousarray(pg_out), np.array(patch, copy=True), _i32(pid0)
    getattr(lib(), "oracle_close_patches_" + _suffix(patch))(_ptr(pg_out), _ptr(patch), pg_out.shape[0], _ptr(pid0))
    return patch


# GAMER's own StoreFlux / CorrectFlux / Flu_Close copy-back (src/Fluid/Flu_Close.cpp) and Flu_FixUp_Flux /
# Flu_FixUp_Restrict (src/Fluid/Flu_FixUp_*.cpp) on small patch trees built from the same slot tables, in the
# compiled reference of the arrays' precision (its flux arrays are `real`: registers pass through that type)
def _ref_sync(a):
    return ref_lib("MHM_RP", "HLLC", _suffix(a))


def _rc(name, rc):
    if rc != 0:
        raise ValueError(f"gamer_ref_{name} returned {rc}")


def ref_store_flux(flux, slot, dt, reg):
    flux, slot, reg = np.ascontiguousarray(flux), _i32(slot), np.array(reg, dtype=np.float64, copy=True)
    _rc("store_flux", _ref_sync(flux).gamer_ref_store_flux(_ptr(flux), flux.shape[0], _ptr(slot), dt, _ptr(reg)))
    return reg


def ref_correct_flux(flux, slot, dt, reg):
    flux, slot, reg = np.ascontiguousarray(flux), _i32(slot), np.array(reg, dtype=np.float64, copy=True)
    _rc("correct_flux", _ref_sync(flux).gamer_ref_correct_flux(_ptr(flux), flux.shape[0], _ptr(slot), dt, _ptr(reg)))
    return reg


def ref_fixup_flux(patch, slot, reg, dh, gamma=5.0 / 3.0, min_dens=0.0, min_pres=0.0, min_eint=0.0):
    patch, slot, reg = np.array(patch, copy=True), _i32(slot), np.ascontiguousarray(reg, dtype=np.float64)
    _rc("fixup_flux", _ref_sync(patch).gamer_ref_fixup_flux(_ptr(patch), patch.shape[0], _ptr(slot), _ptr(reg), dh,
                                                            gamma, min_dens, min_pres, min_eint))
    return patch


def ref_fixup_restrict(son, fa, son0, fa_idx, min_eint=0.0):
    son, fa = np.ascontiguousarray(son), np.array(fa, copy=True)
    son0, fa_idx = _i32(son0), _i32(fa_idx)
    _rc("fixup_restrict", _ref_sync(fa).gamer_ref_fixup_restrict(_ptr(son), _ptr(fa), son0.shape[0], _ptr(son0),
                                                                 _ptr(fa_idx), fa.shape[0], min_eint))
    return fa


def ref_close_patches(pg_out, patch, pid0):
    pg_out, patch, pid0 = np.ascontiguousarray(pg_out), np.array(patch, copy=True), _i32(pid0)
    _rc("close_patches", _ref_sync(patch).gamer_ref_close_patches(_ptr(pg_out), _ptr(patch), pg_out.shape[0],
                                                                  _ptr(pid0), patch.shape[0]))
    return patch


_INT = {"MINMOD3D": 1, "MINMOD1D": 2, "VANLEER": 3, "CQUAD": 4, "QUAD": 5, "CQUAR": 6, "QUAR": 7}


def interpolate(cdata, fdata, cstart, crange, fstart=(0, 0, 0), scheme="CQUAD", monotonic=0x1F, mono_coeff=2.0,
                opp_sign_0th=False, int_prim=True, reduce_mono=True, mono_max_iter=10, gamma=5.0 / 3.0, min_pres=0.0):
    """Restatement of Interpolate(): returns (cdata after the call, fdata, number of failed boxes); (x, y, z) tuples."""
    c, f = np.array(cdata, copy=True), np.array(fdata, copy=True)
    i3 = lambda t: np.ascontiguousarray(t, dtype=np.int32)
    nf = np.zeros(1, dtype=np.int32)
    getattr(lib(), "oracle_interpolate_" + _suffix(f))(
        _ptr(c), _ptr(f), c.shape[0], _ptr(i3(c.shape[:1:-1])), _ptr(i3(cstart)), _ptr(i3(crange)),
        _ptr(i3(f.shape[:1:-1])), _ptr(i3(fstart)), _INT[scheme], int(monotonic), ctypes.c_double(mono_coeff),
        int(opp_sign_0th), int(int_prim), int(reduce_mono), int(mono_max_iter), ctypes.c_double(gamma),
        ctypes.c_double(min_pres), _ptr(nf))
    return c, f, int(nf[0])


def prepare_patches(patch, pg_in, slot, ghost):
    patch, pg_in, slot = np.ascontiguousarray(patch), np.array(pg_in, copy=True), _i32(slot)
    getattr(lib(), "oracle_prepare_patches_" + _suffix(pg_in))(_ptr(patch), _ptr(pg_in), pg_in.shape[0], _ptr(slot),
                                                               int(ghost))
    return pg_in


def gather_boxes(patch, box, slot, origin):
    """Restatement of the coarse-box gather: returns a copy of box (nbox, 5, sz, sy, sx) filled from the patches;
    slot (nbox, tz, ty, tx); origin (x, y, z)."""
    patch, box, slot = np.ascontiguousarray(patch), np.array(box, copy=True), _i32(slot)
    i3 = lambda t: np.ascontiguousarray(t, dtype=np.int32)
    getattr(lib(), "oracle_gather_boxes_" + _suffix(box))(_ptr(patch), _ptr(box), box.shape[0],
                                                          _ptr(i3(box.shape[:1:-1])), _ptr(i3(origin)), _ptr(slot),
                                                          _ptr(i3(slot.shape[:0:-1])))
    return box


def assemble_patch_groups(patch, fbox, pg_in, slot, ghost, fb_origin):
    patch, fbox, pg_in, slot = np.ascontiguousarray(patch), np.ascontiguousarray(fbox), np.array(pg_in, copy=True), _i32(slot)
    i3 = lambda t: np.ascontiguousarray(t, dtype=np.int32)
    getattr(lib(), "oracle_assemble_patch_groups_" + _suffix(pg_in))(
        _ptr(patch), _ptr(fbox), _ptr(pg_in), pg_in.shape[0], _ptr(slot), int(ghost), _ptr(i3(fbox.shape[:1:-1])),
        _ptr(i3(fb_origin)))
    return pg_in
