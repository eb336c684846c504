"""Error metrics shared by the parity tests."""
import numpy as np


def rel_err_cells(out, ref):
    """Per-cell relative error of a (npg, 5, 16, 16, 16) result.

    Density and energy are compared relative to themselves; the three momentum components relative to the
    cell's momentum scale sqrt(rho * E) (>= |m|/sqrt(2)), so a component that happens to cross zero does not
    divide by ~0.  Returns the max over cells of each field.
    """
    out = np.asarray(out, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = np.abs(out - ref)
    scale = np.empty_like(ref)
    scale[:, 0] = np.abs(ref[:, 0])
    scale[:, 4] = np.abs(ref[:, 4])
    ms = np.sqrt(np.abs(ref[:, 0] * ref[:, 4]))
    for v in (1, 2, 3):
        scale[:, v] = ms
    return (d / scale).reshape(ref.shape[0], 5, -1).max(axis=(0, 2))


def rel_l1(out, ref):
    """Per-field relative L1 error: sum|out-ref| / sum|ref| (momentum normalised by sum sqrt(rho E))."""
    out = np.asarray(out, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = np.abs(out - ref).sum(axis=(0, 2, 3, 4))
    s = np.abs(ref).sum(axis=(0, 2, 3, 4))
    ms = np.sqrt(np.abs(ref[:, 0] * ref[:, 4])).sum()
    s[1:4] = ms
    return d / s


def flux_rel_err(flux, ref_flux, ref_out):
    """Max flux error relative to the flux scale sqrt(rho E) * c_s ~ |F| of the patch group."""
    flux = np.asarray(flux, dtype=np.float64)
    ref_flux = np.asarray(ref_flux, dtype=np.float64)
    scale = np.abs(ref_flux).reshape(ref_flux.shape[0], -1).max(axis=1)
    d = np.abs(flux - ref_flux).reshape(flux.shape[0], -1).max(axis=1)
    return float((d / scale).max())
