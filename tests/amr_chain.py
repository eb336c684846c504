"""The two-level step of gamer_amd.amr.TwoLevel composed from the oracle's restatements (CPU): the checker of the
device driver and, through conservation, of the composition itself."""
import numpy as np

import oracle


def int_ghost(scheme):
    return 2 if scheme in ("CQUAR", "QUAR") else 1


def two_level_step(base, fine, tab, dh, dt, int_scheme="CQUAD", rsolver="HLLC", gamma=5.0 / 3.0, fixup=True):
    """One base step + two refined substeps + restrict + flux fix-up; returns (base, fine, registers) (new arrays)."""
    dtype = base.dtype
    nf = len(tab.fathers)
    pg_b = oracle.prepare_patches(base, np.zeros((tab.npg_base, 5, 22, 22, 22), dtype), tab.base_pg_slot, 3)
    out, flux = oracle.fluid_solver(pg_b, dt, dh, rsolver=rsolver, gamma=gamma)
    out, flux, _ = oracle.correct_unphysical(pg_b, out, flux, dt, dh, gamma=gamma)
    reg = oracle.store_flux(flux, tab.base_store_slot, dt, np.zeros((max(tab.nreg, 1), 5, 8, 8)))
    old = base.copy()
    base = oracle.close_patches(out, base, tab.base_pid0)
    g = int_ghost(int_scheme)
    for sub in range(2):
        coarse = old if sub == 0 else 0.5 * old + 0.5 * base
        cbox = oracle.gather_boxes(coarse, np.zeros((nf, 5) + (12 + 2 * g,) * 3, dtype), tab.fine_coarse_slot,
                                   (6 - g,) * 3)
        _, fbox, _ = oracle.interpolate(cbox, np.zeros((nf, 5, 24, 24, 24), dtype), (g,) * 3, (12, 12, 12),
                                        scheme=int_scheme, gamma=gamma)
        pg_f = oracle.assemble_patch_groups(fine, fbox, np.zeros((nf, 5, 22, 22, 22), dtype), tab.fine_pg_slot, 3,
                                            (1, 1, 1))
        out_f, flux_f = oracle.fluid_solver(pg_f, 0.5 * dt, 0.5 * dh, rsolver=rsolver, gamma=gamma)
        out_f, flux_f, _ = oracle.correct_unphysical(pg_f, out_f, flux_f, 0.5 * dt, 0.5 * dh, gamma=gamma)
        reg = oracle.correct_flux(flux_f, tab.fine_correct_slot, 0.5 * dt, reg)
        fine = oracle.close_patches(out_f, fine, (8 * np.arange(nf)).astype(np.int32))
    base = oracle.fixup_restrict(fine, base, tab.son0, tab.fa_idx)
    if fixup:
        base = oracle.fixup_flux(base, tab.fixup_slot, reg, dh, gamma=gamma)
    return base, fine, reg


def setup(fathers=None, seed=21, dtype=np.float64):
    """Base level of 4^3 patches (32^3 cells on the unit box) and the refined patches of ``fathers``, both sampled
    from one smooth periodic field on the 64^3 fine grid; the base level is its restriction, so the two levels
    start consistent.  Returns (tables, base patches, fine patches, dh, dt)."""
    import amr_mesh as M
    from gamer_amd import amr
    if fathers is None:
        fathers = [(i, j, k) for k in (1, 2) for j in (1, 2) for i in (1, 2)]
    tab = amr.build_two_level_tables((4, 4, 4), fathers)
    fg = M.gas((64, 64, 64), np.random.default_rng(seed))
    base = amr.patches_from_grid(fg.reshape(5, 32, 2, 32, 2, 32, 2).mean(axis=(2, 4, 6))).astype(dtype)
    fpg = amr.patches_from_grid(fg)                         # the whole fine grid as patches of an 8^3-patch level
    fine = np.empty((8 * len(fathers), 5, 8, 8, 8), dtype)
    for q, (I, J, K) in enumerate(fathers):
        for l, (a, b, c) in enumerate(amr.LOCAL_XYZ):
            fine[8 * q + l] = fpg[amr.base_patch_index((8, 8, 8), 2 * I + a, 2 * J + b, 2 * K + c)]
    return tab, np.ascontiguousarray(base), fine, 1.0 / 32.0, 5e-3
