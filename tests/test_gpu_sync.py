"""GPU: Flu_Close and the AMR level synchronisation kernels against the oracle restatements (pinned in
test_oracle_sync.py): CorrectUnphysical (3 modes, 3 solvers, both precisions, with flux-plane replacement),
StoreFlux / CorrectFlux into FP64 registers, FixUp_Flux, FixUp_Restrict and the patch copy-back."""
import numpy as np
import pytest

from gamer_amd import ic

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda:0"
G = 5.0 / 3.0


def _dev(a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype=dtype)


def _corrupted_case(seed, dtype):
    """A real MHM_RP solve of the Sod + Sedov workload, then ~1% of the output cells made unphysical."""
    import oracle
    w = ic.random_sod((2, 2, 1), seed=seed, sedov=True)
    u = w.u.astype(dtype)
    out, flux = oracle.fluid_solver(u, w.dt, w.dh)
    rng = np.random.default_rng(seed)
    n = out.shape[0] * 4096
    cells = rng.choice(n, size=n // 100, replace=False)
    kind = rng.integers(0, 3, size=cells.size)
    o = out.reshape(out.shape[0], 5, -1)
    for c, kd in zip(cells, kind):
        p, q = divmod(int(c), 4096)
        if kd == 0:
            o[p, 0, q] = -abs(o[p, 0, q])
        elif kd == 1:
            o[p, 4, q] = 0.1 * o[p, 4, q] - 10.0
        else:
            o[p, 2, q] = np.nan
    return u, out, flux, w.dt, w.dh


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode", ["NONE", "3D", "3D1D"])
@pytest.mark.parametrize("rsolver,main", [("HLLE", None), ("HLLC", None), ("ROE", None), ("ROE", "HLLC")])
def test_correct_unphysical_matches_oracle(dtype, mode, rsolver, main):
    """main: the fluid solver's Riemann solver (None: rsolver); a first-order Roe solve checks its intermediate
    states only when it is ROE, as in GAMER builds with RSOLVER == ROE."""
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] correct_unphysical {np.dtype(dtype).name} {mode} {rsolver} main={main}", flush=True)
    u, out, flux, dt, dh = _corrupted_case(11, dtype)
    kw = dict(mode=mode, rsolver=rsolver, min_dens=1e-6, min_pres=1e-8, min_eint=1e-8, main_rsolver=main)
    ro, rf, rc = oracle.correct_unphysical(u, out, flux, dt, dh, **kw)
    go, gf = _dev(out), _dev(flux)
    cnt = hydro.correct_unphysical(_dev(u), go, gf, dt, dh, **kw)
    torch.cuda.synchronize()
    assert cnt.cpu().tolist() == rc.tolist()
    go, gf = go.cpu().numpy(), gf.cpu().numpy()
    tol = 1e-13 if dtype == np.float64 else 2e-5
    # cells that stay unphysical keep their NaN input in both (assert_allclose treats NaN == NaN as equal)
    np.testing.assert_allclose(go, ro, rtol=tol, atol=tol * np.nanmax(np.abs(ro)))
    np.testing.assert_allclose(gf, rf, rtol=tol, atol=tol * np.nanmax(np.abs(rf)))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode", ["3D", "3D1D"])
def test_correct_unphysical_mhm_scheme_matches_oracle(dtype, mode):
    """The MHM scheme's closure: the solver input has 2 ghost zones (20^3 per patch group), so the first-order
    stencils and the 3D1D neighbourhood index a 20^3 array."""
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] correct_unphysical MHM {np.dtype(dtype).name} {mode}", flush=True)
    w = ic.random_sod((2, 1, 2), seed=23, scheme_ghost=2, sedov=True)
    u = w.u.astype(dtype)
    out, flux = oracle.fluid_solver(u, w.dt, w.dh, scheme="MHM")
    rng = np.random.default_rng(23)
    o = out.reshape(out.shape[0], 5, -1)
    for c in rng.choice(o.shape[0] * 4096, size=o.shape[0] * 60, replace=False):
        p, q = divmod(int(c), 4096)
        o[p, 0, q] = -abs(o[p, 0, q])
    kw = dict(mode=mode, rsolver="HLLC", scheme="MHM", min_dens=1e-6, min_pres=1e-8, min_eint=0.3)
    ro, rf, rc = oracle.correct_unphysical(u, out, flux, w.dt, w.dh, **kw)
    go, gf = _dev(out), _dev(flux)
    cnt = hydro.correct_unphysical(_dev(u), go, gf, w.dt, w.dh, **kw)
    torch.cuda.synchronize()
    assert cnt.cpu().tolist() == rc.tolist() and rc[0] >= out.shape[0] * 60
    tol = 1e-13 if dtype == np.float64 else 2e-5
    np.testing.assert_allclose(go.cpu().numpy(), ro, rtol=tol, atol=tol * np.nanmax(np.abs(ro)))
    np.testing.assert_allclose(gf.cpu().numpy(), rf, rtol=tol, atol=tol * np.nanmax(np.abs(rf)))


def test_correct_unphysical_accumulates_into_given_counters():
    """With ``counters`` the kernel adds to the caller's tensor (the AMR step keeps one running count): two calls
    leave the sum of the oracle's counts, and the outputs equal those of the call without it."""
    import oracle
    from gamer_amd import hydro
    print("[gpu] correct_unphysical running counters", flush=True)
    u, out, flux, dt, dh = _corrupted_case(11, np.float64)
    kw = dict(mode="3D1D", rsolver="HLLE", min_dens=1e-6, min_pres=1e-8, min_eint=1e-8)
    _, _, rc = oracle.correct_unphysical(u, out, flux, dt, dh, **kw)
    run = torch.tensor([5, 7], dtype=torch.int32, device="cuda:0")
    go1, gf1, go2, gf2 = _dev(out), _dev(flux), _dev(out), _dev(flux)
    assert hydro.correct_unphysical(_dev(u), go1, gf1, dt, dh, counters=run, **kw) is run
    hydro.correct_unphysical(_dev(u), _dev(out), _dev(flux), dt, dh, counters=run, **kw)
    hydro.correct_unphysical(_dev(u), go2, gf2, dt, dh, **kw)
    torch.cuda.synchronize()
    assert run.cpu().tolist() == [5 + 2*int(rc[0]), 7 + 2*int(rc[1])]
    assert torch.equal(go1, go2) and torch.equal(gf1, gf2)
    with pytest.raises(hydro.HydroError):
        hydro.correct_unphysical(_dev(u), go1, gf1, dt, dh, counters=run.to(torch.int64), **kw)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_correct_unphysical_split_sweep_matches_oracle(dtype):
    """The 3D1D sweep path (unsplit update rejected by MIN_EINT, split accepted; see test_oracle_sync)."""
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] correct_unphysical split sweep {np.dtype(dtype).name}", flush=True)
    cons, dt, dh = ic.smooth_modes((1, 1, 1), seed=3)
    u = ic.cut_patch_groups(cons, 3)
    out, _ = oracle.fluid_solver(u, dt, dh)
    i, j, k = 6, 7, 8
    kk, jj, ii = k + 3, j + 3, i + 3
    u[0][:, kk, jj, ii] = [1.0, 1.0, 0.0, 0.0, 0.5 + 1e-4]
    for d, (dz, dy, dx) in enumerate([(0, 0, 1), (0, 1, 0), (1, 0, 0)]):
        for sign in (-1, 1):
            z, y, x = kk + sign * dz, jj + sign * dy, ii + sign * dx
            rho = u[0][0, z, y, x]
            u[0][4, z, y, x] += 0.5 * (2.0 * rho) ** 2 / rho - 0.5 * u[0][1 + d, z, y, x] ** 2 / rho
            u[0][1 + d, z, y, x] = sign * 2.0 * rho
    out[0, 0, k, j, i] = -1.0
    u, out = u.astype(dtype), out.astype(dtype)
    kw = dict(mode="3D1D", rsolver="HLLE", last_resort_floor=False, fixup_flux=False, min_eint=0.5)
    ro, _, rc = oracle.correct_unphysical(u, out, None, dt, dh, **kw)
    go = _dev(out)
    cnt = hydro.correct_unphysical(_dev(u), go, None, dt, dh, **kw)
    torch.cuda.synchronize()
    assert rc.tolist() == [1, 0] and cnt.cpu().tolist() == [1, 0]
    tol = 1e-13 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(go.cpu().numpy(), ro, rtol=tol, atol=tol)


def _slots(rng, shape, nreg):
    """Unique register slots in a table of `shape` (about half of the entries -1)."""
    t = -np.ones(int(np.prod(shape)), dtype=np.int32)
    pick = rng.choice(t.size, size=min(nreg, t.size), replace=False)
    t[pick] = np.arange(pick.size, dtype=np.int32)
    return t.reshape(shape)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_store_and_correct_flux_match_oracle(dtype):
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] store/correct flux {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(2)
    npg = 12
    flux = rng.standard_normal((npg, 9, 5, 16, 16)).astype(dtype)
    s_store = _slots(rng, (npg, 8, 6), 200)
    s_corr = _slots(rng, (npg, 6), 40)
    reg0 = rng.standard_normal((200, 5, 8, 8))
    r1 = oracle.correct_flux(flux, s_corr, 0.375, oracle.store_flux(flux, s_store, 0.75, reg0))
    reg = _dev(reg0)
    hydro.store_flux(_dev(flux), _dev(s_store), 0.75, reg)
    hydro.correct_flux(_dev(flux), _dev(s_corr), 0.375, reg)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(reg.cpu().numpy(), r1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fixup_flux_and_restrict_match_oracle(dtype):
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] fixup flux/restrict {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(3)
    npatch = 40
    rho = np.exp(0.3 * rng.standard_normal((npatch, 8, 8, 8)))
    v = 0.3 * rng.standard_normal((3, npatch, 8, 8, 8))
    p = np.exp(0.3 * rng.standard_normal((npatch, 8, 8, 8)))
    patch = np.stack([rho, rho * v[0], rho * v[1], rho * v[2], p / (G - 1) + 0.5 * rho * (v ** 2).sum(0)], axis=1)
    patch = patch.astype(dtype)
    slot = _slots(rng, (npatch, 6), 100)
    reg = 0.01 * rng.standard_normal((100, 5, 8, 8))
    reg[::7, 0] = 5.0                                      # some corrections would make rho < 0: skipped
    ref = oracle.fixup_flux(patch, slot, reg, 0.25, G, min_dens=1e-6, min_pres=1e-8, min_eint=1e-8)
    got = _dev(patch)
    hydro.fixup_flux(got, _dev(slot), _dev(reg), 0.25, G, min_dens=1e-6, min_pres=1e-8, min_eint=1e-8)
    torch.cuda.synchronize()
    tol = 1e-14 if dtype == np.float64 else 1e-6
    np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=tol, atol=tol)

    son0 = np.array([0, 8, 16, 24, 32], dtype=np.int32)
    fa_idx = np.array([3, 0, 4, 1, 2], dtype=np.int32)
    fa0 = rng.standard_normal((5, 5, 8, 8, 8)).astype(dtype)
    ref = oracle.fixup_restrict(patch, fa0, son0, fa_idx, min_eint=0.2)
    got = _dev(fa0)
    hydro.fixup_restrict(_dev(patch), got, _dev(son0), _dev(fa_idx), min_eint=0.2)
    torch.cuda.synchronize()
    np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=tol, atol=tol)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_solve_correct_close_pipeline(dtype):
    """Device pipeline of one Flu_AdvanceDt pass for 4 PGs: solve -> CorrectUnphysical -> copy into patches,
    against the oracle's pipeline."""
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] solve/correct/close {dtype}", flush=True)
    w = ic.random_sod((2, 2, 1), seed=5, sedov=True)
    npdt = np.float64 if dtype == torch.float64 else np.float32
    u = w.u.astype(npdt)
    out, flux = hydro.fluid_solver(_dev(u), w.dt, w.dh)
    cnt = hydro.correct_unphysical(_dev(u), out, flux, w.dt, w.dh, mode="3D1D")
    pid0 = np.array([24, 0, 8, 16], dtype=np.int32)
    patch = hydro.close_patches(out, torch.zeros((32, 5, 8, 8, 8), dtype=dtype, device=DEV), _dev(pid0))
    torch.cuda.synchronize()
    ro, rf = oracle.fluid_solver(u, w.dt, w.dh)
    ro, _, rc = oracle.correct_unphysical(u, ro, rf, w.dt, w.dh, mode="3D1D")
    rp = oracle.close_patches(ro, np.zeros((32, 5, 8, 8, 8), npdt), pid0)
    assert cnt.cpu().tolist() == rc.tolist()
    tol = 1e-12 if dtype == torch.float64 else 1e-4
    np.testing.assert_allclose(patch.cpu().numpy(), rp, rtol=tol, atol=tol * np.abs(rp).max())
