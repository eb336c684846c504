"""GPU: coarse -> fine interpolation (all seven schemes of src/Interpolation except INT_SPECTRAL, with and without the
iterative primitive / monotonic-coefficient retries) and the same-level patch gather, against the oracle restatement
(pinned in test_oracle_interp.py)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda:0"
G = 5.0 / 3.0


def _gas_boxes(rng, nbox, cs):
    shp = (nbox,) + tuple(cs[::-1])
    rho = np.exp(0.3 * rng.standard_normal(shp))
    v = 0.3 * rng.standard_normal((3,) + shp)
    p = np.exp(0.3 * rng.standard_normal(shp))
    c = np.stack([rho, rho * v[0], rho * v[1], rho * v[2], p / (G - 1) + 0.5 * rho * (v ** 2).sum(0)], axis=1)
    # every third box gets a fast cold cell next to a slow one: conserved interpolation fails there
    for q in range(0, nbox, 3):
        c[q, :, 3, 3, 3] = [1.0, 4.0, 0.0, 0.0, 8.0 + 1e-3]
        c[q, :, 3, 3, 4] = [1.0, 0.0, 0.0, 0.0, 1e-3]
    return c


CASES = [("CQUAD", True), ("CQUAD", False), ("MINMOD1D", False), ("MINMOD1D", "prim"),
         ("MINMOD3D", False), ("MINMOD3D", "prim"), ("VANLEER", False), ("VANLEER", "prim"),
         ("QUAD", True), ("QUAD", False), ("CQUAR", True), ("CQUAR", False), ("QUAR", True), ("QUAR", False)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("scheme,iterate", CASES)
def test_interpolate_matches_oracle(dtype, scheme, iterate):
    """iterate: True = primitive retry + monotonic-coefficient reduction, "prim" = primitive retry only."""
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] interpolate {scheme} iterate={iterate} {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(5)
    g = 2 if scheme in ("CQUAR", "QUAR") else 1
    cr = (5, 4, 6)
    cs = tuple(r + 2 * g for r in cr)
    c = _gas_boxes(rng, 12, cs).astype(dtype)
    f0 = rng.standard_normal((12, 5, 12, 8, 10)).astype(dtype)
    kw = dict(scheme=scheme, int_prim=bool(iterate), reduce_mono=(iterate is True), mono_coeff=2.0, mono_max_iter=10)
    rc, rf, rn = oracle.interpolate(c, f0, (g,) * 3, cr, **kw)
    gc, gf = torch.from_numpy(c).to(DEV), torch.from_numpy(f0).to(DEV)
    nf = hydro.interpolate(gc, gf, (g,) * 3, cr, **kw)
    torch.cuda.synchronize()
    assert int(nf.item()) == rn
    tol = 1e-14 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(gf.cpu().numpy(), rf, rtol=tol, atol=tol)
    np.testing.assert_allclose(gc.cpu().numpy(), rc, rtol=tol, atol=tol)      # primitive retry converts in place
    # a caller-held running count (the AMR step's) is added to, with the same result
    run = torch.full((1,), 3, dtype=torch.int32, device=DEV)
    gf2 = torch.from_numpy(f0).to(DEV)
    assert hydro.interpolate(torch.from_numpy(c).to(DEV), gf2, (g,) * 3, cr, nfail=run, **kw) is run
    torch.cuda.synchronize()
    assert int(run.item()) == 3 + rn and torch.equal(gf2, gf)


def test_interpolate_fine_offset_without_iteration():
    import oracle
    from gamer_amd import hydro
    print("[gpu] interpolate with fstart", flush=True)
    rng = np.random.default_rng(6)
    c = _gas_boxes(rng, 4, (6, 6, 6))
    f0 = rng.standard_normal((4, 5, 11, 11, 11))
    kw = dict(scheme="CQUAD", int_prim=False, reduce_mono=False, opp_sign_0th=True, monotonic=0b10101)
    _, rf, _ = oracle.interpolate(c, f0, (1, 2, 1), (4, 3, 4), fstart=(1, 3, 2), **kw)
    gf = torch.from_numpy(f0).to(DEV)
    hydro.interpolate(torch.from_numpy(c).to(DEV), gf, (1, 2, 1), (4, 3, 4), fstart=(1, 3, 2), **kw)
    torch.cuda.synchronize()
    np.testing.assert_allclose(gf.cpu().numpy(), rf, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("ghost", [3, 2])
def test_prepare_patches_matches_oracle(ghost):
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] prepare_patches ghost={ghost}", flush=True)
    rng = np.random.default_rng(8)
    patch = rng.standard_normal((90, 5, 8, 8, 8))
    slot = rng.integers(-1, 90, size=(6, 4, 4, 4)).astype(np.int32)
    n = 16 + 2 * ghost
    pg0 = rng.standard_normal((6, 5, n, n, n))
    ref = oracle.prepare_patches(patch, pg0, slot, ghost)
    got = torch.from_numpy(pg0).to(DEV)
    hydro.prepare_patches(torch.from_numpy(patch).to(DEV), got, torch.from_numpy(slot).to(DEV), ghost)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got.cpu().numpy(), ref)


@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("scheme", ["CQUAD", "QUAR", "MINMOD1D"])
def test_refined_patch_groups_on_device(dtype, scheme, masked):
    """Prepare_PatchData of 8 fine patch groups at a refinement boundary entirely on the device (coarse boxes from the
    father + 26 siblings -> Interpolate -> same-level fine patches), against the oracle chain; then one MHM_RP step
    of the assembled inputs against the oracle solver."""
    import oracle
    import amr_mesh as M
    from gamer_amd import hydro
    print(f"[gpu] refined patch groups {scheme} {np.dtype(dtype).name} masked={masked}", flush=True)
    rng = np.random.default_rng(15)
    cpatch = M.patches_of(M.gas((32, 32, 32), rng, smooth=False)).astype(dtype)
    fpatch = M.patches_of(M.gas((32, 32, 32), rng, smooth=False)).astype(dtype)
    cslot, fslot = M.coarse_slots(), M.fine_slots()
    npg, g = len(M.fathers()), (2 if scheme in ("CQUAR", "QUAR") else 1)
    kw = dict(int_prim=True, reduce_mono=(scheme != "MINMOD1D"))
    cbox = oracle.gather_boxes(cpatch, np.zeros((npg, 5) + (12 + 2 * g,) * 3, dtype), cslot, (6 - g,) * 3)
    _, fbox, rn = oracle.interpolate(cbox, np.zeros((npg, 5, 24, 24, 24), dtype), (g,) * 3, (12, 12, 12), scheme=scheme,
                                     **kw)
    ref = oracle.assemble_patch_groups(fpatch, fbox, np.zeros((npg, 5, 22, 22, 22), dtype), fslot, 3, (1, 1, 1))
    pg = torch.zeros((npg, 5, 22, 22, 22), dtype=torch.from_numpy(cpatch).dtype, device=DEV)
    need = None
    if masked:                          # interpolate only under window cells without a same-level patch
        from gamer_amd import amr
        need = torch.from_numpy(amr.refined_interp_need(fslot)).to(DEV)
    nf = hydro.prepare_refined_patch_groups(torch.from_numpy(cpatch).to(DEV), torch.from_numpy(fpatch).to(DEV),
                                            torch.from_numpy(cslot).to(DEV), torch.from_numpy(fslot).to(DEV), pg,
                                            scheme=scheme, need=need, **kw)
    torch.cuda.synchronize()
    assert int(nf.item()) <= rn if masked else int(nf.item()) == rn
    tol = 1e-14 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(pg.cpu().numpy(), ref, rtol=tol, atol=tol)
    if dtype == np.float64:
        dt, dh = 2e-3, 1.0 / 64.0
        out, _ = hydro.fluid_solver(pg, dt, dh)
        rout, _ = oracle.fluid_solver(ref, dt, dh)
        torch.cuda.synchronize()
        np.testing.assert_allclose(out.cpu().numpy(), rout, rtol=1e-12, atol=1e-12)
