"""GPU: one launch over the patch groups of several AMR levels (per-PG dt, dh), the persistent work-queue mode,
HIP-graph capture of a batched substep, and the launch autotuner — all checked against the oracle per level."""
import numpy as np
import pytest

from parity import rel_err_cells, flux_rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda:0"
COUNTS = (8, 16, 24)          # patch groups on levels 0, 1, 2


def _batch(scheme="MHM_RP", dtype=None, interleave=True, counts=COUNTS):
    from gamer_amd import amr
    return amr.make_level_batch(torch, counts=counts, seed=77, dtype=dtype or torch.float64,
                                scheme_ghost=3 if scheme == "MHM_RP" else 2, interleave=interleave)


def _oracle_per_level(b, scheme, rsolver="HLLC"):
    """Oracle run level by level with that level's scalar (dt, dh), scattered back into launch order."""
    import oracle
    u = b.u.cpu().numpy()
    out = np.empty((u.shape[0], 5, 16, 16, 16), u.dtype)
    flux = np.empty((u.shape[0], 9, 5, 16, 16), u.dtype)
    for l in range(len(b.npg_per_level)):
        sel = np.nonzero(b.level == l)[0]
        o, f = oracle.fluid_solver(u[sel], b.dt[l], b.dh[l], scheme=scheme, rsolver=rsolver)
        out[sel], flux[sel] = o, f
    return out, flux


@pytest.mark.parametrize("scheme", ["MHM_RP", "MHM"])
@pytest.mark.parametrize("persistent", [False, True])
def test_batch_levels_match_oracle_fp64(scheme, persistent):
    from gamer_amd import hydro
    print(f"[gpu] batch levels {scheme} persistent={persistent}", flush=True)
    b = _batch(scheme)
    cnt = torch.zeros(1, dtype=torch.int32, device=DEV) if persistent else None
    out, flux = hydro.solve_batch(b.u, b.pg_dt_dh, queue_counter=cnt, groups_per_cu=2, scheme=scheme)
    torch.cuda.synchronize()
    ref_o, ref_f = _oracle_per_level(b, scheme)
    err = rel_err_cells(out.cpu().numpy(), ref_o)
    assert err.max() <= 1e-12, err.max()
    assert flux_rel_err(flux.cpu().numpy(), ref_f, ref_o) <= 1e-12


@pytest.mark.parametrize("scheme,rsolver", [("MHM_RP", "HLLC"), ("MHM", "HLLC"), ("MHM_RP", "ROE")])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_persistent_equals_grid_bitwise(dtype, scheme, rsolver):
    """The work-queue kernel runs the same per-PG code as the one-workgroup-per-PG kernel: identical bits (HLLC
    kernels march two planes per loop iteration, Roe kernels one)."""
    from gamer_amd import hydro
    print(f"[gpu] persistent vs grid {dtype} {scheme} {rsolver}", flush=True)
    b = _batch(scheme, dtype=dtype, counts=(64, 128, 192))
    kw = dict(scheme=scheme, rsolver=rsolver)
    o1, f1 = hydro.solve_batch(b.u, b.pg_dt_dh, **kw)
    cnt = torch.zeros(1, dtype=torch.int32, device=DEV)
    shapes = (4, 8) if scheme == "MHM_RP" else (4,)       # workgroup shapes of the kernel (waves per group)
    for w in shapes:
        for g, k in ((1, 1), (3, 1), (2, 3), (1, 4), (4, 64)):
            o2, f2 = hydro.solve_batch(b.u, b.pg_dt_dh, queue_counter=cnt, groups_per_cu=g, pgs_per_ticket=k,
                                       waves_per_group=w, **kw)
            torch.cuda.synchronize()
            assert torch.equal(o1, o2) and torch.equal(f1, f2), (w, g, k)


def test_batch_uniform_equals_fluid_solver():
    """pg_dt_dh = None falls back to the scalar dt, dh: same bits as the plain entry point."""
    from gamer_amd import hydro, ic
    print("[gpu] batch scalar dt", flush=True)
    w = ic.random_sod((2, 2, 2), seed=5)
    u = torch.from_numpy(w.u).to(DEV)
    o1, f1 = hydro.fluid_solver(u, w.dt, w.dh)
    o2, f2 = hydro.solve_batch(u, None, dt=w.dt, dh=w.dh)
    dtdh = torch.tensor([[w.dt, w.dh]] * w.npg, dtype=torch.float64, device=DEV)
    o3, f3 = hydro.solve_batch(u, dtdh, queue_counter=torch.zeros(1, dtype=torch.int32, device=DEV))
    torch.cuda.synchronize()
    assert torch.equal(o1, o2) and torch.equal(f1, f2)
    assert torch.equal(o1, o3) and torch.equal(f1, f3)


def test_batch_fused_dt_uses_each_level_dh():
    import oracle
    from gamer_amd import hydro
    print("[gpu] batch fused dt", flush=True)
    b = _batch()
    dtp = torch.empty((b.u.shape[0], 8), dtype=torch.float64, device=DEV)
    out, _ = hydro.solve_batch(b.u, b.pg_dt_dh, queue_counter=torch.zeros(1, dtype=torch.int32, device=DEV),
                               dt_patch=dtp, safety=0.4)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    npg = o.shape[0]
    patches = o.reshape(npg, 5, 2, 8, 2, 8, 2, 8).transpose(0, 2, 4, 6, 1, 3, 5, 7).reshape(npg, 8, 5, 8, 8, 8)
    got = dtp.cpu().numpy()
    for l in range(len(b.npg_per_level)):
        sel = np.nonzero(b.level == l)[0]
        ref = oracle.dt_cfl(np.ascontiguousarray(patches[sel].reshape(-1, 5, 8, 8, 8)), b.dh[l], 0.4)
        np.testing.assert_allclose(got[sel].reshape(-1), ref, rtol=1e-6)


@pytest.mark.parametrize("persistent", [True, False])
def test_graph_replays_match_eager(persistent):
    """Capture queue reset + batched solve once; every replay (with eager work in between) equals eager."""
    from gamer_amd import amr, hydro
    print(f"[gpu] graph replay persistent={persistent}", flush=True)
    b = _batch(dtype=torch.float32, counts=(64, 128, 192))
    g = amr.GraphSubstep(torch, b, groups_per_cu=2, persistent=persistent)
    ref_o, ref_f = hydro.solve_batch(b.u, b.pg_dt_dh)
    torch.cuda.synchronize()
    ref_o, ref_f = ref_o.clone(), ref_f.clone()
    for r in range(4):
        g.out.fill_(float("nan"))
        g.flux.fill_(float("nan"))
        o, f = g.replay()
        torch.cuda.synchronize()
        assert torch.equal(o, ref_o) and torch.equal(f, ref_f), r
        x = hydro.solve_batch(b.u, b.pg_dt_dh)[0]           # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(x, ref_o)


def test_autotune_picks_a_measured_config(tmp_path):
    from gamer_amd import amr, hydro
    print("[gpu] autotune", flush=True)
    b = _batch(dtype=torch.float32, counts=(64, 128, 192))
    cache = str(tmp_path / "tune.json")
    res = amr.autotune(torch, b, groups=(1, 2), tickets=(1, 2), reps=2, cache=cache)
    assert res["best"] in res["trials_ms"] and res["ms"] > 0 and len(res["trials_ms"]) == 10    # (grid + 2x2) x 2 shapes
    assert {k.split(":")[-1] for k in res["trials_ms"]} == {"4", "8"}
    assert amr.autotune(torch, b, groups=(1, 2), tickets=(1, 2), reps=2, cache=cache)["best"] == res["best"]   # cached
    o, _ = hydro.solve_batch(b.u, b.pg_dt_dh, **amr.launch_kwargs(torch, res["best"], DEV))
    ref, _ = hydro.solve_batch(b.u, b.pg_dt_dh)
    torch.cuda.synchronize()
    assert torch.equal(o, ref)
