"""GPU parity at the sizes the benchmarks time.

* The exact throughput workload: ``ic.random_sod_grid((16, 16, 16), seed=20260, sedov=True)`` cut into 4096 patch
  groups (16.8M cells; one x1e3 pressure spike per 16 PGs), MHM_RP + PLM(VL_GMINMOD) + HLLC, StoreFlux.
  FP64: every cell within 1e-12 relative of the oracle, fluxes within 1e-12 of each PG's flux scale, the domain
  sums of the five conserved fields equal to the oracle's (and to the input's: the domain is periodic) to
  round-off, and the fused per-patch CFL dt equal to the oracle's dt solver on the updated cells to 1e-14.
  FP32-bulk: per-field relative L1 <= 1e-5 against the FP64 oracle; per-field L1 and Linf are printed.
* The same grid through the MHM scheme (20^3 inputs): FP64 per cell within 1e-12 with conserved sums to round-off,
  FP32-bulk per-field relative L1 <= 1e-5.
* The uneven 3-level batch at its stated counts 512 / 2048 / 6144 (8704 PGs in one launch, per-PG dt and dh), in
  FP32-bulk, level by level against the FP64 oracle at that level's dt, dh: outputs and coarse-fine flux arrays.
* Every golden fixture written from GAMER's own CPU_FluidSolver (tests/golden/golden.json, the NPG = 64 C1
  configuration included) through the HIP path: FP64 samples within 1e-12, FP32 within the FP32 build's
  rounding, field sums, and the one complete fixture array cell by cell.
"""
import json
import os

import numpy as np
import pytest

from gamer_amd import ic
from parity import rel_err_cells, rel_l1, flux_rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda:0"
HERE = os.path.dirname(os.path.abspath(__file__))

FP64_TOL = 1e-12
FP32_L1_TOL = 1e-5


def rel_linf(out, ref):
    """Per-field max |out - ref| over the field's max |ref| (momenta share sqrt(rho E) as their scale)."""
    out = np.asarray(out, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = np.abs(out - ref).max(axis=(0, 2, 3, 4))
    s = np.abs(ref).max(axis=(0, 2, 3, 4))
    s[1:4] = np.sqrt(np.abs(ref[:, 0] * ref[:, 4])).max()
    return d / s


def flux_rel_l1(flux, ref_flux):
    """Per-field relative L1 of (npg, 9, 5, 16, 16) flux arrays: sum |F - F_ref| / sum |F_ref| over all 9 planes."""
    d = np.abs(np.asarray(flux, np.float64) - ref_flux).sum(axis=(0, 1, 3, 4))
    return d / np.abs(ref_flux).sum(axis=(0, 1, 3, 4))


@pytest.fixture(scope="module")
def bench_workload():
    """The 4096-PG workload and the oracle's FP64 result on it (all host threads)."""
    import oracle
    cons, dt, dh = ic.random_sod_grid((16, 16, 16), seed=20260, sedov=True)
    u = ic.cut_patch_groups(cons, 3)
    print(f"[gpu] full-size workload: npg={u.shape[0]} dt={dt:.6e} dh={dh:.6e}", flush=True)
    ref_out, ref_flux = oracle.fluid_solver(u, dt, dh)
    return cons, u, dt, dh, ref_out, ref_flux


def test_bench_workload_fp64(bench_workload):
    import oracle
    from gamer_amd import hydro
    cons, u, dt, dh, ref_out, ref_flux = bench_workload
    npg = u.shape[0]
    print(f"[gpu] fp64 full-size parity npg={npg}", flush=True)
    ut = torch.from_numpy(u).to(DEV)
    dtp = torch.empty((npg, 8), dtype=torch.float64, device=DEV)
    out, flux = hydro.fluid_solver(ut, dt, dh, dt_patch=dtp, safety=0.5)
    torch.cuda.synchronize()
    o, f = out.cpu().numpy(), flux.cpu().numpy()
    del ut, out, flux
    err = rel_err_cells(o, ref_out)
    ferr = flux_rel_err(f, ref_flux, ref_out)
    print(f"  per-field max rel err {err}, flux {ferr:.3e}", flush=True)
    assert np.all(err <= FP64_TOL), err
    assert ferr <= FP64_TOL
    # conserved sums: equal to the oracle's, and (periodic domain) to the input's, to round-off
    s_gpu = o.sum(axis=(0, 2, 3, 4))
    s_ref = ref_out.sum(axis=(0, 2, 3, 4))
    s_in = cons.reshape(5, -1).sum(axis=1)
    tot = np.abs(cons).reshape(5, -1).sum(axis=1)
    print(f"  conserved-sum |gpu-oracle|/sum|u| {np.abs(s_gpu - s_ref) / tot}, "
          f"|gpu-input|/sum|u| {np.abs(s_gpu - s_in) / tot}", flush=True)
    assert np.all(np.abs(s_gpu - s_ref) <= 1e-13 * tot)
    assert np.all(np.abs(s_gpu - s_in) <= 1e-13 * tot)
    # fused CFL dt of the 8 updated patches of every PG against the oracle's dt solver on the same cells
    patches = o.reshape(npg, 5, 2, 8, 2, 8, 2, 8).transpose(0, 2, 4, 6, 1, 3, 5, 7).reshape(npg * 8, 5, 8, 8, 8)
    ref_dt = oracle.dt_cfl(np.ascontiguousarray(patches), dh, 0.5).reshape(npg, 8)
    d = dtp.cpu().numpy()
    rel = np.abs(d - ref_dt) / ref_dt
    print(f"  fused dt max rel err {rel.max():.3e}", flush=True)
    assert rel.max() <= 1e-14


def test_bench_workload_fp32_bulk(bench_workload):
    from gamer_amd import hydro
    _, u, dt, dh, ref_out, ref_flux = bench_workload
    print(f"[gpu] fp32-bulk full-size accuracy npg={u.shape[0]}", flush=True)
    ut = torch.from_numpy(u).to(DEV, dtype=torch.float32)
    out, flux = hydro.fluid_solver(ut, dt, dh)
    torch.cuda.synchronize()
    o, f = out.cpu().numpy(), flux.cpu().numpy()
    l1, linf = rel_l1(o, ref_out), rel_linf(o, ref_out)
    fl1 = flux_rel_l1(f, ref_flux)
    print(f"  per-field rel L1 {l1}\n  per-field rel Linf {linf}\n  flux per-field rel L1 {fl1}", flush=True)
    assert np.all(l1 <= FP32_L1_TOL), l1
    assert np.all(linf <= 1e-4), linf
    assert np.all(fl1 <= FP32_L1_TOL), fl1


@pytest.fixture(scope="module")
def bench_workload_mhm():
    """The same 4096-PG grid cut for the MHM scheme (ghost 2, 20^3 inputs) and the oracle's FP64 MHM result."""
    import oracle
    cons, dt, dh = ic.random_sod_grid((16, 16, 16), seed=20260, sedov=True)
    u = ic.cut_patch_groups(cons, 2)
    print(f"[gpu] full-size MHM workload: npg={u.shape[0]}", flush=True)
    ref_out, ref_flux = oracle.fluid_solver(u, dt, dh, scheme="MHM")
    return cons, u, dt, dh, ref_out, ref_flux


def test_bench_workload_mhm(bench_workload_mhm):
    """MHM (Hancock half step) on the 4096-PG workload: FP64 per cell <= 1e-12 with conserved sums to round-off,
    FP32-bulk per-field rel L1 <= 1e-5, both against the FP64 oracle."""
    from gamer_amd import hydro
    cons, u, dt, dh, ref_out, ref_flux = bench_workload_mhm
    print(f"[gpu] MHM full-size parity npg={u.shape[0]}", flush=True)
    out, flux = hydro.fluid_solver(torch.from_numpy(u).to(DEV), dt, dh, scheme="MHM")
    torch.cuda.synchronize()
    o, f = out.cpu().numpy(), flux.cpu().numpy()
    err, ferr = rel_err_cells(o, ref_out), flux_rel_err(f, ref_flux, ref_out)
    print(f"  fp64 per-field max rel err {err}, flux {ferr:.3e}", flush=True)
    assert np.all(err <= FP64_TOL), err
    assert ferr <= FP64_TOL
    tot = np.abs(cons).reshape(5, -1).sum(axis=1)
    assert np.all(np.abs(o.sum(axis=(0, 2, 3, 4)) - ref_out.sum(axis=(0, 2, 3, 4))) <= 1e-13 * tot)
    out32, _ = hydro.fluid_solver(torch.from_numpy(u).to(DEV, dtype=torch.float32), dt, dh, scheme="MHM")
    torch.cuda.synchronize()
    o32 = out32.cpu().numpy()
    l1, linf = rel_l1(o32, ref_out), rel_linf(o32, ref_out)
    print(f"  fp32-bulk per-field rel L1 {l1}, rel Linf {linf}", flush=True)
    assert np.all(l1 <= FP32_L1_TOL), l1


def test_amr_batch_fp32_full_counts():
    """C3 at 512/2048/6144 PGs, FP32-bulk, one launch, against the FP64 oracle level by level."""
    import oracle
    from gamer_amd import amr, hydro
    counts = (512, 2048, 6144)
    print(f"[gpu] C3 fp32 batch counts={counts}", flush=True)
    b = amr.make_level_batch(torch, counts=counts, seed=2026, dtype=torch.float64)
    u64 = b.u.cpu().numpy()
    u32 = b.u.float()
    del b.u
    out, flux = hydro.solve_batch(u32, b.pg_dt_dh)
    torch.cuda.synchronize()
    o, f = out.cpu().numpy(), flux.cpu().numpy()
    del u32, out, flux
    for l in range(3):
        sel = np.nonzero(b.level == l)[0]
        ro, rf = oracle.fluid_solver(u64[sel], b.dt[l], b.dh[l])
        l1, linf = rel_l1(o[sel], ro), rel_linf(o[sel], ro)
        fl1 = flux_rel_l1(f[sel], rf)
        print(f"  level {l} npg={len(sel)}: out rel L1 {l1} Linf {linf}; flux rel L1 {fl1}", flush=True)
        assert np.all(l1 <= FP32_L1_TOL), (l, l1)
        assert np.all(fl1 <= FP32_L1_TOL), (l, fl1)
        assert np.all(linf <= 1e-4), (l, linf)


GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_fixture_through_hip(name):
    """Each fixture's input (regenerated from its seed, checked against the recorded SHA-256) through the HIP path,
    against the values GAMER's own CPU_FluidSolver wrote."""
    import hashlib
    from gamer_amd import hydro
    case = GOLDEN[name]
    gen = dict(case["generator"])
    gen["npg_shape"] = tuple(gen["npg_shape"])
    w = ic.random_sod(**gen)
    u = np.ascontiguousarray(w.u if case["precision"] == "f64" else w.u.astype(np.float32))
    assert hashlib.sha256(u.tobytes()).hexdigest() == case["input_sha256"], "input generator drifted"
    print(f"[gpu] golden {name} npg={u.shape[0]}", flush=True)
    out, flux = hydro.fluid_solver(torch.from_numpy(u).to(DEV), w.dt, w.dh, **case["solver"])
    torch.cuda.synchronize()
    o = out.cpu().numpy().astype(np.float64)
    f = flux.cpu().numpy().astype(np.float64)
    os_ = o.ravel()[case["out_sample_idx"]]
    fs = f.ravel()[case["flux_sample_idx"]]
    ro, rf = np.array(case["out_sample"]), np.array(case["flux_sample"])
    sums = np.array(case["out_field_sums"])
    if case["precision"] == "f64":
        # the per-cell scale of rel_err_cells: |rho| and |E| for themselves, sqrt(rho E) for the momenta
        pg, field, cell = np.unravel_index(case["out_sample_idx"], (o.shape[0], 5, 4096))
        flat = o.reshape(o.shape[0], 5, 4096)
        scale = np.where((field == 0) | (field == 4), np.abs(ro),
                         np.sqrt(np.abs(flat[pg, 0, cell] * flat[pg, 4, cell])))
        assert np.all(np.abs(os_ - ro) <= FP64_TOL * scale)
        fpg = np.unravel_index(case["flux_sample_idx"], f.shape)[0]
        fscale = np.abs(f.reshape(f.shape[0], -1)).max(axis=1)[fpg]
        assert np.all(np.abs(fs - rf) <= FP64_TOL * fscale)
        assert np.allclose(o.sum(axis=(0, 2, 3, 4)), sums, rtol=1e-12, atol=1e-12 * np.abs(o).sum())
    else:
        # the FP32 fixture is GAMER's FP32 build: both sides round to float, differently ordered
        assert np.all(np.abs(os_ - ro) <= 1e-5 * np.abs(ro).max())
        assert np.all(np.abs(fs - rf) <= 1e-5 * np.abs(rf).max())
        assert np.allclose(o.sum(axis=(0, 2, 3, 4)), sums, rtol=1e-6, atol=1e-6 * np.abs(o).sum())


def test_full_golden_arrays_through_hip():
    """The complete fixture arrays (MHM_RP + HLLC, FP64) cell by cell."""
    from gamer_amd import hydro
    ref = np.load(os.path.join(HERE, "golden", "golden_rp_hllc_f64.npz"))
    case = GOLDEN["mhm_rp_hllc_f64"]
    gen = dict(case["generator"])
    gen["npg_shape"] = tuple(gen["npg_shape"])
    w = ic.random_sod(**gen)
    out, flux = hydro.fluid_solver(torch.from_numpy(w.u).to(DEV), w.dt, w.dh)
    torch.cuda.synchronize()
    err = rel_err_cells(out.cpu().numpy(), ref["out"])
    assert np.all(err <= FP64_TOL), err
    assert flux_rel_err(flux.cpu().numpy(), ref["flux"], ref["out"]) <= FP64_TOL
