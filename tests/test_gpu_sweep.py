"""GPU: the Roe / HLLE / HLLC sweep on the Kelvin-Helmholtz shear layer at NPG = 8192, every patch group against
the FP64 oracle: FP64 to 1e-12 per cell, FP32-bulk to 1e-5 relative L1 per field, conserved sums to round-off.
Prints one JSON line per solver with the per-field errors (the accuracy half of the sweep; bench.py --riemann-sweep
reports the throughput half)."""
import json

import numpy as np
import pytest

from gamer_amd import ic
from parity import rel_l1, rel_err_cells

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda:0"
_KH = {}


def _kh():
    if not _KH:
        _KH["w"] = ic.kelvin_helmholtz((32, 16, 16), seed=123)
    return _KH["w"]


def _scale(per_field):
    sc = np.array(per_field, dtype=np.float64)
    sc[1:4] = sc[1:4].max()          # the momentum components share one scale (m_z is zero in the 2D layer)
    return sc


@pytest.mark.parametrize("rsolver", ["HLLC", "HLLE", "ROE"])
def test_kh_sweep_accuracy(rsolver):
    import oracle
    from gamer_amd import hydro
    print(f"[gpu] KH sweep {rsolver}", flush=True)
    w = _kh()
    ref, _ = oracle.fluid_solver(w.u, w.dt, w.dh, rsolver=rsolver)
    interior = ic.pg_interior(w.u, 3)
    report = {"solver": rsolver, "npg": w.npg}
    for prec, tdt in (("fp64", torch.float64), ("fp32", torch.float32)):
        out, _ = hydro.fluid_solver(torch.from_numpy(w.u).to(DEV, dtype=tdt), w.dt, w.dh, rsolver=rsolver)
        torch.cuda.synchronize()
        o = out.cpu().numpy().astype(np.float64)
        d = np.abs(o - ref)
        drift = (o.sum(axis=(0, 2, 3, 4)) - interior.sum(axis=(0, 2, 3, 4))) / _scale(np.abs(interior).sum(axis=(0, 2, 3, 4)))
        ref_drift = (ref.sum(axis=(0, 2, 3, 4)) - interior.sum(axis=(0, 2, 3, 4))) / _scale(np.abs(interior).sum(axis=(0, 2, 3, 4)))
        report[prec] = {"rel_l1": rel_l1(o, ref).tolist(),
                        "rel_linf": (d.max(axis=(0, 2, 3, 4)) / _scale(np.abs(ref).max(axis=(0, 2, 3, 4)))).tolist(),
                        "max_cell_rel_err": float(rel_err_cells(o, ref).max()),
                        "conserved_sum_drift": drift.tolist(), "oracle_drift": ref_drift.tolist()}
        if prec == "fp64":
            assert report[prec]["max_cell_rel_err"] <= 1e-12
            assert np.all(np.abs(drift - ref_drift) <= 1e-13)
        else:
            assert max(report[prec]["rel_l1"]) <= 1e-5
    print("SWEEP " + json.dumps(report), flush=True)
