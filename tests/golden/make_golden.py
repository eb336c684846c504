"""Write the golden fixtures of the MHM/MHM_RP solver from GAMER's own CPU_FluidSolver.

Run in a container where the reference tree is present and `make -C oracle ref` has built oracle/_ref:
    python tests/golden/make_golden.py
Each case regenerates its input from the seeded generator gamer_amd.ic (deterministic bytes, recorded as a
SHA-256), runs the compiled reference, and stores: SHA-256 of the out/flux bytes, per-field sums and a fixed
sample of 256 output and 128 flux values at full precision (golden.json).  One case also stores the complete
arrays (golden_rp_hllc_f64.npz).  The fixtures are data only: inputs are regenerated from the seed.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from gamer_amd import ic  # noqa: E402


def cases():
    """(name, generator kwargs, solver kwargs, precision)"""
    out = []
    for scheme, ghost in (("MHM_RP", 3), ("MHM", 2)):
        for rs in ("HLLC", "HLLE", "ROE"):
            for prec in ("f64", "f32"):
                out.append((f"{scheme.lower()}_{rs.lower()}_{prec}",
                            dict(npg_shape=(1, 1, 2), seed=101, scheme_ghost=ghost, sedov=True),
                            dict(scheme=scheme, rsolver=rs), prec))
    for lim in ("VANLEER", "GMINMOD", "ALBADA", "CENTRAL"):
        out.append((f"mhm_rp_hllc_f64_{lim.lower()}", dict(npg_shape=(1, 1, 2), seed=202, scheme_ghost=3),
                    dict(limiter=lim), "f64"))
    # unphysical full-step results at CFL 1.2 -> MinMod coefficient retry loop (CPU_FluidSolver_MHM.cpp:467-584)
    out.append(("mhm_rp_hllc_f64_retry", dict(npg_shape=(1, 1, 2), seed=31, scheme_ghost=3, sedov=True, cfl=1.2,
                                              sigma_v=1.5), dict(minmod_maxiter=4), "f64"))
    out.append(("mhm_rp_roe_f64_retry", dict(npg_shape=(1, 1, 2), seed=31, scheme_ghost=3, sedov=True, cfl=1.2,
                                             sigma_v=1.5), dict(rsolver="ROE", minmod_maxiter=3), "f64"))
    # density / pressure / internal-energy floors active
    out.append(("mhm_rp_hllc_f64_floors", dict(npg_shape=(1, 1, 2), seed=31, scheme_ghost=3, sedov=True, cfl=1.6,
                                               sigma_v=1.5),
                dict(min_dens=1e-2, min_pres=1e-3, min_eint=1e-3), "f64"))
    # the C1 configuration itself: NPG = 64, MHM_RP + VL_GMINMOD + HLLC, FP64
    out.append(("c1_mhm_rp_hllc_f64", dict(npg_shape=(4, 4, 4), seed=1234, scheme_ghost=3, sedov=False),
                dict(), "f64"))
    return out


def make_input(gen, prec):
    w = ic.random_sod(**gen)
    u = w.u if prec == "f64" else w.u.astype(np.float32)
    return w, np.ascontiguousarray(u)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sample_idx(size, n, seed):
    return np.random.Generator(np.random.PCG64(seed)).choice(size, size=n, replace=False)


def main():
    import oracle
    golden = {}
    for name, gen, kw, prec in cases():
        w, u = make_input(gen, prec)
        out, flux = oracle.ref_fluid_solver(u, w.dt, w.dh, **kw)
        oi = sample_idx(out.size, 256, 1)
        fi = sample_idx(flux.size, 128, 2)
        golden[name] = {
            "generator": {k: list(v) if isinstance(v, tuple) else v for k, v in gen.items()},
            "solver": kw, "precision": prec, "dt": w.dt, "dh": w.dh,
            "input_sha256": sha(u), "out_sha256": sha(out), "flux_sha256": sha(flux),
            "out_field_sums": [float(x) for x in out.astype(np.float64).sum(axis=(0, 2, 3, 4))],
            "out_sample_idx": oi.tolist(), "out_sample": [float(x) for x in out.ravel()[oi]],
            "flux_sample_idx": fi.tolist(), "flux_sample": [float(x) for x in flux.ravel()[fi]],
        }
        print(name, golden[name]["out_sha256"][:16])
        if name == "mhm_rp_hllc_f64":
            np.savez_compressed(os.path.join(HERE, "golden_rp_hllc_f64.npz"), out=out, flux=flux)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(golden, f, indent=0)


if __name__ == "__main__":
    main()
