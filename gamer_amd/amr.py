"""Uneven multi-level patch-group batches: all AMR levels in one solver launch.

GAMER advances each level with its own ``CUAPI_Asyn_FluidSolver`` calls (src/Main/EvolveLevel.cpp:203), one
level and substep at a time.  Here the patch groups of several levels, each with its own (dt, dh), go to one
launch of the solver (``hydro_mhm_solve_batch_async``), optionally as a persistent work-queue kernel, and the
whole substep (queue reset + solve) can be captured once into a HIP graph and replayed.  ``autotune`` picks the
launch mode and the number of resident workgroups per compute unit by measurement.
"""
from __future__ import annotations

import dataclasses
import json
import os
import time

import numpy as np

from . import hydro, ic


@dataclasses.dataclass
class LevelBatch:
    u: object              # device tensor (npg, 5, N, N, N)
    pg_dt_dh: object       # device float64 (npg, 2)
    level: np.ndarray      # host int (npg,): level of each patch group
    dt: list               # per level
    dh: list               # per level
    npg_per_level: list


def make_level_batch(torch, counts=(512, 2048, 6144), seed=2026, dtype=None, scheme_ghost=3, interleave=True,
                     device="cuda:0", cfl=0.5):
    """Seeded synthetic hierarchy: level l holds counts[l] patch groups (a periodic block of 16^3-cell PGs,
    dh_l = dh_0 / 2^l) of the random-medium + Sod workload; dt_l = dt_0 / 2^l (refinement ratio 2).  With
    ``interleave`` the PGs of all levels are mixed round-robin in the launch order, otherwise level by level."""
    dtype = dtype or torch.float32
    dh0 = 1.0 / 128.0
    blocks, dh, dt_lim = [], [], []
    for l, n in enumerate(counts):
        shape = _factor3(n)
        cons, _, _ = ic.random_sod_grid(shape, seed=seed + l, sedov=True, cfl=cfl)
        h = dh0 / (2 ** l)
        blocks.append(ic.cut_patch_groups(cons, scheme_ghost))
        dh.append(h)
        dt_lim.append(ic.cfl_dt(cons, h, cfl) * (2 ** l))     # level-0 dt this level's CFL condition allows
    dt0 = min(dt_lim)
    dt = [dt0 / (2 ** l) for l in range(len(counts))]
    level = np.concatenate([np.full(n, l, dtype=np.int32) for l, n in enumerate(counts)])
    u = np.concatenate(blocks)
    order = np.arange(len(level))
    if interleave:
        order = _round_robin(counts)
    u, level = u[order], level[order]
    dtdh = np.stack([np.array(dt)[level], np.array(dh)[level]], axis=1)
    return LevelBatch(torch.from_numpy(np.ascontiguousarray(u)).to(device, dtype=dtype),
                      torch.from_numpy(dtdh).to(device), level, dt, dh, list(counts))


def _factor3(n):
    """(nz, ny, nx) with nz*ny*nx = n, as cubic as possible, nx >= ny >= nz."""
    best = None
    for nz in range(1, int(round(n ** (1 / 3))) + 2):
        if n % nz:
            continue
        m = n // nz
        for ny in range(nz, int(m ** 0.5) + 1):
            if m % ny == 0:
                nx = m // ny
                cand = (nz, ny, nx)
                if best is None or max(cand) < max(best):
                    best = cand
    return best


def _round_robin(counts):
    """Launch order that interleaves the levels in proportion to their sizes."""
    starts = np.cumsum([0] + list(counts[:-1]))
    keys = np.concatenate([(np.arange(n) + 0.5) / n for n in counts])
    idx = np.concatenate([s + np.arange(n) for s, n in zip(starts, counts)])
    return idx[np.argsort(keys, kind="stable")]


class GraphSubstep:
    """A captured substep: queue reset + one batched solver launch, replayed with ``replay()``."""

    def __init__(self, torch, batch: LevelBatch, groups_per_cu=2, persistent=True, **kw):
        self.torch = torch
        self.batch = batch
        npg = batch.u.shape[0]
        self.out = torch.empty((npg, 5, 16, 16, 16), dtype=batch.u.dtype, device=batch.u.device)
        self.flux = torch.empty((npg, 9, 5, 16, 16), dtype=batch.u.dtype, device=batch.u.device)
        self.counter = torch.zeros(1, dtype=torch.int32, device=batch.u.device) if persistent else None
        self.kw = dict(groups_per_cu=groups_per_cu, **kw)
        self.run_eager()                                   # warm-up (module load, first-launch setup)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.run_eager()

    def run_eager(self):
        return hydro.solve_batch(self.batch.u, self.batch.pg_dt_dh, out=self.out, flux=self.flux,
                                 queue_counter=self.counter, **self.kw)

    def replay(self):
        self.graph.replay()
        return self.out, self.flux


def time_launch(torch, fn, reps=5, warmup=1):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(torch, batch: LevelBatch, groups=(1, 2, 3, 4), reps=5, cache=None, **kw):
    """Measure every launch configuration of a batched substep and return the fastest.

    Configurations: one workgroup per patch group ("grid"), and the persistent work queue with g workgroups
    per compute unit for g in ``groups``.  Results are cached in ``cache`` (a JSON path) keyed by the batch
    shape and precision when given."""
    npg = batch.u.shape[0]
    key = f"{torch.cuda.get_device_name(0)}|npg={npg}|{batch.u.dtype}|{kw.get('scheme', 'MHM_RP')}|{kw.get('rsolver', 'HLLC')}"
    if cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                table = json.load(f)
            if key in table:
                return table[key]
        except (OSError, ValueError):
            pass
    out = torch.empty((npg, 5, 16, 16, 16), dtype=batch.u.dtype, device=batch.u.device)
    flux = torch.empty((npg, 9, 5, 16, 16), dtype=batch.u.dtype, device=batch.u.device)
    counter = torch.zeros(1, dtype=torch.int32, device=batch.u.device)
    trials = {}
    trials["grid"] = time_launch(torch, lambda: hydro.solve_batch(batch.u, batch.pg_dt_dh, out=out, flux=flux, **kw),
                                 reps)
    for g in groups:
        trials[f"persistent:{g}"] = time_launch(
            torch, lambda g=g: hydro.solve_batch(batch.u, batch.pg_dt_dh, out=out, flux=flux, queue_counter=counter,
                                                 groups_per_cu=g, **kw), reps)
    best = min(trials, key=trials.get)
    res = {"best": best, "ms": trials[best], "trials_ms": trials, "measured_at": time.time()}
    if cache:
        try:
            table = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    table = json.load(f)
            table[key] = res
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, "w") as f:
                json.dump(table, f, indent=1)
        except (OSError, ValueError):
            pass
    return res
