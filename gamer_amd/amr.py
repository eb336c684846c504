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


def launch_kwargs(torch, config, device):
    """solve_batch keyword arguments of an autotuner configuration name: "grid[:w]" or "persistent:g:k[:w]", with
    w the workgroup shape in waves (absent: the library default)."""
    parts = config.split(":")
    kw = {}
    if parts[0] == "grid":
        if len(parts) > 1:
            kw["waves_per_group"] = int(parts[1])
        return kw
    _, g, k = parts[:3]
    kw = dict(queue_counter=torch.zeros(1, dtype=torch.int32, device=device), groups_per_cu=int(g),
              pgs_per_ticket=int(k))
    if len(parts) > 3:
        kw["waves_per_group"] = int(parts[3])
    return kw


def autotune(torch, batch: LevelBatch, groups=(1, 2, 3, 4), tickets=(1, 2, 4), waves=None, reps=5, rounds=3,
             cache=None, **kw):
    """Measure every launch configuration of a batched substep and return the fastest.

    Configurations: one workgroup per patch group ("grid:w"), and the persistent work queue ("persistent:g:k:w")
    with g workgroups per compute unit for g in ``groups``, each taking k consecutive patch groups per queue
    ticket for k in ``tickets`` (fewer queue atomics and barriers against a coarser tail); each in every workgroup
    shape w of ``waves`` (MHM_RP: 4 waves, every lane marching an output column and a ghost role, or 8 waves, the
    two on separate waves; MHM: 4 only).  The configurations are timed in ``rounds`` interleaved rounds of
    ``reps`` launches and each keeps its best round, so that a clock or thermal drift during the sweep does not
    pick the winner.  Results are cached in ``cache`` (a JSON path) keyed by the batch shape and precision when
    given."""
    npg = batch.u.shape[0]
    if waves is None:
        waves = (4, 8) if kw.get("scheme", "MHM_RP") == "MHM_RP" else (4,)
    key = (f"{torch.cuda.get_device_name(0)}|npg={npg}|{batch.u.dtype}|{kw.get('scheme', 'MHM_RP')}|"
           f"{kw.get('rsolver', 'HLLC')}|groups={tuple(groups)}|tickets={tuple(tickets)}|waves={tuple(waves)}")
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
    configs = {}
    for w in waves:
        configs[f"grid:{w}"] = (lambda w=w: hydro.solve_batch(batch.u, batch.pg_dt_dh, out=out, flux=flux,
                                                              waves_per_group=w, **kw))
        for g in groups:
            for k in tickets:
                configs[f"persistent:{g}:{k}:{w}"] = (lambda g=g, k=k, w=w: hydro.solve_batch(
                    batch.u, batch.pg_dt_dh, out=out, flux=flux, queue_counter=counter, groups_per_cu=g,
                    pgs_per_ticket=k, waves_per_group=w, **kw))
    trials = {name: float("inf") for name in configs}
    for _ in range(max(1, rounds)):
        for name, fn in configs.items():
            trials[name] = min(trials[name], time_launch(torch, fn, reps))
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


# ---------------------------------------------------------------------------------------------------------------
# A refined level on top of a periodic base level, advanced as GAMER's EvolveLevel does for two levels
# (src/Main/EvolveLevel.cpp: Flu_AdvanceDt of the base level, two substeps of the refined level, then
# Flu_FixUp_Restrict (12-1) and Flu_FixUp_Flux (12-3)), with every per-cell operation on the device.
# ---------------------------------------------------------------------------------------------------------------

LOCAL_XYZ = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1))  # Table_02
_LID = {o: l for l, o in enumerate(LOCAL_XYZ)}
_FACE = ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1))                           # s = 0..5


@dataclasses.dataclass
class TwoLevelTables:
    """Patch relations of a two-level mesh (host int32 arrays; what GAMER's AMR tree provides).

    Base level: ``ncp`` = (nx, ny, nz) patches, periodic, stored by patch group: patch 8*pg + LocalID, patch group
    pg = (PZ*npgy + PY)*npgx + PX covering patches (2PX + a, 2PY + b, 2PZ + c).  Refined level: one patch group per
    refined base patch (``fathers``, (x, y, z) patch coordinates), its 8 patches at 8*q + LocalID."""
    ncp: tuple
    fathers: list
    base_pg_slot: np.ndarray        # (npg_b, 4, 4, 4)  base patches around each base patch group (all present)
    fine_coarse_slot: np.ndarray    # (nf, 3, 3, 3)     father + 26 siblings of each refined patch group
    fine_pg_slot: np.ndarray        # (nf, 4, 4, 4)     refined-level patches around each refined patch group, -1 = none
    base_store_slot: np.ndarray     # (npg_b, 8, 6)     StoreFlux register of each base patch face, -1 = none
    fine_correct_slot: np.ndarray   # (nf, 6)           CorrectFlux register of each refined patch-group face, -1 = none
    fixup_slot: np.ndarray          # (npatch_b, 6)     Flu_FixUp_Flux register of each base patch face, -1 = none
    son0: np.ndarray                # (nf,)             first son patch of each father
    fa_idx: np.ndarray              # (nf,)             base patch index of each father
    base_pid0: np.ndarray           # (npg_b,)
    nreg: int
    interp_need: np.ndarray = None  # (nf, 12, 12, 12)  coarse cells whose fine cells no same-level patch covers

    @property
    def npg_base(self):
        return self.base_pg_slot.shape[0]


def base_patch_index(ncp, i, j, k):
    """Index of base patch (i, j, k) (periodic) in the patch-group ordering."""
    nx, ny, nz = ncp
    i, j, k = i % nx, j % ny, k % nz
    pg = ((k // 2) * (ny // 2) + j // 2) * (nx // 2) + i // 2
    return 8 * pg + _LID[(i % 2, j % 2, k % 2)]


def build_two_level_tables(ncp, fathers) -> TwoLevelTables:
    nx, ny, nz = ncp
    if nx % 2 or ny % 2 or nz % 2:
        raise ValueError("the base level needs an even number of patches per axis")
    fathers = [tuple(int(c) for c in f) for f in fathers]
    fidx = {f: q for q, f in enumerate(fathers)}
    if len(fidx) != len(fathers):
        raise ValueError("duplicate fathers")
    npg = (nx // 2) * (ny // 2) * (nz // 2)
    P = lambda i, j, k: base_patch_index(ncp, i, j, k)
    base_pg_slot = np.empty((npg, 4, 4, 4), dtype=np.int32)
    for PZ in range(nz // 2):
        for PY in range(ny // 2):
            for PX in range(nx // 2):
                pg = (PZ * (ny // 2) + PY) * (nx // 2) + PX
                for c in range(4):
                    for b in range(4):
                        for a in range(4):
                            base_pg_slot[pg, c, b, a] = P(2 * PX - 1 + a, 2 * PY - 1 + b, 2 * PZ - 1 + c)
    nf = len(fathers)
    fine_coarse_slot = np.empty((nf, 3, 3, 3), dtype=np.int32)
    fine_pg_slot = np.full((nf, 4, 4, 4), -1, dtype=np.int32)
    refined = {f[0] % nx + nx * (f[1] % ny + ny * (f[2] % nz)) for f in fathers}
    is_ref = lambda i, j, k: (i % nx + nx * (j % ny + ny * (k % nz))) in refined
    for q, (I, J, K) in enumerate(fathers):
        for c in range(3):
            for b in range(3):
                for a in range(3):
                    fine_coarse_slot[q, c, b, a] = P(I + a - 1, J + b - 1, K + c - 1)
        for c in range(4):
            for b in range(4):
                for a in range(4):
                    fx, fy, fz = 2 * I - 1 + a, 2 * J - 1 + b, 2 * K - 1 + c          # refined patch coordinates
                    fa = ((fx // 2) % nx, (fy // 2) % ny, (fz // 2) % nz)
                    if fa in fidx:
                        fine_pg_slot[q, c, b, a] = 8 * fidx[fa] + _LID[(fx % 2, fy % 2, fz % 2)]
    # registers: (base patch, face) with the base patch not refined and its neighbour across the face refined
    reg_of = {}
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if is_ref(i, j, k):
                    continue
                for s, (dx, dy, dz) in enumerate(_FACE):
                    if is_ref(i + dx, j + dy, k + dz):
                        reg_of[(P(i, j, k), s)] = len(reg_of)
    npatch = 8 * npg
    base_store_slot = np.full((npg, 8, 6), -1, dtype=np.int32)
    fixup_slot = np.full((npatch, 6), -1, dtype=np.int32)
    for (p, s), r in reg_of.items():
        base_store_slot[p // 8, p % 8, s] = r
        fixup_slot[p, s] = r
    fine_correct_slot = np.full((nf, 6), -1, dtype=np.int32)
    for q, (I, J, K) in enumerate(fathers):
        for s, (dx, dy, dz) in enumerate(_FACE):
            if not is_ref(I + dx, J + dy, K + dz):
                fine_correct_slot[q, s] = reg_of[(P(I + dx, J + dy, K + dz), s ^ 1)]
    return TwoLevelTables(ncp=tuple(ncp), fathers=fathers, base_pg_slot=base_pg_slot, fine_coarse_slot=fine_coarse_slot,
                          fine_pg_slot=fine_pg_slot, base_store_slot=base_store_slot,
                          fine_correct_slot=fine_correct_slot, fixup_slot=fixup_slot,
                          son0=(8 * np.arange(nf)).astype(np.int32),
                          fa_idx=np.array([P(*f) for f in fathers], dtype=np.int32),
                          base_pid0=(8 * np.arange(npg)).astype(np.int32), nreg=len(reg_of),
                          interp_need=refined_interp_need(fine_pg_slot))


def refined_interp_need(fine_pg_slot, ghost=3):
    """(npg, 12, 12, 12) uint8: coarse cells -2..9 around each father (the range prepare_refined_patch_groups
    interpolates) with at least one fine cell inside the window that has no same-level patch (slot -1)."""
    n = 16 + 2 * ghost
    w = np.arange(n) - ghost + 8
    miss = fine_pg_slot[:, (w >> 3)[:, None, None], (w >> 3)[None, :, None], (w >> 3)[None, None, :]] < 0
    fb = np.zeros((fine_pg_slot.shape[0], 24, 24, 24), dtype=bool)
    fb[:, 1:1 + n, 1:1 + n, 1:1 + n] = miss                 # window cell w sits at fine-box cell w + 1
    return np.ascontiguousarray(fb.reshape(-1, 12, 2, 12, 2, 12, 2).any(axis=(2, 4, 6)).astype(np.uint8))


def patches_from_grid(grid):
    """(5, Z, Y, X) base-level array -> (npatch, 5, 8, 8, 8) patches in the patch-group ordering (numpy or torch)."""
    _, Z, Y, X = grid.shape
    nx, ny, nz = X // 8, Y // 8, Z // 8
    g = grid.reshape(5, nz // 2, 2, 8, ny // 2, 2, 8, nx // 2, 2, 8)
    # -> (PZ, PY, PX, c, b, a, v, z, y, x)
    g = g.permute(1, 4, 7, 2, 5, 8, 0, 3, 6, 9) if hasattr(g, "permute") else g.transpose(1, 4, 7, 2, 5, 8, 0, 3, 6, 9)
    perm = [c * 4 + b * 2 + a for (a, b, c) in LOCAL_XYZ]                 # LocalID -> (c, b, a) position
    g = g.reshape(-1, 8, 5, 8, 8, 8)[:, perm]
    return g.reshape(-1, 5, 8, 8, 8)


def grid_from_patches(patch, ncp):
    """Inverse of patches_from_grid."""
    nx, ny, nz = ncp
    inv = [0] * 8
    for l, (a, b, c) in enumerate(LOCAL_XYZ):
        inv[c * 4 + b * 2 + a] = l
    g = patch.reshape(-1, 8, 5, 8, 8, 8)[:, inv].reshape(nz // 2, ny // 2, nx // 2, 2, 2, 2, 5, 8, 8, 8)
    g = g.permute(6, 0, 3, 7, 1, 4, 8, 2, 5, 9) if hasattr(g, "permute") else g.transpose(6, 0, 3, 7, 1, 4, 8, 2, 5, 9)
    return g.reshape(5, 8 * nz, 8 * ny, 8 * nx)


class TwoLevel:
    """A periodic base level with one refined level, advanced on the device (GAMER's EvolveLevel for two levels):

    base step   prepare (same-level, periodic) -> MHM_RP solve (StoreFlux) -> CorrectUnphysical -> StoreFlux into FP64
                registers -> copy back (Flu_AdvanceDt / Flu_Close, src/Fluid/Flu_Close.cpp:71-85);
    2 substeps  coarse data at the substep's start time (linear in time between the old and new base level,
                src/Main/Prepare_PatchData.cpp:2282-2283) -> coarse boxes -> Interpolate -> same-level patches ->
                solve -> CorrectUnphysical -> CorrectFlux -> copy back;
    sync        Flu_FixUp_Restrict, then Flu_FixUp_Flux (src/Main/EvolveLevel.cpp:757, 815).

    base, fine: device patch arrays (npatch, 5, 8, 8, 8) in the orderings of ``TwoLevelTables``, updated in place."""

    def __init__(self, torch, base, fine, tables: TwoLevelTables, dh_base, rsolver="HLLC", int_scheme="CQUAD",
                 gamma=5.0 / 3.0, **solver_kw):
        self.torch = torch
        self.base, self.fine, self.tab = base, fine, tables
        self.dh = float(dh_base)
        self.rsolver, self.int_scheme, self.gamma, self.kw = rsolver, int_scheme, gamma, solver_kw
        dev = base.device
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.t = {k: T(getattr(tables, k)) for k in ("base_pg_slot", "fine_coarse_slot", "fine_pg_slot",
                                                      "base_store_slot", "fine_correct_slot", "fixup_slot", "son0",
                                                      "fa_idx", "base_pid0", "interp_need")}
        self.reg = torch.zeros((max(tables.nreg, 1), 5, 8, 8), dtype=torch.float64, device=dev)
        nf = len(tables.fathers)
        self.fine_pid0 = T((8 * np.arange(nf)).astype(np.int32))
        self.counters = torch.zeros(2, dtype=torch.int32, device=dev)
        self.interp_fail = torch.zeros(1, dtype=torch.int32, device=dev)    # boxes still unphysical, running count

    def _close(self, pg_in, dt, dh):
        out, flux = hydro.fluid_solver(pg_in, dt, dh, rsolver=self.rsolver, gamma=self.gamma, **self.kw)
        hydro.correct_unphysical(pg_in, out, flux, dt, dh, gamma=self.gamma, counters=self.counters,
                                 main_rsolver=self.rsolver)
        return out, flux

    def step(self, dt_base):
        torch, t = self.torch, self.t
        dh_f, dt_f = 0.5 * self.dh, 0.5 * dt_base
        npg_b = self.tab.npg_base
        pg_b = torch.empty((npg_b, 5, 22, 22, 22), dtype=self.base.dtype, device=self.base.device)
        hydro.prepare_patches(self.base, pg_b, t["base_pg_slot"], 3)
        out, flux = self._close(pg_b, dt_base, self.dh)
        self.reg.zero_()
        hydro.store_flux(flux, t["base_store_slot"], dt_base, self.reg)
        old = self.base.clone()
        hydro.close_patches(out, self.base, t["base_pid0"])
        nf = len(self.tab.fathers)
        for sub in range(2):
            if sub == 0:
                coarse = old
            else:
                # 0.5 old + 0.5 new (GAMER's temporal weights), as (old + new) * 0.5: scaling by 0.5 is exact, so the
                # two agree bit for bit (short of overflow), in two kernels instead of three and into the clone's buffer
                coarse = old.add_(self.base).mul_(0.5)
            pg_f = torch.empty((nf, 5, 22, 22, 22), dtype=self.fine.dtype, device=self.fine.device)
            hydro.prepare_refined_patch_groups(coarse, self.fine, t["fine_coarse_slot"], t["fine_pg_slot"], pg_f,
                                               scheme=self.int_scheme, need=t["interp_need"], gamma=self.gamma,
                                               nfail=self.interp_fail)
            out_f, flux_f = self._close(pg_f, dt_f, dh_f)
            hydro.correct_flux(flux_f, t["fine_correct_slot"], dt_f, self.reg)
            hydro.close_patches(out_f, self.fine, self.fine_pid0)
        hydro.fixup_restrict(self.fine, self.base, t["son0"], t["fa_idx"])
        hydro.fixup_flux(self.base, t["fixup_slot"], self.reg, self.dh, gamma=self.gamma)
        return self

    def capture(self, dt_base):
        """Record one ``step(dt_base)`` into a HIP graph and return it; ``graph.replay()`` then advances base and fine
        in place.  The warm-up step before the capture runs on the live state, which is saved and restored around
        it.  Every clear inside the step is a fill kernel (no memset nodes)."""
        torch = self.torch
        saved = (self.base.clone(), self.fine.clone(), self.counters.clone(), self.interp_fail.clone())
        self.step(dt_base)
        torch.cuda.synchronize()
        for dst, src in zip((self.base, self.fine, self.counters, self.interp_fail), saved):
            dst.copy_(src)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.step(dt_base)
        return graph
