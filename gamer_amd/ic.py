"""Seeded synthetic patch-group inputs for the MHM / MHM_RP solver.

A workload is a periodic uniform grid of (n_x*16) x (n_y*16) x (n_z*16) cells cut into patch groups of
16^3 cells.  Each patch group is handed to the solver as one (5, N, N, N) block with the ghost zones
filled from the periodic neighbours (N = 16 + 2*ghost: 22 for MHM_RP, 20 for MHM), exactly the
layout ``Flu_Prepare`` builds for ``CUAPI_Asyn_FluidSolver`` (src/Fluid/Flu_Prepare.cpp:23-280,
layout ``[PG][var][k][j][i]``, include/Macro.h:1197).  Because neighbouring blocks share their
ghost data, fluxes on shared patch-group faces agree bit for bit and the domain sums of the conserved
fields are conserved to round-off.

Every transcendental used here is evaluated with plain IEEE +,-,*,/ (``det_exp``, ``det_sin``), so the
same seed gives the same bytes on any host, whatever SIMD paths numpy picks.
"""
from __future__ import annotations

import dataclasses
import numpy as np

GAMMA = 5.0 / 3.0
PS2 = 16                      # patch-group interior width (include/Macro.h:863-873)
NCOMP = 5                     # rho, mx, my, mz, E

_LN2_HI = 6.93147180369123816490e-01
_LN2_LO = 1.90821492927058770002e-10
_INV_LN2 = 1.44269504088896338700e+00
_TWO_PI_HI = 6.28318530717958623200e+00
_TWO_PI_LO = 2.44929359829470635445e-16


def det_exp(x: np.ndarray) -> np.ndarray:
    """exp(x) from Cody-Waite reduction (|r| <= ln2/2) and an order-16 Taylor polynomial, basic IEEE ops only."""
    x = np.asarray(x, dtype=np.float64)
    k = np.rint(x * _INV_LN2)
    r = x - k * _LN2_HI
    r -= k * _LN2_LO
    p = np.ones_like(r)
    for n in range(16, 0, -1):
        p *= r
        p /= n
        p += 1.0
    return np.ldexp(p, k.astype(np.int32))


def det_sin(x: np.ndarray) -> np.ndarray:
    """sin(x) from reduction to [-pi, pi] and an odd Taylor series to order 39 (basic IEEE ops only)."""
    x = np.asarray(x, dtype=np.float64)
    k = np.rint(x / _TWO_PI_HI)
    r = (x - k * _TWO_PI_HI) - k * _TWO_PI_LO
    r2 = r * r
    p = np.ones_like(r)
    for n in range(39, 1, -2):            # sin r = r (1 - r^2/(2*3) (1 - r^2/(4*5) (...)))
        p = 1.0 - p * r2 / ((n - 1) * n)
    return r * p


@dataclasses.dataclass
class Workload:
    """Patch-group inputs plus the scalars of one solver call."""
    u: np.ndarray              # (npg, 5, N, N, N) float64, conserved
    dt: float
    dh: float
    npg_shape: tuple           # (n_z, n_y, n_x) patch groups per axis
    ghost: int
    gamma: float = GAMMA

    @property
    def npg(self) -> int:
        return self.u.shape[0]


def prim_to_cons(rho, vx, vy, vz, p, gamma=GAMMA):
    """Conserved fields from primitive ones, E = P/(gamma-1) + rho |v|^2 / 2."""
    out = np.empty((NCOMP,) + np.shape(rho))
    out[0] = rho
    np.multiply(rho, vx, out=out[1])
    np.multiply(rho, vy, out=out[2])
    np.multiply(rho, vz, out=out[3])
    ke = vx * vx
    ke += vy * vy
    ke += vz * vz
    ke *= rho
    ke *= 0.5
    np.divide(p, gamma - 1.0, out=out[4])
    out[4] += ke
    return out


def cfl_dt(cons: np.ndarray, dh: float, cfl: float, gamma: float = GAMMA) -> float:
    """dt = cfl * dh / max_cells max_d(|v_d| + c_s), the hydro CFL condition
    (src/Model_Hydro/CPU_Hydro/CPU_dtSolver_HydroCFL.cpp:57-209)."""
    rho = cons[0]
    inv = 1.0 / rho
    p = (gamma - 1.0) * (cons[4] - 0.5 * (cons[1] ** 2 + cons[2] ** 2 + cons[3] ** 2) * inv)
    cs = np.sqrt(gamma * p * inv)
    vmax = np.maximum(np.maximum(np.abs(cons[1]), np.abs(cons[2])), np.abs(cons[3])) * inv
    return float(cfl * dh / np.max(vmax + cs))


def cut_patch_groups(cons: np.ndarray, ghost: int) -> np.ndarray:
    """Cut a periodic (5, Z, Y, X) grid into (npg, 5, N, N, N) blocks, PG index = x + n_x*(y + n_y*z)."""
    _, nz, ny, nx = cons.shape
    assert nz % PS2 == 0 and ny % PS2 == 0 and nx % PS2 == 0
    pad = np.pad(cons, ((0, 0), (ghost, ghost), (ghost, ghost), (ghost, ghost)), mode="wrap")
    n = PS2 + 2 * ghost
    win = np.lib.stride_tricks.sliding_window_view(pad, (n, n, n), axis=(1, 2, 3))[:, ::PS2, ::PS2, ::PS2]
    gz, gy, gx = nz // PS2, ny // PS2, nx // PS2
    # always a fresh writable array (for a single patch group the strided window view is already contiguous)
    return win.transpose(1, 2, 3, 0, 4, 5, 6).copy(order="C").reshape(gz * gy * gx, NCOMP, n, n, n)


def random_sod(npg_shape=(4, 4, 4), seed=1234, scheme_ghost=3, sedov=False, cfl=0.5, gamma=GAMMA,
               sigma_v=0.3) -> Workload:
    """``random_sod_grid`` cut into patch-group blocks."""
    cons, dt, dh = random_sod_grid(npg_shape, seed, sedov, cfl, gamma, sigma_v)
    return Workload(cut_patch_groups(cons, scheme_ghost), dt, dh, tuple(npg_shape), scheme_ghost, gamma)


def random_sod_grid(npg_shape=(4, 4, 4), seed=1234, sedov=False, cfl=0.5, gamma=GAMMA, sigma_v=0.3):
    """Random medium plus a Sod plane through half of the patch groups, as a periodic cell grid.

    rho, P ~ lognormal(0, 0.5); v_d ~ N(0, 1) * sigma_v * c_s.  In the lower half of the z range every
    patch group holds a Sod discontinuity at its x centre (rho 1 | 0.125, P 1 | 0.1, v = 0).  With
    ``sedov`` one random interior cell of every 16th patch group gets P x 1e3.
    Returns (cons (5, Z, Y, X), dt, dh).
    """
    gz, gy, gx = npg_shape
    nz, ny, nx = gz * PS2, gy * PS2, gx * PS2
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (nz, ny, nx)
    rho = det_exp(0.5 * rng.standard_normal(shape))
    p = det_exp(0.5 * rng.standard_normal(shape))
    cs = np.sqrt(gamma * p / rho)
    vx = sigma_v * cs * rng.standard_normal(shape)
    vy = sigma_v * cs * rng.standard_normal(shape)
    vz = sigma_v * cs * rng.standard_normal(shape)

    ix = np.arange(nx)[None, None, :]
    left = np.broadcast_to((ix % PS2) < PS2 // 2, shape)
    half = np.broadcast_to(np.arange(nz)[:, None, None] < nz // 2, shape)
    rho = np.where(half, np.where(left, 1.0, 0.125), rho)
    p = np.where(half, np.where(left, 1.0, 0.1), p)
    vx = np.where(half, 0.0, vx)
    vy = np.where(half, 0.0, vy)
    vz = np.where(half, 0.0, vz)

    if sedov:
        npg = gz * gy * gx
        for pg in range(0, npg, 16):
            kz, rem = divmod(pg, gy * gx)
            ky, kx = divmod(rem, gx)
            c = rng.integers(0, PS2, size=3)
            p[kz * PS2 + c[0], ky * PS2 + c[1], kx * PS2 + c[2]] *= 1.0e3

    cons = prim_to_cons(rho, vx, vy, vz, p, gamma)
    dh = 1.0 / nx
    dt = cfl_dt(cons, dh, cfl, gamma)
    return cons, dt, dh


def kelvin_helmholtz(npg_shape=(8, 16, 16), seed=123, scheme_ghost=3, cfl=0.5, gamma=GAMMA, amp=0.01) -> Workload:
    """Shear layer: rho 1 (outer) / 2 (inner band), v_x -0.5 / +0.5, P 2.5, and a seeded perturbation
    delta v_y = amp * sum_{k=1..4} sin(2 pi k x + phi_k) (example/test_problem/Hydro/KelvinHelmholtzInstability)."""
    gz, gy, gx = npg_shape
    nz, ny, nx = gz * PS2, gy * PS2, gx * PS2
    rng = np.random.Generator(np.random.PCG64(seed))
    phi = 2.0 * np.pi * rng.random(4)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    inner = (np.abs(y - 0.5) < 0.25)[None, :, None]
    shape = (nz, ny, nx)
    rho = np.broadcast_to(np.where(inner, 2.0, 1.0), shape).astype(np.float64)
    vx = np.broadcast_to(np.where(inner, 0.5, -0.5), shape).astype(np.float64)
    dvy = np.zeros(nx)
    for k in range(1, 5):
        dvy = dvy + det_sin(2.0 * np.pi * k * x + phi[k - 1])
    vy = np.broadcast_to((amp * dvy)[None, None, :], shape).astype(np.float64)
    vz = np.zeros(shape)
    p = np.full(shape, 2.5)
    cons = prim_to_cons(rho, vx, vy, vz, p, gamma)
    dh = 1.0 / nx
    dt = cfl_dt(cons, dh, cfl, gamma)
    return Workload(cut_patch_groups(cons, scheme_ghost), dt, dh, tuple(npg_shape), scheme_ghost, gamma)


def smooth_modes(npg_shape=(4, 4, 4), seed=7, scheme_ghost=3, cfl=0.5, gamma=GAMMA):
    """Smooth periodic state: rho = 1 + 0.1 sum of 3 seeded sine modes, P = 1, v_d = 0.1 sin(...).
    Returns the cell grid (5, Z, Y, X), dt and dh (the rollout refills ghost zones between steps)."""
    gz, gy, gx = npg_shape
    nz, ny, nx = gz * PS2, gy * PS2, gx * PS2
    rng = np.random.Generator(np.random.PCG64(seed))
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx,
                          indexing="ij")
    rho = np.ones((nz, ny, nx))
    for _ in range(3):
        k = rng.integers(1, 4, size=3)
        ph = 2.0 * np.pi * rng.random()
        rho = rho + 0.1 / 3.0 * det_sin(2.0 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + ph)
    vel = []
    for _ in range(3):
        k = rng.integers(1, 3, size=3)
        ph = 2.0 * np.pi * rng.random()
        vel.append(0.1 * det_sin(2.0 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + ph))
    p = np.ones_like(rho)
    cons = prim_to_cons(rho, vel[0], vel[1], vel[2], p, gamma)
    dh = 1.0 / nx
    dt = cfl_dt(cons, dh, cfl, gamma)
    return cons, dt, dh


def pg_interior(u_pg: np.ndarray, ghost: int) -> np.ndarray:
    """The 16^3 interior of (npg, 5, N, N, N) blocks."""
    g = ghost
    return u_pg[:, :, g:g + PS2, g:g + PS2, g:g + PS2]


def assemble(out_pg: np.ndarray, npg_shape) -> np.ndarray:
    """Inverse of the interior cut: (npg, 5, 16, 16, 16) -> periodic grid (5, Z, Y, X)."""
    gz, gy, gx = npg_shape
    o = out_pg.reshape(gz, gy, gx, NCOMP, PS2, PS2, PS2)
    return o.transpose(3, 0, 4, 1, 5, 2, 6).reshape(NCOMP, gz * PS2, gy * PS2, gx * PS2)
