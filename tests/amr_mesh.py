"""A two-level test mesh: a periodic coarse level of 4^3 patches (32^3 cells) whose central 2x2x2 patches are refined
into 8 fine patch groups, with the host-side tables the device preparation reads (the patch relations GAMER keeps in
its AMR tree: father + 26 siblings of each fine patch group, and the same-level fine patches around it)."""
import numpy as np

NPC = 4                      # coarse patches per axis
REF = (1, 2)                 # refined coarse patches per axis: I, J, K in REF


def patches_of(grid):
    """(5, Z, Y, X) grid -> (npatch, 5, 8, 8, 8) patches, patch index (K*n + J)*n + I."""
    n = grid.shape[-1] // 8
    return np.ascontiguousarray(grid.reshape(5, n, 8, n, 8, n, 8).transpose(1, 3, 5, 0, 2, 4, 6).reshape(-1, 5, 8, 8, 8))


def fathers():
    return [(i, j, k) for k in REF for j in REF for i in REF]


def coarse_slots():
    """(npg, 3, 3, 3) indices of the father patch and its 26 siblings (periodic), (z, y, x) order."""
    t = np.empty((len(fathers()), 3, 3, 3), dtype=np.int32)
    for q, (i, j, k) in enumerate(fathers()):
        for c in range(3):
            for b in range(3):
                for a in range(3):
                    t[q, c, b, a] = (((k + c - 1) % NPC) * NPC + (j + b - 1) % NPC) * NPC + (i + a - 1) % NPC
    return t


def fine_slots():
    """(npg, 4, 4, 4) fine patch indices around each fine patch group (-1 outside the refined region); the fine
    patches are those of the 4^3 fine patch grid covering REF, indexed ((fk*4 + fj)*4 + fi) with f = 2*I + d - 2."""
    t = np.full((len(fathers()), 4, 4, 4), -1, dtype=np.int32)
    lo = 2 * REF[0]
    for q, (i, j, k) in enumerate(fathers()):
        for c in range(4):
            for b in range(4):
                for a in range(4):
                    fi, fj, fk = 2 * i + a - 1 - lo, 2 * j + b - 1 - lo, 2 * k + c - 1 - lo
                    if 0 <= fi < 4 and 0 <= fj < 4 and 0 <= fk < 4:
                        t[q, c, b, a] = (fk * 4 + fj) * 4 + fi
    return t


def gas(shape, rng, smooth=True):
    """Conserved (5, *shape) gas with density/pressure in [0.6, 1.6]."""
    z, y, x = np.meshgrid(*(np.linspace(0, 2 * np.pi, n, endpoint=False) for n in shape), indexing="ij")
    rho = 1.0 + 0.3 * np.sin(x + 0.5) * np.cos(y) + 0.1 * np.sin(2 * z)
    vx, vy, vz = 0.3 * np.sin(y), 0.2 * np.cos(z + x), 0.1 * np.sin(x)
    p = 1.0 + 0.2 * np.cos(x - y) + 0.1 * np.sin(z)
    if not smooth:
        rho = rho * np.exp(0.1 * rng.standard_normal(shape))
        p = p * np.exp(0.1 * rng.standard_normal(shape))
    return np.stack([rho, rho * vx, rho * vy, rho * vz, p / (2.0 / 3.0) + 0.5 * rho * (vx ** 2 + vy ** 2 + vz ** 2)])
