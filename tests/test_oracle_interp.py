"""Oracle restatement of the coarse -> fine ghost-zone interpolation (Interpolate / Int_CQuadratic / Int_MinMod1D)
and of the same-level patch gather, pinned by what the mathematics fixes:

* both schemes are conservative: the mean of the 2x2x2 fine cells equals the coarse cell,
* both reproduce a linear field exactly; INT_CQUAD without limiting reproduces the cell averages of a separable
  quadratic exactly (fine average over [i-1/2, i] of x^2 = C_i - (C_{i+1} - C_{i-1})/8),
* the monotonic limiter keeps the fine values inside the range of the coarse neighbours at a discontinuity,
* Interpolate_Iterate: a block that fails in conserved variables is redone in primitive variables (the result equals a
  primitive interpolation converted back), other blocks keep the conserved result,
* the patch gather copies the right patch octants and leaves slots of -1 untouched.
"""
import numpy as np
import pytest

import oracle

G = 5.0 / 3.0


def _box(f, cs=(8, 8, 8)):
    """(1, 5, z, y, x) coarse box whose 5 fields are f(x, y, z) + field offsets, at cell centres 0..cs-1."""
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in cs[::-1]), indexing="ij")
    base = f(x, y, z)
    return np.stack([base + 10.0 * (v + 1) for v in range(5)])[None]


def _fine_centres(n):
    """Fine cell centres in coarse index units for fine cells 0..2n-1 of coarse cells 1..n."""
    return 0.75 + 0.5 * np.arange(2 * n)


@pytest.mark.parametrize("scheme", ["CQUAD", "MINMOD1D"])
def test_conservative_and_linear_exact(scheme):
    c = _box(lambda x, y, z: 0.5 * x - 0.25 * y + 0.125 * z)
    f0 = np.zeros((1, 5, 12, 12, 12))
    _, f, nf = oracle.interpolate(c, f0, (1, 1, 1), (6, 6, 6), scheme=scheme, int_prim=False, reduce_mono=False)
    assert nf == 0
    xf = _fine_centres(6)
    zf, yf, xf3 = np.meshgrid(xf, xf, xf, indexing="ij")
    exp = 0.5 * xf3 - 0.25 * yf + 0.125 * zf
    for v in range(5):
        np.testing.assert_allclose(f[0, v], exp + 10.0 * (v + 1), rtol=0, atol=1e-13)
    mean = f[0].reshape(5, 6, 2, 6, 2, 6, 2).mean(axis=(2, 4, 6))
    np.testing.assert_allclose(mean, c[0][:, 1:7, 1:7, 1:7], rtol=0, atol=1e-13)


def test_cquad_reproduces_separable_quadratic_averages():
    # coarse cell averages of q(x) = x^2 over [i-1/2, i+1/2] are i^2 + 1/12
    avg = lambda a: a ** 2 + 1.0 / 12.0
    c = _box(lambda x, y, z: avg(x) + 2.0 * avg(y) - avg(z))
    _, f, nf = oracle.interpolate(c, np.zeros((1, 5, 12, 12, 12)), (1, 1, 1), (6, 6, 6), monotonic=0,
                                  int_prim=False, reduce_mono=False)
    assert nf == 0
    xf = _fine_centres(6)
    favg = xf ** 2 + 1.0 / 48.0                      # fine averages over half-width cells
    zf, yf, xf3 = np.meshgrid(favg, favg, favg, indexing="ij")
    np.testing.assert_allclose(f[0, 0], xf3 + 2.0 * yf - zf + 10.0, rtol=0, atol=1e-12)


def test_monotonic_limiter_stays_in_range_at_a_jump():
    c = _box(lambda x, y, z: np.where(x < 4, 1.0, 5.0))
    _, f, _ = oracle.interpolate(c, np.zeros((1, 5, 12, 12, 12)), (1, 1, 1), (6, 6, 6), int_prim=False,
                                 reduce_mono=False)
    assert f[0, 0].min() >= 11.0 - 1e-14 and f[0, 0].max() <= 15.0 + 1e-14
    _, fu, _ = oracle.interpolate(c, np.zeros((1, 5, 12, 12, 12)), (1, 1, 1), (6, 6, 6), monotonic=0,
                                  int_prim=False, reduce_mono=False)
    assert fu[0, 0].min() < 11.0 - 0.1 or fu[0, 0].max() > 15.0 + 0.1        # unlimited over/undershoots


def _gas_box(rng, cs=(6, 6, 6)):
    rho = np.exp(0.2 * rng.standard_normal(cs[::-1]))
    v = 0.2 * rng.standard_normal((3,) + cs[::-1])
    p = np.exp(0.2 * rng.standard_normal(cs[::-1]))
    return np.stack([rho, rho * v[0], rho * v[1], rho * v[2], p / (G - 1) + 0.5 * rho * (v ** 2).sum(0)])[None]


def _con2pri(c):
    rho = c[:, 0]
    v = c[:, 1:4] / rho[:, None]
    eint = c[:, 4] - 0.5 * (c[:, 1] ** 2 + c[:, 2] ** 2 + c[:, 3] ** 2) / rho
    return np.concatenate([rho[:, None], v, (eint * (G - 1))[:, None]], axis=1)


def test_iterate_redoes_failing_blocks_in_primitive_variables():
    rng = np.random.default_rng(7)
    c = _gas_box(rng)
    # a fast, cold cell next to a slow one: conserved interpolation of E and m gives Eint < 0 in its fine cells
    z, y, x = 3, 3, 3
    c[0, :, z, y, x] = [1.0, 4.0, 0.0, 0.0, 8.0 + 1e-3]
    c[0, :, z, y, x + 1] = [1.0, 0.0, 0.0, 0.0, 1e-3]
    f0 = np.zeros((1, 5, 8, 8, 8))
    cons_only = oracle.interpolate(c, f0, (1, 1, 1), (4, 4, 4), int_prim=False, reduce_mono=False)[1]
    fine_bad = lambda f: (f[0, 4] - 0.5 * (f[0, 1] ** 2 + f[0, 2] ** 2 + f[0, 3] ** 2) / f[0, 0]) < 0
    assert fine_bad(cons_only).any()
    c_after, f, nf = oracle.interpolate(c, f0, (1, 1, 1), (4, 4, 4), int_prim=True, reduce_mono=True)
    assert nf == 0 and not fine_bad(f).any()
    np.testing.assert_allclose(c_after, _con2pri(c), rtol=1e-14, atol=1e-14)      # converted in place
    prim = oracle.interpolate(_con2pri(c), f0, (1, 1, 1), (4, 4, 4), int_prim=False, reduce_mono=False)[1]
    w = prim[0]
    cons = np.stack([w[0], w[0] * w[1], w[0] * w[2], w[0] * w[3],
                     0.5 * ((w[0] * w[1]) ** 2 + (w[0] * w[2]) ** 2 + (w[0] * w[3]) ** 2) / w[0] + w[4] / (G - 1)])
    bad_blocks = fine_bad(cons_only).reshape(4, 2, 4, 2, 4, 2).any(axis=(1, 3, 5))
    blk = np.repeat(np.repeat(np.repeat(bad_blocks, 2, 0), 2, 1), 2, 2)
    np.testing.assert_allclose(f[0][:, blk], cons[:, blk], rtol=1e-14, atol=1e-14)
    np.testing.assert_array_equal(f[0][:, ~blk], cons_only[0][:, ~blk])


def test_prepare_patches_window_and_keep():
    rng = np.random.default_rng(9)
    patch = rng.standard_normal((70, 5, 8, 8, 8))
    slot = rng.integers(-1, 70, size=(2, 4, 4, 4)).astype(np.int32)
    ghost = 3
    n = 16 + 2 * ghost
    pg0 = rng.standard_normal((2, 5, n, n, n))
    got = oracle.prepare_patches(patch, pg0, slot, ghost)
    for p in range(2):
        for z in range(n):
            for y in range(0, n, 5):
                for x in range(n):
                    X, Y, Z = x - ghost + 8, y - ghost + 8, z - ghost + 8
                    s = slot[p, Z // 8, Y // 8, X // 8]
                    exp = pg0[p, :, z, y, x] if s < 0 else patch[s, :, Z % 8, Y % 8, X % 8]
                    np.testing.assert_array_equal(got[p, :, z, y, x], exp)


# ---- the other tensor-product schemes: INT_MINMOD3D, INT_VANLEER, INT_QUAD, INT_CQUAR, INT_QUAR ----

def _ghost(scheme):
    return 2 if scheme in ("CQUAR", "QUAR") else 1


def _scheme_box(f, scheme, n=6):
    """Coarse box with the scheme's ghost width around n^3 interpolated cells; returns (box, c0, cr)."""
    g = _ghost(scheme)
    cs = (n + 2 * g,) * 3
    z, y, x = np.meshgrid(*(np.arange(m, dtype=np.float64) for m in cs), indexing="ij")
    base = f(x - g + 1, y - g + 1, z - g + 1)          # interpolated coarse cells sit at coordinates 1..n
    return np.stack([base + 10.0 * (v + 1) for v in range(5)])[None], (g,) * 3, (n,) * 3


@pytest.mark.parametrize("scheme", ["MINMOD3D", "VANLEER", "CQUAD", "CQUAR", "MINMOD1D"])
def test_conservative_schemes_keep_coarse_means(scheme):
    rng = np.random.default_rng(11)
    g = _ghost(scheme)
    c = 5.0 + rng.standard_normal((2, 5, 6 + 2 * g, 7 + 2 * g, 8 + 2 * g))
    _, f, _ = oracle.interpolate(c, np.zeros((2, 5, 12, 14, 16)), (g,) * 3, (8, 7, 6), scheme=scheme,
                                 int_prim=False, reduce_mono=False)
    mean = f.reshape(2, 5, 6, 2, 7, 2, 8, 2).mean(axis=(3, 5, 7))
    np.testing.assert_allclose(mean, c[:, :, g:g + 6, g:g + 7, g:g + 8], rtol=0, atol=1e-13)


@pytest.mark.parametrize("scheme", ["MINMOD3D", "VANLEER", "QUAD", "CQUAR", "QUAR"])
def test_linear_fields_are_reproduced(scheme):
    c, c0, cr = _scheme_box(lambda x, y, z: 0.5 * x - 0.25 * y + 0.125 * z, scheme)
    _, f, nf = oracle.interpolate(c, np.zeros((1, 5, 12, 12, 12)), c0, cr, scheme=scheme, int_prim=False,
                                  reduce_mono=False)
    xf = _fine_centres(6)
    zf, yf, xf3 = np.meshgrid(xf, xf, xf, indexing="ij")
    for v in range(5):
        np.testing.assert_allclose(f[0, v], 0.5 * xf3 - 0.25 * yf + 0.125 * zf + 10.0 * (v + 1), rtol=0, atol=1e-12)


def test_quad_interpolates_quadratic_point_values():
    q = lambda x, y, z: x ** 2 + 2.0 * y ** 2 - z ** 2 + 0.5 * x * y
    c, c0, cr = _scheme_box(q, "QUAD")
    _, f, _ = oracle.interpolate(c, np.zeros((1, 5, 12, 12, 12)), c0, cr, scheme="QUAD", monotonic=0,
                                 int_prim=False, reduce_mono=False)
    xf = _fine_centres(6)
    zf, yf, xf3 = np.meshgrid(xf, xf, xf, indexing="ij")
    np.testing.assert_allclose(f[0, 0], q(xf3, yf, zf) + 10.0, rtol=0, atol=1e-12)


def test_quar_interpolates_quartic_point_values():
    q = lambda x, y, z: 0.01 * x ** 4 - 0.05 * y ** 3 + z ** 2 + 0.1 * x * z ** 2
    c, c0, cr = _scheme_box(q, "QUAR")
    _, f, _ = oracle.interpolate(c, np.zeros((1, 5, 12, 12, 12)), c0, cr, scheme="QUAR", monotonic=0,
                                 int_prim=False, reduce_mono=False)
    xf = _fine_centres(6)
    zf, yf, xf3 = np.meshgrid(xf, xf, xf, indexing="ij")
    np.testing.assert_allclose(f[0, 0], q(xf3, yf, zf) + 10.0, rtol=0, atol=1e-11)


def test_cquar_reproduces_quartic_cell_averages():
    # cell averages of a^4 over [i-1/2, i+1/2] and over the fine halves [i-1/2, i], [i, i+1/2]
    A = lambda i: ((i + 0.5) ** 5 - (i - 0.5) ** 5) / 5.0
    c, c0, cr = _scheme_box(lambda x, y, z: 0.01 * A(x) + 0.02 * A(y) - 0.01 * A(z), "CQUAR")
    _, f, _ = oracle.interpolate(c, np.zeros((1, 5, 12, 12, 12)), c0, cr, scheme="CQUAR", monotonic=0,
                                 int_prim=False, reduce_mono=False)
    lo = np.arange(12) * 0.5 + 0.5                       # fine cell [lo, lo + 1/2] in coarse units
    fa = 2.0 * ((lo + 0.5) ** 5 - lo ** 5) / 5.0
    zf, yf, xf = np.meshgrid(fa, fa, fa, indexing="ij")
    np.testing.assert_allclose(f[0, 0], 0.01 * xf + 0.02 * yf - 0.01 * zf + 10.0, rtol=0, atol=1e-10)


def _row_box(vals, g):
    """A box whose field 0 varies along x only: vals along x, the same in every (y, z) row."""
    nx = len(vals)
    c = np.ones((1, 5, 2 * g + 1, 2 * g + 1, nx))
    c[0, 0] = np.asarray(vals, dtype=np.float64)[None, None, :]
    return c


@pytest.mark.parametrize("scheme,expect", [("MINMOD3D", (0.75, 1.25)), ("VANLEER", (1.0 - 1.0 / 3.0, 1.0 + 1.0 / 3.0))])
def test_slope_formulas_by_hand(scheme, expect):
    # neighbours 0, 1, 3: left slope 1, right slope 2 -> minmod 1/4, van Leer 0.5*1*2/3
    _, f, _ = oracle.interpolate(_row_box([0.0, 1.0, 3.0], 1), np.zeros((1, 5, 2, 2, 2)), (1, 1, 1), (1, 1, 1),
                                 scheme=scheme, int_prim=False, reduce_mono=False)
    np.testing.assert_allclose(f[0, 0, :, :, 0], expect[0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(f[0, 0, :, :, 1], expect[1], rtol=0, atol=1e-15)


def test_quartic_monotonic_order_check():
    # L2 L1 C R1 R2 = -10 0 0.5 1 0: both quartic values lie inside [0, 1] but in the wrong order
    # (hi - lo = -128/2048 < 0), which only INT_QUAR's extra test rejects; the limited slope is then
    # min(0.125 (R1 - L1), coeff/4 (C - L1), coeff/4 (R1 - C)) = 0.125
    c = _row_box([-10.0, 0.0, 0.5, 1.0, 0.0], 2)
    kw = dict(int_prim=False, reduce_mono=False, mono_coeff=2.0)
    _, f, _ = oracle.interpolate(c, np.zeros((1, 5, 2, 2, 2)), (2, 2, 2), (1, 1, 1), scheme="QUAR", **kw)
    np.testing.assert_allclose(f[0, 0, :, :, 0], 0.375, rtol=0, atol=1e-15)
    np.testing.assert_allclose(f[0, 0, :, :, 1], 0.625, rtol=0, atol=1e-15)
    _, fu, _ = oracle.interpolate(c, np.zeros((1, 5, 2, 2, 2)), (2, 2, 2), (1, 1, 1), scheme="QUAR", monotonic=0, **kw)
    np.testing.assert_allclose(fu[0, 0, 0, 0], [1143.0 / 2048.0, 1015.0 / 2048.0], rtol=0, atol=1e-15)


@pytest.mark.parametrize("scheme", ["MINMOD3D", "VANLEER", "QUAD", "CQUAR", "QUAR"])
def test_limited_schemes_stay_in_range_at_a_jump(scheme):
    c, c0, cr = _scheme_box(lambda x, y, z: np.where(x < 4, 1.0, 5.0), scheme)
    _, f, _ = oracle.interpolate(c, np.zeros((1, 5, 12, 12, 12)), c0, cr, scheme=scheme, int_prim=False,
                                 reduce_mono=False)
    assert f[0, 0].min() >= 11.0 - 1e-14 and f[0, 0].max() <= 15.0 + 1e-14


# ---- coarse-box gather and patch-group assembly at a refinement boundary ----

def test_gather_boxes_brute_force():
    rng = np.random.default_rng(12)
    patch = rng.standard_normal((30, 5, 8, 8, 8))
    slot = rng.integers(-1, 30, size=(3, 2, 3, 3)).astype(np.int32)       # (nbox, tz, ty, tx)
    box0 = rng.standard_normal((3, 5, 9, 10, 11))                          # (z, y, x) = 9, 10, 11
    origin = (7, 5, 4)
    got = oracle.gather_boxes(patch, box0, slot, origin)
    for q in range(3):
        for z in range(9):
            for y in range(10):
                for x in range(0, 11, 3):
                    X, Y, Z = x + origin[0], y + origin[1], z + origin[2]
                    s = slot[q, Z // 8, Y // 8, X // 8]
                    exp = box0[q, :, z, y, x] if s < 0 else patch[s, :, Z % 8, Y % 8, X % 8]
                    np.testing.assert_array_equal(got[q, :, z, y, x], exp)


def test_assemble_takes_patches_then_the_fine_box():
    rng = np.random.default_rng(13)
    patch = rng.standard_normal((40, 5, 8, 8, 8))
    slot = rng.integers(-1, 40, size=(2, 4, 4, 4)).astype(np.int32)
    fbox = rng.standard_normal((2, 5, 24, 24, 24))
    got = oracle.assemble_patch_groups(patch, fbox, np.zeros((2, 5, 22, 22, 22)), slot, 3, (1, 1, 1))
    for p in range(2):
        for z in range(22):
            for y in range(0, 22, 7):
                for x in range(22):
                    X, Y, Z = x + 5, y + 5, z + 5
                    s = slot[p, Z // 8, Y // 8, X // 8]
                    exp = fbox[p, :, z + 1, y + 1, x + 1] if s < 0 else patch[s, :, Z % 8, Y % 8, X % 8]
                    np.testing.assert_array_equal(got[p, :, z, y, x], exp)


@pytest.mark.parametrize("scheme", ["CQUAD", "MINMOD3D", "CQUAR", "QUAR"])
def test_refined_patch_group_geometry(scheme):
    """On a linear coarse field (cell averages = centre values) every interpolated ghost cell of the assembled
    22^3 window equals the linear function at its fine cell centre, and every same-level cell its fine patch: pins
    the coarse-box origin (6-G), the 3x3x3 sibling table, the interpolated range and the 24^3 -> 22^3 offset."""
    import amr_mesh as M
    n = 8 * M.NPC
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for _ in range(3)), indexing="ij")
    lin = lambda X, Y, Z: 0.5 * X - 0.25 * Y + 0.125 * Z
    grid = np.stack([lin(x, y, z) + 10.0 * (v + 1) for v in range(5)])
    cpatch = M.patches_of(grid)
    fpatch = np.random.default_rng(14).standard_normal((64, 5, 8, 8, 8))
    g = 2 if scheme in ("CQUAR", "QUAR") else 1
    npg = len(M.fathers())
    cbox = oracle.gather_boxes(cpatch, np.zeros((npg, 5) + (12 + 2 * g,) * 3), M.coarse_slots(), (6 - g,) * 3)
    _, fbox, _ = oracle.interpolate(cbox, np.zeros((npg, 5, 24, 24, 24)), (g,) * 3, (12, 12, 12), scheme=scheme,
                                    int_prim=False, reduce_mono=False)
    fslot = M.fine_slots()
    pg = oracle.assemble_patch_groups(fpatch, fbox, np.zeros((npg, 5, 22, 22, 22)), fslot, 3, (1, 1, 1))
    w = np.arange(22)
    for q, (I, J, K) in enumerate(M.fathers()):
        cx, cy, cz = ((16 * np.array([I, J, K])[:, None] + w[None, :] - 3) + 0.5) / 2.0 - 0.5   # coarse units
        Zc, Yc, Xc = np.meshgrid(cz, cy, cx, indexing="ij")
        W = (w - 3 + 8)
        has = fslot[q][np.ix_(W // 8, W // 8, W // 8)] >= 0
        for v in range(5):
            np.testing.assert_allclose(pg[q, v][~has], (lin(Xc, Yc, Zc) + 10.0 * (v + 1))[~has], rtol=0, atol=1e-12)
        s, Wm = fslot[q][np.ix_(W // 8, W // 8, W // 8)], W % 8
        same = fpatch[s, 0, Wm[:, None, None], Wm[None, :, None], Wm[None, None, :]]
        np.testing.assert_array_equal(pg[q, 0][has], same[has])
