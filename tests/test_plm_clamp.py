"""The PLM face clamps under VL_GMINMOD cannot act for a MinMod coefficient <= 1.9999 (CPU, numpy).

The device PLM (gamer_amd/csrc/hydro_math.hpp, plm_face) builds the faces C -/+ s/2 with the branch-free van
Leer + generalised MinMod slope s and, for the compile-time VL_GMINMOD limiter, leaves out the two clamps of
DataReconstruction.cpp:415-430 (left face into [min(C,L), max(C,L)], right face 2C - left into [min(C,R),
max(C,R)]).  Its claim: for coefficients up to 1.9999 both clamps are no-ops in FP32 and FP64 arithmetic, so the
clamp-free faces carry the same bits.  Here the same arithmetic runs in numpy on inputs built to stress the
rounding (neighbours a few ulps apart, binade edges, subnormal differences, mixed scales, extrema), with the
product C - s/2 both rounded twice and fused; coefficient 2.0, which the launchers send to the clamped
run-time kernel, must show that the clamps do act there.
"""
import numpy as np
import pytest


def _inputs(dt, n, seed):
    rng = np.random.default_rng(seed)
    lo = -140 if dt is np.float32 else -1060
    e = rng.integers(lo, 120, n).astype(np.float64)
    kind = rng.integers(0, 3, n)
    base = np.where(kind == 0, 2.0 ** e, np.where(kind == 1, np.nextafter(2.0 ** e, 0), (1 + rng.random(n)) * 2.0 ** e))
    C = (base * np.where(rng.random(n) < 0.5, -1, 1)).astype(dt)
    C[rng.random(n) < 0.02] = 0
    sgn = np.where(rng.random(n) < 0.5, 1, -1).astype(dt)
    L, R = C.copy(), C.copy()
    near = rng.random(n) < 0.6                    # neighbours 0..4 ulps away
    kl, kr = rng.integers(0, 5, n), rng.integers(0, 5, n)
    for k in range(4):
        ml, mr = near & (kl > k), near & (kr > k)
        L[ml] = np.nextafter(L[ml], (-sgn[ml] * np.inf).astype(dt))
        R[mr] = np.nextafter(R[mr], (sgn[mr] * np.inf).astype(dt))
    far = ~near                                   # relative distances 1e-9 .. 10, some extrema
    m = int(far.sum())
    tiny = float(np.finfo(dt).tiny)
    d1 = np.abs(C[far]).astype(np.float64) * 10.0 ** rng.uniform(-9, 1, m) + tiny
    d2 = np.abs(C[far]).astype(np.float64) * 10.0 ** rng.uniform(-9, 1, m) + tiny
    turn = np.where(rng.random(m) < 0.2, -1, 1)
    L[far] = (C[far] - sgn[far] * d1).astype(dt)
    R[far] = (C[far] + turn * sgn[far] * d2).astype(dt)
    keep = np.isfinite(L) & np.isfinite(R) & (np.abs(C.astype(np.float64)) < float(np.finfo(dt).max) / 4)
    return C[keep], L[keep], R[keep]


def _faces(C, L, R, theta, fused):
    """(clamped left, clamped right, clamp-free left, clamp-free right), the device's operation order."""
    dt = C.dtype.type
    with np.errstate(all="ignore"):
        sl, sr = C - L, R - C
        t = sl + sr
        h = dt(2) * (sl * sr) * (dt(1) / np.abs(t))
        m = np.fmin(np.fmin(np.abs(sl) * dt(theta), np.abs(sr) * dt(theta)), np.fmax(h, dt(0)))
        s = np.copysign(m, t).astype(dt)
        if fused:
            x = (C.astype(np.longdouble) - np.longdouble(0.5) * s.astype(np.longdouble)).astype(dt)
        else:
            x = C - dt(0.5) * s
        fl = np.minimum(np.maximum(x, np.minimum(C, L)), np.maximum(C, L))
        fr = np.minimum(np.maximum(dt(2) * C - fl, np.minimum(C, R)), np.maximum(C, R))
        return fl, fr, x, dt(2) * C - x


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("fused", [False, True])
def test_clamps_inactive_below_two(dtype, fused):
    C, L, R = _inputs(dtype, 120_000, 31 + fused)
    for theta in (0.0, 1.0, 1.5, 1.9, 1.999, 1.9999):
        fl, fr, xl, xr = _faces(C, L, R, theta, fused)
        bad = (fl.view(np.uint8) != xl.view(np.uint8)).reshape(len(C), -1).any(1) | \
              (fr.view(np.uint8) != xr.view(np.uint8)).reshape(len(C), -1).any(1)
        assert not bad.any(), (theta, C[bad][:3], L[bad][:3], R[bad][:3])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_clamps_act_at_two(dtype):
    C, L, R = _inputs(dtype, 120_000, 5)
    fl, fr, xl, xr = _faces(C, L, R, 2.0, False)
    assert ((fl != xl) | (fr != xr)).any()


def test_subnormal_extremum_has_zero_slope():
    """Both slopes subnormal with opposite signs: sl*sr underflows to -0 and 1/|sl+sr| overflows, so h is NaN;
    max(h, 0) turns that into a zero slope, as the reference's sl*sr > 0 test does."""
    f32 = np.float32
    C, L, R = f32(-1.1815557e-37), f32(-1.2991057e-37), f32(-1.3160595e-37)
    fl, fr, xl, xr = _faces(np.array([C]), np.array([L]), np.array([R]), 1.5, False)
    assert xl[0] == C and xr[0] == C
