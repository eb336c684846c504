"""The oracle restatement against golden fixtures written from GAMER's own compiled CPU_FluidSolver
(tests/golden/make_golden.py): bit-for-bit output/flux hashes, plus full-precision samples and field sums."""
import hashlib
import json
import os

import numpy as np
import pytest

from gamer_amd import ic

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _input(case):
    gen = dict(case["generator"])
    gen["npg_shape"] = tuple(gen["npg_shape"])
    w = ic.random_sod(**gen)
    u = w.u if case["precision"] == "f64" else w.u.astype(np.float32)
    return w, np.ascontiguousarray(u)


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_oracle_matches_golden(name):
    import oracle
    case = GOLDEN[name]
    w, u = _input(case)
    assert _sha(u) == case["input_sha256"], "input generator drifted"
    assert w.dt == case["dt"] and w.dh == case["dh"]
    out, flux = oracle.fluid_solver(u, w.dt, w.dh, **case["solver"])
    np.testing.assert_array_equal(out.ravel()[case["out_sample_idx"]], np.array(case["out_sample"], dtype=out.dtype))
    np.testing.assert_array_equal(flux.ravel()[case["flux_sample_idx"]], np.array(case["flux_sample"], dtype=flux.dtype))
    np.testing.assert_allclose(out.astype(np.float64).sum(axis=(0, 2, 3, 4)), case["out_field_sums"], rtol=1e-13)
    assert _sha(out) == case["out_sha256"]
    assert _sha(flux) == case["flux_sha256"]


def test_oracle_matches_full_golden_arrays():
    import oracle
    ref = np.load(os.path.join(HERE, "golden", "golden_rp_hllc_f64.npz"))
    w, u = _input(GOLDEN["mhm_rp_hllc_f64"])
    out, flux = oracle.fluid_solver(u, w.dt, w.dh)
    np.testing.assert_array_equal(out, ref["out"])
    np.testing.assert_array_equal(flux, ref["flux"])


def test_retry_case_exercises_the_retry_loop():
    """The retry fixtures differ from a run without retries: the MinMod loop really ran in the reference."""
    import oracle
    case = GOLDEN["mhm_rp_hllc_f64_retry"]
    w, u = _input(case)
    out0, _ = oracle.fluid_solver(u, w.dt, w.dh, minmod_maxiter=0)
    assert _sha(out0) != case["out_sha256"]
