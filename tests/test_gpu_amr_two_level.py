"""GPU: the two-level step on the device (gamer_amd.amr.TwoLevel: base step with StoreFlux, two refined substeps with
time-interpolated coarse ghost data and CorrectFlux, restriction and flux fix-up) against the same step composed from
the oracle (tests/amr_chain.py), and conservation of the device result."""
import numpy as np
import pytest

import amr_chain

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("int_scheme", ["CQUAD", "QUAR"])
def test_two_level_step_matches_oracle_chain(int_scheme):
    from gamer_amd import amr
    print(f"[gpu] two-level step {int_scheme}", flush=True)
    tab, base, fine, dh, dt = amr_chain.setup()
    rb, rf, rreg = amr_chain.two_level_step(base, fine, tab, dh, dt, int_scheme=int_scheme)
    lv = amr.TwoLevel(torch, torch.from_numpy(base).to(DEV), torch.from_numpy(fine).to(DEV), tab, dh,
                      int_scheme=int_scheme)
    lv.step(dt)
    torch.cuda.synchronize()
    gb, gf = lv.base.cpu().numpy(), lv.fine.cpu().numpy()
    assert lv.counters.cpu().tolist() == [0, 0]
    # registers are near-cancelling sums of dt * flux terms: compare on the scale of the largest register
    np.testing.assert_allclose(lv.reg.cpu().numpy(), rreg, rtol=0, atol=1e-10 * np.abs(rreg).max())
    np.testing.assert_allclose(gf, rf, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(gb, rb, rtol=1e-12, atol=1e-13)
    t0, s0 = base.sum(axis=(0, 2, 3, 4)), np.abs(base).sum(axis=(0, 2, 3, 4))
    assert np.all(np.abs(gb.sum(axis=(0, 2, 3, 4)) - t0) <= 1e-13 * s0)


def test_two_level_steps_fp32_stay_close():
    """Three steps in FP32-bulk against the FP64 oracle chain: relative L1 of the base level below 1e-5."""
    from gamer_amd import amr
    print("[gpu] two-level steps fp32", flush=True)
    tab, base, fine, dh, dt = amr_chain.setup()
    lv = amr.TwoLevel(torch, torch.from_numpy(base.astype(np.float32)).to(DEV),
                      torch.from_numpy(fine.astype(np.float32)).to(DEV), tab, dh)
    rb, rf = base, fine
    for _ in range(3):
        lv.step(dt)
        rb, rf, _ = amr_chain.two_level_step(rb, rf, tab, dh, dt)
    torch.cuda.synchronize()
    gb = lv.base.cpu().numpy().astype(np.float64)
    l1 = np.abs(gb - rb).sum(axis=(0, 2, 3, 4)) / np.abs(rb).sum(axis=(0, 2, 3, 4))
    l1[1:4] = np.abs(gb - rb)[:, 1:4].sum() / np.abs(rb[:, 1:4]).sum()
    assert np.all(l1 <= 1e-5), l1


def test_two_level_step_graph_replays_match_eager():
    """One captured step replayed three times, with eager work on the device between the replays, equals three eager
    steps bit for bit."""
    from gamer_amd import amr, hydro
    print("[gpu] two-level graph replay", flush=True)
    tab, base, fine, dh, dt = amr_chain.setup()
    ref = amr.TwoLevel(torch, torch.from_numpy(base).to(DEV), torch.from_numpy(fine).to(DEV), tab, dh)
    states = []
    for _ in range(3):
        ref.step(dt)
        states.append((ref.base.clone(), ref.fine.clone()))
    lv = amr.TwoLevel(torch, torch.from_numpy(base).to(DEV), torch.from_numpy(fine).to(DEV), tab, dh)
    graph = lv.capture(dt)
    assert torch.equal(lv.base, torch.from_numpy(base).to(DEV))          # capturing executes nothing
    for k in range(3):
        graph.replay()
        torch.cuda.synchronize()
        hydro.prepare_patches(torch.from_numpy(base).to(DEV), torch.zeros((8, 5, 22, 22, 22), dtype=torch.float64,
                              device=DEV), torch.from_numpy(tab.base_pg_slot).to(DEV), 3)   # eager work in between
        torch.cuda.synchronize()
        assert torch.equal(lv.base, states[k][0]) and torch.equal(lv.fine, states[k][1]), k
