"""The calibration form: many parameter sets on one grid, each fitted to its own target field
by gradient descent with the adjoint gradient (amcafe_heat_loss_grad_batch_f32).

Bars: each instance of a batch gives the single-instance loss and gradient bit for bit (the
sums run in a fixed order); a calibration reduces every instance's loss at least tenfold and
recovers the identifiable parameters (eta*P, v, sigma; P and eta enter the field only as their
product) within 5%, 5% and 10%, on a small case and at the stated size (32 instances on
128x128x48 voxels).
"""
import numpy as np
import pytest

import amcafe_amd
from oracle_util import read

pytestmark = pytest.mark.gpu

START = [200.0, 0.8, 0.35, 50e-6]


def _targets(n, seed=7):
    rng = np.random.default_rng(seed)
    return [[rng.uniform(150, 300), rng.uniform(0.5, 1.5), rng.uniform(0.3, 0.5), rng.uniform(40e-6, 60e-6)]
            for _ in range(n)]


def test_batch_matches_single_instances():
    desc = amcafe_amd.track_desc((48, 40, 16), steps=120, v_ref=1.0, checkpoint=50)
    tg = _targets(5)
    thetas = [[p * (1 + 0.05 * (b - 2)) for p in START] for b in range(5)]
    loss, grad, _ = amcafe_amd.HeatBatch(desc, tg).loss_grad_f32(thetas, 0, 120)
    for b in range(5):
        l1, g1, _ = amcafe_amd.HeatField(desc, tg[b]).loss_grad_f32(thetas[b], 0, 120)
        assert loss[b] == l1 and grad[b] == g1


def test_small_calibration_recovers_the_identifiable_parameters():
    desc = amcafe_amd.track_desc((64, 64, 24), steps=200, v_ref=1.0, checkpoint=100)
    tg = _targets(6)
    thetas, hist = amcafe_amd.calibrate(desc, tg, START, iters=60, lr=0.08, nsteps=200)
    for b in range(6):
        assert hist[-1][b] < 0.1 * hist[0][b], (b, hist[0][b], hist[-1][b])
        th, t = thetas[b], tg[b]
        assert abs(th[0] * th[2] / (t[0] * t[2]) - 1) < 0.05, (b, th, t)
        assert abs(th[1] / t[1] - 1) < 0.05, (b, th, t)
        assert abs(th[3] / t[3] - 1) < 0.10, (b, th, t)


@pytest.mark.slow
def test_calibration_at_stated_size():
    """32 instances on 128x128x48 voxels, one +x track of 400 steps, targets drawn with seed 7
    from the stated ranges: every instance ends within 5% on eta*P and v and 10% on sigma
    (instances that stall are restarted from the spread starts of calibrate())."""
    n = 32
    desc = amcafe_amd.track_desc((128, 128, 48), steps=400, v_ref=1.0, checkpoint=100)
    tg = _targets(n)
    thetas, hist = amcafe_amd.calibrate(desc, tg, START, iters=60, lr=0.08)
    for b in range(n):
        assert hist[-1][b] < 0.1 * hist[0][b], (b, hist[0][b], hist[-1][b])
        th, t = thetas[b], tg[b]
        assert abs(th[0] * th[2] / (t[0] * t[2]) - 1) < 0.05, (b, th, t)
        assert abs(th[1] / t[1] - 1) < 0.05, (b, th, t)
        assert abs(th[3] / t[3] - 1) < 0.10, (b, th, t)


def test_calibration_drives_the_production_run():
    """The calibration-and-production chain end to end: two instances fitted on 64x64x24 voxels,
    the first one's fitted (P, v, eta, sigma) turned into the production deck (1024x512x128,
    8 layers, 67 deg rotation, 5000-grain plate) by amcafe_amd.production_deck -- speed, time
    step, solidus and the melt-pool semi-axes of the fitted source -- and 40 production steps run
    from it."""
    desc = amcafe_amd.track_desc((64, 64, 24), steps=200, v_ref=1.0, checkpoint=100)
    tg = _targets(2)
    thetas, hist = amcafe_amd.calibrate(desc, tg, START, iters=60, lr=0.08)
    th = thetas[0]
    assert abs(th[0] * th[2] / (tg[0][0] * tg[0][2]) - 1) < 0.05 and abs(th[1] / tg[0][1] - 1) < 0.05
    deck = amcafe_amd.production_deck(th, read("c4_production.in"))
    kv = {ln.split()[0].lower(): ln.split()[1:] for ln in deck.split("\n") if ln.split() and ln[0] != "#"}
    assert float(kv["bmv"][0]) == round(th[1], 9)
    mp = amcafe_amd.semi_axes(*th)
    assert [float(q) for q in kv["meltparam"]] == [float(f"{round(q * 1e6)}e-6") for q in mp]
    sim = amcafe_amd.Sim(deck)
    events = 0
    for _ in range(40):
        sim.step()
        events += sim.stats(0)["n_events"]
    assert sim.info(0)["nX"] == [1024, 512, 128] and events > 100
