"""The calibration form: many parameter sets on one grid, each fitted to its own target field
by gradient descent with the adjoint gradient (amcafe_heat_loss_grad_batch_f32).

Bars: each instance of a batch gives the single-instance loss and gradient (FP32 block sums
in a different order: 1e-5 relative); a small calibration reduces every instance's loss at
least tenfold and recovers the identifiable parameters (eta*P, v, sigma; P and eta enter the
field only as their product) within 5%, 5% and 10%.
"""
import numpy as np
import pytest

import amcafe_amd

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
        assert abs(loss[b] - l1) <= 1e-5 * abs(l1)
        np.testing.assert_allclose(grad[b], g1, rtol=1e-5)


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
