"""The differentiable temperature field (an extension without a reference counterpart):
the FP32 loss against the FP64 one, the FP32 adjoint gradient against FP64 central
differences (relative step 1e-4), the torch.autograd.Function against the adjoint, and the
closed forms: zero loss and gradient at the target, and the loss's invariance to the launch
segmentation (the field of a step depends only on the parameters and the time)."""
import numpy as np
import pytest

import amcafe_amd
from amcafe_amd.heat import HeatField, track_desc, heat_loss_function

pytestmark = pytest.mark.gpu

TARGET = (220.0, 0.9, 0.4, 45e-6)
START = (200.0, 0.8, 0.35, 50e-6)


def _field(steps=60, nX=(64, 32, 16), checkpoint=100):
    return HeatField(track_desc(nX, steps=steps, checkpoint=checkpoint), TARGET)


def test_fp32_loss_matches_fp64_and_is_zero_at_the_target():
    f = _field()
    l32, _ = f.loss_f32(START, 0, 60)
    l64, _ = f.loss_f64(START, 0, 60)
    assert l64 > 1.0
    assert abs(l32 - l64) <= 1e-4 * l64, (l32, l64)
    l0, g0, _ = f.loss_grad_f32(TARGET, 0, 60)
    assert l0 == 0.0 and g0 == [0.0, 0.0, 0.0, 0.0]


def test_loss_does_not_depend_on_the_segmentation():
    a = _field(checkpoint=7).loss_f64(START, 0, 60)[0]
    b = _field(checkpoint=100).loss_f64(START, 0, 60)[0]
    c = _field().loss_f64(START, 0, 25)[0] + _field().loss_f64(START, 25, 35)[0]
    assert abs(a - b) <= 1e-12 * b and abs(c - b) <= 1e-12 * b


def test_adjoint_gradient_matches_fp64_central_differences():
    f = _field()
    _, g, _ = f.loss_grad_f32(START, 0, 60)
    fd = f.central_differences(START, 0, 60, h=1e-4)
    for q in range(4):
        assert abs(g[q] - fd[q]) <= 1e-4 * abs(fd[q]), (amcafe_amd.heat.PARAMS[q], g[q], fd[q])


def test_autograd_function_uses_the_adjoint():
    import torch
    f = _field()
    theta = torch.tensor(START, dtype=torch.float64, requires_grad=True)
    loss = heat_loss_function().apply(theta, f, 0, 60)
    loss.backward()
    l, g, _ = f.loss_grad_f32(START, 0, 60)
    # the sums run in a fixed order: the autograd path returns the adjoint's bits
    assert float(loss.detach()) == l
    assert theta.grad.numpy().tolist() == g


def test_loss_and_gradient_are_bit_reproducible():
    """Fixed-order sums: the same call twice gives the same bits, for any segmentation."""
    f = _field(steps=60, checkpoint=13)
    a = f.loss_grad_f32(START, 0, 60)
    b = f.loss_grad_f32(START, 0, 60)
    assert a[0] == b[0] and a[1] == b[1]
    assert f.loss_f32(START, 0, 60)[0] == f.loss_f32(START, 0, 60)[0]


def test_pruned_voxels_contribute_nothing():
    """On a domain much wider (640 um in y) and deeper than the heated band around the track,
    voxels far from it are skipped; the FP32 loss still matches the FP64 loss, which evaluates
    every voxel, and the gradient still matches central differences."""
    nX = (48, 256, 64)
    f = HeatField(track_desc(nX, steps=40, checkpoint=20), TARGET)
    l32, g, _ = f.loss_grad_f32(START, 0, 40)
    evals = amcafe_amd.lib().amcafe_heat_last_evaluations()
    l64, _ = f.loss_f64(START, 0, 40)
    total = nX[0] * nX[1] * nX[2] * 40
    print(f"evaluated {evals} of {total} voxel-steps ({evals / total:.1%})")
    assert 0 < evals < 0.8 * total
    assert abs(l32 - l64) <= 1e-4 * l64, (l32, l64)
    fd = f.central_differences(START, 0, 40, h=1e-4)
    for q in range(4):
        assert abs(g[q] - fd[q]) <= 1e-4 * abs(fd[q]), (q, g[q], fd[q])


@pytest.mark.slow
def test_full_size_adjoint_against_central_differences():
    """The differentiation form at its stated size: 256x256x64 voxels, 2000 steps of one track,
    checkpoints of 100 steps; the FP32 adjoint gradient within 1e-4 relative of FP64 central
    differences (h = 1e-4 relative) for P, v, eta and sigma."""
    f = HeatField(track_desc((256, 256, 64), steps=2000, checkpoint=100), TARGET)
    loss, g, ms_g = f.loss_grad_f32(START, 0, 2000)
    _, ms_f = f.loss_f32(START, 0, 2000)
    fd = f.central_differences(START, 0, 2000, h=1e-4)
    print(f"forward {ms_f:.1f} ms, forward + adjoint {ms_g:.1f} ms, loss {loss:.6g}")
    for q in range(4):
        assert abs(g[q] - fd[q]) <= 1e-4 * abs(fd[q]), (amcafe_amd.heat.PARAMS[q], g[q], fd[q])
    assert ms_g <= 3.0 * ms_f
