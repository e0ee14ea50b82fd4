"""amcafe_heat_loss_grad_batch_f32_noalloc: the differentiable field's batch loss and adjoint
gradient with every device buffer in the library's static scratch array.  Checked through the
C ABI against amcafe_heat_loss_grad_batch_f32 (the same kernels and the same fixed-order sums,
so the same bits), on two calls in a row, the second smaller than the first (the array keeps
the first call's contents), and a request beyond the array's size is refused before any
device work."""
import ctypes

import numpy as np
import pytest

import amcafe_amd
from amcafe_amd.heat import track_desc

START = [200.0, 0.8, 0.35, 50e-6]


def _targets(n, seed=11):
    rng = np.random.default_rng(seed)
    return [[rng.uniform(150, 300), rng.uniform(0.5, 1.5), rng.uniform(0.3, 0.5), rng.uniform(40e-6, 60e-6)]
            for _ in range(n)]


def _call(name, desc, thetas, targets, nsteps):
    L = amcafe_amd.lib()
    n = len(thetas)
    th = (ctypes.c_double * (4 * n))(*[float(v) for t in thetas for v in t])
    tg = (ctypes.c_double * (4 * n))(*[float(v) for t in targets for v in t])
    loss, grad = (ctypes.c_double * n)(), (ctypes.c_double * (4 * n))()
    if name == "noalloc":
        rc = L.amcafe_heat_loss_grad_batch_f32_noalloc(ctypes.byref(desc), n, th, tg, 0, nsteps, loss, grad)
    else:
        rc = L.amcafe_heat_loss_grad_batch_f32(ctypes.byref(desc), n, th, tg, 0, nsteps, loss, grad, None)
    assert rc == 0, L.amcafe_last_error().decode()
    return list(loss), list(grad)


@pytest.mark.gpu
def test_noalloc_matches_the_allocating_entry_point_on_two_calls():
    big = track_desc((96, 80, 32), steps=150, v_ref=1.0, checkpoint=40)
    small = track_desc((48, 40, 16), steps=60, v_ref=1.0, checkpoint=25)
    t6, t3 = _targets(6), _targets(3, seed=12)
    th6 = [[p * (1 + 0.04 * (b - 3)) for p in START] for b in range(6)]
    th3 = [[p * (1 - 0.03 * b) for p in START] for b in range(3)]
    for desc, th, tg, ns in ((big, th6, t6, 150), (small, th3, t3, 60)):
        got = _call("noalloc", desc, th, tg, ns)
        want = _call("batch", desc, th, tg, ns)
        assert got == want
        assert all(v > 0 for v in got[0])


def test_noalloc_refuses_requests_beyond_its_array():
    L = amcafe_amd.lib()
    desc = track_desc((256, 256, 64), steps=10)
    mx = L.amcafe_heat_noalloc_max_bytes()
    assert L.amcafe_heat_noalloc_bytes(ctypes.byref(desc), 25) <= mx < L.amcafe_heat_noalloc_bytes(ctypes.byref(desc), 26)
    n = 26
    th = (ctypes.c_double * (4 * n))(*(START * n))
    loss, grad = (ctypes.c_double * n)(), (ctypes.c_double * (4 * n))()
    rc = L.amcafe_heat_loss_grad_batch_f32_noalloc(ctypes.byref(desc), n, th, th, 0, 10, loss, grad)
    assert rc == -1 and b"scratch" in L.amcafe_last_error()
