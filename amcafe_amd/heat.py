"""Differentiable temperature field over the C ABI (include/amcafe.h, section 3).

HeatField evaluates the sigmoid-indicator loss of the constant-property moving Gaussian source
(the form of Utilities::Temp_Fun_G, S/Utilities.C:228-249) against a target parameter set, in
FP32 with its hand-written adjoint, or in FP64 for finite differences.  HeatLoss is the
torch.autograd.Function over it: forward runs the FP32 loss kernel, backward re-runs the field
with the adjoint kernel (nothing is stored between the two).
"""
import ctypes

from ._lib import check, lib

PARAMS = ("P", "v", "eta", "sigma")


class HeatDesc(ctypes.Structure):
    _fields_ = [("nX", ctypes.c_int * 3), ("dX", ctypes.c_double), ("x_start", ctypes.c_double),
                ("y_track", ctypes.c_double), ("z_track", ctypes.c_double), ("dt", ctypes.c_double),
                ("t_hist", ctypes.c_double), ("T0", ctypes.c_double), ("tL", ctypes.c_double),
                ("width", ctypes.c_double), ("k", ctypes.c_double), ("alpha", ctypes.c_double),
                ("checkpoint", ctypes.c_int)]


def track_desc(nX, dX=2.5e-6, steps=2000, v_ref=0.8, checkpoint=100, t_hist=8e-4):
    """A straight +x track along the top surface through the middle of the domain, with the
    step length chosen so that a source at v_ref crosses the domain in `steps` steps."""
    d = HeatDesc()
    d.nX[:] = list(nX)
    d.dX = dX
    d.x_start = 0.0
    d.y_track = nX[1] * dX / 2
    d.z_track = nX[2] * dX
    d.dt = nX[0] * dX / (v_ref * steps)
    d.t_hist = t_hist
    d.T0, d.tL, d.width = 300.0, 1733.0, 10.0
    d.k, d.alpha = 20.0, 5e-6
    d.checkpoint = checkpoint
    d.steps = steps  # the track's length in steps (the default of calibrate's nsteps)
    return d


def _arr(x):
    return (ctypes.c_double * 4)(*[float(v) for v in x])


class HeatField:
    def __init__(self, desc, target):
        self.desc = desc
        self.target = list(target)

    def loss_f32(self, theta, step0, nsteps):
        out, ms = ctypes.c_double(), ctypes.c_float()
        check(lib().amcafe_heat_loss_f32(ctypes.byref(self.desc), _arr(theta), _arr(self.target), step0, nsteps,
                                         ctypes.byref(out), ctypes.byref(ms)))
        return out.value, ms.value

    def loss_grad_f32(self, theta, step0, nsteps):
        out, ms = ctypes.c_double(), ctypes.c_float()
        g = (ctypes.c_double * 4)()
        check(lib().amcafe_heat_loss_grad_f32(ctypes.byref(self.desc), _arr(theta), _arr(self.target), step0,
                                              nsteps, ctypes.byref(out), g, ctypes.byref(ms)))
        return out.value, list(g), ms.value

    def loss_f64(self, theta, step0, nsteps):
        out, ms = ctypes.c_double(), ctypes.c_float()
        check(lib().amcafe_heat_loss_f64(ctypes.byref(self.desc), _arr(theta), _arr(self.target), step0, nsteps,
                                         ctypes.byref(out), ctypes.byref(ms)))
        return out.value, ms.value

    def central_differences(self, theta, step0, nsteps, h=1e-4):
        """FP64 central differences with relative step h."""
        g = []
        for q in range(4):
            tp, tm = list(theta), list(theta)
            tp[q] *= 1 + h
            tm[q] *= 1 - h
            g.append((self.loss_f64(tp, step0, nsteps)[0] - self.loss_f64(tm, step0, nsteps)[0]) /
                     (2 * h * theta[q]))
        return g


class HeatBatch:
    """Many parameter sets on one grid and track, each against its own target: loss and
    adjoint gradient of all of them in one launch per checkpoint segment."""

    def __init__(self, desc, targets):
        self.desc = desc
        self.targets = [list(t) for t in targets]

    def loss_grad_f32(self, thetas, step0, nsteps):
        n = len(self.targets)
        if len(thetas) != n:
            raise ValueError("one parameter set per target")
        th = (ctypes.c_double * (4 * n))(*[float(v) for t in thetas for v in t])
        tg = (ctypes.c_double * (4 * n))(*[float(v) for t in self.targets for v in t])
        loss, g, ms = (ctypes.c_double * n)(), (ctypes.c_double * (4 * n))(), ctypes.c_float()
        check(lib().amcafe_heat_loss_grad_batch_f32(ctypes.byref(self.desc), n, th, tg, step0, nsteps, loss, g,
                                                    ctypes.byref(ms)))
        return list(loss), [list(g[4 * b:4 * b + 4]) for b in range(n)], ms.value


def _adam(batch, thetas0, iters, lr, step0, nsteps):
    """Adam on the log-parameters of every instance of the batch; (thetas, losses per iteration)."""
    import math

    n = len(thetas0)
    z = [[math.log(v) for v in t] for t in thetas0]
    m = [[0.0] * 4 for _ in range(n)]
    s2 = [[0.0] * 4 for _ in range(n)]
    b1, b2, eps = 0.9, 0.999, 1e-12
    history = []
    for it in range(1, iters + 1):
        thetas = [[math.exp(v) for v in zb] for zb in z]
        loss, grad, _ = batch.loss_grad_f32(thetas, step0, nsteps)
        history.append(loss)
        for b in range(n):
            for q in range(4):
                gq = grad[b][q] * thetas[b][q]  # d loss / d log theta
                m[b][q] = b1 * m[b][q] + (1 - b1) * gq
                s2[b][q] = b2 * s2[b][q] + (1 - b2) * gq * gq
                mh, vh = m[b][q] / (1 - b1 ** it), s2[b][q] / (1 - b2 ** it)
                z[b][q] -= lr * mh / (math.sqrt(vh) + eps)
    thetas = [[math.exp(v) for v in zb] for zb in z]
    loss, _, _ = batch.loss_grad_f32(thetas, step0, nsteps)
    history.append(loss)
    return thetas, history


# alternative starts for instances that stop in a local minimum: speed and radius, which shape
# the melt pool most, spread over the calibration ranges (power and absorptivity kept)
RESTART_V = (0.6, 1.0, 1.4)
RESTART_SIGMA = (42e-6, 58e-6)


def calibrate(desc, targets, theta0, iters=40, lr=0.08, step0=0, nsteps=None, restart_below=1e-2):
    """Fit one parameter set (P, v, eta, sigma) per target field by gradient descent with the
    adjoint gradient: Adam on the logarithms of the parameters (all positive), every instance
    advanced in the same launches.  P and eta enter the field only as the product eta*P, so the
    fit determines eta*P, v and sigma; the split of eta*P stays where theta0 puts it.
    Instances whose loss has not fallen below restart_below times its first value are run again
    from the starts RESTART_V x RESTART_SIGMA (all of them in one batch) and keep the best fit.
    nsteps defaults to the track's length recorded by track_desc; a missing or non-positive step
    count is an error (a fit over zero steps has a zero gradient and would return theta0).
    Returns (thetas, losses per iteration of the first run)."""
    n = len(targets)
    if nsteps is None:
        nsteps = getattr(desc, "steps", None)
    if nsteps is None or int(nsteps) <= 0:
        raise ValueError("calibrate needs a positive step count: pass nsteps or build desc with track_desc")
    nsteps = int(nsteps)
    if n == 0:
        raise ValueError("calibrate needs at least one target")
    thetas, history = _adam(HeatBatch(desc, targets), [list(theta0)] * n, iters, lr, step0, nsteps)
    best = list(history[-1])
    stuck = [b for b in range(n) if not (history[-1][b] < restart_below * history[0][b])]
    if stuck and restart_below > 0:
        starts = [[theta0[0], v, theta0[2], sg] for v in RESTART_V for sg in RESTART_SIGMA]
        tg = [targets[b] for b in stuck for _ in starts]
        th0 = [st for _ in stuck for st in starts]
        th, hist = _adam(HeatBatch(desc, tg), th0, iters, lr, step0, nsteps)
        for i, b in enumerate(stuck):
            for j in range(len(starts)):
                k = i * len(starts) + j
                if hist[-1][k] < best[b]:
                    best[b] = hist[-1][k]
                    thetas[b] = th[k]
    history.append(best)
    return thetas, history


def heat_loss_function():
    """torch.autograd.Function over HeatField (imported lazily: torch is optional)."""
    import torch

    class HeatLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, field, step0, nsteps):
            th = theta.detach().double().cpu().tolist()
            ctx.field, ctx.args, ctx.th = field, (step0, nsteps), th
            loss, _ = field.loss_f32(th, step0, nsteps)
            return theta.new_tensor(loss)

        @staticmethod
        def backward(ctx, gout):
            _, g, _ = ctx.field.loss_grad_f32(ctx.th, *ctx.args)
            return gout * gout.new_tensor(g), None, None, None

    return HeatLoss
