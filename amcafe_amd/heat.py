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
