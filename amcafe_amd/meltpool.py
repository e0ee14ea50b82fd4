"""Melt-pool semi-axes for the analytic temperature model, and production decks from fitted
beam parameters (the calibration-and-production form).

The reference's analytic model (TempField::AnalyticTempCurr, src/src-mpi/TempField.C:71-76,
175-190) needs the melt pool's extents as `meltparam` (front, tail, half-width, depth).  They
are taken as the extent of the T >= tL isosurface of the constant-property moving Gaussian
source -- the form of Utilities::Temp_Fun_G (src/src-mpi/Utilities.C:228-249) with constant
conductivity k and diffusivity alpha, the same field the differentiable heat kernels evaluate --
in the source's frame, integrated over its whole history with scipy quadrature.
"""
import math
import re

import numpy as np

AVEL, NVEL = 2e-5, 3.0  # dendritegrowth2 of the decks: v = Avel dT^nvel


def temperature(x, y, z, P, v, eta, sigma, k=20.0, alpha=5e-6, T0=300.0):
    """Steady temperature at (x, y, z) from the source (x along the track, the source at the
    origin moving towards +x, z depth below the surface)."""
    from scipy.integrate import quad
    C = eta * P / (np.pi * (k / alpha) * np.sqrt(4 * np.pi * alpha))
    f = lambda tp: C * tp ** -0.5 / (2 * alpha * tp + sigma ** 2) * np.exp(
        -((x + v * tp) ** 2 + y ** 2) / (4 * alpha * tp + 2 * sigma ** 2) - z ** 2 / (4 * alpha * tp))
    val, _ = quad(f, 1e-12, 2e-3, limit=400, points=[1e-7, 1e-6, 1e-5, 1e-4])
    return T0 + val


def _edge(f, a, b, tol=1e-7):
    """Bisection for the boundary between f(a) True and f(b) False."""
    while abs(b - a) > tol:
        m = .5 * (a + b)
        if f(m):
            a = m
        else:
            b = m
    return a


def semi_axes(P, v, eta=0.35, sigma=50e-6, tL=1733.0, k=20.0, alpha=5e-6):
    """(front, tail, half-width, depth) in metres of the T >= tL region, or None if the source
    does not melt: along the track at the surface, the widest cross-section, and the depth
    under the point of widest cross-section."""
    z0 = 1e-9
    hot = lambda x, y, z: temperature(x, y, z, P, v, eta, sigma, k, alpha) >= tL
    xs = np.linspace(-800e-6, 200e-6, 251)
    h = [x for x in xs if hot(x, 0, z0)]
    if not h:
        return None
    front = _edge(lambda x: hot(x, 0, z0), max(h), max(h) + 4e-6)
    tail = -_edge(lambda x: hot(x, 0, z0), min(h), min(h) - 4e-6)
    best = (0, 0)
    for x in np.linspace(-tail, front, 40):
        if hot(x, 0, z0):
            w = _edge(lambda y: hot(x, y, z0), 0.0, 400e-6)
            if w > best[0]:
                best = (w, x)
    hw, xw = best
    depth = _edge(lambda z: hot(xw, 0, z), 1e-7, 400e-6)
    return front, tail, hw, depth


def _set(text, key, value):
    pat = re.compile(rf"^{key}\b.*$", re.MULTILINE | re.IGNORECASE)
    if not pat.search(text):
        raise ValueError(f"deck has no {key} line")
    return pat.sub(f"{key} {value}", text, count=1)


def production_deck(theta, template, tL=1733.0):
    """The production deck for fitted beam parameters theta = (P, v, eta, sigma): the template
    deck (grid, layers, raster, nucleation) with the beam's speed, time step (one voxel per step),
    solidus (largest growth per step half a voxel at the largest undercooling) and the melt-pool
    semi-axes of the fitted source; the laser grid is widened, if needed, so that every track
    overhangs the domain by more than half a hatch plus the new tail."""
    P, v, eta, sigma = (float(q) for q in theta)
    kv = {}
    for line in template.split("\n"):
        p = line.split("#")[0].split()
        if p:
            kv[p[0].lower()] = p[1:]
    dx = float(kv["dx"][0])
    nX = [int(q) for q in kv["nx"]]
    mp = semi_axes(P, v, eta, sigma, tL)
    if mp is None:
        raise ValueError(f"the fitted source (eta P = {eta * P:.4g} W, v = {v:.4g} m/s) does not melt")
    front, tail, hw, depth = (round(q * 1e6) for q in mp)
    tS = tL - (0.5 * v / AVEL) ** (1.0 / NVEL)
    half_diag = 0.5 * math.hypot(nX[0] * dx, nX[1] * dx)
    bhatch = float(kv["bhatch"][0])
    gx_need = 2 * (half_diag + bhatch / 2 + tail * 1e-6 + 40e-6)
    gx = max(float(kv["gridsize"][0]), gx_need)
    out = template
    out = _set(out, "bmP", repr(round(P, 6)))
    out = _set(out, "bmV", repr(round(v, 9)))
    out = _set(out, "bmDelT", repr(dx / round(v, 9)))
    out = _set(out, "tS", f"{tS:.3f}")
    out = _set(out, "meltparam", f"{front}e-6 {tail}e-6 {hw}e-6 {depth}e-6")
    out = _set(out, "gridsize", f"{round(gx * 1e6)}e-6 {kv['gridsize'][1]}")
    head = (f"# production deck of the fitted beam P {P:.6g} W, v {v:.6g} m/s, eta {eta:.6g}, sigma {sigma:.6g} m "
            f"(amcafe_amd.meltpool.production_deck)\n")
    return head + out
