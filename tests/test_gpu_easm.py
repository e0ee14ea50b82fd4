"""EASM temperature on the device (amcafe_easm_temp_curr) against the oracle's restatement of
TempField::EASM with glibc math (S/TempField.C:79-137, S/Utilities.C:60-140, 228-249).

Bar: temperatures within 1e-9 relative (the device's FP64 exp and pow are not glibc's, so the
last bits may differ), BuildjID bit-exact, the end-of-scan flag and the number of source
integrals equal -- the latter shows that no Picard stopping decision (|dT| > 10 K) flipped.
"""
import math

import numpy as np
import pytest

import amcafe_amd
from oracle_util import Oracle, read

pytestmark = pytest.mark.gpu


def _domain(deck):
    kv = {}
    for line in deck.split("\n"):
        p = line.split("#")[0].split()
        if p:
            kv[p[0].lower()] = p[1:]
    nX = [int(v) for v in kv["nx"]]
    dom = amcafe_amd.Domain(nX, float(kv["dx"][0]), float(kv["tl"][0]), float(kv["ts"][0]), 300.0,
                            [float(v) for v in kv["meltparam"]], float(kv["bmv"][0]), float(kv["bmdelt"][0]),
                            float(kv["r"][0]), float(kv["dendritegrowth2"][0]), float(kv["dendritegrowth2"][1]))
    return dom, nX, float(kv["dx"][0]), float(kv["bmp"][0]), float(kv["lt"][0])


@pytest.mark.parametrize("deck_name", ["t_small.in", "c0_single_track.in"])
def test_easm_matches_the_oracle(deck_name):
    deck = read(deck_name)
    dom, nX, dx, bmP, lT = _domain(deck)
    orc = Oracle(deck, crmath=False)
    Lx, Ly = nX[0] * dx, nX[1] * dx
    il = nX[2] - 2
    cases = [(0.5 * Lx, 0.5 * Ly, 0.0), (0.2 * Lx, 0.4 * Ly, 0.0), (0.7 * Lx, 0.5 * Ly, math.pi / 2),
             (0.45 * Lx, 0.6 * Ly, 1.1693705988362009), (-2e-3, 0.5 * Ly, 0.0)]
    worst = 0.0
    for (x, y, gth) in cases:
        z = il * dx
        allm_d, n_d = dom.easm_temperature(x, y, z, 1.0, gth, il, bmP, lT)
        allm_o, n_o = orc.easm(x, y, z, gth)
        g, o = dom.download(), orc.fields()
        assert n_d == n_o, (x, y, gth, n_d, n_o)
        assert allm_d == allm_o
        assert np.array_equal(g["BuildjID"], o["BuildjID"]), (x, y, gth)
        Td, To = g["TempCurr"], o["TempCurr"]
        assert np.array_equal(Td == 0.0, To == 0.0)
        rel = np.abs(Td - To) / np.maximum(np.abs(To), 1.0)
        worst = max(worst, float(rel.max()))
        np.testing.assert_allclose(Td, To, rtol=1e-9, atol=0)
    print(f"{deck_name}: largest relative temperature difference {worst:.3g}")


def test_easm_temperatures_drive_update_voxels():
    """EASM -> UpdateVoxels at the seam level: every growing cell sits at its own EASM
    temperature, so each growth velocity Avel (tL - T)^nvel is evaluated on the device
    (pow rounded to nearest).  The device's EASM field is handed to both oracles, which then
    run UpdateVoxels on the same temperatures: every field and the grain table bit-exact
    against the correctly rounded oracle; grain ids and states bit-exact and octahedra
    within 1e-12 relative against the glibc oracle.  The laser advances one voxel a step
    along the deck's first track height from a state with a solidified analytic track."""
    deck = read("t_small.in")
    dom, nX, dx, bmP, lT = _domain(deck)
    cr, gl = Oracle(deck, crmath=True), Oracle(deck, crmath=False)
    for _ in range(40):
        cr.step()
        gl.step()
    f = cr.fields()
    dom.upload(gID=f["gID"], vState=f["vState"], extents=f["extents"], centroidOct=f["centroidOct"],
               BuildjID=f["BuildjID"])
    dom.set_grains(cr.grains())
    sc = cr.scalars()
    il, t0, isp = int(sc["ilaserLoc"]), int(sc["tInd"]), int(sc["isp"])
    y = 0.5 * nX[1] * dx
    events = 0
    odd_T = 0
    for k in range(16):
        x = (6 + k) * dx
        tind = t0 + k
        dom.easm_temperature(x, y, il * dx, 1.0, 0.0, il, bmP, lT)
        g = dom.download()
        for o in (cr, gl):
            o.set_fields(TempCurr=g["TempCurr"], BuildjID=g["BuildjID"])
            o.L.orc_set_laser(o.h, isp, il, tind)
        st = dom.update_voxels(tind, il)
        cr.update_voxels()
        gl.update_voxels()
        d, c, lg = dom.download(), cr.fields(), gl.fields()
        for key in ("gID", "vState", "extents", "centroidOct", "BuildjID"):
            assert np.array_equal(d[key], c[key], equal_nan=d[key].dtype.kind == "f"), f"step {k}: {key}"
        assert np.array_equal(d["gID"], lg["gID"]) and np.array_equal(d["vState"], lg["vState"]), f"step {k}"
        np.testing.assert_allclose(d["extents"], lg["extents"], rtol=1e-12, atol=1e-21)
        assert np.array_equal(dom.grains(), cr.grains()), f"step {k}: grain table"
        so = cr.stats()
        assert st["n_events"] == so["n_events"] and st["draws"] == 2 * so["draws"]
        events += st["n_events"]
        mushy = d["vState"] == 2
        odd_T += int(np.sum(mushy & (d["TempCurr"] < 1733.0) & (d["TempCurr"] != 300.0)))
    assert events > 20 and odd_T > 20
