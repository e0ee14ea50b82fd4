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
