"""Host-side logic of the Python layer that needs no GPU: argument checks of calibrate() and the
production deck built from fitted beam parameters."""
import math

import pytest

import amcafe_amd
from amcafe_amd.heat import HeatDesc
from oracle_util import Oracle, read


def test_calibrate_rejects_a_missing_step_count():
    d = HeatDesc()  # built without track_desc: no recorded step count
    with pytest.raises(ValueError, match="step count"):
        amcafe_amd.calibrate(d, [[200.0, 0.8, 0.35, 50e-6]], [200.0, 0.8, 0.35, 50e-6])
    d = amcafe_amd.track_desc((8, 8, 8), steps=10)
    with pytest.raises(ValueError, match="step count"):
        amcafe_amd.calibrate(d, [[200.0, 0.8, 0.35, 50e-6]], [200.0, 0.8, 0.35, 50e-6], nsteps=0)
    with pytest.raises(ValueError, match="target"):
        amcafe_amd.calibrate(d, [], [200.0, 0.8, 0.35, 50e-6])
    assert d.steps == 10


def test_production_deck_from_fitted_parameters():
    """The deck keeps the template's grid and raster and takes the fitted beam's speed, one voxel
    per step, the solidus for half a voxel of growth per step and the fitted melt pool; the
    reference's own parser (restated in the oracle) reads it."""
    th = (231.0, 1.12, 0.37, 47e-6)
    deck = amcafe_amd.production_deck(th, read("c4_production.in"))
    kv = {ln.split()[0].lower(): ln.split()[1:] for ln in deck.split("\n") if ln.split() and ln[0] != "#"}
    assert kv["nx"] == ["1024", "512", "128"] and kv["gridtheta"] == ["1.1693705988362009"]
    assert float(kv["bmv"][0]) == 1.12 and float(kv["bmdelt"][0]) == 2.5e-6 / 1.12
    assert abs(float(kv["ts"][0]) - (1733 - (0.5 * 1.12 / 2e-5) ** (1 / 3))) < 1e-3
    mp = amcafe_amd.semi_axes(*th)
    assert [float(q) for q in kv["meltparam"]] == [float(f"{round(q * 1e6)}e-6") for q in mp]
    tail = float(kv["meltparam"][1])
    half_diag = 0.5 * math.hypot(1024 * 2.5e-6, 512 * 2.5e-6)
    assert float(kv["gridsize"][0]) / 2 - half_diag > 30e-6 + tail
    o = Oracle(deck, bare=True)
    assert int(o.scalars()["Ntd"]) == 22
