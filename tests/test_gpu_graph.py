"""The driver's step as HIP graphs (the default) against the same step launched stage by stage
(AMCAFE_GRAPH=0).  Every step is one replay of the step graph, with the host's own work (laser,
layers, pools, output records) and a synchronisation between replays; every step's fields and
counters must equal the stage-by-stage run's, including steps where the graph is captured
again (capture tier changes, active-list capacity growth) and steps that suspend for a pool
refill.
"""
import ctypes

import numpy as np
import pytest

import amcafe_amd
from oracle_util import read

pytestmark = pytest.mark.gpu

FIELDS = ("gID", "vState", "TempCurr", "extents", "centroidOct", "BuildjID")
COUNTERS = ("n_active", "n_events", "n_captures", "n_nucleations", "n_growing", "nGrain", "draws", "tinc",
            "interface_events")


def _launches():
    f = amcafe_amd.lib().amcd_graph_launches
    f.restype = ctypes.c_longlong
    return f()


def _lockstep(monkeypatch, names, steps, env=()):
    decks = [read(n) for n in names]
    for k, v in env:
        monkeypatch.setenv(k, v)
    monkeypatch.setenv("AMCAFE_AUTOTUNE", "0")  # a tuning step runs stage by stage
    monkeypatch.delenv("AMCAFE_GRAPH", raising=False)
    g = amcafe_amd.Sim(decks)
    monkeypatch.setenv("AMCAFE_GRAPH", "0")
    e = amcafe_amd.Sim(decks)
    l0 = _launches()
    for step in range(steps):
        rg, re_ = g.step(), e.step()
        assert rg == re_, step
        for i in range(len(decks)):
            sg, se = g.stats(i), e.stats(i)
            for k in COUNTERS:
                assert sg[k] == se[k], (names[i], step, k)
            fg, fe = g.fields(i, FIELDS), e.fields(i, FIELDS)
            for k in FIELDS:
                assert np.array_equal(fg[k], fe[k], equal_nan=fg[k].dtype.kind == "f"), (names[i], step, k)
        if rg == 0:
            break
    launched = _launches() - l0
    for i in range(len(decks)):
        assert np.array_equal(g.grains(i), e.grains(i)), names[i]
    g.close()
    e.close()
    return launched


@pytest.mark.parametrize("tier", ["0", "4", "5"])
def test_graph_step_equals_stage_by_stage(monkeypatch, tier):
    """Register tiers from the smallest up, the memory-resident tier, and the memory-resident
    tier with the liveness bitmap in memory (AMCAFE_CAPTURE_TIER_MIN floors the tier)."""
    n = _lockstep(monkeypatch, ["t_small.in"], 120, env=(("AMCAFE_CAPTURE_TIER_MIN", tier),))
    assert n >= 2 * 120  # the step graph and the finishing graph, every step


def test_graph_recaptured_on_capacity_and_tier_changes(monkeypatch):
    """A small initial active capacity makes early steps overflow and grow it (new block counts,
    a new capture of both graphs); the batch's growing lists cross the smaller tiers' limits;
    one spare nucleation orientation per step makes steps suspend for refills."""
    names = ["t_small.in", "c2_batched/inst_05.in", "c0_single_track.in"]
    n = _lockstep(monkeypatch, names, 90, env=(("AMCAFE_ACTIVE_CAP_INIT", "64"), ("AMCAFE_POOL_CHUNK", "1")))
    assert n > 0
