"""Batched runs: many decks of uneven size in one Sim share every launch (one grid per
stage over all instances, one capture-loop workgroup per instance).  Batching changes
the execution, not the arithmetic, so every instance of a batch must end bit-identical
to the same deck run alone, and a batched-workload deck run alone must match the oracle.
"""
import numpy as np
import pytest

import amcafe_amd
from oracle_util import Oracle, read

pytestmark = pytest.mark.gpu

FIELDS = ("gID", "vState", "TempCurr", "extents", "centroidOct", "BuildjID", "Build")


def _same(a, b):
    return np.array_equal(a, b, equal_nan=a.dtype.kind == "f")


def test_batch_of_uneven_decks_matches_single_runs():
    names = ["c2_batched/inst_05.in", "t_small.in", "c2_batched/inst_24.in", "c2_batched/inst_44.in",
             "c0_single_track.in"]
    decks = [read(n) for n in names]
    batch = amcafe_amd.Sim(decks)
    n_batch = batch.run()
    assert n_batch > 100
    for i, (name, deck) in enumerate(zip(names, decks)):
        one = amcafe_amd.Sim(deck)
        n_one = one.run()
        assert batch.info(i)["done"] == 1 and one.info(0)["done"] == 1
        assert batch.info(i)["tInd"] == one.info(0)["tInd"], name
        fb, fo = batch.fields(i), one.fields(0)
        for k in FIELDS:
            assert _same(fb[k], fo[k]), f"{name}: {k} differs between the batched and the single run"
        assert np.array_equal(batch.grains(i), one.grains(0)), name
        assert n_one <= n_batch
        one.close()


def test_batched_workload_instance_against_oracle():
    """The smallest deck of the batched workload (64x70x72), its first 120 steps (the track
    across the domain) bit-identical to the correctly rounded oracle."""
    deck = read("c2_batched/inst_05.in")
    sim = amcafe_amd.Sim(deck)
    cr = Oracle(deck, crmath=True)
    n = 0
    while n < 120:
        running = sim.step()
        done = cr.step()
        n += 1
        f, o = sim.fields(0), cr.fields()
        for k in ("gID", "vState", "BuildjID", "TempCurr", "extents", "centroidOct"):
            assert _same(f[k], o[k]), f"step {n}: {k}"
        if done:
            assert running == 0
            break
    assert np.array_equal(sim.grains(0), cr.grains())
    assert n == 120


def _run_batch(decks):
    sim = amcafe_amd.Sim(decks)
    refills = 0
    last = [0] * len(decks)
    while True:
        running = sim.step()
        for i in range(len(decks)):
            nuc = sim.stats(i)["n_nucleations"]
            refills += int(nuc > 1 + 2 * last[i])  # more than a one-entry margin holds
            last[i] = nuc
        if running == 0:
            break
    out = [(sim.fields(i), sim.grains(i), sim.info(i)["tInd"]) for i in range(len(decks))]
    sim.close()
    return out, refills


def test_pool_refills_leave_the_batch_unchanged(monkeypatch):
    """Nucleation pools staged for all instances in one copy: with one spare orientation per
    step (AMCAFE_POOL_CHUNK=1) steps exhaust their pools and the capture loop resumes after a
    refill; every instance ends bit-identical to the batch with the default pools."""
    names = ["t_small.in", "c2_batched/inst_05.in", "c0_single_track.in"]
    decks = [read(n) for n in names]
    ref, _ = _run_batch(decks)
    monkeypatch.setenv("AMCAFE_POOL_CHUNK", "1")
    got, refills = _run_batch(decks)
    assert refills > 0
    for name, (fa, ga, ta), (fb, gb, tb) in zip(names, got, ref):
        assert ta == tb, name
        for k in FIELDS:
            assert _same(fa[k], fb[k]), f"{name}: {k} differs with minimal pools"
        assert np.array_equal(ga, gb), name
