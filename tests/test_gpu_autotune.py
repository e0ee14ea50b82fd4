"""The launch autotuner (the candidate-list pass's chunking, timed between steps) changes only
launch shapes: a batch run with it, which tunes as the interface grows, equals the same batch
with the default chunking (AMCAFE_AUTOTUNE=0) in every field and counter after every step."""
import numpy as np
import pytest

import amcafe_amd
from oracle_util import read

pytestmark = pytest.mark.gpu

FIELDS = ("gID", "vState", "extents", "centroidOct", "BuildjID")


def test_autotuned_batch_equals_default_chunking(monkeypatch):
    names = ["c2_batched/inst_24.in", "c0_single_track.in", "c2_batched/inst_44.in"]
    decks = [read(n) for n in names]
    monkeypatch.delenv("AMCAFE_AUTOTUNE", raising=False)
    tuned = amcafe_amd.Sim(decks)
    monkeypatch.setenv("AMCAFE_AUTOTUNE", "0")
    plain = amcafe_amd.Sim(decks)
    biggest = 0
    for step in range(150):
        ra, rb = tuned.step(), plain.step()
        assert ra == rb, step
        for i in range(len(decks)):
            sa, sb = tuned.stats(i), plain.stats(i)
            assert sa["n_events"] == sb["n_events"] and sa["n_growing"] == sb["n_growing"], (names[i], step)
            biggest = max(biggest, sa["n_growing"])
            fa, fb = tuned.fields(i, FIELDS), plain.fields(i, FIELDS)
            for k in FIELDS:
                assert np.array_equal(fa[k], fb[k], equal_nan=fa[k].dtype.kind == "f"), (names[i], step, k)
        if ra == 0:
            break
    t, p = tuned.tuning(), plain.tuning()
    print("tuning:", t, "largest growing list", biggest)
    assert p["tunings"] == 0 and p["list_blocks"] == 0
    assert biggest >= 512 and t["tunings"] >= 2
    assert t["list_blocks"] in t["candidates_ms"] and all(m > 0 for m in t["candidates_ms"].values())
    tuned.close()
    plain.close()
