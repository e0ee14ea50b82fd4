"""The driver's output files against the oracle: a small deck with output every 25 steps
and at every new layer (outint 25, outNL 0) writes CA3D<stem><tInd>.h5 at the reference's
output steps (S/main1.C:76-84) and once more after the final CleanLayer (S/main1.C:94-97).
Every file holds exactly what the oracle's WriteToHDF1 would write at that moment:
gID, vState and Build bit-exact, Temperature and angleAxis as float32 of the oracle's
doubles, the IPF maps the colours of the oracle's grain table."""
import os

import numpy as np
import pytest

import amcafe_amd
import h5read
from oracle_util import Oracle, read

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(h5read.hdf5() is None, reason="HDF5 shared library not available")]


def _colours(ct):
    import test_output
    return test_output.ipf(ct)


def test_output_files_match_the_oracle(tmp_path):
    deck = read("t_small.in").replace("outint 100000", "outint 25").replace("outNL 1", "outNL 0")
    assert "outint 25" in deck and "outNL 0" in deck
    prefix = str(tmp_path / "CA3Dt_small")
    sim = amcafe_amd.Sim(deck)
    sim.set_output(prefix)
    sim.run()
    orc = Oracle(deck, crmath=True)
    orc.record_outputs()
    while not orc.step():
        pass
    want = dict(orc.outputs())          # the last write of a tInd wins, as the file does
    files = sorted(f for f in os.listdir(tmp_path) if f.endswith(".h5"))
    assert files == sorted(f"CA3Dt_small{t}.h5" for t in want), (files, sorted(want))
    assert len(want) >= 4
    for t, o in want.items():
        f = h5read.File(f"{prefix}{t}.h5")
        try:
            for k in ("gID", "vState", "Build"):
                assert np.array_equal(f.dataset("/Step0/" + k)[3], o[k]), (t, k)
            assert np.array_equal(f.dataset("/Step0/Temperature")[3], o["TempCurr"].astype(np.float32)), t
            assert np.array_equal(f.dataset("/Step0/angleAxis")[3], o["cTheta"].astype(np.float32)), t
            col = _colours(o["cTheta"])
            g = o["gID"]
            for q, name in enumerate(("IPFz", "IPFx", "IPFy")):
                exp = np.zeros((g.size, 3), np.float32)
                exp[g >= 1] = col[g[g >= 1] - 1, q]
                assert np.array_equal(f.dataset("/Step0/" + name)[3].reshape(-1, 3), exp), (t, name)
        finally:
            f.close()
