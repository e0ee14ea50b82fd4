"""GPU parity of the whole CAFE step against the oracle.

The product (libamcafe.so through its C ABI) and the oracle run the same deck in
lockstep.  Bar: grain ids, voxel states, melt flags and Build bit-exact; the
temperature bit-exact (it is one of three constants per voxel, so the 1e-9 relative
tolerance of the FP64 oracle is met with zero error); extents and octahedron
centres bit-exact against the oracle with correctly rounded math (the device's own
math model) and within 1e-12 relative of the glibc oracle.
"""
import numpy as np
import pytest

import amcafe_amd
from oracle_util import Oracle, read

pytestmark = pytest.mark.gpu


def _same(a, b):
    """Bitwise equality; NaN equals NaN (the CPU and the GPU build different NaN payloads)."""
    if a.dtype.kind == "f":
        return np.array_equal(a, b, equal_nan=True)
    return np.array_equal(a, b)


def _diff(name, a, b):
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b)))) if a.dtype.kind == "f" else np.flatnonzero(a != b)
    return f"{name}: {bad.size} differ, first at {bad[:5].tolist()}: {a[bad[:5]].tolist()} vs {b[bad[:5]].tolist()}"


def _lockstep(deck_name, steps, exact_fp=True, glibc_check=True):
    deck = read(deck_name)
    sim = amcafe_amd.Sim(deck)
    cr = Oracle(deck, crmath=True)
    gl = Oracle(deck, crmath=False) if glibc_check else None
    n = 0
    while n < steps:
        running = sim.step()
        done_cr = cr.step()
        if gl is not None:
            gl.step()
        n += 1
        f = sim.fields(0)
        o = cr.fields()
        for k in ("gID", "vState", "BuildjID", "TempCurr"):
            assert _same(f[k], o[k]), f"{deck_name} step {n}: " + _diff(k, f[k], o[k])
        if exact_fp:
            for k in ("extents", "centroidOct"):
                assert _same(f[k], o[k]), f"{deck_name} step {n}: " + _diff(k, f[k], o[k])
        s, so = sim.stats(0), cr.stats()
        for k in ("n_active", "n_mushy0", "n_events", "n_captures", "n_nucleations", "draws"):
            want = so[k] * 2 if k == "draws" else so[k]
            assert s[k] == want, f"{deck_name} step {n}: stat {k} {s[k]} vs oracle {want}"
        if gl is not None:
            g = gl.fields()
            for k in ("gID", "vState"):
                assert np.array_equal(f[k], g[k]), f"{deck_name} step {n} (glibc oracle): " + _diff(k, f[k], g[k])
            for k in ("extents", "centroidOct"):
                np.testing.assert_allclose(f[k], g[k], rtol=1e-12, atol=1e-21)
        if done_cr:
            assert running == 0, "oracle finished but the product did not"
            break
    assert np.array_equal(sim.grains(0), cr.grains()), "grain orientations differ"
    if done_cr:
        f = sim.fields(0, which=("Build",))
        assert np.array_equal(f["Build"], cr.fields()["Build"]), "Build differs"
    return n


def test_small_deck_first_steps():
    _lockstep("t_small.in", 40)


def test_small_deck_full_run():
    n = _lockstep("t_small.in", 10**6)
    assert n > 200


def test_odd_sized_deck_full_run():
    """37x29x20 voxels: rows and the voxel count are not multiples of 16, so the grid
    passes that work in runs of 16 voxels take their voxel-by-voxel row paths and their
    partial last run."""
    n = _lockstep("t_odd.in", 10**6)
    assert n > 200


@pytest.mark.parametrize("deck,tier", [("t_small.in", 2), ("t_small.in", 3), ("t_small.in", 4),
                                       ("c0_single_track.in", 4), ("t_small.in", 5), ("c0_single_track.in", 5)])
def test_capture_tiers_full_run(monkeypatch, deck, tier):
    """Small decks run with the capture loop held in a larger tier (AMCAFE_CAPTURE_TIER_MIN):
    2 = eight cells per owner thread with the LDS head cache, 3 = sixteen cells per thread
    with head distances in LDS and the head's ruqs read from the list, 4 = every cell's state
    in memory, 5 = the same with the liveness bitmap in memory (interfaces beyond the LDS
    bitmap).  Every step against the oracle."""
    monkeypatch.setenv("AMCAFE_CAPTURE_TIER_MIN", str(tier))
    n = _lockstep(deck, 10**6)
    assert n > 100


def test_c0_single_track_full_run():
    """Workload c0 (128x64x32, one track, one layer): every step bit-identical."""
    n = _lockstep("c0_single_track.in", 10**6)
    assert n >= 128


def _deck_values(text):
    kv = {}
    for line in text.split("\n"):
        p = line.split("#")[0].split()
        if p:
            kv[p[0].lower()] = p[1:]
    return kv


def test_seam_level_api_per_step():
    """The seam entry points (amcafe_domain_*: upload, AnalyticTempCurr, UpdateVoxels,
    download, grains) against the oracle, one step at a time from the oracle's state,
    over the whole small deck including its layer changes."""
    assert seam_run("t_small.in", 400) > 100


def seam_run(deck_name, max_steps, upload_after_temperature=False):
    """Drive the seam level from the oracle's state every step; returns the events seen.
    With upload_after_temperature the state is uploaded between the temperature and the
    UpdateVoxels calls (SetLiquid3 must still see the melt pool of the device's field)."""
    deck = read(deck_name)
    kv = _deck_values(deck)
    nX = [int(v) for v in kv["nx"]]
    tL = float(kv["tl"][0])
    dom = amcafe_amd.Domain(nX, float(kv["dx"][0]), tL, float(kv["ts"][0]), 300.0,
                            [float(v) for v in kv["meltparam"]], float(kv["bmv"][0]), float(kv["bmdelt"][0]),
                            float(kv["r"][0]), float(kv["dendritegrowth2"][0]), float(kv["dendritegrowth2"][1]))
    o = Oracle(deck, crmath=True)
    events = 0
    for step in range(max_steps):
        f = o.fields()
        s = o.scalars()
        x, y, z, dirp, gth, il = o.laser_inputs()
        if upload_after_temperature:
            any_melt = dom.temperature(x, y, z, dirp, gth, il)
        dom.upload(gID=f["gID"], vState=f["vState"], extents=f["extents"], centroidOct=f["centroidOct"],
                   BuildjID=None if upload_after_temperature else f["BuildjID"])
        dom.set_grains(o.grains())
        if not upload_after_temperature:
            any_melt = dom.temperature(x, y, z, dirp, gth, il)
        o.temperature()
        st = dom.update_voxels(int(s["tInd"]), il)
        o.update_voxels()
        g, of = dom.download(), o.fields()
        assert any_melt == int(np.any(of["TempCurr"] >= tL)), f"step {step}: end-of-scan flag"
        for k in ("gID", "vState", "BuildjID", "TempCurr", "extents", "centroidOct"):
            assert _same(g[k], of[k]), f"seam step {step}: " + _diff(k, g[k], of[k])
        assert np.array_equal(dom.grains(), o.grains()), f"seam step {step}: grain table"
        so = o.stats()
        assert st["n_events"] == so["n_events"] and st["n_nucleations"] == so["n_nucleations"]
        events += st["n_events"]
        if o.finish_step():
            break
    return events


def _mid_run_handover(deck_name, warm, label, min_events):
    """The GPU's state after `warm` steps handed to both oracles (correctly rounded and glibc);
    the next two steps field for field: every field bit-exact against the correctly rounded
    oracle, grain ids and states bit-exact and octahedra within 1e-12 relative against glibc's."""
    deck = read(deck_name)
    sim = amcafe_amd.Sim(deck)
    for _ in range(warm):
        sim.step()
    f = sim.fields(0, which=("gID", "vState", "extents", "centroidOct", "BuildjID"))
    cr, gl = Oracle(deck, crmath=True, bare=True), Oracle(deck, crmath=False, bare=True)
    for o in (cr, gl):
        o.load_state(f, sim.grains(0), sim.info(0))
    for n in range(2):
        sim.step()
        for o in (cr, gl):
            o.temperature()
            o.update_voxels()
            o.finish_step()
        g, of, lg = sim.fields(0), cr.fields(), gl.fields()
        for k in ("gID", "vState", "BuildjID", "TempCurr", "extents", "centroidOct"):
            assert _same(g[k], of[k]), f"{label} step {warm + 1 + n}: " + _diff(k, g[k], of[k])
        for k in ("gID", "vState"):
            assert np.array_equal(g[k], lg[k]), f"{label} step {warm + 1 + n} (glibc oracle): " + _diff(k, g[k], lg[k])
        np.testing.assert_allclose(g["extents"], lg["extents"], rtol=1e-12, atol=1e-21)
        s, so = sim.stats(0), cr.stats()
        print(f"{label} step {warm + 1 + n}: {s['n_events']} events, {s['n_active']} interface cells")
        assert s["n_events"] == so["n_events"] == gl.stats()["n_events"] >= min_events
        assert s["n_active"] == so["n_active"]
    assert np.array_equal(sim.grains(0), cr.grains())


def test_c4_production_lockstep_from_mid_run_state():
    """The production run of the calibration-and-production form at full size (1024x512x128,
    5000-grain base plate, 67 deg layer rotation) from the GPU's state after 120 steps."""
    _mid_run_handover("c4_production.in", 120, "c4", 1)


@pytest.mark.parametrize("tier", [None, 3])
def test_c1_full_size_lockstep_from_mid_run_state(monkeypatch, tier):
    """Workload c1 at full size (512x512x128) from the GPU's state after 150 steps; with tier 3
    the capture loop runs throughout in the sixteen-cells-per-thread tier (head distances in
    LDS, ruqs from the lists) instead of the eight-cell tier c1's ~3,200 growing cells pick."""
    if tier is not None:
        monkeypatch.setenv("AMCAFE_CAPTURE_TIER_MIN", str(tier))
    _mid_run_handover("c1_raster.in", 150, "c1" if tier is None else f"c1 tier {tier}", 100)
