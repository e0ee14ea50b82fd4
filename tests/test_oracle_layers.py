"""Pins of the oracle's layer routines: the base plate against the reference's own
BasePlate::GenerateVoxel (S/BasePlate.C:31-97, compiled from the reference sources into
oracle/_ref by oracle/Makefile), and AddLayer1 (S/VoxelsCA.C:887-982) by a brute-force
Voronoi labelling written here:
  * base plate: the Poisson grain count from the default-seeded engine, the sites, the
    labels of every plate voxel (voxel centres, first minimum) and the orientations equal
    the reference build's; InitializeVoxels' states (S/VoxelsCA.C:134-140): solid below the
    plate's top plane, mushy in it, centroids at voxel centres;
  * AddLayer1: every voxel of the new layer takes the nearest site measured from its corner
    (i dx, j dx, k dx), the sites that label some voxel are renumbered 1.. in site order after
    the grains already there, nGrain grows by (labels used) + 1, the orientations appended are
    the reference build's GenerateSamples(labels used + 1, sdloc) with sdloc from the layer
    engine's draw (minstd_rand0 seeded 2912351 for the first layer), states 3 in the layer and
    0 above it, centroids at voxel centres.
Nothing here needs a GPU.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle_util import REF_HARNESS, Oracle, deck_path, read
from test_oracle_pins import M31, minstd_rand0

need_ref = pytest.mark.skipif(not os.path.exists(REF_HARNESS),
                              reason="oracle/_ref not built (the reference sources are not on this machine)")


def _ref_plate(deck_name):
    out = subprocess.run([REF_HARNESS, "plate", deck_path(deck_name)], capture_output=True, text=True,
                         check=True, timeout=600).stdout.split("\n")
    head = out[0].split()
    ng, nzh, nl = int(head[1]), int(head[3]), int(head[5])
    labels = np.array([int(v) for v in out[1:1 + nl]], np.int32)
    ct = np.array([[float.fromhex(v) for v in line.split()] for line in out[1 + nl:1 + nl + ng]])
    return ng, nzh, labels, ct


def _ref_samples(n, seed):
    out = subprocess.run([REF_HARNESS, "samples", str(n), str(seed)], capture_output=True, text=True, check=True)
    return np.array([[float.fromhex(v) for v in line.split()] for line in out.stdout.split("\n") if line])


@need_ref
@pytest.mark.parametrize("deck_name", ["t_small.in", "t_odd.in", "c0_single_track.in"])
def test_base_plate_matches_reference_build(deck_name):
    ng, nzh, labels, ct = _ref_plate(deck_name)
    o = Oracle(read(deck_name))
    nx, ny, nz = o.nX
    nxy = nx * ny
    f = o.fields()
    assert int(o.scalars()["Nzh"]) == nzh
    assert np.array_equal(f["gID"][:nzh * nxy], labels)
    assert np.array_equal(o.grains().reshape(-1, 4)[:ng], ct)
    vs = f["vState"][:nzh * nxy]
    assert np.all(vs[:(nzh - 1) * nxy] == 3) and np.all(vs[(nzh - 1) * nxy:] == 2)
    k, r = np.divmod(np.arange(nzh * nxy), nxy)
    j, i = np.divmod(r, nx)
    cen = f["centroidOct"][:3 * nzh * nxy].reshape(-1, 3)
    dx = 2.5e-6
    assert np.array_equal(cen, np.stack([(i + .5) * dx, (j + .5) * dx, (k + .5) * dx], axis=1))
    assert np.all(f["Build"][:nzh * nxy] == 1)


def _check_layer(o, info_before_ngrain=None):
    """Brute-force check of the last AddLayer1 of oracle o against its recorded sites."""
    nx, ny, nz = o.nX
    nxy = nx * ny
    L = o.last_layer()
    iz1, il, ng2, i2, sdloc, g0 = (L[k] for k in ("iz1", "ilaserLoc", "ng2", "i2", "sdloc", "nGrain_before"))
    xs = o.last_sites()
    assert xs.shape == (ng2, 3)
    dx = 2.5e-6
    k, r = np.divmod(np.arange(iz1 * nxy, il * nxy), nxy)
    j, i = np.divmod(r, nx)
    corners = np.stack([i * dx, j * dx, k * dx], axis=1)
    d2 = ((corners[:, None, :] - xs[None, :, :]) ** 2).sum(axis=2)
    lab = np.argmin(d2, axis=1)                      # first minimum
    used = np.unique(lab)
    rank = np.full(ng2, -1)
    rank[used] = np.arange(len(used))
    f = o.fields()
    assert len(used) == i2
    assert np.array_equal(f["gID"][iz1 * nxy:il * nxy], g0 + 1 + rank[lab])
    assert np.all(f["vState"][iz1 * nxy:il * nxy] == 3)
    assert np.all(f["vState"][il * nxy:] == 0) and np.all(f["gID"][il * nxy:] == 0)
    cen = f["centroidOct"][3 * iz1 * nxy:3 * il * nxy].reshape(-1, 3)
    assert np.array_equal(cen, np.stack([(i + .5) * dx, (j + .5) * dx, (k + .5) * dx], axis=1))
    ct = o.grains().reshape(-1, 4)
    assert len(ct) == g0 + i2 + 1                     # one orientation more than labels used
    assert np.array_equal(ct[g0:], _ref_samples(i2 + 1, sdloc))
    # the sites lie in the layer's slab
    assert np.all((xs[:, 2] >= iz1 * dx) & (xs[:, 2] < il * dx * (1 + 1e-12)))
    # measured from voxel centres instead, some labels would differ
    centres = corners + dx / 2
    lab_c = np.argmin(((centres[:, None, :] - xs[None, :, :]) ** 2).sum(axis=2), axis=1)
    assert np.any(lab_c != lab)
    return L


@need_ref
@pytest.mark.parametrize("deck_name", ["t_small.in", "c0_single_track.in"])
def test_first_powder_layer_brute_force(deck_name):
    o = Oracle(read(deck_name))
    L = _check_layer(o)
    # the first layer's orientation seed: the layer engine (seed 2912351) draws once before
    # its Poisson count: sdloc = unsigned(x / (2^31 - 2) * 2^32)
    x = next(minstd_rand0(2912351))
    assert L["sdloc"] == int(float(x) / float(M31 - 1) * 2.0 ** 32)
    ng_plate = len(_ref_plate(deck_name)[3])
    assert L["nGrain_before"] == ng_plate     # the first layer's ids follow the plate's grains


@need_ref
def test_second_powder_layer_brute_force():
    """t_small: after the first layer's scan, CleanLayer then AddLayer1 at the next plane."""
    o = Oracle(read("t_small.in"))
    first = o.last_layer()
    for _ in range(2000):
        if o.step():
            break
        if o.last_layer()["ilaserLoc"] != first["ilaserLoc"]:
            break
    L = _check_layer(o)
    assert L["iz1"] == first["ilaserLoc"] and L["ilaserLoc"] > first["ilaserLoc"]
    # grain ids below the new layer are contiguous after CleanLayer: 1..nGrain_before all occur
    g = o.fields()["gID"][:L["iz1"] * o.nX[0] * o.nX[1]]
    assert set(np.unique(g[g > 0]).tolist()) == set(range(1, L["nGrain_before"] + 1))
