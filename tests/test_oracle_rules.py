"""Pins of the oracle's whole-grid rules by brute force on small random states.

Each expected result below is computed from the rule as the reference states it, with
numpy operations of its own (neighbourhood tests by array shifts, set membership by
np.isin), not by replaying the oracle's loops:
  * SetLiquid3              S/VoxelsCA.C:1098-1116 (the z = 0 plane turns mushy and keeps
                            its grain; below the laser plane liquid with grain 0; at or
                            above it the state stays; extents 0 wherever T >= tL)
  * ConvertSolid1(1)        S/VoxelsCA.C:775-797  solid with a liquid Moore neighbour -> mushy
  * ConvertSolid1(0)        S/VoxelsCA.C:738-760  mushy whose Moore neighbours are all >= 2
                            -> solid, extents 0 (both modes decide on the state before the
                            pass and then apply)
  * active set + CSR        S/VoxelsCA.C:262-335
  * SetBuild                S/VoxelsCA.C:76-115
  * CleanLayer              S/VoxelsCA.C:800-839
Nothing here needs a GPU.
"""
import numpy as np
import pytest

from oracle_util import Oracle, tiny_deck

TL, TS, T0 = 1733.0, 1705.856, 300.0
XI = TL * (1 + np.finfo(float).eps)


def _shifts(nz, ny, nx):
    """The 26 Moore offsets (dz, dy, dx)."""
    return [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if (a, b, c) != (0, 0, 0)]


def _neighbour_view(a, off, fill):
    """b[z, y, x] = a[z+dz, y+dy, x+dx], `fill` outside the grid."""
    dz, dy, dx = off
    nz, ny, nx = a.shape
    b = np.full_like(a, fill)
    zs, ys, xs = slice(max(0, -dz), nz - max(0, dz)), slice(max(0, -dy), ny - max(0, dy)), \
        slice(max(0, -dx), nx - max(0, dx))
    zd, yd, xd = slice(max(0, dz), nz - max(0, -dz)), slice(max(0, dy), ny - max(0, -dy)), \
        slice(max(0, dx), nx - max(0, -dx))
    b[zs, ys, xs] = a[zd, yd, xd]
    return b


def _random_state(rng, shape, pT=(0.3, 0.4, 0.3), pV=(0.25, 0.25, 0.25, 0.25)):
    nz, ny, nx = shape
    N = nx * ny * nz
    vs = rng.choice(4, size=N, p=pV).astype(np.int32)
    gid = np.where(vs > 0, rng.integers(1, 9, N), 0).astype(np.int32)
    T = rng.choice([T0, TS, XI], size=N, p=pT)
    ext = rng.uniform(0, 5e-6, N)
    cen = rng.uniform(0, 1e-5, 3 * N)
    bj = rng.integers(0, 2, N).astype(np.int32)
    return dict(gID=gid, vState=vs, TempCurr=T, extents=ext, centroidOct=cen, BuildjID=bj)


def _bare(shape, state, ilaserLoc):
    nz, ny, nx = shape
    o = Oracle(tiny_deck((nx, ny, nz)), bare=True)
    o.set_fields(**state)
    o.L.orc_set_laser(o.h, 0, ilaserLoc, 0)
    return o


SHAPES = [(4, 4, 4), (4, 5, 6), (5, 3, 7), (3, 6, 4)]


@pytest.mark.parametrize("seed", range(6))
def test_set_liquid3_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    shape = SHAPES[seed % len(SHAPES)]
    nz, ny, nx = shape
    st = _random_state(rng, shape)
    il = int(rng.integers(1, nz + 1))
    o = _bare(shape, st, il)
    o.set_liquid()
    got = o.fields()
    vs, gid, ext = st["vState"].copy(), st["gID"].copy(), st["extents"].copy()
    ids = np.arange(nx * ny * nz)
    hot = st["TempCurr"] >= TL
    plane0 = hot & (ids < nx * ny)
    below = hot & (ids >= nx * ny) & (ids < il * nx * ny)
    vs[plane0] = 2
    vs[below] = 1
    gid[below] = 0
    ext[hot] = 0.0
    assert np.array_equal(got["vState"], vs)
    assert np.array_equal(got["gID"], gid)
    assert np.array_equal(got["extents"], ext)
    # a plausible misreading (every hot voxel liquid) would differ on this state
    assert np.any(plane0) or np.any(hot & (ids >= il * nx * ny)) or seed > 3


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("iswitch", [0, 1])
def test_convert_solid1_brute_force(seed, iswitch):
    rng = np.random.default_rng(200 + seed)
    shape = SHAPES[seed % len(SHAPES)]
    # mostly mushy and solid, so that mushy cells with no liquid or empty neighbour occur
    st = _random_state(rng, shape, pV=(0.1, 0.1, 0.4, 0.4) if iswitch else (0.01, 0.02, 0.57, 0.4))
    o = _bare(shape, st, shape[0])
    o.convert_solid(iswitch)
    got = o.fields()
    v3 = st["vState"].reshape(shape)
    if iswitch == 1:
        liq_nbr = np.zeros(shape, bool)
        for off in _shifts(*shape):
            liq_nbr |= _neighbour_view(v3, off, 0) == 1
        flip = (v3 == 3) & liq_nbr
        want = np.where(flip, 2, v3).ravel()
        want_ext = st["extents"]
    else:
        all_ge2 = np.ones(shape, bool)
        for off in _shifts(*shape):
            all_ge2 &= _neighbour_view(v3, off, 9) >= 2   # outside the grid: no neighbour
        flip = (v3 == 2) & all_ge2
        want = np.where(flip, 3, v3).ravel()
        want_ext = np.where(flip.ravel(), 0.0, st["extents"])
    assert flip.any(), "the random state must exercise the rule"
    assert np.array_equal(got["vState"], want)
    assert np.array_equal(got["extents"], want_ext)
    assert np.array_equal(got["gID"], st["gID"])


def test_convert_solid1_decides_on_the_old_state():
    """A chain solid-solid-liquid along x: only the solid next to the liquid turns mushy;
    applying the rule in place would not change that, but the mushy->solid rule would
    cascade if applied in place (a mushy cell whose neighbour just became solid)."""
    shape = (1, 1, 5)
    st = dict(gID=np.array([1, 1, 1, 1, 1], np.int32), vState=np.array([2, 2, 2, 2, 1], np.int32),
              TempCurr=np.full(5, TS), extents=np.ones(5), centroidOct=np.zeros(15),
              BuildjID=np.zeros(5, np.int32))
    o = _bare(shape, st, 1)
    o.convert_solid(0)
    # all neighbours >= 2 holds for voxels 0, 1, 2 (voxel 3 touches the liquid voxel 4)
    assert o.fields()["vState"].tolist() == [3, 3, 3, 2, 1]
    st["vState"] = np.array([3, 3, 3, 1, 1], np.int32)
    o = _bare(shape, st, 1)
    o.convert_solid(1)
    assert o.fields()["vState"].tolist() == [3, 3, 2, 1, 1]


@pytest.mark.parametrize("seed", range(8))
def test_active_set_and_csr_brute_force(seed):
    rng = np.random.default_rng(300 + seed)
    shape = SHAPES[seed % len(SHAPES)]
    nz, ny, nx = shape
    st = _random_state(rng, shape, pT=(0.2, 0.5, 0.3))
    o = _bare(shape, st, nz)
    a, p, x = o.active_set()
    v3, T3 = st["vState"].reshape(shape), st["TempCurr"].reshape(shape)
    seedset = (v3 == 2) | ((v3 == 1) & (T3 < TL))
    near = np.zeros(shape, bool)
    for off in _shifts(*shape):
        # voxel n is added when some Moore neighbour of n is in the seed set and T[n] >= tL
        near |= _neighbour_view(seedset, off, False)
    act = seedset | (near & (T3 >= TL))
    want = np.flatnonzero(act.ravel())
    assert np.array_equal(a, want)
    pos = -np.ones(nx * ny * nz, np.int64)
    pos[want] = np.arange(len(want))
    wp, wx = [0], []
    for v in want:
        k, r = divmod(int(v), nx * ny)
        j, i = divmod(r, nx)
        nb = sorted((k + dz) * nx * ny + (j + dy) * nx + (i + dx) for dz, dy, dx in _shifts(*shape)
                    if 0 <= k + dz < nz and 0 <= j + dy < ny and 0 <= i + dx < nx)
        wx += [int(pos[n]) for n in nb if pos[n] >= 0]
        wp.append(len(wx))
    assert np.array_equal(p, wp)
    assert np.array_equal(x, wx)


@pytest.mark.parametrize("seed", range(5))
def test_set_build_brute_force(seed):
    rng = np.random.default_rng(400 + seed)
    shape = (4, 5, 6)
    st = _random_state(rng, shape)
    st["gID"] = np.where(st["vState"] > 0, rng.integers(1, 61, st["gID"].size), 0).astype(np.int32)
    st["BuildjID"] = (rng.random(st["gID"].size) < 0.1).astype(np.int32)
    build0 = (rng.random(st["gID"].size) < 0.05).astype(np.int32)
    g = st["gID"]
    counts = np.bincount(g, minlength=61)
    cold = np.flatnonzero((g > 0) & ~np.isin(g, g[st["BuildjID"] == 1]) & (counts[g] >= 2))
    build0[cold[0]] = 1  # a Build voxel in a grain of several voxels that never melted
    o = _bare(shape, st, 4)
    o.set_build_flags(build0)
    o.set_build()
    got = o.fields()["Build"]
    g = st["gID"]
    melted = np.unique(g[(st["BuildjID"] == 1) & (g > 0)])
    b = build0.copy()
    b[(g > 0) & np.isin(g, melted)] = 1
    built = np.unique(g[(b == 1) & (g > 0)])
    b[(g > 0) & np.isin(g, built)] = 1
    assert np.array_equal(got, b)
    # the second pass matters: a grain with a Build voxel but no melted voxel is marked whole
    assert np.any(b != np.where((g > 0) & np.isin(g, melted), 1, build0))


@pytest.mark.parametrize("seed", range(4))
def test_clean_layer_brute_force(seed):
    """Grains with no voxel are dropped from the orientation table; the remaining ids become
    1..n in their old order; gID 0 stays 0."""
    rng = np.random.default_rng(500 + seed)
    shape = (3, 4, 5)
    N = 60
    ngrain = 30
    st = _random_state(rng, shape)
    st["gID"] = np.where(rng.random(N) < 0.8, rng.integers(1, ngrain + 1, N), 0).astype(np.int32)
    ct = rng.normal(size=4 * ngrain)
    o = _bare(shape, st, 3)
    o.set_grains(ct)
    o.clean_layer()
    used = np.unique(st["gID"][st["gID"] > 0])
    assert len(used) < ngrain
    remap = np.zeros(ngrain + 1, np.int32)
    remap[used] = np.arange(1, len(used) + 1)
    assert np.array_equal(o.fields()["gID"], remap[st["gID"]])
    assert np.array_equal(o.grains(), ct.reshape(ngrain, 4)[used - 1].ravel())
    assert o.L.orc_ngrain(o.h) == len(used)
