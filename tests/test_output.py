"""The output file of VoxelsCA::WriteToHDF1 (S/VoxelsCA.C:1592-1856), written by the
library's host writer and read back through the HDF5 C library.  No GPU needed.

What pins what:
  * layout: group Step0 with the ten datasets, their HDF5 classes, element sizes and
    shapes, and the root attribute NumSteps, against the reference's Put calls
    (S/VoxelsCA.C:1822-1853) -- names and types only: the reference goes through ADIOS2,
    which is not on this machine, so byte-level identity of the file is unpinned;
  * values: every dataset equals the input buffers narrowed as the reference narrows them
    (int32 unchanged, float32 of the doubles);
  * IPF colours: closed forms the reference's formulas fix -- the [001] corner of the
    standard triangle is pure red (its hue wraps to 0), rotations about z leave the build
    direction there, and the y map of the identity grain takes the hue carried over from
    the x map (S/VoxelsCA.C:1763) instead of red.
"""
import ctypes
import math
import os

import numpy as np
import pytest

import h5read

pytestmark = pytest.mark.skipif(h5read.hdf5() is None, reason="HDF5 shared library not available")


def _lib():
    import amcafe_amd
    return amcafe_amd.lib()


def ipf(ct):
    ct = np.ascontiguousarray(ct, dtype=np.float64)
    n = len(ct) // 4
    out = np.zeros(9 * n, np.float32)
    assert _lib().amcafe_ipf_colours(ct.ctypes.data, n, out.ctypes.data) == 0
    return out.reshape(n, 3, 3)


def _triangle_hue_offset():
    """The constant hue offset of S/VoxelsCA.C:1599-1612 from the standard-triangle corners."""
    t02 = 2 / math.sqrt(3) / (1 + 1 / math.sqrt(3))
    t01 = 2 / math.sqrt(2) / (1 + 1 / math.sqrt(2))
    t12 = t02
    m = math.tan(0.5 * math.atan2(t12, t02))
    a = math.hypot(t12, t02 - t01)
    b = t01
    c = math.hypot(t02, t12)
    y0 = 0.5 * math.sqrt((b + c - a) * (c + a - b) * (a + b - c) / (a + b + c))
    x0 = y0 / m
    return math.degrees(math.atan((t12 - y0) / (t02 - x0)))


def test_ipf_closed_forms():
    c = ipf([0.0, 0.0, 0.0, 1.0,            # identity
             math.pi / 2, 0.0, 0.0, 1.0,    # 90 deg about z
             1.234, 0.0, 0.0, -1.0])        # arbitrary angle about -z
    red = np.array([1, 0, 0], np.float32)
    # build direction stays [001]; x goes to a <100> direction (exactly red for the
    # identity, within the rounding of cos(pi/2) for the quarter turn)
    for g in range(3):
        np.testing.assert_allclose(c[g, 0], red, rtol=0, atol=1e-15)
    assert np.array_equal(c[0, 0], red) and np.array_equal(c[0, 1], red)
    np.testing.assert_allclose(c[1, 1], red, rtol=0, atol=1e-15)
    # identity, y map: S = 1 (the projected point is the corner), hue = 240 - offset
    H = 240 - _triangle_hue_offset()
    icase = math.floor(H / 60)
    ff = H / 6.0 - icase
    assert icase == 2
    np.testing.assert_allclose(c[0, 2], [0.0, 1.0, ff], rtol=1e-6)
    assert not np.array_equal(c[0, 2], red)


def test_hdf5_file_layout_and_values(tmp_path):
    rng = np.random.default_rng(7)
    nX = np.array([5, 4, 3], np.int32)
    dX = np.array([2.5e-6, 2.5e-6, 2.5e-6])
    N = int(nX.prod())
    ng = 6
    ax = rng.normal(size=(ng, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    ct = np.concatenate([rng.uniform(0, math.pi, (ng, 1)), ax], axis=1).ravel()
    gid = rng.integers(0, ng + 1, N).astype(np.int32)
    vs = rng.integers(0, 4, N).astype(np.int32)
    build = rng.integers(0, 2, N).astype(np.int32)
    T = rng.uniform(300, 2000, N)
    path = str(tmp_path / "CA3Dtest12.h5")
    rc = _lib().amcafe_write_hdf5_fields(path.encode(), nX.ctypes.data, dX.ctypes.data, gid.ctypes.data,
                                         vs.ctypes.data, build.ctypes.data, T.ctypes.data, ct.ctypes.data, ng)
    assert rc == 0, _lib().amcafe_last_error()
    f = h5read.File(path)
    try:
        assert f.n_links("/Step0") == 10
        assert f.attribute_int("/", "NumSteps") == 1
        want = {"dims": (0, (3,), nX), "VoxelDX": (1, (3,), dX.astype(np.float32)),
                "gID": (0, (N,), gid), "vState": (0, (N,), vs), "Build": (0, (N,), build),
                "Temperature": (1, (N,), T.astype(np.float32)), "angleAxis": (1, (4 * ng,), ct.astype(np.float32))}
        for name, (cls, shape, vals) in want.items():
            c, size, shp, got = f.dataset("/Step0/" + name)
            assert (c, size, shp) == (cls, 4, shape), name
            assert np.array_equal(got, vals), name
        col = ipf(ct)
        for q, name in enumerate(("IPFz", "IPFx", "IPFy")):
            c, size, shp, got = f.dataset("/Step0/" + name)
            assert (c, size, shp) == (1, 4, (3 * N,)), name
            exp = np.zeros((N, 3), np.float32)
            has = gid >= 1
            exp[has] = col[gid[has] - 1, q]
            assert np.array_equal(got.reshape(N, 3), exp), name
    finally:
        f.close()


def test_unwritable_path_is_an_io_error(tmp_path):
    one = np.zeros(1, np.int32)
    d = np.zeros(3)
    nX = np.ones(3, np.int32)
    rc = _lib().amcafe_write_hdf5_fields(str(tmp_path / "no" / "such" / "dir.h5").encode(), nX.ctypes.data,
                                         d.ctypes.data, one.ctypes.data, one.ctypes.data, one.ctypes.data,
                                         d.ctypes.data, None, 0)
    assert rc == -5  # AMCAFE_ERR_IO
