"""Pins of the oracle's capture loop (S/VoxelsCA.C:381-607) and analytic temperature
(S/TempField.C:139-217) by closed forms and hand-placed configurations.

Capture loop, on single-seed states set up by hand in a bare oracle (one mushy seed of
grain 1 with extent 0, liquid neighbours at tS, so vhat = Avel (tL - tS)^nvel; dx = 2^-18 m
so that voxel centres are exact):
  * the first event's time is D / vhat with D = dx (face), 2^(3/4) dx (edge), 3^(3/4) dx
    (corner) for an unrotated grain and a centred octahedron;
  * among the six tied face neighbours the capture goes to the smallest
    ruqs = (ruq[seed] + ruq[n]) / 2, ruq drawn from minstd_rand0(30 tInd + 2912351) in
    active (voxel id) order before any event (the libstdc++ stream pinned in
    test_oracle_pins.py);
  * a capture along a face, edge or corner re-centres onto the capturer's centroid with
    extent dx, 2^(3/4) dx, 3^(3/4) dx;
  * the off-centre case: capturer centroid (-dx/2, -dx/2, 0) from its voxel centre and the
    +x face captured gives extent sqrt(3/sqrt(10)) (1 + sqrt 6)/2 dx and the new centroid
    (3 - sqrt 6)/2 dx along +x from the capturer's (worked out below from the decentred-
    octahedron construction of Gandin and Rappaz);
  * mirror images of a capture in all eight octants give mirror-image results, and a
    rotation of pi/2 about z leaves a face capture's result unchanged;
  * a randomly oriented grain captures the neighbour minimising f(d) |R^T d|_1 (scipy's
    rotation matrix, an independent construction of loadRotMat's Rodrigues rotation);
  * nucleation: with rX >= 1 every event nucleates, and the grain table gains, in event
    order, the reference's own GenerateSamples(1, seed0 + 32 tInd + 64 nGrain + 10 k)
    (oracle/_ref); the event sequence equals a brute-force re-run of the scan of
    S/VoxelsCA.C:386-440 written here; with 0 < rX < 1 event k nucleates exactly when draw
    Na + k of the stream is below rX; with rX = 0 none does.
Temperature, on a 41x41x16 grid with the laser placed on voxel centres: each semi-axis of
meltparam (tail b = meltparam[1], half-width c = meltparam[2], depth d = meltparam[3],
S/TempField.C:71-76) probed just inside and just outside, the front box at 0.99 and 1.01 of
bmV*bmDelT, T0 above the laser plane, the pi/2 grid rotation, dirp looking forward at a
scan start, and BuildjID never cleared by the analytic model.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle_util import REF_HARNESS, Oracle, tiny_deck
from test_oracle_pins import canonical, minstd_rand0

need_ref = pytest.mark.skipif(not os.path.exists(REF_HARNESS),
                              reason="oracle/_ref not built (the reference sources are not on this machine)")

# dx a power of two: voxel centres (i + 1/2) dx and their differences are exact, so
# symmetric neighbours of a cell tie exactly and only the rules decide between them
DX = 2.0 ** -18
TL, TS = 1733.0, 1705.856
AVEL, NVEL = 2e-5, 3.0
VHAT = AVEL * (TL - TS) ** NVEL
SEED0, SEED1 = 2132512, 2912351


def _ulps(a, b):
    return abs(a - b) / math.ulp(b)


def _ruq(tind, n):
    g = minstd_rand0(30 * tind + SEED1)
    return np.array([(canonical(g) - .5) * DX / 100. for _ in range(n)])


def _uniforms(tind, n):
    g = minstd_rand0(30 * tind + SEED1)
    return np.array([canonical(g) for _ in range(n)])


def seed_state(n=7, liquid="all", cen_off=(0.0, 0.0, 0.0), ct=(0.0, 1.0, 0.0, 0.0), tind=7, r=0.0,
               delt_events=None, crmath=False):
    """A bare oracle on an n^3 grid: the centre voxel mushy (grain 1, orientation ct, extent 0,
    centroid at its centre + cen_off*dx), the voxels at the offsets in `liquid` (or all) liquid,
    the rest empty (state 0), T = tS everywhere.  bmDelT = 0.999 * delt_events * dx / vhat,
    so that the step ends after the first event whose clock passes it."""
    N = n ** 3
    c = n // 2
    sid = c * n * n + c * n + c
    vs = np.zeros(N, np.int32)
    if liquid == "all":
        vs[:] = 1
    else:
        for dz, dy, dx in liquid:
            vs[(c + dz) * n * n + (c + dy) * n + (c + dx)] = 1
    vs[sid] = 2
    gid = np.where(vs == 2, 1, 0).astype(np.int32)
    cen = np.zeros(3 * N)
    cen[3 * sid:3 * sid + 3] = (np.array([c, c, c]) + .5 + np.array(cen_off)) * DX
    delt = 0.999 * (delt_events if delt_events is not None else 1.0) * DX / VHAT
    deck = tiny_deck((n, n, n), dx=repr(DX), r=repr(r), bmDelT=repr(delt))
    o = Oracle(deck, crmath=crmath, bare=True)
    o.set_fields(gID=gid, vState=vs, TempCurr=np.full(N, TS), extents=np.zeros(N), centroidOct=cen,
                 BuildjID=np.zeros(N, np.int32))
    o.set_grains(np.array(ct, dtype=np.float64))
    o.L.orc_set_laser(o.h, 0, n, tind)
    return o, sid


def _offset_id(n, off):
    c = n // 2
    return (c + off[0]) * n * n + (c + off[1]) * n + (c + off[2])


FACES = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]  # (dz, dy, dx)


@pytest.mark.parametrize("crmath", [False, True])
def test_first_capture_face_tie_broken_by_ruq(crmath):
    n, tind = 7, 7
    o, sid = seed_state(n, "all", tind=tind, crmath=crmath)
    o.update_voxels()
    vox, src, dt = o.events()
    assert len(vox) == 1 and src[0] == sid
    assert _ulps(dt[0], DX / VHAT) <= 4        # D = dx: f = 1, |locX|_1 = dx (centres rounded)
    ruq = _ruq(tind, n ** 3)                   # every voxel is active: active index = voxel id
    cand = sorted(_offset_id(n, f) for f in FACES)
    rs = [.5 * (ruq[sid] + ruq[v]) for v in cand]
    assert vox[0] == cand[int(np.argmin(rs))]
    f = o.fields()
    v = vox[0]
    assert _ulps(f["extents"][v], DX) <= 4                      # re-centred extent = dx
    d = f["centroidOct"][3 * v:3 * v + 3] - f["centroidOct"][3 * sid:3 * sid + 3]
    assert np.all(np.abs(d) <= 4 * math.ulp(DX))                 # onto the capturer's centroid
    assert _ulps(f["extents"][sid], DX) <= 2                    # the seed grew by vhat * Dt = dx
    st = o.stats()
    assert st["draws"] == n ** 3 + 1 and st["n_active"] == n ** 3


@pytest.mark.parametrize("kind,off,k", [("edge", (0, 1, 1), 2), ("edge", (1, 0, -1), 2), ("edge", (-1, -1, 0), 2),
                                        ("corner", (1, 1, 1), 3), ("corner", (-1, 1, -1), 3)])
def test_edge_and_corner_capture_closed_forms(kind, off, k):
    """D = f |d|_1 with f = k^(-1/4) (the shape factor pinned in test_oracle_pins.py) and
    |d|_1 = k dx, so D = k^(3/4) dx; the captured octahedron is the capturer's."""
    n = 5
    o, sid = seed_state(n, [off])
    o.update_voxels()
    vox, src, dt = o.events()
    assert vox.tolist() == [_offset_id(n, off)] and src.tolist() == [sid]
    D = k ** 0.75 * DX
    assert _ulps(dt[0], D / VHAT) <= 6
    f = o.fields()
    v = vox[0]
    assert _ulps(f["extents"][v], D) <= 8
    d = f["centroidOct"][3 * v:3 * v + 3] - f["centroidOct"][3 * sid:3 * sid + 3]
    assert np.all(np.abs(d) <= 8 * math.ulp(DX))


def test_off_centre_face_capture_closed_form():
    """Capturer centroid at (-dx/2, -dx/2, 0) from its voxel centre, +x face neighbour captured:
    d = (3/2, 1/2, 0) dx, |d|_1 = 2 dx, octant (+,+,+) with vertices 2dx e_x, 2dx e_y, 2dx e_z;
    nearest vertex v0 = 2dx e_x (squared distance dx^2/2), farthest 2dx e_z, so v1 = 2dx e_y.
    d lies on the edge v0-v1: xI = d, |v0-xI| = dx/sqrt2, |v1-xI| = 3dx/sqrt2 (capped at
    sqrt3 l = sqrt3 dx), L12 = (1/sqrt2 + sqrt3)/2 dx.  Projection on v0-v2: xJ = (7/4, 0, 1/4)
    dx, |v0-xJ| = dx/(2 sqrt2), |v2-xJ| = 7dx/(2 sqrt2) -> capped, L13 < L12.  Lmud =
    sqrt3 sqrt(2/3) L12 = (1 + sqrt6)/2 dx; extent = f(d) Lmud with f = sqrt(cos(atan(1/3))) =
    sqrt(3/sqrt10); centroid = capturer's + v0 - Lmud e_x = capturer's + (3 - sqrt6)/2 dx e_x."""
    o, sid = seed_state(5, [(0, 0, 1)], cen_off=(-0.5, -0.5, 0.0))
    o.update_voxels()
    vox, src, dt = o.events()
    v = vox[0]
    f = o.fields()
    fs = math.sqrt(3 / math.sqrt(10))
    assert abs(dt[0] - 2 * fs * DX / VHAT) <= 1e-14 * dt[0]
    assert abs(f["extents"][v] - fs * (1 + math.sqrt(6)) / 2 * DX) <= 1e-14 * DX
    d = f["centroidOct"][3 * v:3 * v + 3] - f["centroidOct"][3 * sid:3 * sid + 3]
    np.testing.assert_allclose(d, [(3 - math.sqrt(6)) / 2 * DX, 0, 0], rtol=0, atol=1e-14 * DX)


def _capture_result(cen_off, nb_off, ct=(0.0, 1.0, 0.0, 0.0)):
    o, sid = seed_state(5, [nb_off], cen_off=cen_off, ct=ct)
    o.update_voxels()
    vox, _, dt = o.events()
    f = o.fields()
    v = vox[0]
    return (f["centroidOct"][3 * v:3 * v + 3] - f["centroidOct"][3 * sid:3 * sid + 3], f["extents"][v], dt[0])


def test_capture_octants_mirror():
    """The same capture mirrored into each of the eight octants of the local frame (which
    selects the octant's three vertices, S/VoxelsCA.h:56-77) gives the mirrored centroid
    shift and the same extent."""
    base = np.array([0.3, 0.2, 0.1])
    d0, e0, t0 = _capture_result(tuple(-base), (0, 0, 1))
    assert np.linalg.norm(d0) > 1e-2 * DX      # a genuine re-centring, not the capturer's centroid
    for sx in (1, -1):
        for sy in (1, -1):
            for sz in (1, -1):
                s = np.array([sx, sy, sz])     # (x, y, z) signs
                d, e, t = _capture_result(tuple(-s * base), (0, 0, sx))
                np.testing.assert_allclose(d, s * d0, rtol=0, atol=1e-12 * DX)
                assert abs(e - e0) <= 1e-12 * e0 and abs(t - t0) <= 1e-12 * t0


def test_rotation_by_quarter_turn_about_z_leaves_face_capture():
    """The octahedron is invariant under a quarter turn about its z axis."""
    d0, e0, t0 = _capture_result((-0.5, -0.5, 0.0), (0, 0, 1))
    d1, e1, t1 = _capture_result((-0.5, -0.5, 0.0), (0, 0, 1), ct=(math.pi / 2, 0.0, 0.0, 1.0))
    np.testing.assert_allclose(d1, d0, rtol=0, atol=1e-12 * DX)
    assert abs(e1 - e0) <= 1e-12 * e0 and abs(t1 - t0) <= 1e-12 * t0


def _shape_factor(d):
    th = math.atan2(abs(d[1]), abs(d[0]))
    th = math.pi / 2 - th if th > math.pi / 4 else th
    ph = math.atan2(math.hypot(d[0], d[1]), abs(d[2]))
    ph = math.pi / 2 - ph if ph < math.pi / 4 else ph
    return math.sqrt(math.cos(th) * math.sin(ph))


def _rotation(ct):
    from scipy.spatial.transform import Rotation
    return Rotation.from_rotvec(ct[0] * np.asarray(ct[1:], dtype=np.float64)).as_matrix()


@pytest.mark.parametrize("sample", range(4))
def test_first_capture_of_a_rotated_grain(sample):
    """A grain with a sampled orientation: the first capture goes to one of the pair of
    opposite neighbours +o, -o minimising f(d) |R^T d|_1 (R local -> global; the pair ties
    exactly), the one with the smaller ruqs, at time D_min / vhat.  The orientation is picked
    so that the untransposed R would choose a different pair."""
    L = Oracle(tiny_deck(), bare=True).L
    q = np.zeros(4 * 32)
    L.orc_generate_samples(32, 1000 + sample, q.ctypes.data)
    n, tind = 5, 5
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if (a, b, c) != (0, 0, 0)]
    vec = [np.array(o[::-1], dtype=np.float64) * DX for o in offs]       # (dz, dy, dx) -> (x, y, z)

    def ranking(R):
        D = np.array([_shape_factor(v) * np.abs(R @ v).sum() for v in vec])
        order = np.argsort(D, kind="stable")
        return D, order

    for k in range(32):
        ct = q[4 * k:4 * k + 4]
        R = _rotation(ct)
        Dt, ot = ranking(R.T)
        Dn, on = ranking(R)
        pair_t = {offs[ot[0]], offs[ot[1]]}
        gap = Dt[ot[2]] - Dt[ot[1]]
        if pair_t != {offs[on[0]], offs[on[1]]} and gap > 1e-9 * Dt[ot[0]] and \
                abs(Dt[ot[1]] - Dt[ot[0]]) <= 1e-15 * Dt[ot[0]]:
            break
    else:
        pytest.fail("no suitable orientation among the samples")
    o, sid = seed_state(n, "all", ct=tuple(ct), tind=tind)
    o.update_voxels()
    vox, src, dt = o.events()
    ruq = _ruq(tind, n ** 3)
    a, b = (_offset_id(n, offs[ot[0]]), _offset_id(n, offs[ot[1]]))
    want = a if .5 * (ruq[sid] + ruq[a]) < .5 * (ruq[sid] + ruq[b]) else b
    assert vox[0] == want
    assert abs(dt[0] - Dt[ot[0]] / VHAT) <= 1e-13 * dt[0]


def _ref_nucs(n, seed, stride):
    out = subprocess.run([REF_HARNESS, "nucs", str(n), str(seed), str(stride)], capture_output=True, text=True,
                         check=True)
    return np.array([[float.fromhex(v) for v in line.split()] for line in out.stdout.split("\n") if line])


def _brute_force_nucleation_sequence(n, ct_table, ruq, cands_order):
    """The scan of S/VoxelsCA.C:386-440 restated directly: growing cells in the order they
    became mushy, their liquid neighbours in ascending id order, candidate (t, ruqs)
    replaces the best when t < best or (t == best and ruqs < best ruqs).  All extents stay
    0 (every event nucleates, Dt = 0); each nucleus is its own grain."""
    N = n ** 3
    vs = np.ones(N, np.int32)
    c = n // 2
    sid = c * n * n + c * n + c
    vs[sid] = 2
    grain = {sid: 0}
    order = [sid]
    seq = []
    cen = {}

    def centre(v):
        k, r = divmod(v, n * n)
        j, i = divmod(r, n)
        return np.array([(i + .5) * DX, (j + .5) * DX, (k + .5) * DX])

    cen[sid] = centre(sid)
    rots = [_rotation(ct_table[g]) for g in range(len(ct_table))]
    while True:
        best, bt, br = None, 1e6, 1000.0
        for cell in order:
            k, r = divmod(cell, n * n)
            j, i = divmod(r, n)
            nbs = sorted((k + a) * n * n + (j + b) * n + (i + d) for a in (-1, 0, 1) for b in (-1, 0, 1)
                         for d in (-1, 0, 1) if (a, b, d) != (0, 0, 0) and 0 <= k + a < n and 0 <= j + b < n
                         and 0 <= i + d < n)
            if not any(vs[m] == 1 for m in nbs):
                continue
            R = rots[grain[cell]]
            for m in nbs:
                if vs[m] != 1:
                    continue
                dv = centre(m) - cen[cell]
                t = _shape_factor(dv) * np.abs(R.T @ dv).sum() / VHAT
                rq = .5 * (ruq[cell] + ruq[m])
                if t < bt or (t == bt and rq < br):
                    best, bt, br = m, t, rq
        if best is None:
            break
        vs[best] = 2
        grain[best] = len(seq) + 1
        cen[best] = centre(best)
        order.append(best)
        seq.append(best)
    return seq


@need_ref
def test_every_event_nucleates_when_rx_at_least_one():
    n, tind = 5, 3
    o, sid = seed_state(n, "all", tind=tind, r=1.0 / (DX * 1e6) ** 3)   # rX = r (dx/um)^3 = 1
    ct0 = o.grains()
    o.update_voxels()
    vox, src, dt = o.events()
    st = o.stats()
    assert st["n_nucleations"] == len(vox) == n ** 3 - 1 and np.all(src == -1) and np.all(dt == 0)
    sd0 = (SEED0 + 32 * tind + 64 * 1) % 2 ** 32
    want = _ref_nucs(len(vox), sd0, 10)
    got = o.grains().reshape(-1, 4)
    assert np.array_equal(got[0], ct0) and np.array_equal(got[1:], want)
    f = o.fields()
    for k, v in enumerate(vox):
        kk, r = divmod(int(v), n * n)
        j, i = divmod(r, n)
        assert f["gID"][v] == 2 + k and f["extents"][v] == 0.0
        assert np.array_equal(f["centroidOct"][3 * v:3 * v + 3], [(i + .5) * DX, (j + .5) * DX, (kk + .5) * DX])
    seq = _brute_force_nucleation_sequence(n, got, _ruq(tind, n ** 3), None)
    assert seq == vox.tolist()


def test_nucleation_draws_follow_the_ruq_draws():
    """0 < rX < 1: event k nucleates exactly when canonical draw Na + k of the step's stream
    is below rX (Na draws for ruq come first, S/VoxelsCA.C:380, 483)."""
    n, tind = 5, 11
    rX = 0.5
    o, sid = seed_state(n, "all", tind=tind, r=rX / (DX * 1e6) ** 3, delt_events=40.0)
    o.update_voxels()
    vox, src, dt = o.events()
    st = o.stats()
    Na = st["n_active"]
    assert Na == n ** 3 and len(vox) > 20
    u = _uniforms(tind, Na + len(vox))[Na:]
    assert np.array_equal(src == -1, u < rX)
    assert 0 < st["n_nucleations"] < len(vox)


def test_no_nucleation_when_r_is_zero():
    o, sid = seed_state(5, "all", r=0.0, delt_events=40.0)
    o.update_voxels()
    vox, src, dt = o.events()
    assert len(vox) > 20 and np.all(src >= 0) and o.stats()["n_nucleations"] == 0


# ---------------------------------------------------------------------------------------
# AnalyticTempCurr
XI = TL * (1 + np.finfo(float).eps)
NT = 41           # grid 41 x 41 x 16: the single track's y, LY/2 = 20.5 dx, is voxel 20's centre
IL = 12           # laser plane z = 12 dx; the top plane below it, k = 11, has r2 = -dx/2


def _temp_oracle(meltparam, step=1.0, pattern=1, ntrack=1, gth=0.0, gcells=NT - 1):
    """Bare oracle on the 41x41x16 grid; the laser's scan points are (k*step + x0) dx along x
    with x0 = (41 - gcells)/2, so with step = 1 and gcells = 40 the laser at scan point k sits
    on voxel k's centre."""
    gx = gcells * DX
    over = dict(dx=repr(DX), bmV="1.0", bmDelT=repr(step * DX),
                meltparam=" ".join(repr(v * DX) for v in meltparam), patternID=str(pattern),
                gridsize=f"{gx!r} {(ntrack - 1) * 8 * DX!r}", bhatch=repr(8 * DX), gridtheta=repr(gth),
                lT=repr(4 * DX))
    return Oracle(tiny_deck((NT, NT, 16), **over), bare=True)


def _temperature_at(o, isp, gth=0.0, tind=0):
    o.L.orc_set_laser(o.h, isp, IL, tind)
    o.L.orc_set_grid(o.h, 0.0, gth, 1, 0, 0)
    o.temperature()
    return o.fields()


def _T(f, i, j, k):
    return f["TempCurr"][k * NT * NT + j * NT + i]


# semi-axes in dx: tail b, half-width c, depth d; probes sit at levels 0.98 / 1.02 of the
# ellipsoid (the r2 = -dx/2 offset of the top plane adds (0.5/d)^2 = 0.0025)
_B, _C, _D = 10.0, 6.0, 10.0


def _axes(scale_b=1.0, scale_c=1.0, scale_d=1.0):
    return (1.0, _B * scale_b, _C * scale_c, _D * scale_d)


@pytest.mark.parametrize("axis", ["tail", "width", "depth"])
def test_melt_pool_semi_axes_probed_inside_and_outside(axis):
    isp = 20                                   # laser on voxel (20, 20, plane 12)
    inside = {"tail": 10.0 / math.sqrt(0.9775), "width": 6.0 / math.sqrt(0.9775), "depth": 9.5 / 0.99}
    outside = {"tail": 10.0 / math.sqrt(1.0175), "width": 6.0 / math.sqrt(1.0175), "depth": 9.5 / 1.01}
    probe = {"tail": (10, 20, 11), "width": (20, 26, 11), "depth": (20, 20, 2)}[axis]  # r = -10dx / +6dx / -9.5dx
    for val, want in ((inside[axis], XI), (outside[axis], TS)):
        mp = {"tail": (1.0, val, _C, _D), "width": (1.0, _B, val, _D), "depth": (1.0, _B, _C, val)}[axis]
        f = _temperature_at(_temp_oracle(mp), isp)
        assert _T(f, *probe) == want, (axis, val)
        # the other two semi-axes do not reach this far: a swapped meltparam would melt it
        assert _T(f, 20, 20, 11) == XI
    # permutation check: with tail b long and depth d short, a point 10 dx behind the laser
    # melts while a point 9.5 dx below it does not
    f = _temperature_at(_temp_oracle((1.0, 12.0, 6.0, 5.0)), isp)
    assert _T(f, 10, 20, 11) == XI and _T(f, 20, 20, 2) == TS and _T(f, 20, 26, 11) == TS


@pytest.mark.parametrize("frac,want", [(1.01, XI), (0.99, TS)])
def test_front_box_length_is_bmV_bmDelT(frac, want):
    """Ahead of the laser (dirp * r0 > 0) a voxel melts inside |r0| < bmV bmDelT and the
    width/depth ellipse; at scan point 0 the laser is on voxel 0's centre and voxel 1 is
    one dx ahead."""
    o = _temp_oracle((1.0, 0.5, _C, _D), step=frac)
    f = _temperature_at(o, 0)
    assert _T(f, 1, 20, 11) == want
    assert _T(f, 0, 20, 11) == XI     # r0 = 0 takes the rear ellipsoid: the laser's own voxel


def test_above_the_laser_plane_is_T0_and_buildjid_is_kept():
    o = _temp_oracle(_axes())
    N = NT * NT * 16
    bj = np.zeros(N, np.int32)
    bj[15 * NT * NT + 5] = 1                   # above the laser plane
    bj[3 * NT * NT + 7] = 1                    # below it, far from the melt pool
    o.set_fields(BuildjID=bj)
    f = _temperature_at(o, 20)
    T = f["TempCurr"].reshape(16, NT, NT)
    assert np.all(T[IL:] == 300.0) and np.all(T[:IL] >= TS)
    melted1 = f["TempCurr"] > TL
    f2 = _temperature_at(o, 30)
    melted2 = f2["TempCurr"] > TL
    want = (bj == 1) | melted1 | melted2       # T never drops below T0: BuildjID stays set
    assert np.array_equal(f2["BuildjID"] == 1, want)
    assert melted1.sum() > 50 and (melted1 & ~melted2).sum() > 10


def test_grid_rotation_quarter_turn():
    """gth = pi/2: voxel (x0, y0) is evaluated at xp = (y0 - h) + h, yp = -(x0 - h) + h
    (S/TempField.C:159-166) while the laser stays at its unrotated scan point, so the melt
    pool's tail, behind the laser along x, lies along -y in the grid."""
    f = _temperature_at(_temp_oracle(_axes()), 20, gth=math.pi / 2)
    # laser at (20.5 dx, 20.5 dx): in the grid the point 8 dx behind it is voxel (20, 12)
    assert _T(f, 20, 12, 11) == XI and _T(f, 12, 20, 11) == TS
    assert _T(f, 20, 9, 11) == TS           # 11 dx behind: beyond the tail


def test_dirp_looks_forward_at_a_scan_start():
    """Serpentine raster (patternID 2), two tracks 8 dx apart, scan points 1.25 dx apart from
    x = 5.5 dx: scan point Nsd = 25 starts the second track, which runs towards -x from
    x = 35.5 dx (voxel 35's centre), y = 24.5 dx.  There dirp = lcoor[ispvec[isp+1]] -
    lcoor[ispvec[isp]] < 0 (S/TempField.C:167-171): voxel 34, 1 dx ahead in -x, lies in the
    front box (length 1.25 dx) and melts; voxel 36, 1 dx behind, lies outside the tail
    ellipsoid (b = dx) and does not; nor does voxel 33, beyond the box.  A backward-looking
    dirp (zero at this point) or one of the wrong sign gives the opposite for 34 or 36."""
    o = _temp_oracle((1.0, 1.0, _C, _D), step=1.25, pattern=2, ntrack=2, gcells=30)
    nsd = int(o.scalars()["Nsd"])
    assert nsd == 25
    f = _temperature_at(o, nsd)
    assert _T(f, 35, 24, 11) == XI
    assert _T(f, 34, 24, 11) == XI
    assert _T(f, 36, 24, 11) == TS and _T(f, 33, 24, 11) == TS
    # mid-track (not a scan start) dirp looks backward and is also negative on this track
    f = _temperature_at(o, nsd + 3)           # laser at 35.5 - 3 * 1.25 = 31.75 dx
    assert _T(f, 31, 24, 11) == XI             # r0 = -0.25 dx, in the front box
