"""Pins of the oracle (oracle/cafe_oracle.cpp, a CPU restatement of the reference path)
and of the device arithmetic the product runs on the host side of its library.

Nothing here needs a GPU.  What pins what:
  * libstdc++ minstd_rand0 + generate_canonical<double,53>: the published algorithm
    (GCC 11 bits/random.tcc:3345-3381) restated in Python with exact integers, plus the
    known answer minstd_rand0(1) -> 16807, 282475249.
  * SampleOrientation::GenerateSamples and Grid (deck parser + UpdateLaser state
    machine): the reference's own S/SampleOrientation.C and S/Grid.C compiled from
    /root/reference by oracle/Makefile into oracle/_ref/ref_harness (skipped where the
    reference sources are absent).
  * the shape factor sqrt(cos(th) sin(ph)) of S/VoxelsCA.C:420-425: closed forms for
    face, edge and corner directions, glibc vs correctly rounded within 2 ulp, and the
    device's double-double evaluation (amcd_host_shape_factor, the same source the
    kernels compile) bit-exact against the correctly rounded oracle.
  * UpdateVoxels bookkeeping: identities that hold for every step of the reference
    (draws = Na + events, every event turns one liquid voxel mushy, states in 0..3).
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

from oracle_util import REF_HARNESS, ROOT, Oracle, deck_path, oracle_lib, read

need_ref = pytest.mark.skipif(not os.path.exists(REF_HARNESS),
                              reason="oracle/_ref not built (the reference sources are not on this machine)")
M31 = 2147483647


# ---------------------------------------------------------------------------------------
# random numbers
def minstd_rand0(seed):
    """std::minstd_rand0: x <- 16807 x mod (2^31 - 1); a seed that is 0 mod m becomes 1."""
    x = seed % M31
    if x == 0:
        x = 1
    while True:
        x = (16807 * x) % M31
        yield x


def canonical(gen):
    """generate_canonical<double, 53>: two engine calls, sum in double, divide by r^2.
    r = max - min + 1 = 2147483646; r^2 is formed in long double (exact) and stored in a
    double, which rounds it to 4611686009837453312."""
    r = 2147483646.0
    s = 0.0
    s += float(next(gen) - 1) * 1.0
    s += float(next(gen) - 1) * r
    tmp = float(2147483646 * 2147483646)  # exact integer, rounded once to double
    u = s / tmp
    return u if u < 1.0 else math.nextafter(1.0, 0.0)


def test_minstd_known_answer():
    g = minstd_rand0(1)
    assert [next(g), next(g)] == [16807, 282475249]
    assert float(2147483646 * 2147483646) == 4611686009837453312.0


@pytest.mark.parametrize("seed", [1, 2912351, 2912351 + 30 * 7, 30 * 400 + 2912351, 0, M31, 123456789])
def test_oracle_uniform_stream_is_libstdcxx(seed):
    L = oracle_lib(False)
    n = 2000
    out = np.zeros(n)
    L.orc_uniform_stream(seed, n, out.ctypes.data)
    g = minstd_rand0(seed)
    want = np.array([canonical(g) for _ in range(n)])
    assert np.array_equal(out, want)


def _product_lib():
    import amcafe_amd
    L = amcafe_amd.lib()
    L.amcd_uniform_stream.argtypes = [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_void_p]
    L.amcd_uniform_jump.argtypes = [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_void_p]
    L.amcd_host_shape_factor.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    return L


@pytest.mark.parametrize("seed", [1, 2912351, 2912351 + 30 * 99])
def test_device_rng_sequential_and_jump_ahead(seed):
    """The device draws ruq[a] from jump-ahead positions 16807^(2a+1); both forms must
    give the libstdc++ stream."""
    L = _product_lib()
    n = 3000
    a, b = np.zeros(n), np.zeros(n)
    L.amcd_uniform_stream(seed, n, a.ctypes.data)
    L.amcd_uniform_jump(seed, n, b.ctypes.data)
    g = minstd_rand0(seed)
    want = np.array([canonical(g) for _ in range(n)])
    assert np.array_equal(a, want)
    assert np.array_equal(b, want)


# ---------------------------------------------------------------------------------------
# orientation sampler (S/SampleOrientation.C:124-161) against the reference build
def _ref_samples(n, seed):
    out = subprocess.run([REF_HARNESS, "samples", str(n), str(seed)], capture_output=True, text=True, check=True)
    return np.array([[float.fromhex(v) for v in line.split()] for line in out.stdout.split("\n") if line])


@need_ref
@pytest.mark.parametrize("seed", [0, 1, 531231, 2912351, 4294967295, 1234567 + 64 * 2000 + 32 * 17])
def test_orientation_sampler_matches_reference_build(seed):
    n = 64
    L = oracle_lib(False)
    got = np.zeros(4 * n)
    L.orc_generate_samples(n, seed, got.ctypes.data)
    want = _ref_samples(n, seed).ravel()
    assert np.array_equal(got, want)


def test_orientation_sampler_invariants():
    L = oracle_lib(False)
    n = 4000
    got = np.zeros(4 * n)
    L.orc_generate_samples(n, 99, got.ctypes.data)
    q = got.reshape(n, 4)
    assert np.all((q[:, 0] >= 0) & (q[:, 0] <= math.pi + 1e-12))
    np.testing.assert_allclose(np.linalg.norm(q[:, 1:], axis=1), 1.0, rtol=0, atol=1e-12)
    # uniform on SO(3): the rotation angle has density (1 - cos w)/pi, mean pi/2 + 2/pi
    assert abs(q[:, 0].mean() - (math.pi / 2 + 2 / math.pi)) < 0.03


# ---------------------------------------------------------------------------------------
# deck parser + laser / scan / layer state machine (S/Grid.C:110-164, 165-339)
def _oracle_trajectory(deck_name, max_steps):
    """(isp, ilaserLoc, indlayer, newscan, newlayer, gth) before each step, and the
    end-of-scan input any(T >= tL) of each step."""
    o = Oracle(read(deck_name))
    tL = None
    for line in read(deck_name).split("\n"):
        parts = line.split()
        if parts and parts[0].lower() == "tl":
            tL = float(parts[1])
    states, flags = [], []
    for _ in range(max_steps):
        s = o.scalars()
        states.append((int(s["isp"]), int(s["ilaserLoc"]), int(s["indlayer"]), int(s["inewscanflg"]),
                       int(s["inewlayerflg"]), s["gth"]))
        done = o.step()
        flags.append(int(np.any(o.fields()["TempCurr"] >= tL)))
        if done:
            break
    return states, flags


@need_ref
@pytest.mark.parametrize("deck_name", ["t_small.in", "c0_single_track.in"])
def test_laser_state_machine_matches_reference_build(deck_name, tmp_path):
    states, flags = _oracle_trajectory(deck_name, 400)
    ff = tmp_path / "flags.txt"
    ff.write_text("\n".join(str(f) for f in flags) + "\n")
    out = subprocess.run([REF_HARNESS, "grid", deck_path(deck_name), str(ff)], capture_output=True, text=True,
                         check=True, cwd=str(tmp_path)).stdout.split("\n")
    ref = []
    for line in out:
        if line.startswith("step "):
            p = line.split()
            ref.append((int(p[3]), int(p[5]), int(p[7]), int(p[9]), int(p[11]), float.fromhex(p[13])))
    assert len(ref) == len(states)
    for st, (a, b) in enumerate(zip(states, ref)):
        assert a == b, f"step {st}: oracle {a} vs reference build {b}"


@need_ref
def test_deck_parse_matches_reference_build(tmp_path):
    ff = tmp_path / "flags.txt"
    ff.write_text("")
    out = subprocess.run([REF_HARNESS, "grid", deck_path("c1_raster.in"), str(ff)], capture_output=True, text=True,
                         check=True, cwd=str(tmp_path)).stdout
    kv = {line.split()[0]: line.split()[1:] for line in out.split("\n") if line and not line.startswith("step")}
    assert [int(v) for v in kv["nX"]] == [512, 512, 128]
    o = Oracle(read("c1_raster.in"), bare=True)
    s = o.scalars()
    words = out.split()
    val = lambda k: words[words.index(k) + 1]
    assert int(val("NpT")) == int(s["NpT"]) and int(val("Nsd")) == int(s["Nsd"]) and int(val("Ntd")) == int(s["Ntd"])
    assert int(val("ilaserLoc")) == int(s["ilaserLoc"])
    assert float.fromhex(val("gth")) == s["gth"]


# ---------------------------------------------------------------------------------------
# shape factor sqrt(cos th * sin ph) (S/VoxelsCA.C:420-425)
def _shape(L, d):
    d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(len(d))
    L.orc_shape_factor_n(d.ctypes.data, len(d), out.ctypes.data)
    return out


@pytest.mark.parametrize("crmath", [False, True])
def test_shape_factor_closed_forms(crmath):
    L = oracle_lib(crmath)
    dx = 2.5e-6
    face = _shape(L, [[dx, 0, 0], [0, dx, 0], [0, 0, dx], [-dx, 0, 0], [0, 0, -dx]])
    assert np.all(face == 1.0)
    edge_xy = _shape(L, [[dx, dx, 0]])[0]      # th = pi/4, ph = pi/2: sqrt(cos(pi/4))
    edge_xz = _shape(L, [[dx, 0, dx]])[0]      # th = 0, ph = pi/4: sqrt(sin(pi/4))
    corner = _shape(L, [[dx, dx, dx]])[0]      # cos(pi/4) sin(atan(sqrt 2)) = 1/sqrt 3
    for got, want in ((edge_xy, 2 ** -0.25), (edge_xz, 2 ** -0.25), (corner, 3 ** -0.25)):
        assert abs(got - want) <= 2 * math.ulp(want)


def test_shape_factor_glibc_vs_correctly_rounded():
    rng = np.random.default_rng(7)
    d = rng.normal(size=(20000, 3)) * 2.5e-6
    a = _shape(oracle_lib(False), d)
    b = _shape(oracle_lib(True), d)
    ulps = np.abs(a - b) / np.spacing(np.abs(b))
    assert ulps.max() <= 2


def test_device_double_double_math_is_correctly_rounded():
    """The kernels' shape factor (ddmath.h) evaluated on the host against the quadmath
    correctly rounded oracle: bit-exact, including lattice directions and near-ties of
    the pi/4 folds."""
    P = _product_lib()
    rng = np.random.default_rng(11)
    dx = 2.5e-6
    lattice = np.array([[i, j, k] for i in range(-3, 4) for j in range(-3, 4) for k in range(-3, 4)
                        if (i, j, k) != (0, 0, 0)], dtype=np.float64) * dx
    offs = rng.uniform(-0.5, 0.5, size=(6000, 3)) * dx
    near = np.repeat(lattice, 20, axis=0)[:6000] + offs * 1e-9
    d = np.concatenate([lattice, rng.normal(size=(20000, 3)) * dx, near])
    want = _shape(oracle_lib(True), d)
    got = np.zeros(len(d))
    P.amcd_host_shape_factor(np.ascontiguousarray(d).ctypes.data, len(d), got.ctypes.data)
    assert np.array_equal(got, want)


# ---------------------------------------------------------------------------------------
# UpdateVoxels bookkeeping on a small deck, both math modes
@pytest.mark.parametrize("crmath", [False, True])
def test_update_voxels_identities(crmath):
    o = Oracle(read("t_small.in"), crmath=crmath)
    events = 0
    for _ in range(80):
        before = o.fields()["vState"].copy()
        done = o.step()
        f = o.fields()
        st = o.stats()
        vox, src, dt = o.events()
        assert st["n_events"] == len(vox) == st["n_captures"] + st["n_nucleations"]
        assert st["draws"] == st["n_active"] + st["n_events"]   # ruq draws, then one per event
        assert np.all((f["vState"] >= 0) & (f["vState"] <= 3))
        assert len(set(vox.tolist())) == len(vox)                # no voxel captured twice
        assert np.all(f["vState"][vox] >= 2)
        assert np.all((src >= 0) | (src == -1))
        ext = f["extents"]
        assert np.all((ext >= 0) | np.isnan(ext))
        events += len(vox)
        if done:
            break
    assert events > 50


def test_glibc_and_correctly_rounded_runs_agree_on_grain_fields():
    """Measured property the parity bar relies on: glibc's misrounded libm results move
    octahedron geometry by an ulp but change no grain id or state on the small deck."""
    a = Oracle(read("t_small.in"), crmath=False)
    b = Oracle(read("t_small.in"), crmath=True)
    for _ in range(120):
        da, db = a.step(), b.step()
        fa, fb = a.fields(), b.fields()
        assert np.array_equal(fa["gID"], fb["gID"]) and np.array_equal(fa["vState"], fb["vState"])
        np.testing.assert_allclose(fa["extents"], fb["extents"], rtol=1e-12, atol=1e-21)
        if da or db:
            assert da and db
            break


def test_bare_oracle_takes_over_a_running_state():
    """The hand-over the full-size GPU test and the CPU baseline use: a bare oracle loaded
    with another run's fields, grain table and laser scalars continues it bit for bit
    (it must not run the end-of-build CleanLayer of S/main1.C:94 before the last layer)."""
    deck = read("c0_single_track.in")
    a = Oracle(deck, crmath=True)
    for _ in range(60):
        a.step()
    sc = a.scalars()
    info = {k: sc[k] for k in ("time", "gth")}
    info.update({k: int(sc[k]) for k in ("isp", "ilaserLoc", "tInd", "indlayer", "inewscanflg", "inewlayerflg")})
    b = Oracle(deck, crmath=True, bare=True)
    b.load_state(a.fields(), a.grains(), info)
    for _ in range(2):
        a.step()
        b.temperature()
        b.update_voxels()
        b.finish_step()
        fa, fb = a.fields(), b.fields()
        for k in ("gID", "vState", "BuildjID", "TempCurr", "extents", "centroidOct"):
            assert np.array_equal(fa[k], fb[k], equal_nan=fa[k].dtype.kind == "f"), k
        assert np.array_equal(a.grains(), b.grains())


# ---------------------------------------------------------------------------------------
# EASM temperature (S/TempField.C:79-137, S/Utilities.C:60-140, 228-249)
def _gl25_table():
    import re
    text = open(os.path.join(ROOT, "amcafe_amd", "csrc", "gl25.h")).read()
    vals = [float.fromhex(v) for v in re.findall(r"-?0x[0-9a-fp.+-]+", text)]
    return np.array(vals[:25]), np.array(vals[25:])


def test_gauss_legendre_rule_of_easm():
    """The 25-point rule the device compiles in (gl25.h, mpmath at 60 digits) equals the
    oracle's own (Newton on P25 in quad precision) bit for bit, and integrates every
    monomial of degree <= 49 on [-1, 1] exactly (to rounding)."""
    x, w = np.zeros(25), np.zeros(25)
    oracle_lib(False).orc_gl25(x.ctypes.data, w.ctypes.data)
    tx, tw = _gl25_table()
    assert np.array_equal(x, tx) and np.array_equal(w, tw)
    assert np.all(np.diff(x) > 0) and np.array_equal(x, -x[::-1]) and np.array_equal(w, w[::-1])
    for k in range(50):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(math.fsum(w * x ** k) - exact) < 1e-14, k


def test_easm_regions_and_far_field():
    """Closed forms of TempField::EASM on the small deck's grid: 0 K above the laser plane,
    0.99 T0 in the top layer ahead of the laser, and with the source 2 mm away the
    ambient 300 K (the Picard iteration stops after two integrals); with the source on the
    grid the melt pool exceeds tL under the beam and marks BuildjID."""
    o = Oracle(read("t_small.in"), crmath=False)
    dx, (nx, ny, nz), il = 2.5e-6, (40, 24, 20), 12
    z = il * dx
    allm, nint = o.easm(-2e-3, ny * dx / 2, z, 0.0)
    T = o.fields()["TempCurr"].reshape(nz, ny, nx)
    top = int(round((z - 20e-6) / dx))          # planes with zp > z - layerT
    assert allm == 0
    assert np.all(T[il:] == 0.0)
    assert np.all(T[top:il] == 0.99 * 300.0)
    assert np.all((T[:top] >= 300.0) & (T[:top] < 300.0 + 1e-9))
    assert nint == 2 * top * nx * ny
    allm, nint = o.easm(nx * dx / 2, ny * dx / 2, z, 0.0)
    f = o.fields()
    T = f["TempCurr"].reshape(nz, ny, nx)
    assert T[il - 1, ny // 2, nx // 2 - 1] > 1733.0
    assert np.all(f["BuildjID"][f["TempCurr"] > 1733.0] == 1)


# ---------------------------------------------------------------------------------------
# growth velocity pow(tL - T, nvel) (S/VoxelsCA.C:405)
def _pow(L, name, x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(len(x))
    f = getattr(L, name)
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    f(x.ctypes.data, y.ctypes.data, len(x), out.ctypes.data)
    return out


def test_device_pow_is_correctly_rounded():
    """The device's double-double pow (ddmath.h, host build) against quadmath's powq rounded
    to double, on undercoolings tL - T of any temperature an EASM field gives (0 K to tL),
    exponents around the decks' 3, exact cases and values near 1."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(1e-6, 1733.0, 20000), 10 ** rng.uniform(-8, 3.3, 20000),
                        1 + rng.normal(size=2000) * 1e-9, [2.0, 10.0, 1733.0, 0.5, 1.0, 27.144, 29.24]])
    y = np.concatenate([np.full(20000, 3.0), rng.uniform(0.5, 4.0, 20000), np.full(2000, 3.0),
                        [3.0, 3.0, 3.0, 3.0, 3.0, 3.0, 3.0]])
    got = _pow(_product_lib(), "amcd_host_pow_n", x, y)
    want = _pow(oracle_lib(True), "orc_pow_n", x, y)
    assert np.array_equal(got, want)
    assert got[-7] == 8.0 and got[-6] == 1000.0 and got[-3] == 1.0


@pytest.mark.parametrize("deck_name", ["t_small.in", "t_odd.in", "c0_single_track.in", "c1_raster.in",
                                       "c4_production.in"] + [f"c2_batched/inst_{i:02d}.in" for i in range(64)])
def test_analytic_velocities_agree_in_both_math_modes(deck_name):
    """The analytic model's growing cells sit at tS or T0; for every committed deck glibc's
    pow gives the correctly rounded value there, so the device's velocities equal the
    reference arithmetic's exactly (the bit-exact bar against the glibc oracle relies on it)."""
    kv = {}
    for line in read(deck_name).split("\n"):
        p = line.split("#")[0].split()
        if p:
            kv[p[0].lower()] = p[1:]
    tL, tS, nvel = float(kv["tl"][0]), float(kv["ts"][0]), float(kv["dendritegrowth2"][1])
    x = np.array([tL - tS, tL - 300.0])
    assert np.array_equal(_pow(oracle_lib(False), "orc_pow_n", x, [nvel, nvel]),
                          _pow(_product_lib(), "amcd_host_pow_n", x, [nvel, nvel]))
