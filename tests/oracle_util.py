"""ctypes access to the test-only oracle (oracle/_build/liboracle.so)."""
import ctypes
import math
import os
import shutil
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
_P, _I = ctypes.c_void_p, ctypes.c_int


def _load(path):
    L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    L.orc_create.restype = _P
    L.orc_create.argtypes = [ctypes.c_char_p]
    L.orc_create_bare.restype = _P
    L.orc_create_bare.argtypes = [ctypes.c_char_p]
    L.orc_set_grid.argtypes = [_P, ctypes.c_double, ctypes.c_double, _I, _I, _I]
    L.orc_last_error.restype = ctypes.c_char_p
    for f in ("orc_destroy", "orc_temperature", "orc_update_voxels", "orc_state_in", "orc_state_out"):
        getattr(L, f).argtypes = [_P]
    for f in ("orc_step", "orc_finish_step", "orc_done", "orc_ngrain"):
        getattr(L, f).argtypes = [_P]
        getattr(L, f).restype = _I
    L.orc_dims.argtypes = [_P, _P]
    L.orc_get_ctheta.argtypes = [_P, _P]
    L.orc_set_ctheta.argtypes = [_P, _P, _I]
    L.orc_get_fields.argtypes = [_P] + [_P] * 7
    L.orc_set_fields.argtypes = [_P] + [_P] * 6
    L.orc_get_scalars.argtypes = [_P, _P]
    L.orc_set_laser.argtypes = [_P, _I, _I, _I]
    L.orc_stats.argtypes = [_P, _P, _P]
    L.orc_events.argtypes = [_P, _P, _P, _P, ctypes.c_long]
    L.orc_events.restype = ctypes.c_long
    L.orc_laser_inputs.argtypes = [_P, _P]
    L.orc_easm.argtypes = [_P] + [ctypes.c_double] * 5 + [ctypes.POINTER(ctypes.c_long)]
    L.orc_easm.restype = _I
    L.orc_gl25.argtypes = [_P, _P]
    L.orc_record_outputs.argtypes = [_P, _I]
    L.orc_n_outputs.argtypes = [_P]
    L.orc_n_outputs.restype = _I
    L.orc_output.argtypes = [_P, _I] + [_P] * 5 + [ctypes.POINTER(_I)]
    L.orc_output.restype = _I
    L.orc_generate_samples.argtypes = [_I, ctypes.c_uint, _P]
    L.orc_uniform_stream.argtypes = [ctypes.c_ulong, _I, _P]
    L.orc_rot_mat.argtypes = [ctypes.c_double, _P, _P]
    L.orc_shape_factor_n.argtypes = [_P, _I, _P]
    L.orc_set_crmath.argtypes = [_I]
    L.orc_active_set.argtypes = [_P, _P, _P, _P, ctypes.POINTER(ctypes.c_long)]
    L.orc_active_set.restype = ctypes.c_long
    L.orc_last_sites.argtypes = [_P, _P]
    L.orc_last_sites.restype = ctypes.c_long
    L.orc_last_layer.argtypes = [_P, _P]
    for f in ("orc_set_build", "orc_clean_layer", "orc_set_liquid"):
        getattr(L, f).argtypes = [_P]
    L.orc_convert_solid.argtypes = [_P, _I]
    L.orc_set_build_flags.argtypes = [_P, _P]
    L.orc_initial_sites.argtypes = [ctypes.c_char_p, _P, _P, _P, _P]
    return L


_copies = {}


def oracle_lib(crmath):
    """A private copy of the oracle per math mode (the mode is a library global)."""
    key = bool(crmath)
    if key not in _copies:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built; run __graft_entry__.build()")
        d = tempfile.mkdtemp(prefix="amcafe_orc_")
        p = os.path.join(d, f"liboracle_{int(key)}.so")
        shutil.copy(ORACLE_SO, p)
        L = _load(p)
        L.orc_set_crmath(1 if key else 0)
        _copies[key] = L
    return _copies[key]


class Oracle:
    """The reference's driver loop restated on the CPU (glibc or correctly rounded math)."""

    def __init__(self, deck_text, crmath=False, bare=False):
        self.L = oracle_lib(crmath)
        self.h = (self.L.orc_create_bare if bare else self.L.orc_create)(deck_text.encode())
        if not self.h:
            raise RuntimeError(self.L.orc_last_error().decode())
        d = (ctypes.c_int * 3)()
        self.L.orc_dims(self.h, d)
        self.nX = list(d)
        self.N = d[0] * d[1] * d[2]

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def step(self):
        return self.L.orc_step(self.h)

    def temperature(self):
        self.L.orc_temperature(self.h)

    def update_voxels(self):
        self.L.orc_update_voxels(self.h)

    def finish_step(self):
        return self.L.orc_finish_step(self.h)

    def fields(self):
        N = self.N
        o = dict(gID=np.zeros(N, np.int32), vState=np.zeros(N, np.int32), TempCurr=np.zeros(N),
                 extents=np.zeros(N), centroidOct=np.zeros(3 * N), BuildjID=np.zeros(N, np.int32),
                 Build=np.zeros(N, np.int32))
        self.L.orc_get_fields(self.h, *[o[k].ctypes.data for k in
                                         ("gID", "vState", "TempCurr", "extents", "centroidOct", "BuildjID", "Build")])
        return o

    def set_fields(self, gID=None, vState=None, TempCurr=None, extents=None, centroidOct=None, BuildjID=None):
        c = lambda a, t: None if a is None else np.ascontiguousarray(a, dtype=t)
        arrs = [c(gID, np.int32), c(vState, np.int32), c(TempCurr, np.float64), c(extents, np.float64),
                c(centroidOct, np.float64), c(BuildjID, np.int32)]
        self.L.orc_set_fields(self.h, *[None if a is None else a.ctypes.data for a in arrs])
        self._keep = arrs

    def grains(self):
        n = self.L.orc_ngrain(self.h)
        ct = np.zeros(4 * n)
        self.L.orc_get_ctheta(self.h, ct.ctypes.data)
        return ct

    def set_grains(self, ct):
        ct = np.ascontiguousarray(ct, dtype=np.float64)
        self.L.orc_set_ctheta(self.h, ct.ctypes.data, len(ct) // 4)

    def scalars(self):
        o = np.zeros(14)
        self.L.orc_get_scalars(self.h, o.ctypes.data)
        keys = ("time", "tInd", "isp", "ilaserLoc", "indlayer", "inewscanflg", "inewlayerflg", "gth", "nlayerTot",
                "Nzh", "NpT", "Nsd", "Ntd", "TtInd")
        return dict(zip(keys, o.tolist()))

    def stats(self):
        o = np.zeros(6, np.int64)
        t = np.zeros(1)
        self.L.orc_stats(self.h, o.ctypes.data, t.ctypes.data)
        return dict(n_active=int(o[0]), n_mushy0=int(o[1]), n_events=int(o[2]), n_captures=int(o[3]),
                    n_nucleations=int(o[4]), draws=int(o[5]), tinc=float(t[0]))

    def events(self):
        """Event log of the last UpdateVoxels: captured voxel, capturing voxel (-1 for a
        nucleation) and Dt, in event order (grid indices)."""
        n = self.L.orc_events(self.h, None, None, None, 0)
        vox, src, dt = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
        self.L.orc_events(self.h, vox.ctypes.data, src.ctypes.data, dt.ctypes.data, n)
        return vox, src, dt

    def laser_inputs(self):
        """(x, y, z, dirp, gth, ilaserLoc) of the next temperature evaluation."""
        o = np.zeros(6)
        self.L.orc_laser_inputs(self.h, o.ctypes.data)
        return o[0], o[1], o[2], o[3], o[4], int(o[5])

    def easm(self, x, y, z, gth):
        """TempField::EASM on the oracle's grid (TempCurr, BuildjID) for a laser at (x, y, z)
        and grid rotation gth; returns (all_of(T >= tL), source integrals evaluated)."""
        n = ctypes.c_long()
        r = self.L.orc_easm(self.h, x, y, z, math.cos(gth), math.sin(gth), ctypes.byref(n))
        return r, n.value

    def record_outputs(self, on=True):
        self.L.orc_record_outputs(self.h, 1 if on else 0)

    def outputs(self):
        """What WriteToHDF1 writes at each output of the run so far: a list of
        (tInd, dict(gID, vState, Build, TempCurr, cTheta)), later entries win per tInd."""
        n = self.L.orc_n_outputs(self.h)
        N = self.N
        res = []
        for k in range(n):
            ng = ctypes.c_int()
            self.L.orc_output(self.h, k, None, None, None, None, None, ctypes.byref(ng))
            o = dict(gID=np.zeros(N, np.int32), vState=np.zeros(N, np.int32), Build=np.zeros(N, np.int32),
                     TempCurr=np.zeros(N), cTheta=np.zeros(4 * ng.value))
            t = self.L.orc_output(self.h, k, o["gID"].ctypes.data, o["vState"].ctypes.data, o["Build"].ctypes.data,
                                  o["TempCurr"].ctypes.data, o["cTheta"].ctypes.data, ctypes.byref(ng))
            res.append((t, o))
        return res

    def done(self):
        return self.L.orc_done(self.h)

    # ---- single routines of the step, for the rule pins ----
    def state_in(self):
        """SetLiquid3 then ConvertSolid1(1), as UpdateVoxels starts."""
        self.L.orc_state_in(self.h)

    def set_liquid(self):
        self.L.orc_set_liquid(self.h)

    def convert_solid(self, iswitch):
        self.L.orc_convert_solid(self.h, iswitch)

    def active_set(self):
        """(indA, nptr, nidx) of the current state: the active voxels in ascending id order and
        the CSR of each one's active neighbours (as active indices)."""
        nnz = ctypes.c_long()
        na = self.L.orc_active_set(self.h, None, None, None, ctypes.byref(nnz))
        a, p, x = np.zeros(na, np.int32), np.zeros(na + 1, np.int32), np.zeros(max(nnz.value, 1), np.int32)
        self.L.orc_active_set(self.h, a.ctypes.data, p.ctypes.data, x.ctypes.data, ctypes.byref(nnz))
        return a, p, x[:nnz.value]

    def last_sites(self):
        n = self.L.orc_last_sites(self.h, None)
        xs = np.zeros(3 * n)
        self.L.orc_last_sites(self.h, xs.ctypes.data)
        return xs.reshape(n, 3)

    def last_layer(self):
        o = np.zeros(6, np.int64)
        self.L.orc_last_layer(self.h, o.ctypes.data)
        return dict(zip(("iz1", "ilaserLoc", "ng2", "i2", "sdloc", "nGrain_before"), o.tolist()))

    def set_build(self):
        self.L.orc_set_build(self.h)

    def set_build_flags(self, build):
        b = np.ascontiguousarray(build, dtype=np.int32)
        self.L.orc_set_build_flags(self.h, b.ctypes.data)

    def clean_layer(self):
        self.L.orc_clean_layer(self.h)

    def load_state(self, fields, grains, info):
        """Take over a state produced elsewhere (fields in the reference layouts, the grain
        table and the laser / layer scalars of the driver)."""
        self.set_fields(gID=fields["gID"], vState=fields["vState"], TempCurr=fields.get("TempCurr"),
                        extents=fields["extents"], centroidOct=fields["centroidOct"],
                        BuildjID=fields["BuildjID"])
        self.set_grains(grains)
        self.L.orc_set_laser(self.h, info["isp"], info["ilaserLoc"], info["tInd"])
        self.L.orc_set_grid(self.h, info["time"], info["gth"], info["indlayer"], info["inewscanflg"],
                            info["inewlayerflg"])


def deck_path(name):
    return os.path.join(ROOT, "decks", name)


def read(name):
    with open(deck_path(name)) as f:
        return f.read()


def tiny_deck(nX=(6, 5, 4), **over):
    """A deck text for bare oracles whose state a test sets itself.  Keys as in the
    reference's deck grammar (S/Grid.C:195-317); `over` replaces or adds keys."""
    kv = {"dx": "2.5e-6", "nX": " ".join(str(v) for v in nX), "bmV": "0.8", "bmP": "200", "bmDelT": "3.125e-6",
          "r": "0", "lT": "5e-6", "tL": "1733", "tS": "1705.856", "bhatch": "30e-6", "mu": "1e-3",
          "bpH": "2.5e-6", "lrate": "1e-3", "meltparam": "20e-6 60e-6 15e-6 12e-6", "patternID": "1",
          "outint": "100000", "outNL": "1", "offset": "0.0 0.0 0.0",
          "gridsize": f"{nX[0] * 2.5e-6!r} 0", "gridtheta": "0", "dendritegrowth2": "2e-5 3"}
    for k, v in over.items():
        kv[k] = v if isinstance(v, str) else " ".join(repr(float(x)) for x in v) if hasattr(v, "__len__") \
            else repr(v)
    return "\n".join(f"{k} {v}" for k, v in kv.items()) + "\n"


def initial_sites(deck_text):
    """The initial state's random draws of a deck, without the Voronoi labelling: (plate sites
    (n, 3), plate orientation seed, first layer sites (m, 3), layer orientation seed, Nzh,
    first layer's iz1, its ilaserLoc)."""
    L = oracle_lib(False)
    c = np.zeros(5, np.int64)
    seeds = np.zeros(2, np.uint32)
    L.orc_initial_sites(deck_text.encode(), None, None, c.ctypes.data, seeds.ctypes.data)
    if c[0] < 0:
        raise RuntimeError(L.orc_last_error().decode())
    xp, xl = np.zeros(3 * c[0]), np.zeros(3 * c[1])
    L.orc_initial_sites(deck_text.encode(), xp.ctypes.data, xl.ctypes.data, c.ctypes.data, seeds.ctypes.data)
    return xp.reshape(-1, 3), int(seeds[0]), xl.reshape(-1, 3), int(seeds[1]), int(c[2]), int(c[3]), int(c[4])


def samples(n, seed):
    """GenerateSamples(n, seed) of the oracle's sampler (pinned against the reference build)."""
    out = np.zeros(4 * n)
    oracle_lib(False).orc_generate_samples(n, seed, out.ctypes.data)
    return out


def voronoi_labels(points, sites):
    """0-based index of the nearest site of each point, with the reference's squared distance
    ax*ax + ay*ay + az*az and its first-minimum rule: a k-d tree proposes the two nearest
    sites, which are compared with that expression (ties go to the lower index)."""
    from scipy.spatial import cKDTree
    _, idx = cKDTree(sites).query(points, k=2)
    d = [((points - sites[idx[:, q]]) ** 2) for q in range(2)]
    d = [(e[:, 0] + e[:, 1]) + e[:, 2] for e in d]
    a, b = idx[:, 0], idx[:, 1]
    pick_b = (d[1] < d[0]) | ((d[1] == d[0]) & (b < a))
    return np.where(pick_b, b, a)
