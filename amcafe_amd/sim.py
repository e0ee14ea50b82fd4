"""Thin Python mirror of the two levels of include/amcafe.h.

Sim    -- the reference's driver loop (src/src-mpi/main1.C:22-98) over one or more decks.
Domain -- the per-step seams TempField::AnalyticTempCurr / VoxelsCA::UpdateVoxels on a
          device-resident domain, state exchanged in the reference's layouts.
"""
import ctypes

import numpy as np

from ._lib import DomainDesc, Info, Laser, Stats, Tuning, check, lib


def read_deck(path):
    with open(path) as f:
        return f.read()


def _ptr(a):
    return None if a is None else a.ctypes.data


class Sim:
    def __init__(self, decks):
        if isinstance(decks, str):
            decks = [decks]
        L = lib()
        arr = (ctypes.c_char_p * len(decks))(*[d.encode() for d in decks])
        h = ctypes.c_void_p()
        check(L.amcafe_sim_create(arr, len(decks), ctypes.byref(h)))
        self.h, self.n = h, len(decks)

    def close(self):
        if self.h:
            lib().amcafe_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self):
        """One main-loop iteration for every unfinished deck; returns decks still running."""
        r = ctypes.c_int()
        check(lib().amcafe_sim_step(self.h, ctypes.byref(r)))
        return r.value

    def run(self, max_steps=10**9):
        n = 0
        while n < max_steps and self.step() > 0:
            n += 1
        return n

    def info(self, i=0):
        o = Info()
        check(lib().amcafe_sim_info(self.h, i, ctypes.byref(o)))
        return {k: (list(getattr(o, k)) if k == "nX" else getattr(o, k)) for k, _ in Info._fields_}

    def stats(self, i=0):
        o = Stats()
        check(lib().amcafe_sim_stats(self.h, i, ctypes.byref(o)))
        return {k: getattr(o, k) for k, _ in Stats._fields_}

    def tuning(self):
        """The launch autotuner's state: chunking in use, tunings, list size, candidates' times."""
        t = Tuning()
        check(lib().amcafe_sim_tuning(self.h, ctypes.byref(t)))
        return {"list_blocks": t.list_blocks, "tunings": t.tunings, "growing_cells": t.growing_cells,
                "candidates_ms": {int(c): round(float(m), 4) for c, m in zip(t.candidates, t.ms)}}

    def fields(self, i=0, which=("gID", "vState", "TempCurr", "extents", "centroidOct", "BuildjID", "Build")):
        nx, ny, nz = self.info(i)["nX"]
        N = nx * ny * nz
        spec = {"gID": (N, np.int32), "vState": (N, np.int32), "TempCurr": (N, np.float64),
                "extents": (N, np.float64), "centroidOct": (3 * N, np.float64), "BuildjID": (N, np.int32),
                "Build": (N, np.int32)}
        out = {k: np.zeros(spec[k][0], spec[k][1]) for k in which}
        g = lambda k: _ptr(out.get(k))
        check(lib().amcafe_sim_get_fields(self.h, i, g("gID"), g("vState"), g("TempCurr"), g("extents"),
                                          g("centroidOct"), g("BuildjID"), g("Build")))
        return out

    def grains(self, i=0):
        n = ctypes.c_int()
        check(lib().amcafe_sim_get_grains(self.h, i, None, ctypes.byref(n)))
        ct = np.zeros(4 * n.value)
        check(lib().amcafe_sim_get_grains(self.h, i, ct.ctypes.data, ctypes.byref(n)))
        return ct

    def write_hdf5(self, path, i=0):
        """WriteToHDF1 of instance i's current state to the file `path`."""
        check(lib().amcafe_sim_write_hdf5(self.h, i, path.encode()))

    def set_output(self, prefix, i=0):
        """The reference's output schedule: <prefix><tInd>.h5 at every output step (None: off)."""
        check(lib().amcafe_sim_set_output(self.h, i, None if prefix is None else prefix.encode()))

    def set_timing(self, on=True):
        check(lib().amcafe_sim_set_timing(self.h, 1 if on else 0))

    def timings(self):
        ms = np.zeros(6)
        check(lib().amcafe_sim_timings(self.h, ms.ctypes.data, 6))
        return ms


class Domain:
    """Device-resident domain for the TempField/VoxelsCA seams."""

    def __init__(self, nX, dX, tL, tS, T0, meltparam, bmV, bmDelT, rNmax, Avel, nvel):
        d = DomainDesc()
        d.nX[:] = list(nX)
        d.dX, d.tL, d.tS, d.T0 = dX, tL, tS, T0
        d.meltparam[:] = list(meltparam)
        d.bmV, d.bmDelT, d.rNmax, d.Avel, d.nvel = bmV, bmDelT, rNmax, Avel, nvel
        h = ctypes.c_void_p()
        check(lib().amcafe_domain_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.N = int(nX[0]) * int(nX[1]) * int(nX[2])

    def close(self):
        if self.h:
            lib().amcafe_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload(self, gID=None, vState=None, extents=None, centroidOct=None, BuildjID=None):
        c = lambda a, t: None if a is None else np.ascontiguousarray(a, dtype=t)
        arrs = [c(gID, np.int32), c(vState, np.int32), c(extents, np.float64), c(centroidOct, np.float64),
                c(BuildjID, np.int32)]
        check(lib().amcafe_domain_upload(self.h, *[_ptr(a) for a in arrs]))

    def download(self):
        N = self.N
        o = dict(gID=np.zeros(N, np.int32), vState=np.zeros(N, np.int32), extents=np.zeros(N),
                 centroidOct=np.zeros(3 * N), BuildjID=np.zeros(N, np.int32), TempCurr=np.zeros(N))
        check(lib().amcafe_domain_download(self.h, _ptr(o["gID"]), _ptr(o["vState"]), _ptr(o["extents"]),
                                           _ptr(o["centroidOct"]), _ptr(o["BuildjID"]), _ptr(o["TempCurr"])))
        return o

    def set_grains(self, cTheta):
        ct = np.ascontiguousarray(cTheta, dtype=np.float64)
        check(lib().amcafe_domain_set_grains(self.h, ct.ctypes.data, len(ct) // 4))

    def grains(self):
        n = ctypes.c_int()
        check(lib().amcafe_domain_get_grains(self.h, None, ctypes.byref(n)))
        ct = np.zeros(4 * n.value)
        check(lib().amcafe_domain_get_grains(self.h, ct.ctypes.data, ctypes.byref(n)))
        return ct

    def temperature(self, x, y, z, dirp, gth, ilaserLoc):
        las = Laser(x, y, z, dirp, gth, ilaserLoc)
        m = ctypes.c_int()
        check(lib().amcafe_analytic_temp_curr(self.h, ctypes.byref(las), ctypes.byref(m)))
        return m.value

    def easm_temperature(self, x, y, z, dirp, gth, ilaserLoc, bmP, layerT):
        """TempField::EASM on the device; returns (all voxels at T >= tL, source integrals evaluated)."""
        las = Laser(x, y, z, dirp, gth, ilaserLoc)
        m, n = ctypes.c_int(), ctypes.c_longlong()
        check(lib().amcafe_easm_temp_curr(self.h, ctypes.byref(las), bmP, layerT, ctypes.byref(m), ctypes.byref(n)))
        return m.value, n.value

    def update_voxels(self, tInd, ilaserLoc):
        s = Stats()
        check(lib().amcafe_update_voxels(self.h, tInd, ilaserLoc, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}
