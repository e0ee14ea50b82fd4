"""Derivation of the decks' `meltparam` (front, tail, half-width, depth of the melt pool).

The workload texts give beam power, speed, absorptivity 0.35, Gaussian radius 50 um,
k = 20 W/mK, alpha = 5e-6 m^2/s and T0 = 300 K; the reference's analytic temperature
model needs the melt-pool semi-axes instead (S/TempField.C:71-76, 175-190).  They are the
extent of the T >= tL (1733 K) isosurface of the constant-property Gaussian moving source,
computed by amcafe_amd.meltpool (the same routine the calibration-to-production chain uses).
Run: python decks/derive_meltparam.py
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from amcafe_amd.meltpool import semi_axes as dims_bisect, temperature as T  # noqa: E402,F401

if __name__ == "__main__":
    for name, P, v in [("c0", 200, 0.8), ("c1", 250, 1.0)]:
        print(name, ["%.1f um" % (1e6 * q) for q in dims_bisect(P, v)])
