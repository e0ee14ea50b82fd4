"""Generator of the 64 decks of the batched workload "c2_batched" (decks/c2_batched/inst_NN.in).

Workload text: 64 independent CA instances; domains drawn uniformly between 64^3 and
256x128x96 voxels, dx 2.5 um; per instance uniform draws of power 150-300 W, speed
0.5-1.5 m/s and nucleation density 1e13-1e15 m^-3; 1-2 single-track layers; FP64.

Mapping onto the reference deck grammar (S/Grid.C:195-317), per instance:
  nX ~ U{64..256}, nY ~ U{64..128}, nZ ~ U{64..96}; dx 2.5e-6
  layers L ~ U{1,2} of 30 um (lT 30e-6, 12 voxels); bpH = (nZ - 12 L) dx so that
      nlayerTot = ceil((nZ - Nzh)/12) = L (S/main1.C:59)
  one straight track along x through the domain centre: patternID 1, gridsize (nX*dx + 2*(bhatch/2
      + tail + 20 um)) 0: the track overhangs the domain by more than bhatch/2 plus the melt
      pool's tail, so its melt pool clears before the pattern ends and a second layer really
      starts one layer higher (S/Grid.C:137-151) instead of rescanning the first
  bmP ~ U[150,300] W (0.1 W), bmV ~ U[0.5,1.5] m/s (1 mm/s); bmDelT = dx/bmV (one voxel per step);
      tS = tL - (0.5 bmV / Avel)^(1/3) (max growth per step 0.5 dx at the maximum undercooling)
  nucleation density n: log-uniform on [1e13,1e15] m^-3 (uniform in the exponent), r = n*1e-18 um^-3
  base plate: 200 Voronoi grains (mu = 200 / plate volume in um^3), lrate = mu
  meltparam: T = tL isosurface of the constant-property Gaussian source (absorptivity .35,
      radius 50 um, k 20, alpha 5e-6), derive_meltparam.dims_bisect
The per-instance seeds 100..163 of the workload text have no deck keyword: the reference's
seeds are fixed in its source (S/BasePlate.C:28-29, S/VoxelsCA.C:47-49), so every instance
uses them and the instances differ by their draws above.

Run: python decks/gen_c2_batched.py   (rng numpy default_rng(20261202); rewrites the decks)
"""
import os

import numpy as np

import derive_meltparam

DX = 2.5e-6
TL = 1733.0
AVEL, NVEL = 2e-5, 3


def deck(i, nx, ny, nz, layers, P, v, n_m3):
    nzh = nz - 12 * layers
    bpH = nzh * DX
    mu = 200.0 / ((nx * 2.5) * (ny * 2.5) * (nzh * 2.5))
    dt = DX / v
    tS = TL - (0.5 * v / AVEL) ** (1.0 / NVEL)
    r = n_m3 * 1e-18
    mp = derive_meltparam.dims_bisect(P, v)
    if mp is None:
        raise RuntimeError(f"instance {i}: no melt pool at P={P} v={v}")
    front, tail, hw, depth = (round(q * 1e6) for q in mp)
    gx = round(nx * 2.5 + 2 * (30 + tail + 20), 1) * 1e-6
    return f"""# Batched workload "c2" instance {i}: {nx}x{ny}x{nz} voxels, {layers} layer(s), P {P} W, v {v} m/s,
# nucleation density {n_m3:.4g} m^-3 (written by decks/gen_c2_batched.py)
dx 2.5e-6
nX {nx} {ny} {nz}
bmV {v!r}
bmP {P!r}
bmDelT {dt!r}
r {r!r}
lT 30e-6
tL 1733
tS {tS:.3f}
bhatch 60e-6
mu {mu!r}
bpH {bpH!r}
lrate {mu!r}
meltparam {front}e-6 {tail}e-6 {hw}e-6 {depth}e-6
patternID 1
outint 100000
outNL 1
offset 0.0 0.0 0.0
gridsize {gx!r} 0
gridtheta 0
dendritegrowth2 2e-5 3
"""


def main():
    rng = np.random.default_rng(20261202)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_batched")
    os.makedirs(out, exist_ok=True)
    for i in range(64):
        nx, ny, nz = int(rng.integers(64, 257)), int(rng.integers(64, 129)), int(rng.integers(64, 97))
        layers = int(rng.integers(1, 3))
        P = round(float(rng.uniform(150, 300)), 1)
        v = round(float(rng.uniform(0.5, 1.5)), 3)
        n_m3 = float(10 ** rng.uniform(13, 15))
        with open(os.path.join(out, f"inst_{i:02d}.in"), "w") as f:
            f.write(deck(i, nx, ny, nz, layers, P, v, n_m3))
        print(i, nx, ny, nz, layers, P, v, f"{n_m3:.3g}", flush=True)


if __name__ == "__main__":
    main()
