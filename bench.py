#!/usr/bin/env python3
"""bench.py -- CA cell-updates/s of the CAFE time step on one MI355X.

A step is one iteration of the reference's main loop (src/src-mpi/main1.C:62-93):
temperature, UpdateVoxels, laser / layer bookkeeping.  One cell-update is one voxel
advanced one step, whatever its state, so value = nX*nY*nZ * steps / wall time of the
timed steps, with the domain resident in HBM before the timed region starts.

Default workload: c1_raster (512x512x128 voxels, 4 layers of 30 um, 3-track
bidirectional raster, 90 deg layer rotation), the largest single-GPU configuration.

The timed window starts at a fixed point of the run, --start-step (default: step 150 of
the workload, where the laser is on its first raster track inside the domain and the melt
pool has its steady size), not at the laser's entry: the run is advanced to that step
(untimed, reported as config.start_step), then --warmup untimed steps, then --steps timed
steps.  The timed steps run as a caller runs them, with no events between the stages.  The
next as many steps run with HIP events between the stages and give the per-stage breakdown
and the dominant kernel's launch time.  full_run is the whole workload run from its first
step to its last in a fresh simulation (layer changes included), with its own wall time.

Also reported, in the same JSON line:
  roofline      the dominant kernel's algorithmic bytes per launch / its mean launch
                duration (HIP events on the library's stream, over the stage-timed steps),
                against 8 TB/s HBM3E
  cpu_baseline  the oracle (a one-rank CPU port of the reference path) timed on the
                host for a few steps starting from the GPU's state before the timed region
                (one thread; the reference's N-rank MPI figure is not measured)
  batched       the 64 decks of decks/c2_batched (64^3 .. 256x128x96 voxels, one track,
                1-2 layers) run to completion as one batch and then one by one: aggregate
                cell-updates/s of each and the ratio
  adjoint       the differentiable temperature on 256x256x64 voxels over 2000 steps: FP32
                forward and forward + adjoint device times, their ratio, and the adjoint
                gradient against FP64 central differences
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CA cell-updates/s on 1 MI355X + %HBM peak; batched vs single; adjoint cost"
HBM_PEAK_GBS = 8000.0
STAGES = ("temperature+SetLiquid3", "solid->mushy", "active list", "growing setup", "capture loop",
          "scatter+mushy->solid")
KERNEL_NAMES = ("k_temperature", "k_solid_to_mushy", "k_active_flags/k_scan/k_compact", "k_active_setup",
                "k_event_loop", "k_scatter/k_mushy_to_solid")
WORKLOADS = {
    "c1_raster": "512x512x128 voxels, dx 2.5 um, 4 layers of 30 um, 3-track bidirectional raster, "
                 "hatch 60 um, 90 deg rotation, 250 W, 1.0 m/s, FP64",
    "c0_single_track": "128x64x32 voxels, dx 2.5 um, one layer, one track at 0.8 m/s, FP64",
    "c4_production": "1024x512x128 voxels, dx 2.5 um, 8 layers of 30 um over a 5000-grain base plate, "
                     "22-track bidirectional raster, hatch 60 um, 67 deg rotation, 250 W, 1.0 m/s, FP64",
}


def note(msg):
    """A progress line on stderr (the result line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stage, nvox, st):
    """Bytes one launch of a stage must move (each array element read or written once)."""
    if stage == 0:   # T write 8 + vState r/w 2 + BuildjID r/w 2 (device layout)
        return 12.0 * nvox
    if stage == 1:   # vState read + conditional write
        return 2.0 * nvox
    if stage == 4:   # capture loop, per growing-interface cell per event (the reference scans every
        # growing cell at every event, S/VoxelsCA.C:386-440): ExtA r/w 16, vhat 8, ruq 8, state and
        # grain 4, one cached candidate distance 8.  The cells stay on chip here, so this is the
        # algorithmic figure; the HBM bytes actually moved are `traffic`.
        return 44.0 * st["interface_events"]
    return None


def _digest(sim, i):
    """SHA-256 of an instance's fields and grain table (outside every timed region)."""
    import hashlib
    h = hashlib.sha256()
    f = sim.fields(i)
    for key in sorted(f):
        h.update(key.encode())
        h.update(np.ascontiguousarray(f[key]).tobytes())
    h.update(sim.grains(i).tobytes())
    return h.hexdigest()


def batched_vs_single(deck_dir):
    """The batched workload: every deck of deck_dir run to completion in one Sim (one launch per
    stage for all instances), then each deck alone in its own Sim, one after the other.  Host
    set-up (deck parse, base plate, first layer) is outside both timed regions; so is the
    comparison of every instance's final fields and grain table between the two."""
    import glob
    import amcafe_amd
    names = sorted(glob.glob(os.path.join(deck_dir, "*.in")))
    decks = [amcafe_amd.read_deck(n) for n in names]
    sim = amcafe_amd.Sim(decks)
    nv = []
    t0i = []
    for i in range(len(decks)):
        inf = sim.info(i)
        nv.append(inf["nX"][0] * inf["nX"][1] * inf["nX"][2])
        t0i.append(inf["tInd"])
    t0 = time.perf_counter()
    bsteps = sim.run()
    tb = time.perf_counter() - t0
    upd = sum(nv[i] * (sim.info(i)["tInd"] - t0i[i]) for i in range(len(decks)))
    want = [_digest(sim, i) for i in range(len(decks))]
    sim.close()
    ts, sup, same = 0.0, 0, 0
    note(f"batched: {len(decks)} instances in {tb:.2f} s; now each alone")
    for i, d in enumerate(decks):
        if i % 16 == 0:
            note(f"single runs: {i} of {len(decks)}")
        one = amcafe_amd.Sim(d)
        inf = one.info(0)
        t0 = time.perf_counter()
        one.run()
        ts += time.perf_counter() - t0
        sup += inf["nX"][0] * inf["nX"][1] * inf["nX"][2] * (one.info(0)["tInd"] - inf["tInd"])
        same += int(_digest(one, 0) == want[i])
        one.close()
    assert sup == upd
    return {"workload": os.path.basename(deck_dir.rstrip("/")), "instances": len(decks), "voxels": int(sum(nv)),
            "cell_updates": int(upd), "batched_steps": bsteps,
            "batched_value": round(upd / tb, 1), "batched_s": round(tb, 3),
            "sequential_value": round(sup / ts, 1), "sequential_s": round(ts, 3),
            "speedup": round(ts / tb, 3), "unit": "cell-updates/s",
            "instances_bit_identical_to_single_runs": same}


def adjoint_cost(fd_steps=2000):
    """The differentiation form: loss of the sigmoid melt indicator of the constant-property
    moving source on 256x256x64 voxels (dx 2.5 um) over 2000 steps of one +x track, target
    (P, v, eta, sigma) = (220 W, 0.9 m/s, 0.4, 45 um), evaluated at (200, 0.8, 0.35, 50 um):
    FP32 forward time, FP32 forward + adjoint time, and the adjoint gradient against FP64
    central differences (relative step 1e-4) over the first fd_steps steps."""
    import amcafe_amd
    from amcafe_amd.heat import HeatField, track_desc
    target, start = (220.0, 0.9, 0.4, 45e-6), (200.0, 0.8, 0.35, 50e-6)
    f = HeatField(track_desc((256, 256, 64), steps=2000, checkpoint=100), target)
    f.loss_f32(start, 0, 100)  # warm-up
    t0 = time.perf_counter()
    loss, ms_f = f.loss_f32(start, 0, 2000)
    tf = time.perf_counter() - t0
    evals = int(amcafe_amd.lib().amcafe_heat_last_evaluations())
    t0 = time.perf_counter()
    loss_g, grad, ms_g = f.loss_grad_f32(start, 0, 2000)
    tg = time.perf_counter() - t0
    again = f.loss_grad_f32(start, 0, 2000)
    reproducible = again[0] == loss_g and again[1] == grad
    fd = f.central_differences(start, 0, fd_steps, h=1e-4)
    if fd_steps != 2000:
        _, grad, _ = f.loss_grad_f32(start, 0, fd_steps)
    rel = [abs(g - d) / abs(d) for g, d in zip(grad, fd)]
    # FP32 operations per evaluated voxel-step, counted in the kernel's source: per quadrature
    # node 6 for the field and 6 for the target's (x offset, square, scale, exp, weight, sum),
    # 10 more for the two gradient sums; 25 nodes, plus the indicator and accumulation per step
    ops_f, ops_g = 25 * 12 + 12, 25 * 22 + 20
    return {"workload": "c3_heat", "voxels": 256 * 256 * 64, "steps": 2000, "dtype": "f32",
            "loss": loss, "forward_ms": round(ms_f, 2), "forward_and_adjoint_ms": round(ms_g, 2),
            "cost_ratio": round(ms_g / ms_f, 3), "wall_forward_s": round(tf, 3), "wall_adjoint_s": round(tg, 3),
            "evaluated_voxel_steps": evals, "all_voxel_steps": 256 * 256 * 64 * 2000,
            "forward_tflops": round(evals * ops_f / (ms_f * 1e-3) / 1e12, 2),
            "adjoint_tflops": round(evals * ops_g / (ms_g * 1e-3) / 1e12, 2), "fp32_peak_tflops": 157.3,
            "bit_reproducible": bool(reproducible),
            "grad": grad, "fd_grad": fd, "fd_steps": fd_steps, "max_rel_err_vs_fd": max(rel)}


def calibration_run(n=32, iters=60, steps=400):
    """The calibration form (opt-in): n instances on 128x128x48 voxels (dx 2.5 um), one +x track
    of `steps` steps, each fitted by Adam on the log-parameters with the adjoint gradient of
    all instances in the same launches, to its own target drawn uniformly (seed 7) from power
    150-300 W, speed 0.5-1.5 m/s, absorptivity 0.3-0.5, radius 40-60 um, from (200 W, 0.8 m/s,
    0.35, 50 um).  eta*P, v and sigma are identifiable (P and eta enter as their product)."""
    import amcafe_amd
    rng = np.random.default_rng(7)
    targets = [[rng.uniform(150, 300), rng.uniform(0.5, 1.5), rng.uniform(0.3, 0.5), rng.uniform(40e-6, 60e-6)]
               for _ in range(n)]
    desc = amcafe_amd.track_desc((128, 128, 48), steps=steps, v_ref=1.0, checkpoint=100)
    t0 = time.perf_counter()
    thetas, hist = amcafe_amd.calibrate(desc, targets, [200.0, 0.8, 0.35, 50e-6], iters=iters, lr=0.08,
                                        nsteps=steps)
    wall = time.perf_counter() - t0
    red = [hist[-1][b] / hist[0][b] for b in range(n)]
    err = lambda f: [abs(f(thetas[b]) / f(targets[b]) - 1) for b in range(n)]
    e_pe, e_v, e_s = err(lambda t: t[0] * t[2]), err(lambda t: t[1]), err(lambda t: t[3])
    return {"workload": "c4_calibration", "instances": n, "voxels_per_instance": 128 * 128 * 48, "steps": steps,
            "iterations": iters, "dtype": "f32", "wall_s": round(wall, 2),
            "ms_per_iteration": round(1e3 * wall / (iters + 1), 1),
            "loss_ratio_median": float(np.median(red)), "loss_ratio_max": float(max(red)),
            "rel_err_etaP_median": float(np.median(e_pe)), "rel_err_etaP_max": float(max(e_pe)),
            "rel_err_v_median": float(np.median(e_v)), "rel_err_v_max": float(max(e_v)),
            "rel_err_sigma_median": float(np.median(e_s)), "rel_err_sigma_max": float(max(e_s)),
            "within_5pct": int(sum(max(a, b, c) < 0.05 for a, b, c in zip(e_pe, e_v, e_s))),
            "restarted": int(sum(hist[-2][b] >= 1e-2 * hist[0][b] for b in range(n)))}


def full_run(deck):
    """The whole workload from its first step to its last in a fresh Sim (layer changes, host
    layer routines and output bookkeeping included), set-up excluded."""
    import amcafe_amd
    sim = amcafe_amd.Sim(deck)
    info = sim.info()
    nvox = info["nX"][0] * info["nX"][1] * info["nX"][2]
    t0 = time.perf_counter()
    steps, planes, last = 0, [], None
    while sim.step() > 0:
        steps += 1
        inf = sim.info()
        if inf["ilaserLoc"] != last:   # a new layer: (step, laser plane, grid rotation)
            planes.append((inf["tInd"], inf["ilaserLoc"], round(inf["gth"], 6)))
            last = inf["ilaserLoc"]
    steps += 1
    wall = time.perf_counter() - t0
    sim.close()
    return {"steps": steps, "layers": planes, "wall_s": round(wall, 3), "ms_per_step": round(1e3 * wall / steps, 4),
            "value": round(nvox * steps / wall, 1), "unit": "cell-updates/s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--start-step", type=int, default=150,
                    help="step the run is advanced to (untimed) before the warm-up and timed steps")
    ap.add_argument("--workload", default="c1_raster", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=16, help="oracle steps of the CPU baseline (about 14 s on c1)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-full-run", action="store_true", help="skip the whole-run measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-single comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation-form measurement")
    ap.add_argument("--calibration", action="store_true",
                    help="also run the calibration form (32 instances, 128x128x48, about a minute)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (scale for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the batched form shards "
              "instances over GPUs with no communication, so N GPUs is N independent runs.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 5}), flush=True)
        return

    import amcafe_amd
    deck = amcafe_amd.read_deck(os.path.join(ROOT, "decks", args.workload + ".in"))
    t_setup = time.perf_counter()
    sim = amcafe_amd.Sim(deck)
    t_setup = time.perf_counter() - t_setup
    info = sim.info()
    nvox = info["nX"][0] * info["nX"][1] * info["nX"][2]
    t_ff = time.perf_counter()
    for _ in range(args.start_step):
        sim.step()
    t_ff = time.perf_counter() - t_ff
    for _ in range(args.warmup):
        sim.step()
    first_timed = sim.info()["tInd"]
    snap = None
    if not args.no_cpu:
        snap = (sim.fields(0, which=("gID", "vState", "extents", "centroidOct", "BuildjID")), sim.grains(0),
                sim.info(0))
    # timed steps as a caller runs them; the stage breakdown and the capture kernel's launch
    # time come from the next as many steps, run with HIP events between the stages, with
    # their own counters
    steps = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        running = sim.step()
        steps += 1
        if running == 0:
            break
    wall = time.perf_counter() - t0
    value = nvox * steps / wall
    sim.set_timing(True)
    stage_ms = np.zeros(6)
    tot = dict(slot_visits=0, cand_evals=0, interface_events=0, n_events=0, n_active=0, n_growing=0,
               loop_cycles=0)
    tsteps = 0
    for _ in range(args.steps):
        if running == 0:
            break
        running = sim.step()
        tsteps += 1
        stage_ms += sim.timings()
        st = sim.stats()
        for k in tot:
            tot[k] += st[k]
    sim.set_timing(False)
    tsteps = max(tsteps, 1)

    dom = int(np.argmax(stage_ms))
    launch_ms = stage_ms[dom] / tsteps
    abytes = algorithmic_bytes(dom, nvox, tot)
    roof = None
    if abytes is not None:
        per_launch = abytes / tsteps
        achieved = per_launch / (launch_ms * 1e-3) / 1e9
        # HBM bytes per launch come from rocprofv3 counter runs, which cannot run inside this
        # process: a committed measurement is used only when it was taken on this workload over
        # the same steps (the kernel's dispatches of the stage-timed steps of this command)
        window = f"steps {first_timed + steps}-{first_timed + steps + tsteps - 1}"
        traffic, tsrc = None, None
        tfile = os.path.join(ROOT, "profiles", f"traffic_{KERNEL_NAMES[dom]}_{args.workload}.json")
        if os.path.exists(tfile):
            with open(tfile) as f:
                tj = json.load(f)
            if tj.get("window") == window and tj.get("workload") == args.workload:
                traffic = tj.get("hbm_bytes_per_launch")
                tsrc = os.path.relpath(tfile, ROOT)
        roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "traffic_source": tsrc,
                "kernel": KERNEL_NAMES[dom], "launch_ms": round(launch_ms, 4),
                "algorithmic_bytes_per_launch": round(per_launch, 1), "window": window + " (HIP events)"}
    tuning = sim.tuning()
    ev = tot["n_events"] / tsteps
    cap_us_per_event = (stage_ms[4] / tsteps) * 1e3 / ev if ev > 0 else None
    # the loop's own clock (shader cycles between its first decision and its exit, engine wave)
    cyc_per_event = tot["loop_cycles"] / tot["n_events"] if tot["n_events"] > 0 else None
    loop_clock_ghz = tot["loop_cycles"] / (stage_ms[4] * 1e6) if stage_ms[4] > 0 else None

    note(f"c1 window timed: {1e3 * wall / max(steps, 1):.3f} ms/step")
    cpu = None
    if snap is not None:
        note("cpu baseline sample")
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_util import Oracle
        orc = Oracle(deck, crmath=False, bare=True)
        f, g, inf = snap
        orc.load_state(f, g, inf)
        tc = time.perf_counter()
        for _ in range(args.cpu_steps):
            orc.temperature()
            orc.update_voxels()
            orc.finish_step()
        tc = time.perf_counter() - tc
        cpu = {"value": round(nvox * args.cpu_steps / tc, 1), "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "nproc": os.cpu_count(),
               "sample": f"{args.cpu_steps} steps of {args.workload} from step {first_timed} (the GPU's state at "
                         f"the start of the timed window), one thread, {tc:.1f} s; the reference's N-rank MPI "
                         f"figure is not measured"}

    if not args.no_full_run:
        note("full c1 run")
    whole = None if args.no_full_run else full_run(deck)
    batched = None
    if not args.no_batched:
        batched = batched_vs_single(os.path.join(ROOT, "decks", "c2_batched"))
    if not args.no_adjoint:
        note("adjoint cost")
    adjoint = None if args.no_adjoint else adjoint_cost()
    calibration = calibration_run() if args.calibration else None

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * wall / steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "description": WORKLOADS[args.workload], "voxels": nvox,
                   "instances": 1, "setup_s": round(t_setup, 2),
                   "start_step": args.start_step, "fast_forward_s": round(t_ff, 3),
                   "step_graph": os.environ.get("AMCAFE_GRAPH", "1") != "0",
                   "autotune": tuning,
                   "timed_steps": f"{first_timed}-{first_timed + steps - 1}",
                   "stage_ms_per_step": dict(zip(STAGES, np.round(stage_ms / tsteps, 4).tolist())),
                   "stage_steps": tsteps,
                   "events_per_step": round(ev, 1),
                   "capture_us_per_event": None if cap_us_per_event is None else round(cap_us_per_event, 3),
                   "capture_cycles_per_event": None if cyc_per_event is None else round(cyc_per_event, 1),
                   "capture_loop_clock_GHz": None if loop_clock_ghz is None else round(loop_clock_ghz, 3),
                   "active_cells_per_step": round(tot["n_active"] / tsteps, 1),
                   "growing_cells_per_step": round(tot["n_growing"] / tsteps, 1)},
        "roofline": roof, "cpu_baseline": cpu, "full_run": whole, "batched": batched, "adjoint": adjoint,
    }
    if calibration is not None:
        out["calibration"] = calibration
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
