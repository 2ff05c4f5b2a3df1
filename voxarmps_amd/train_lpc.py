"""Train the learned PPE preconditioner on systems the solver dumps.

    python -m voxarmps_amd.train_lpc [--out voxarmps_amd/data/lpc_weights.npy]

Training set (every system assembled by the library itself, then dumped on
the GPU with vox_lpc_dump):
  * the 2D dam break (l0 = 0.01, water 25 x 50) with jittered initial
    positions (0.1 l0, seeds 1..50), batched side by side in one context: each
    step's PPE is one block-diagonal system holding 50 scenario systems;
    `--steps-2d` steps give 50 * steps scenario systems;
  * the 3D dam break at 1/4 linear scale (water 25 x 25 x 20, l0 = 0.005),
    jittered the same way (seeds 1..`--seeds-3d`), a system every
    `--every-3d` steps.
Jitter raises the maximum initial PND, which the reference takes as n0, so
far that most particles would start below the free-surface threshold; every
jittered run therefore takes n0 from the same case without jitter.
Then Adam over all of them; the weights (float32, packed order of
voxarmps.h) and a JSON summary are written next to each other.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import voxarmps_amd as vx  # noqa: E402

DEFAULT_WEIGHTS = os.path.join(HERE, "data", "lpc_weights.npy")
STAGES = ("PNUM", "NEIGHBOR", "SPS", "V_FRACTION", "VISCOSITY", "PREDICTION", "PNUMSTAR", "BCON",
          "MATRIX")


def load_weights(path=DEFAULT_WEIGHTS):
    w = np.load(path, allow_pickle=False)
    if w.dtype != np.float32 or w.shape != (vx.lpc.NPARAM,):
        raise ValueError(f"{path}: expected {vx.lpc.NPARAM} float32 weights")
    return w


def _step_and_dump(sim, net, steps, every):
    """Run `steps` time steps; after every `every`-th, assemble the next
    step's PPE (the stages up to MATRIX, as the step does) and dump it."""
    n = 0
    for k in range(1, steps + 1):
        sim.step(1)
        if k % every == 0:
            for st in STAGES:
                getattr(sim, st)()
            net.dump(sim)
            n += 1
    return n


def lattice_n0(case, params):
    """n0 (lists S and L) of a case's unjittered lattice."""
    s = vx.Simulation(params, case)
    s.init()
    n0 = s.n0()
    s.close()
    return n0


def build_training_set(net, steps_2d=400, seeds_2d=50, seeds_3d=4, steps_3d=100, every_3d=10,
                       log=print):
    t0 = time.time()
    c0 = vx.dam_break_2d_c0()
    n0 = lattice_n0(c0, vx.dam2d_params(c0))
    case, region, boxes, _ = vx.jittered_dam_breaks_2d(range(1, seeds_2d + 1))
    p = vx.dam2d_params(case, tol=1e-8)
    s = vx.Simulation(p, case)
    s.set_regions(region, boxes)
    s.init()
    s.set_n0(*n0)
    n2 = _step_and_dump(s, net, steps_2d, 1)
    s.close()
    log(f"2D: {n2} batched systems of {seeds_2d} jittered dam breaks ({case.tp} particles), "
        f"{time.time() - t0:.1f} s")
    n3 = 0
    c3l = vx.dam_break_3d_c1(4)
    n03 = lattice_n0(c3l, vx.dam3d_params(c3l))
    for seed in range(1, seeds_3d + 1):
        c3 = vx.jitter(c3l, 0.1 * 0.005, seed=seed)
        s = vx.Simulation(vx.dam3d_params(c3), c3)
        s.init()
        s.set_n0(*n03)
        n3 += _step_and_dump(s, net, steps_3d, every_3d)
        s.close()
    log(f"3D: {n3} systems of the 1/4-scale dam break, {time.time() - t0:.1f} s")
    return n2, n3


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--steps-2d", type=int, default=400)
    ap.add_argument("--seeds-2d", type=int, default=50)
    ap.add_argument("--seeds-3d", type=int, default=4)
    ap.add_argument("--steps-3d", type=int, default=100)
    ap.add_argument("--every-3d", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    net = vx.LearnedPreconditioner(seed=a.seed)
    n2, n3 = build_training_set(net, a.steps_2d, a.seeds_2d, a.seeds_3d, a.steps_3d, a.every_3d, log)
    nsys, rows = net.count()
    curve = []
    t0 = time.time()
    for ep in range(a.epochs):
        l0, l1 = net.train(1, a.lr, seed=a.seed + 7919 * ep)
        curve.append(l1)
        log(f"epoch {ep + 1}/{a.epochs}: mean loss {l1:.5e} ({time.time() - t0:.1f} s)")
    w = net.weights()
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.save(a.out, w)
    summary = {"systems_dumped": nsys, "rows": rows, "batched_2d_systems": n2,
               "scenario_systems_2d": n2 * a.seeds_2d, "systems_3d": n3, "epochs": a.epochs,
               "lr": a.lr, "loss_per_epoch": curve, "train_s": time.time() - t0}
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(json.dumps(summary))
    net.close()


if __name__ == "__main__":
    main()
