"""Benchmark: particle-steps/s of the MPS time step on one MI355X.

Workload: 3D dam break, plain MPS, FP64 on the GPU — l0 = 0.005 m, water
100 x 100 x 80 particles in a 1.6 x 0.6 x 0.405 m open tank with one wall and
two dummy layers (1,175,450 particles), radii 2.1 l0 (PND / gradient) and
4.0 l0 (Laplacian), dt = 5e-5 s, Jacobi-PCG to relative residual 1e-8, the
neighbour lists rebuilt every step. A step is one full pass of the time loop
(PND, neighbour grid, viscosity, predictor, PND*, bcon, PPE assembly + solve,
p-hat, pressure gradient, integration, collision, prepdata) with the state
resident in HBM.

Prints one JSON line. The roofline entry is for the PPE SpMV kernel (the CG
K1 pass): algorithmic bytes per launch / mean launch time from HIP events
recorded on the library's stream during the timed steps.

`--solver learned` runs the step loop with flexible CG and the learned
preconditioner (weights from voxarmps_amd/data/lpc_weights.npy, written by
python -m voxarmps_amd.train_lpc). With the weights present, the line also
carries "ppe_solve": the last step's assembled system solved cold to the same
relative residual by Jacobi-CG and by the learned flexible CG (network
included), wall time between device synchronisations, median of 3.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the product kernel measured with rocprofv3 --pmc
# (FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy); see the file for the
# commands. Used when its workload matches the one measured here.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_cg_spmv_dam3d_1175k.json")


def measured_traffic(precision, nu, nnz, cbytes):
    """HBM bytes per k_cg_spmv launch from the committed counter summary, or
    None when no summary matches this workload (same unknown and nonzero
    counts within 1%, same column format)."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(precision)
    if not e or abs(e["unknowns"] - nu) > 0.01 * nu or abs(e["nnz"] - nnz) > 0.01 * nnz:
        return None
    if e.get("column_bytes", 4) != cbytes:
        return None
    return e["bytes_per_launch"]


def list_counts(sim, which):
    """Entries per particle of a neighbour list (counts only, no ids)."""
    import numpy as np
    cnt = np.empty(sim.tp, np.int32)
    sim.be.check(sim.be.fn("get_neighbors")(sim.ctx, which, cnt.ctypes.data_as(C.POINTER(C.c_int32)),
                                            None, 0), "get_neighbors")
    return cnt


def step_bytes(sim, case, nu, nnz_total, iters, s=8, cb=4, wc=0):
    """Compulsory HBM bytes of one time step, stage by stage: every array read
    once and written once, neighbour-gathered fields counted once, the index
    list of the radius used read once per consuming stage (4 bytes an entry
    plus its count). Lists, unknowns, nonzeros and CG iterations are those of
    the run. FP64 values (s = 8); cb bytes per PPE column the product reads
    (2: 16-bit steps and a 4-byte first column per row, or 16-bit ranks in
    LDS windows whose wc columns in all take 4 bytes each)."""
    import numpy as np
    import voxarmps_amd as vx
    N, d = case.tp, case.dim
    fl = case.ptype > 0
    F = int(np.count_nonzero(fl))
    cs = list_counts(sim, vx.LIST_S).astype(np.int64) + 1
    cl = list_counts(sim, vx.LIST_L).astype(np.int64) + 1
    bc = sim.get("BCON")
    S, Lsum = int(cs.sum()), int(cl.sum())
    SF, LF, LU = int(cs[fl].sum()), int(cl[fl].sum()), int(cl[bc == 0].sum())
    off = nnz_total - nu
    rb = 4 if cb == 2 and not wc else 0      # first column per row
    parts = {
        "neighbour lists": 2 * N * d * 8 + 4 * (S + Lsum),
        "pnd (n, n*)": 2 * (4 * S + N * (d * s + s)),
        # without a second fluid C is written directly (every fluid-2 sum is 0)
        "volume fraction": (4 * S + N * (d * s + 4) + N * s) if np.any(case.ptype == 2) else N * s,
        "viscosity": N * (4 + s),
        "predictor": 4 * LF + N * (2 * d * s + 2 * s + 4) + F * (3 * s + 2 * d * s),
        "bcon": N * (8 + s),
        "assembly": 4 * LU + N * (d * s + 4) + nu * 2 * s + off * (cb + s) + (4 + rb) * nu + 4 + 4 * wc,
        "cg": iters * (off * (cb + s) + (4 + rb) * nu + 4 + 4 * wc + 13 * s * nu),
        "phat": 4 * S + 2 * N * s,
        "gradient": 4 * SF + N * (d * s + 2 * s) + F * (2 * d * s + 2 * s),
        "euler": 4 * F * d * s,
        "collision": 4 * S + 2 * N * d * s + 2 * F * d * s,
        "prepdata": N * (2 * s + 2 * d * s),
    }
    return sum(parts.values()), parts


def spmv_bytes(nu, nnz_total, vbytes=8, cbytes=4, window_cols=0):
    """Compulsory bytes of one q = A p launch: off-diagonal values (vbytes)
    and columns (cbytes: 4 for 32-bit ids; 2 for 16-bit steps plus the
    first column of each row, 4 B, or for 16-bit ranks in the blocks' LDS
    windows plus each window's column list, 4 B per column), diagonal (8 B,
    FP64 in both modes), row length (4 B), p read once and q written
    (vbytes each) per row; gathers of p are counted once. vbytes is 8 in
    FP64 and 4 for the FP32 inner solves of the mixed mode."""
    off = nnz_total - nu
    per_row = 8 + 4 + 2 * vbytes + (4 if cbytes == 2 and not window_cols else 0)
    return off * (cbytes + vbytes) + nu * per_row + 4 * window_cols


def cpu_baseline(sample_scale=2, min_s=10.0, max_steps=20):
    """Time the CPU oracle on a bounded sample of the same workload: the 3D
    dam break at 1/sample_scale linear scale, stepped until min_s seconds have
    passed (at least 2 steps, at most max_steps). The oracle runs its
    per-particle loops and the CG's dot products on OMP_NUM_THREADS threads
    (collision stays sequential); the thread count is reported as cores."""
    import voxarmps_amd as vx
    from oracle.pyoracle import oracle_backend

    be = oracle_backend()
    threads = be.lib.orc_threads()
    case = vx.dam_break_3d_c1(scale=sample_scale)
    p = vx.dam3d_params(case)
    s = vx.Simulation(p, case, backend=be)
    s.init()
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and (steps < 2 or time.perf_counter() - t0 < min_s):
        s.step(1)
        steps += 1
    el = time.perf_counter() - t0
    s.close()
    return {
        "value": case.tp * steps / el,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} time steps of the 3D dam break at 1/{sample_scale} linear scale "
                  f"({case.tp} particles), C oracle with OpenMP, {threads} threads",
    }


def calibrate_copy(mib):
    """Five device-to-device copies of a known size through torch's copy
    kernel: each reads and writes exactly mib MiB."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.rand(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "copies": 5}))


def stream_bandwidth(mib):
    """Bandwidth plain torch streams reach over MIB MiB of float64 on this GPU:
    a read-only sum and a device copy (read + write), 10 timed repetitions
    each between HIP events."""
    import torch

    n = mib * (1 << 20) // 8
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()

    def timed(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    tc = timed(lambda: b.copy_(a))
    ts = timed(lambda: a.sum())
    print(json.dumps({"bytes": n * 8, "copy_GBps": 2 * n * 8 / tc / 1e9,
                      "sum_read_GBps": n * 8 / ts / 1e9}))


STATE_FIELDS = ("X", "Y", "Z", "U", "V", "W", "P", "PNEW")


def precision_report(checkpoints, scale=1, relaxp=None):
    """The mixed-precision run against the all-FP64 run of the same workload,
    both on the GPU from the same initial state: relative L2 and Linf
    differences of the pressure field (and of the fluid positions) at each
    checkpoint step, and the error of a single step: both runs restart from
    the FP64 run's state at the checkpoint and take one step.
    Prints one JSON line per checkpoint."""
    import numpy as np

    import voxarmps_amd as vx

    be = vx.hip_backend()
    case = vx.dam_break_3d_c1(scale=scale)
    p64 = vx.dam3d_params(case)
    if relaxp is not None:
        p64.relaxp = relaxp
    p32 = p64.copy()
    p32.precision = vx._abi.PREC_MIXED
    a = vx.Simulation(p64, case, backend=be)
    b = vx.Simulation(p32, case, backend=be)
    a.init()
    b.init()
    done = 0
    for cp in sorted(checkpoints):
        while done < cp:
            n = min(cp - done, 10)
            t0 = time.perf_counter()
            a.step(n)
            t1 = time.perf_counter()
            b.step(n)
            t2 = time.perf_counter()
            done += n
            sa, sb = a.status(), b.status()
            print(f"# step {done}: fp64 {1e3 * (t1 - t0) / n:.1f} ms/step iters {sa.last_iters} "
                  f"coll {sa.collisions}/{sa.collision_cand} | mixed {1e3 * (t2 - t1) / n:.1f} ms/step iters "
                  f"{sb.last_iters} coll {sb.collisions}/{sb.collision_cand}", file=sys.stderr, flush=True)
        pa, pb = a.get("P"), b.get("P")
        d = pb - pa
        f = case.ptype > 0
        dx = max(float(np.max(np.abs(b.get(k)[f] - a.get(k)[f]))) for k in ("X", "Y", "Z"))
        rec = {
            "step": cp, "particles": case.tp,
            "p_l2_rel": float(np.linalg.norm(d) / max(np.linalg.norm(pa), 1e-300)),
            "p_linf_rel": float(np.max(np.abs(d)) / max(np.max(np.abs(pa)), 1e-300)),
            "p_linf_abs_pa": float(np.max(np.abs(d))),
            "p_max_pa": float(np.max(np.abs(pa))),
            "p_mean_fluid_pa": float(np.mean(pa[f])),
            "front_m": float(a.get("X")[f].max()), "height_m": float(a.get("Y")[f].max()),
            "fluid_pos_linf_m": dx,
            "iters_fp64": a.status().last_iters,
            "iters_mixed": b.status().last_iters,
            "relres_mixed": b.status().last_relres,
        }
        # The trajectories part after a few hundred steps (collisions, the CG
        # tolerance), so the error of one step is measured as well: both runs
        # restart from the FP64 run's state, with the neighbour lists (which
        # PNUM reads before the step rebuilds them) built from it, and take
        # one step each.
        state = {k: a.get(k) for k in STATE_FIELDS}
        for sim in (a, b):
            for k in STATE_FIELDS:
                sim.set(k, state[k])
            sim.NEIGHBOR()
        a.step(1)
        b.step(1)
        done += 1
        for k in ("P", "U", "V", "W", "X", "Y", "Z"):
            va, vb = a.get(k), b.get(k)
            if k == "P":
                rec["step1_p_l2_rel"] = float(np.linalg.norm(vb - va) / max(np.linalg.norm(va), 1e-300))
                rec["step1_p_linf_rel"] = float(np.max(np.abs(vb - va)) / max(np.max(np.abs(va)), 1e-300))
            else:
                key = "step1_vel_linf_rel" if k in ("U", "V", "W") else "step1_pos_linf_m"
                sc = max(float(np.max(np.abs(va[f]))), 1e-300) if k in ("U", "V", "W") else 1.0
                rec[key] = max(rec.get(key, 0.0), float(np.max(np.abs(vb[f] - va[f]))) / sc)
        print(json.dumps(rec), flush=True)
    a.close()
    b.close()


PPE_SOLVERS = (("jacobi_cg", "SOLVER_JACOBI_CG", "PREC_FP64"),
               ("learned_fcg", "SOLVER_LEARNED_FCG", "PREC_FP64"),
               ("jacobi_mixed", "SOLVER_JACOBI_CG", "PREC_MIXED"),
               ("learned_mixed", "SOLVER_LEARNED_FCG", "PREC_MIXED"))


def ppe_compare(sim, p, net, reps=3):
    """The assembled system of the current step solved cold, to the same
    relative residual, by FP64 Jacobi-CG, the learned flexible CG, and both
    inside the mixed-precision refinement (FP32 inner solves): median wall ms
    between synchronisations, iterations, relative residuals, and each
    variant's speed-up over FP64 Jacobi-CG."""
    import voxarmps_amd as vx
    be = sim.be
    out = {}
    for name, solver, prec in PPE_SOLVERS:
        q = p.copy()
        q.solver = getattr(vx._abi, solver)
        q.precision = getattr(vx._abi, prec)
        sim.set_params(q)
        sim.attach_lpc(net)          # resets the baseline: the next learned solve fixes it
        sim.CGM()
        times, info = [], None
        be.lib.vox_prof_reset(sim.ctx)
        be.lib.vox_prof_enable(sim.ctx, 1)
        for _ in range(reps):
            be.lib.vox_sync(sim.ctx)
            t0 = time.perf_counter()
            info = sim.CGM()
            be.lib.vox_sync(sim.ctx)
            times.append((time.perf_counter() - t0) * 1e3)
        be.lib.vox_prof_enable(sim.ctx, 0)
        prof = {}
        for k in ("lpc_net", "lpc_apply", "spmv"):
            t = C.c_double(0.0)
            n = C.c_int64(0)
            be.lib.vox_prof_get(sim.ctx, k.encode(), C.byref(t), C.byref(n))
            prof[k + "_ms"] = t.value / reps
        times.sort()
        out[name] = {"ms": times[len(times) // 2], "iterations": info.iterations,
                     "unknowns": info.unknowns,
                     "relres": info.relres, "converged": bool(info.converged),
                     "fallbacks": sim.status().lpc_fallbacks, **prof}
    sim.set_params(p)
    sim.attach_lpc(net)
    base = out["jacobi_cg"]["ms"]
    out["speedup"] = base / out["learned_fcg"]["ms"]
    out["speedup_vs_jacobi_cg"] = {k: base / out[k]["ms"] for k, _, _ in PPE_SOLVERS}
    out["system"] = "PPE of the current step, cold start, rel-res %g" % p.cg_tol
    return out


STAGES_TO_MATRIX = ("PNUM", "NEIGHBOR", "SPS", "V_FRACTION", "VISCOSITY", "PREDICTION", "PNUMSTAR",
                    "BCON", "MATRIX")


def ppe_report(checkpoints, scale=1, weights=None, relaxp=None):
    """The FP64 run of the 3D dam break; at each checkpoint step the next
    step's PPE is assembled (the stages up to MATRIX, as the step runs them)
    and solved by every variant of ppe_compare. One JSON line per checkpoint.
    The out-of-band assembly leaves the run itself unchanged: the next step
    recomputes every field it wrote."""
    import voxarmps_amd as vx
    from voxarmps_amd.train_lpc import DEFAULT_WEIGHTS, load_weights

    be = vx.hip_backend()
    case = vx.dam_break_3d_c1(scale=scale)
    p = vx.dam3d_params(case)
    if relaxp is not None:
        p.relaxp = relaxp
    net = vx.LearnedPreconditioner(backend=be)
    net.set_weights(load_weights(weights or DEFAULT_WEIGHTS))
    sim = vx.Simulation(p, case, backend=be)
    sim.init()
    sim.attach_lpc(net)
    done = 0
    for cp in sorted(checkpoints):
        while done < cp:
            n = min(cp - done, 50)
            sim.step(n)
            done += n
            print(f"# step {done}: {sim.status().last_iters} iterations", file=sys.stderr, flush=True)
        for st in STAGES_TO_MATRIX:
            getattr(sim, st)()
        r = ppe_compare(sim, p, net)
        print(json.dumps({"step": cp, "particles": case.tp, **r}), flush=True)
    sim.close()
    net.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=1, help="linear down-scaling of the workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("dam3d", "ensemble"), default="dam3d",
                    help="dam3d: the 3D dam break (1,175,450 particles); ensemble: 64 uneven 2D "
                         "two-phase dam breaks batched in one context (mixed precision, SPS, "
                         "Bingham fluid 2, neighbour-search walks chosen by the autotuner)")
    ap.add_argument("--compare-solo", action="store_true",
                    help="ensemble only: also run every scenario alone for the same steps and "
                         "report the summed time")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default=None,
                    help="fp64: every kernel in FP64 (the reference's arithmetic); mixed: FP32 "
                         "velocity stages and FP32 inner CG inside FP64 iterative refinement "
                         "(default: fp64 for dam3d, mixed for ensemble)")
    ap.add_argument("--solver", choices=("jacobi", "learned"), default="jacobi",
                    help="PPE solver of the step loop: Jacobi-PCG, or flexible CG with the "
                         "learned preconditioner (Jacobi-CG fallback)")
    ap.add_argument("--lpc-weights", default=None, help="weights of the learned preconditioner")
    ap.add_argument("--no-ppe-compare", action="store_true",
                    help="skip the Jacobi-CG / learned flexible CG comparison on the last system")
    ap.add_argument("--warm-start", action="store_true",
                    help="start each PPE solve from the previous step's pressure")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (neighbour walks, predictor and product variants)")
    ap.add_argument("--relaxp", type=float, default=None,
                    help="dam3d: pressure relaxation of the predictor / corrector split "
                         "(default 0.2, the reference's; 1.0 drops the explicit pressure term "
                         "of PREDICTION, which keeps the 3D dam break physical)")
    ap.add_argument("--cg-tol", type=float, default=None,
                    help="relative residual of every PPE solve (default 1e-8; the 2D dam "
                         "break's configuration states 1e-9)")
    ap.add_argument("--precision-report", type=str, default="", metavar="STEPS",
                    help="instead of the benchmark, run the FP64 and the mixed-precision "
                         "simulations side by side and print the pressure-field L2 and Linf "
                         "differences at the listed steps (comma separated, e.g. 100,500,2000)")
    ap.add_argument("--ppe-report", type=str, default="", metavar="STEPS",
                    help="instead of the benchmark, run the FP64 simulation and solve the PPE "
                         "of each listed step (comma separated) with every solver variant")
    ap.add_argument("--stream-bandwidth", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark, time a read-only sum and a device copy of "
                         "MIB MiB (what plain streams reach on this GPU)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark, run 5 plain device copies of MIB MiB "
                         "(reference reading for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.stream_bandwidth:
        stream_bandwidth(args.stream_bandwidth)
        return
    if args.precision_report:
        precision_report([int(s) for s in args.precision_report.split(",")], args.scale,
                         args.relaxp)
        return
    if args.ppe_report:
        ppe_report([int(s) for s in args.ppe_report.split(",")], args.scale, args.lpc_weights,
                   args.relaxp)
        return

    import voxarmps_amd as vx

    be = vx.hip_backend()
    nd = C.c_int(0)
    be.lib.vox_device_count(C.byref(nd))
    if nd.value < 1:
        raise SystemExit("no HIP device visible")

    ensemble = args.workload == "ensemble"
    if ensemble:
        case, region, boxes, scen = vx.ensemble_c4()
        p = vx.ensemble_params(case)
        precision = args.precision or "mixed"
    else:
        case = vx.dam_break_3d_c1(scale=args.scale)
        p = vx.dam3d_params(case)
        if args.relaxp is not None:
            p.relaxp = args.relaxp
        precision = args.precision or "fp64"
    if args.cg_tol is not None:
        p.cg_tol = args.cg_tol
    mixed = precision == "mixed"
    p.precision = vx._abi.PREC_MIXED if mixed else vx._abi.PREC_FP64
    p.warm_start = 1 if args.warm_start else 0
    net = None
    wpath = args.lpc_weights or os.path.join(ROOT, "voxarmps_amd", "data", "lpc_weights.npy")
    if os.path.exists(wpath) or args.solver == "learned":
        from voxarmps_amd.train_lpc import load_weights
        net = vx.LearnedPreconditioner(backend=be)
        net.set_weights(load_weights(wpath))
    if args.solver == "learned":
        p.solver = vx._abi.SOLVER_LEARNED_FCG
    sim = vx.Simulation(p, case, backend=be)
    if net is not None:
        sim.attach_lpc(net)
    modes = None
    if ensemble:
        sim.set_regions(region, boxes)
    sim.init()
    if not args.no_autotune:
        sim.autotune(reps=3)
        modes = sim.tuning()
    print(f"# init done: {case.tp} particles", file=sys.stderr, flush=True)
    if args.warmup:
        sim.step(args.warmup)
    be.lib.vox_sync(sim.ctx)
    print("# warmup done", file=sys.stderr, flush=True)

    # The timed steps carry HIP events around the product's launches only
    # (the roofline's kernel); events around every stage cost about 0.5 ms per
    # step, so the stage breakdown comes from a second pass of as many steps.
    be.lib.vox_prof_reset(sim.ctx)
    be.lib.vox_prof_only(sim.ctx, b"spmv")
    be.lib.vox_prof_enable(sim.ctx, 1)
    iters = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step(1)
        iters.append(sim.status().last_iters)
    be.lib.vox_sync(sim.ctx)
    el = time.perf_counter() - t0
    be.lib.vox_prof_enable(sim.ctx, 0)
    ms = C.c_double(0.0)
    nl = C.c_int64(0)
    be.lib.vox_prof_get(sim.ctx, b"spmv", C.byref(ms), C.byref(nl))
    avg_ms = ms.value / max(nl.value, 1)
    print("# timed steps done", file=sys.stderr, flush=True)

    be.lib.vox_prof_only(sim.ctx, None)
    be.lib.vox_prof_reset(sim.ctx)
    be.lib.vox_prof_enable(sim.ctx, 1)
    t1 = time.perf_counter()
    sim.step(args.steps)
    be.lib.vox_sync(sim.ctx)
    el_prof = time.perf_counter() - t1
    be.lib.vox_prof_enable(sim.ctx, 0)
    print("# profiled steps done", file=sys.stderr, flush=True)

    st = sim.status()
    info = sim.ppe()
    nu = int(info["rowptr"].shape[0] - 1)
    nnz = int(info["rowptr"][-1])
    del info
    cbytes = st.product_col_bytes or 4
    step_total, step_parts = (None, None) if (ensemble or mixed) else \
        step_bytes(sim, case, nu, nnz, sum(iters) / len(iters), cb=cbytes, wc=st.product_window_cols)
    bytes_launch = spmv_bytes(nu, nnz, 4 if mixed else 8, cbytes, st.product_window_cols)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None
    per_kernel = {}
    for k in ("list", "list_l", "cell", "sort", "pnd", "sps", "pred", "assemble", "spmv", "cg_update", "cg_p",
              "ir_residual", "lpc_net", "lpc_apply",
              "collision", "pgrad", "phat", "prep", "euler", "bcon", "visc", "vfrac", "misc"):
        t = C.c_double(0.0)
        n = C.c_int64(0)
        be.lib.vox_prof_get(sim.ctx, k.encode(), C.byref(t), C.byref(n))
        per_kernel[k] = round(t.value / max(args.steps, 1), 4)
    ppe = ppe_compare(sim, p, net) if (net is not None and not ensemble and not args.no_ppe_compare) else None
    fallbacks = st.lpc_fallbacks
    sim.close()

    solo = None
    if ensemble and args.compare_solo:
        tot = 0.0
        for k, s in enumerate(scen):
            ps = vx.ensemble_params(s, box=boxes[k])
            ps.precision = p.precision
            a = vx.Simulation(ps, s, backend=be)
            a.init()
            a.step(1)
            be.lib.vox_sync(a.ctx)
            t1 = time.perf_counter()
            a.step(args.steps)
            be.lib.vox_sync(a.ctx)
            tot += time.perf_counter() - t1
            a.close()
        solo = {"scenarios": len(scen), "ms_per_step_sum": tot / args.steps * 1e3,
                "value": case.tp * args.steps / tot}
        print("# solo runs done", file=sys.stderr, flush=True)

    cpu = None if (args.no_cpu_baseline or ensemble) else cpu_baseline()
    value = case.tp * args.steps / el
    if ensemble:
        metric = "particle-steps/s, 64-scenario 2D two-phase ensemble"
        workload = ("64 uneven 2D two-phase dam breaks in one context (columns 20-200 x 40-200 "
                    "particles, l0=0.01 m), fluid 2 Bingham 1 Pa, SPS-LES Cs 0.12, " +
                    ("mixed FP32/FP64" if mixed else "FP64"))
    else:
        metric = "particle-steps/s at 3D 1M"
        workload = ("3D dam break, plain MPS, " + ("mixed FP32/FP64" if mixed else "FP64") +
                    ", l0=0.005 m, water 100x100x80 in 1.6x0.6x0.405 m tank, "
                    "1 wall + 2 dummy layers")
    out = {
        "metric": metric,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "particles": case.tp,
            "fluid": case.fp,
            "re_s": "2.1 l0", "re_l": "4.0 l0", "dt": p.dt, "relaxp": p.relaxp,
            "ppe": (f"FP64 iterative refinement around FP32 Jacobi-PCG (eta {p.ir_eta:g}) to "
                    f"rel-res {p.cg_tol:g}" if mixed else f"Jacobi-PCG FP64 to rel-res {p.cg_tol:g}")
                   + (", warm start" if args.warm_start else ", cold start"),
            "cg_iters_mean": sum(iters) / len(iters),
            "solver": ("flexible CG, learned preconditioner (Jacobi-CG fallback: "
                       f"{fallbacks} fallbacks)") if args.solver == "learned" else "Jacobi-PCG",
            "parallelism": "1 GPU",
            "tuning": None if modes is None else {
                "walk_s": modes[0], "walk_l": modes[1], "predictor_variant": modes[2],
                "product_variant": modes[3]},
            "solo_runs": solo,
        },
        "roofline": {
            "kernel": "k_cg_spmv (PPE q = A p + p.q)" + (", FP32" if mixed else ", FP64"),
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": measured_traffic("mixed" if mixed else "fp64", nu, nnz, cbytes),
            "bytes_per_launch": bytes_launch,
            "column_bytes": cbytes,
            "window_columns": st.product_window_cols,
            "avg_launch_ms": avg_ms,
            "launches": nl.value,
            "unknowns": nu,
            "nnz": nnz,
        },
        "fluid_particle_steps_per_s": case.fp * args.steps / el,
        "step_roofline": None if step_total is None else {
            "bytes_per_step": step_total,
            "achieved": step_total / (el / args.steps) / 1e9,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": step_total / (el / args.steps) / 1e9 / HBM_PEAK_GBS,
            "bytes_by_stage": step_parts,
        },
        "stage_ms_per_step": per_kernel,
        "stage_pass_ms_per_step": el_prof / args.steps * 1e3,
        "ppe_solve": ppe,
        "cpu_baseline": cpu,
        "status": {"collisions": st.collisions, "nan": st.nan_count, "out_of_box": st.out_of_box,
                   "max_count_s": st.max_count_s, "max_count_l": st.max_count_l},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
