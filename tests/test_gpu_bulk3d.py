"""Long 3D runs against the oracle: the dam break at a quarter of the
1,175,450-particle size (water 25 x 25 x 20, l0 = 5 mm, 38,030 particles,
radii 2.1 l0 / 4.0 l0, dt 5e-5 s, PPE to 1e-8), where single particles part
after a few hundred steps (collisions, the CG tolerance), so bulk metrics are
compared every 50 steps: the surge front (largest fluid x), the column height
(largest fluid y) and the pressure of a probe region at the bottom of the
column (mean over the fluid within 2 spacings of the floor and 5 of the back
wall), averaged over the run.

With the configuration's relaxp = 0.2 the explicit (1 - relaxp) pressure term
of PREDICTION (functions.cu:2354-2358) on the 2.1 l0 radius drives the 3D
pressure up without bound, on the GPU and on the oracle alike; with relaxp = 1
the run stays physical. Both are checked."""
import numpy as np
import pytest

import voxarmps_amd as vx

pytestmark = pytest.mark.gpu


def _bulk(sim, case):
    x, y, P = sim.get("X"), sim.get("Y"), sim.get("P")
    f = case.ptype > 0
    l0 = case.l0
    probe = f & (y < 2.5 * l0) & (x < 5.5 * l0)
    return x[f].max(), y[f].max(), float(P[probe].mean()), float(P[f].mean()), float(P[f].max())


def _run(hip, oracle, relaxp, nsteps, every=50):
    case = vx.dam_break_3d_c1(scale=4)
    p = vx.dam3d_params(case, tol=1e-8)
    p.relaxp = relaxp
    g = vx.Simulation(p, case, backend=hip)
    c = vx.Simulation(p, case, backend=oracle)
    g.init()
    c.init()
    rows = []
    for k in range(every, nsteps + 1, every):
        g.step(every)
        c.step(every)
        mg, mc = _bulk(g, case), _bulk(c, case)
        rows.append((k, mg, mc))
        print(f"relaxp {relaxp} step {k}: front {mg[0]:.4f} / {mc[0]:.4f} m, height {mg[1]:.4f} / "
              f"{mc[1]:.4f} m, probe p {mg[2]:.0f} / {mc[2]:.0f} Pa, mean p {mg[3]:.0f} / {mc[3]:.0f}, "
              f"max p {mg[4]:.0f} / {mc[4]:.0f}", flush=True)
    assert g.status().nan_count == 0
    g.close()
    c.close()
    return case, rows


def test_long_run_bulk_metrics_3d(hip, oracle):
    """relaxp = 1, 600 steps (30 ms): front and height to one spacing every 50
    steps, the probe pressure averaged over the run to 5%, and the run
    bounded: the mean fluid pressure stays between 0.2 and 2 times the
    column's mean hydrostatic pressure and its maximum below 20 times the
    hydrostatic pressure at the floor."""
    case, rows = _run(hip, oracle, 1.0, 600)
    l0 = case.l0
    for k, mg, mc in rows:
        assert abs(mg[0] - mc[0]) <= l0 and abs(mg[1] - mc[1]) <= l0, k
    a = np.mean([r[1][2] for r in rows])
    b = np.mean([r[2][2] for r in rows])
    print(f"probe pressure averaged over the run: {a:.1f} / {b:.1f} Pa")
    assert abs(a - b) <= 0.05 * abs(b)
    floor = 1000.0 * 9.81 * 0.125               # hydrostatic pressure at the floor
    for k, mg, mc in rows:
        assert 0.2 * floor / 2 <= mg[3] <= 2.0 * floor / 2, k
        assert mg[4] <= 20 * floor, k


def test_parity_mode_3d_pressure_growth(hip, oracle):
    """relaxp = 0.2 (the configuration), 250 steps: the maximum fluid
    pressure grows far past hydrostatic on both sides, and the GPU follows the
    oracle: mean and maximum fluid pressure within 10% every 50 steps, front
    and height to one spacing."""
    case, rows = _run(hip, oracle, 0.2, 250)
    l0 = case.l0
    for k, mg, mc in rows:
        assert abs(mg[0] - mc[0]) <= l0 and abs(mg[1] - mc[1]) <= l0, k
        assert abs(mg[3] - mc[3]) <= 0.1 * abs(mc[3]), k
        assert abs(mg[4] - mc[4]) <= 0.1 * abs(mc[4]), k
    assert rows[-1][2][4] > 20 * 1000.0 * 9.81 * 0.125     # the oracle's own growth
