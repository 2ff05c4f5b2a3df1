"""VTU output and restart, and the PPE dump/load pair, on the GPU.

* vox_write_vtu writes full precision: vox_vtu_read gives back the context's
  fields bit for bit (every particle, and the fluid only), in 2D and 3D.
* A run restarted from that file (getVTU, inOut.cpp:43-102, then the
  initialisation of main.cu:966-1012) matches the oracle restarted from the
  same file.
* vox_ppe_load is the inverse of vox_ppe_export: a loaded system exports back
  unchanged, and vox_ppe_solve solves it (residual recomputed in numpy),
  through the product on 32-bit columns (an assembled system's product
  gathers from LDS windows)."""
import numpy as np
import pytest
import scipy.sparse as sp

import voxarmps_amd as vx

pytestmark = pytest.mark.gpu


def _run(hip, dim, steps):
    if dim == 2:
        case = vx.dam_break_2d_c0()
        p = vx.dam2d_params(case)
    else:
        case = vx.dam_break(3, 0.005, (12, 12, 10), (30, 16, 11))
        p = vx.dam3d_params(case)
    s = vx.Simulation(p, case, backend=hip)
    s.init()
    s.step(steps)
    return s, case, p


@pytest.mark.parametrize("dim", [2, 3])
def test_vtu_round_trip_bit_exact(hip, tmp_path, dim):
    s, case, p = _run(hip, dim, 15)
    path = tmp_path / "out.vtu"
    s.write_vtu(path)
    back = vx.read_vtu(path, dim, case.l0, backend=hip)
    assert back.tp == case.tp
    for k, f in (("x", "X"), ("y", "Y"), ("u", "U"), ("v", "V"), ("p", "P")) + \
            ((("z", "Z"), ("w", "W")) if dim == 3 else ()):
        assert np.array_equal(getattr(back, k), s.get(f)), k
    assert np.array_equal(back.ptype, case.ptype)
    assert back.meta["counts"] == (case.gp, case.wp, case.fp)
    fl = tmp_path / "fluid.vtu"
    s.write_vtu(fl, fluid_only=True)
    only = vx.read_vtu(fl, dim, case.l0, backend=hip)
    f = case.ptype > 0
    assert only.tp == case.fp and np.array_equal(only.x, s.get("X")[f])
    s.close()


def test_vtu_restart_matches_oracle(hip, oracle, tmp_path):
    s, case, p = _run(hip, 2, 40)
    path = tmp_path / "restart.vtu"
    s.write_vtu(path)
    s.close()
    rc = vx.read_vtu(path, 2, case.l0, backend=hip)
    assert np.any(rc.u != 0) and np.any(rc.p != 0)
    q = vx.dam2d_params(rc)
    q.xmin, q.xmax, q.ymin, q.ymax = p.xmin, p.xmax, p.ymin, p.ymax
    g = vx.Simulation(q, rc, backend=hip)
    c = vx.Simulation(q, rc, backend=oracle)
    for sim in (g, c):
        sim.init()
        sim.step(20)
    for n in ("X", "Y", "U", "V", "P"):
        a, b = g.get(n), c.get(n)
        assert np.max(np.abs(a - b)) <= 1e-6 * max(np.max(np.abs(b)), 1.0), n
    for w in (vx.LIST_S, vx.LIST_L):
        cg, ig = g.neighbors(w)
        cc, ic = c.neighbors(w)
        assert np.array_equal(cg, cc) and np.array_equal(ig, ic)
    g.close()
    c.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_ppe_load_is_the_inverse_of_export(hip, dim):
    s, case, p = _run(hip, dim, 10)
    for st in ("PNUM", "NEIGHBOR", "V_FRACTION", "VISCOSITY", "PREDICTION", "PNUMSTAR", "BCON",
               "MATRIX"):
        getattr(s, st)()
    A = s.ppe()
    nu = A["rhs"].shape[0]
    assert nu > 50
    # the assembled rows ascend (unknowns in the list's cell order): the
    # product gathers p from its blocks' LDS windows by 16-bit ranks
    s.CGM()
    assert s.status().product_col_bytes == 2 and s.status().product_window_cols > 0
    # a second context of the same case, never assembled: load, export, solve
    t, _, _ = _run(hip, dim, 0)
    rng = np.random.default_rng(dim)
    B = dict(A)
    B["rhs"] = rng.normal(0, 1e5, nu)
    t.load_ppe(B)
    E = t.ppe()
    for k in ("rowptr", "cols", "vals", "rhs", "ids"):
        assert np.array_equal(E[k], B[k]), k
    info = t.CGM()
    st = t.status()                                   # loaded rows: 32-bit columns
    assert st.product_col_bytes == 4 and st.product_window_cols == 0
    x = t.get("PNEW")
    M = sp.csr_matrix((A["vals"], A["cols"], A["rowptr"]), shape=(nu, nu))
    xu = x[A["ids"]]
    rel = np.linalg.norm(B["rhs"] - M @ xu) / np.linalg.norm(B["rhs"])
    assert info.converged and info.unknowns == nu and rel <= 1.01 * p.cg_tol
    others = np.ones(case.tp, bool)
    others[A["ids"]] = False
    assert np.all(x[others] == 0.0)
    # malformed systems are refused before anything is copied
    bad = dict(B)
    bad["cols"] = B["cols"].copy()
    bad["cols"][B["rowptr"][3]] = 0 if 3 != 0 else 1       # row 3 no longer starts with its diagonal
    with pytest.raises(vx.VoxError):
        t.load_ppe(bad)
    dup = dict(B)
    dup["ids"] = B["ids"].copy()
    dup["ids"][1] = dup["ids"][0]
    with pytest.raises(vx.VoxError):
        t.load_ppe(dup)
    s.close()
    t.close()
