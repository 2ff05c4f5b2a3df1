"""Run parameters: the reference's constants (main.cu:211-494) and the
workload presets of the two dam breaks.

`default_params` mirrors vox_params_default() of the C ABI value by value; a
test keeps the two in agreement.
"""
import math

from ._abi import PREC_FP64, SOLVER_JACOBI_CG, TEST_DAM, VoxParams


def default_params(dim, dl):
    p = VoxParams()
    p.dim = dim
    p.ktype = 2
    p.dl = dl
    re = 3.2 * dl if dim == 3 else 4.0 * dl
    p.re_s = re
    p.re_l = re
    p.delta = dl / 5.0
    p.lambda_ = 0.16832 * (re * re)
    p.rho1, p.nu1, p.rho2, p.nu2 = 1000.0, 0.000001, 500.0, 0.001
    p.gx, p.gy, p.gz = 0.0, -9.8, 0.0
    p.courant, p.c01, p.c02, p.dt_max = 1.0, 20.0, 20.0, 0.005
    p.dt = p.courant * dl / 20.0
    p.dt_mode = 0
    p.turb = 0
    p.dirichlet = 0.97
    p.relaxp = 0.2
    p.coll = math.sqrt(dl * dl * 0.5 * 0.5)
    p.cc = 0.20
    p.vel_cor = 1.0
    p.cs = 0.18
    p.ncorrection = 0.99
    p.vmax = 2.50
    p.pmin = 0.0
    p.pmax = -1.0
    p.fluid2_type = 0
    p.kh_correction = 1
    p.n_pl, p.meu0, p.phi, p.cohes = 1.3, 0.04, 0.0, 0.0
    p.test_kind = TEST_DAM
    p.coef2d = 1
    p.precision = PREC_FP64
    p.solver = SOLVER_JACOBI_CG
    p.cg_tol = 1e-9
    p.cg_maxit = 5000
    p.cg_check = 1
    p.cap_s = 0
    p.cap_l = 0
    p.ir_eta = 1e-4
    p.ir_maxouter = 8
    p.warm_start = 0
    p.list_mode = 0
    p.lpc_cap_pct = 200
    p.method = 1
    p.matopt = 0
    p.srcopt = 0
    p.ref_gpu = 0
    p.wc_beta = 0.92
    return p


def set_box(p, case):
    """Grid box = particle bounding box expanded by DL (main.cu:692-715)."""
    p.xmin = float(case.x.min()) - p.dl
    p.xmax = float(case.x.max()) + p.dl
    p.ymin = float(case.y.min()) - p.dl
    p.ymax = float(case.y.max()) + p.dl
    if p.dim == 3:
        p.zmin = float(case.z.min()) - p.dl
        p.zmax = float(case.z.max()) + p.dl
    else:
        p.zmin = 0.0
        p.zmax = 0.0
    return p


def two_radius(p, re_s_factor=2.1, re_l_factor=4.0):
    """Separate radii for gradient/PND (re_s) and Laplacian (re_l) stages;
    lambda keeps the reference's formula 0.16832 re^2 on the Laplacian radius."""
    p.re_s = re_s_factor * p.dl
    p.re_l = re_l_factor * p.dl
    p.lambda_ = 0.16832 * (p.re_l * p.re_l)
    return p


def dam2d_params(case, dt=1e-4, tol=1e-9):
    """2D dam break, plain MPS, incompressible, FP64: re 2.1 l0 / 4.0 l0,
    dt 1e-4 s, rho 1000, nu 1e-6, g -9.81, CG to rel-res `tol`."""
    p = default_params(2, case.l0)
    two_radius(p)
    p.gy = -9.81
    p.dt = dt
    p.cg_tol = tol
    return set_box(p, case)


def dam3d_params(case, dt=5e-5, tol=1e-8):
    """3D dam break, plain MPS, FP64 on GPU: same radii and fluid as the 2D
    case, dt 5e-5 s."""
    p = default_params(3, case.l0)
    two_radius(p)
    p.gy = -9.81
    p.dt = dt
    p.cg_tol = tol
    return set_box(p, case)


def ensemble_params(case, dt=1e-4, tol=1e-8, box=None):
    """The batched two-phase ensemble (also each scenario run alone, with
    `box` = its own box): 2D, re 2.1 l0 / 4.0 l0, dt 1e-4 s; fluid 1 water
    (rho 1000, nu 1e-6), fluid 2 rho 850, nu 1e-3 with a Bingham yield stress
    of 1 Pa (Herschel-Bulkley with N = 1, MEU0 = rho2 nu2, cohesion 1 Pa, phi
    0); SPS-LES on with Cs 0.12; mixed precision, PPE to rel-res `tol`."""
    from ._abi import PREC_MIXED
    p = default_params(2, case.l0)
    two_radius(p)
    p.gy = -9.81
    p.dt = dt
    p.cg_tol = tol
    p.rho2, p.nu2 = 850.0, 1e-3
    p.fluid2_type = 1
    p.cohes = 1.0
    p.phi = 0.0
    p.n_pl = 1.0
    p.meu0 = p.rho2 * p.nu2
    p.turb = 1
    p.cs = 0.12
    p.precision = PREC_MIXED
    if box is None:
        return set_box(p, case)
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = (float(v) for v in box)
    return p


def wc_params(case, dt=None):
    """Weakly compressible MPS (PRESSURECALC method 3, main.cu:249-291 with
    `comp`): cubic kernel (KTYPE 6), relaxp 0.5, collision coefficient 0.5,
    equation-of-state free-surface ratio 0.92, sound speeds 20 m/s; one radius
    4.0 l0 in 2D and 3.2 l0 in 3D (main.cu:252-259)."""
    p = default_params(case.dim, case.l0)
    k = 4.0 if case.dim == 2 else 3.2
    p.re_s = p.re_l = k * case.l0
    p.lambda_ = 0.16832 * (p.re_l * p.re_l)
    p.ktype = 6
    p.method = 3
    p.relaxp = 0.5
    p.cc = 0.5
    p.wc_beta = 0.92
    p.gy = -9.81
    p.dt = dt if dt is not None else 1e-4
    return set_box(p, case)
