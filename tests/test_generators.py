"""The case generators and the command-line presets, on the CPU.

The reference's inputs are its generator's output: the restated generators
must write them byte for byte (the shear cavity's inner lid layer excepted:
the input moves it at 0.5, the generator source at 1.0). The oil spill has no
input in the reference, so its restatement is checked by construction: row
counts per block, type order, geometry. vox_params_main must set the
constants main.cu derives from its 16 arguments."""
import ctypes as C
import gzip
import os

import numpy as np
import pytest

import voxarmps_amd as vx
from voxarmps_amd import generators as gen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(name):
    path = os.path.join(GOLDEN, name)
    if os.path.exists(path):
        return open(path).read()
    return gzip.open(path + ".gz", "rt").read()


@pytest.mark.parametrize("name,make", [("input-dam.grd", lambda: gen.dam_break_2d(0.01, 1)),
                                       ("input-drop.grd", lambda: gen.water_drop_2d(0.01, 1)),
                                       ("input-dam3d.grd", lambda: gen.dam_break_3d(0.01, 1))])
def test_generator_writes_the_reference_input(name, make):
    assert make() == _golden(name)


def test_shear_generator_matches_but_for_the_inner_lid():
    ref = _golden("input-shear.grd").split("\n")
    txt = gen.shear_cavity_2d(0.01, 1).split("\n")
    assert len(ref) == len(txt)
    diff = [(a, b) for a, b in zip(ref, txt) if a != b]
    assert len(diff) == 36 - 4          # the mp - 4 rows of the inner lid layer (y = 0.34)
    for a, b in diff:
        fa, fb = a.split(), b.split()
        assert fa[:3] == fb[:3] and fa[4:] == fb[4:]
        assert (fa[1], fa[3], fb[3]) == ("0.340000", "0.5", "1.0")


def test_oil_spill_generator_by_construction():
    dl = 0.004
    text = gen.oil_spill_2d(dl, 1.0)
    case = gen.case_from_text(text, 2, dl)
    mp, npart = 200, 25
    t = case.ptype
    # dummy, wall, water, oil blocks in that order
    assert np.all(np.diff(np.where(t == -1, 0, np.where(t == 0, 1, t + 1))) >= 0)
    assert (t == -1).sum() == (2 * mp + 5) + (mp // 4 + 25) + (mp // 4 + 26)
    walls = (2 * mp + 3) + (mp // 4 + 24) + (mp // 4 + 25) + (2 * mp + 1) + (mp // 4 + 23) + \
        (mp // 4 + 24) + 3 * (mp // 20) + 3 * ((mp // 10) * 3)
    assert (t == 0).sum() == walls
    size = int(200.0 * 465 * dl)                        # 372 columns at 4 mm
    assert size == 372
    assert (t == 1).sum() == (size - 1) * npart
    assert (t == 2).sum() == (size + 28 - (size + 2)) * (2 * npart + 13)
    # the partition: three columns at 2 xl - 28 dl + (2..4) dl with a gap of
    # three rows above the floor's mp/20 rows
    part = (t == 0) & (case.x > 1.494) & (case.x < 1.506) & (case.y > 2.5 * dl)
    assert np.unique(np.round(case.x[part] / dl)).size == 3
    ys = np.unique(np.round(case.y[part] / dl).astype(int))
    assert set(range(3, 3 + mp // 20)) <= set(ys) and not ({13, 14, 15} & set(ys))
    # water left of the partition, oil right of it, both inside the tank
    assert case.x[t == 1].max() < case.x[part].min() < case.x[part].max() < case.x[t == 2].min()
    assert case.x[t > 0].max() < case.x[t == -1].max()
    # no two particles on one lattice site
    key = np.round(case.x / dl).astype(np.int64) * 100000 + np.round(case.y / dl).astype(np.int64)
    assert np.unique(key).size == case.tp
    assert int(text.split("\n")[0]) == case.tp


def _preset(test, visc=0, turb=0, comp=0, matopt=0, srcopt=0, code=b"sequential", three=0,
            dt=0.0005, case=None):
    be = vx.hip_backend()
    p = vx.VoxParams()
    dp = C.POINTER(C.c_double)
    args = (0, None, None, None) if case is None else (
        case.tp, case.x.ctypes.data_as(dp), case.y.ctypes.data_as(dp), case.z.ctypes.data_as(dp))
    rc = be.lib.vox_params_main(C.byref(p), test, visc, turb, comp, matopt, srcopt, code, three,
                                1000.0, 1e-6, 500.0, 1e-3, dt, *args)
    return rc, p


def test_cli_presets_follow_main_cu():
    case = vx.cases.read_grd(os.path.join(GOLDEN, "input-dam.grd"), 2)
    rc, p = _preset(b"dam", case=case)
    assert rc == 0
    ref = vx.set_box(vx.default_params(2, 0.01), case)
    for k in ("dl", "re_s", "re_l", "lambda_", "relaxp", "cc", "coll", "delta", "xmin", "xmax",
              "ymin", "ymax"):
        assert getattr(p, k) == getattr(ref, k), k
    assert (p.ktype, p.method, p.dt_mode, p.turb, p.gy, p.pmax) == (2, 1, 1, 0, -9.8, 4.5 * case.tp)
    rc, p = _preset(b"oil", code=b"openmp")               # oil forces the compressible model
    assert rc == 0 and p.ktype == 6 and p.method == 3 and p.dl == 0.004 and p.relaxp == 0.5
    assert p.cc == 0.5 and p.lambda_ == 0.14288 * (4.0 * 0.004) ** 2
    rc, p = _preset(b"turb")
    assert (p.turb, p.gy, p.vel_cor, p.test_kind) == (1, 0.0, 0.8, vx._abi.TEST_TURB)
    rc, p = _preset(b"shear")
    assert p.gy == 0.0 and p.turb == 0 and p.test_kind == vx._abi.TEST_DAM
    rc, p = _preset(b"dam", three=1)
    assert p.dim == 3 and p.re_s == p.re_l == 3.2 * 0.01
    rc, p = _preset(b"drop", three=1)                     # 3D only for the dam test
    assert p.dim == 2 and p.test_kind == vx._abi.TEST_DROP
    rc, p = _preset(b"dam", visc=1, code=b"gpu")
    assert (p.fluid2_type, p.n_pl, p.meu0, p.ref_gpu) == (1, 1.3, 0.04, 1)
    drop = vx.cases.read_grd(os.path.join(GOLDEN, "input-drop.grd"), 2)
    rc, p = _preset(b"drop", case=drop)
    assert p.xmin == drop.x.min() - 100 * 0.01 and p.ymin == drop.y.min() - 10000 * 0.01
    assert p.ymax == drop.y.max() + 100 * 0.01
    assert _preset(b"dam", dt=0.006)[0] == vx._abi.VOX_ERR_ARG
    assert _preset(b"dam", code=b"cuda")[0] == vx._abi.VOX_ERR_ARG
    assert _preset(b"dam", three=1, code=b"gpu")[0] == vx._abi.VOX_ERR_ARG
    be = vx.hip_backend()
    assert be.lib.vox_main_input(b"dam", 1) == b"input-dam3d.grd"
    assert be.lib.vox_main_input(b"oil", 0) == b"input-oil.grd"
    assert be.lib.vox_main_input(b"anything", 0) == b"input-dam.grd"
