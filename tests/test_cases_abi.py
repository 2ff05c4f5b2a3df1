"""CPU tests of the case generator, the .grd format and the C ABI surface
(library loads, exports every declared symbol, struct layouts agree, argument
errors are reported) — no GPU work is launched."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import voxarmps_amd as vx
from voxarmps_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_splitmix64_known_answers():
    # first outputs of SplitMix64 seeded with 0 (Vigna's reference generator)
    z = vx.cases.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_jitter_bounds_and_determinism():
    c = vx.dam_break_2d_c0()
    a = vx.jitter(c, 0.001, seed=0)
    b = vx.jitter(c, 0.001, seed=0)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    d = np.concatenate([a.x - c.x, a.y - c.y])
    assert d.min() >= -0.001 and d.max() < 0.001 and abs(d.mean()) < 1e-4


def test_dam_break_counts_and_order():
    c0 = vx.dam_break_2d_c0()
    assert c0.counts() == (2105, 574, 281, 1250)
    assert np.all(np.diff(c0.ptype) >= 0)                # dummy -> wall -> fluid
    c1 = vx.dam_break_3d_c1()
    assert c1.fp == 800000 and c1.tp == 1175450
    assert np.all(np.diff(c1.ptype) >= 0)
    f = c1.ptype > 0
    assert np.isclose(c1.x[f].min(), 0.005) and np.isclose(c1.x[f].max(), 0.5)
    assert np.isclose(c1.z[f].max(), 0.4)
    # no two particles share a lattice site
    key = np.round(np.stack([c1.x, c1.y, c1.z], 1) / 0.005).astype(np.int64)
    assert np.unique(key, axis=0).shape[0] == c1.tp


def test_default_params_match_c_abi():
    be = vx.hip_backend()
    for dim, dl in ((2, 0.01), (3, 0.005)):
        a = vx.VoxParams()
        assert be.lib.vox_params_default(C.byref(a), dim, dl) == 0
        b = vx.default_params(dim, dl)
        assert a.as_dict() == b.as_dict()


def test_params_box_matches_python():
    be = vx.hip_backend()
    c = vx.jitter(vx.dam_break(3, 0.005, (8, 8, 8), (20, 12, 9)), 0.0005)
    a = vx.default_params(3, 0.005)
    xd, yd, zd = (np.ascontiguousarray(v) for v in (c.x, c.y, c.z))
    dp = C.POINTER(C.c_double)
    assert be.lib.vox_params_box(C.byref(a), c.tp, xd.ctypes.data_as(dp), yd.ctypes.data_as(dp),
                                 zd.ctypes.data_as(dp)) == 0
    b = vx.set_box(vx.default_params(3, 0.005), c)
    for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
        assert getattr(a, k) == getattr(b, k)


def test_library_exports_every_declared_symbol():
    be = vx.hip_backend()
    hdr = open(os.path.join(ROOT, "include", "voxarmps.h")).read()
    names = set(re.findall(r"\b(vox_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(be.lib, n)]
    assert not missing, missing
    assert set(_abi.PROTOTYPES) <= names | {"vox_last_error", "vox_version"}
    _abi.check_sizes(be.lib)
    assert b"gfx950" in be.lib.vox_version()


def test_create_rejects_bad_arguments():
    be = vx.hip_backend()
    p = vx.default_params(2, 0.01)
    ctx = C.c_void_p()
    # counts that do not add up
    rc = be.lib.vox_create(C.byref(ctx), C.byref(p), 10, 1, 1, 1)
    assert rc == _abi.VOX_ERR_ARG and b"gp + wp + fp" in be.lib.vox_last_error()
    # empty grid box
    rc = be.lib.vox_create(C.byref(ctx), C.byref(p), 3, 1, 1, 1)
    assert rc == _abi.VOX_ERR_ARG and b"box" in be.lib.vox_last_error()
    # unsupported precision mode is refused, not silently run
    p.xmin, p.xmax, p.ymin, p.ymax = 0.0, 1.0, 0.0, 1.0
    p.precision = 7
    rc = be.lib.vox_create(C.byref(ctx), C.byref(p), 3, 1, 1, 1)
    assert rc != 0


@pytest.mark.parametrize("field,value,code,text", [
    ("method", 2, _abi.VOX_ERR_NOLIB, b"method"),          # Method 2 (M-MPS) is not built
    ("matopt", 2, _abi.VOX_ERR_ARG, b"matopt"),
    ("srcopt", -1, _abi.VOX_ERR_ARG, b"srcopt"),
    ("solver", 5, _abi.VOX_ERR_NOLIB, b"solver"),
    ("lpc_cap_pct", 0, _abi.VOX_ERR_ARG, b"lpc_cap_pct"),
    ("list_mode", 3, _abi.VOX_ERR_ARG, b"list_mode"),
    ("ir_maxouter", 0, _abi.VOX_ERR_ARG, b"ir_maxouter"),
    ("ref_gpu", 2, _abi.VOX_ERR_ARG, b"ref_gpu"),
])
def test_create_validates_options(field, value, code, text):
    """Options outside their documented values are refused before any device
    work (so this runs without a GPU)."""
    be = vx.hip_backend()
    p = vx.default_params(2, 0.01)
    p.xmin, p.xmax, p.ymin, p.ymax = 0.0, 1.0, 0.0, 1.0
    setattr(p, field, value)
    ctx = C.c_void_p()
    rc = be.lib.vox_create(C.byref(ctx), C.byref(p), 3, 1, 1, 1)
    assert rc == code and text in be.lib.vox_last_error()


@pytest.mark.parametrize("dim,precision", [(3, 0), (2, 1)])
def test_ref_gpu_is_2d_fp64_only(dim, precision):
    """The reference's CUDA path exists in 2D only; its mode runs in FP64."""
    be = vx.hip_backend()
    p = vx.default_params(dim, 0.01)
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = 0.0, 1.0, 0.0, 1.0, 0.0, 1.0
    p.ref_gpu = 1
    p.precision = precision
    ctx = C.c_void_p()
    rc = be.lib.vox_create(C.byref(ctx), C.byref(p), 3, 1, 1, 1)
    assert rc == _abi.VOX_ERR_ARG and b"ref_gpu" in be.lib.vox_last_error()


def test_lpc_entry_points_check_arguments():
    """The learned-preconditioner entry points refuse null handles and
    arguments without touching a device."""
    be = vx.hip_backend()
    assert be.lib.vox_lpc_create(None, 0) == _abi.VOX_ERR_ARG
    assert be.lib.vox_lpc_get_weights(None, None) == _abi.VOX_ERR_ARG
    assert be.lib.vox_lpc_train(None, 1, 1e-3, 0, None, None) == _abi.VOX_ERR_ARG
    assert be.lib.vox_attach_lpc(None, None) == _abi.VOX_ERR_ARG
    assert be.lib.vox_lpc_destroy(None) == 0


def test_profiling_entry_points_check_arguments():
    """Profiling calls refuse a null context before any device work."""
    be = vx.hip_backend()
    assert be.lib.vox_prof_only(None, b"spmv") != 0
    assert be.lib.vox_prof_enable(None, 1) != 0
    assert be.lib.vox_prof_reset(None) != 0


def test_grd_round_trip_through_c_reader(tmp_path):
    be = vx.hip_backend()
    for c in (vx.jitter(vx.dam_break_2d_c0(), 0.001),
              vx.jitter(vx.dam_break(3, 0.005, (6, 6, 6), (14, 10, 7)), 0.0005)):
        c.u = np.linspace(-1, 1, c.tp)
        c.p = np.linspace(0, 5000, c.tp)
        path = str(tmp_path / f"case{c.dim}.grd")
        vx.cases.write_grd(path, c)
        n = C.c_int()
        assert be.lib.vox_grd_count(path.encode(), C.byref(n)) == 0 and n.value == c.tp
        arrs = [np.zeros(c.tp) for _ in range(7)]
        pt = np.zeros(c.tp, np.int32)
        g, w, f = C.c_int(), C.c_int(), C.c_int()
        dp = C.POINTER(C.c_double)
        rc = be.lib.vox_grd_read(path.encode(), c.dim, c.tp, *[a.ctypes.data_as(dp) for a in arrs],
                                 pt.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(g), C.byref(w),
                                 C.byref(f))
        assert rc == 0
        x, y, z, u, v, ww, p = arrs
        assert np.array_equal(x, c.x) and np.array_equal(y, c.y) and np.array_equal(u, c.u)
        assert np.array_equal(p, c.p) and np.array_equal(pt, c.ptype)
        if c.dim == 3:
            assert np.array_equal(z, c.z)
        assert (g.value, w.value, f.value) == (c.gp, c.wp, c.fp)
        back = vx.cases.read_grd(path, c.dim)
        assert np.array_equal(back.x, c.x) and np.array_equal(back.ptype, c.ptype)


def test_reference_input_fixtures_parse(tmp_path):
    """The reference's 2D inputs, committed as fixtures: 6 columns per row,
    particles ordered dummy, wall, fluid."""
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    c = vx.cases.read_grd(os.path.join(golden, "input-dam.grd"), 2)
    assert c.tp == 1122 and (c.gp, c.wp, c.fp) == (162, 312, 648)
    assert np.all(np.diff(c.ptype) >= 0)
    d = vx.cases.read_grd(os.path.join(golden, "input-drop.grd"), 2)
    assert d.tp == 1005 and (d.gp, d.wp, d.fp) == (0, 0, 1005)
    import gzip
    import hashlib
    for name, dim, counts in (("input-turb.grd", 2, (280, 560, 8970)),
                              ("input-dam3d.grd", 3, (2732, 4384, 1024))):
        with gzip.open(os.path.join(golden, name + ".gz"), "rb") as f:
            data = f.read()
        path = os.path.join(tmp_path, name)
        with open(path, "wb") as f:
            f.write(data)
        e = vx.cases.read_grd(path, dim)
        assert (e.gp, e.wp, e.fp) == counts and np.all(np.diff(e.ptype) >= 0)
        assert hashlib.sha256(data).hexdigest() in open(os.path.join(golden, "README.md")).read()


def test_reference_input_files_parse(tmp_path):
    """The reference's own .grd inputs hold 6 (2D) / 8 (3D) columns per row."""
    ref = "/root/reference/inputs"
    if not os.path.isdir(ref):
        pytest.skip("reference inputs not present on this machine")
    c = vx.cases.read_grd(os.path.join(ref, "input-dam.grd"), 2)
    assert c.tp == 1122 and (c.gp, c.wp, c.fp) == (162, 312, 648)
    c3 = vx.cases.read_grd(os.path.join(ref, "input-dam3d.grd"), 3)
    assert c3.tp == 8140 and (c3.gp, c3.wp, c3.fp) == (2732, 4384, 1024)


def test_bench_refuses_several_gpus():
    """bench.py runs on one GPU; asked for more it says so and exits 2
    before touching a device."""
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_header_is_plain_c(tmp_path):
    """include/voxarmps.h compiles as C99 and links against the library: a C
    caller (the reference's C++ loop, or an FFI) needs no C++ types."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler on this machine")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "caller.c"
    src.write_text('#include "voxarmps.h"\n'
                   "int main(void) {\n"
                   "    vox_params p;\n"
                   "    if (vox_params_default(&p, 3, 0.005) != VOX_OK) return 2;\n"
                   "    return (p.dim == 3 && p.re_s > 0.0) ? 0 : 1;\n"
                   "}\n")
    exe = tmp_path / "caller"
    libdir = os.path.join(root, "voxarmps_amd")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(root, "include"),
                        str(src), "-L", libdir, "-lvoxarmps", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert subprocess.run([str(exe)], timeout=60).returncode == 0
