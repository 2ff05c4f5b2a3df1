"""voxmps, the reference's 16-argument command line over the library: the
dam-break run it writes to output/ matches the oracle run set up with the
same preset (vox_params_main) from the same input; the other tests of
main.cu (drop, shear, turb, and oil from the restated generator) run."""
import os
import subprocess

import numpy as np
import pytest

import voxarmps_amd as vx
from voxarmps_amd import generators as gen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "voxarmps_amd", "voxmps")


def _voxmps(args, tmp, inputs=GOLDEN, steps=20, every=None):
    cmd = [CLI, *args, "--inputs", str(inputs), "--outdir", str(tmp), "--max-steps", str(steps),
           "--every", str(every or steps)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-600:], r.stderr[-600:])
    return r


def test_cli_dam_break_matches_oracle(oracle, tmp_path):
    r = _voxmps(["dam", "novisc", "noturb", "nocomp", "nomatopt", "nosrcopt", "sequential", "2d",
                 "printall", "no", "1000.0", "0.000001", "500.0", "0.001", "1", "0.0005"],
                tmp_path)
    assert r.returncode == 0
    out = vx.read_vtu(tmp_path / "output" / "out20.vtu", 2, 0.01)
    assert (tmp_path / "output" / "out0.vtu").exists()
    case = vx.cases.read_grd(os.path.join(GOLDEN, "input-dam.grd"), 2)
    case.l0 = 0.01
    be = vx.hip_backend()
    import ctypes as C
    p = vx.VoxParams()
    dp = C.POINTER(C.c_double)
    assert be.lib.vox_params_main(C.byref(p), b"dam", 0, 0, 0, 0, 0, b"sequential", 0, 1000.0,
                                  1e-6, 500.0, 1e-3, 0.0005, case.tp, case.x.ctypes.data_as(dp),
                                  case.y.ctypes.data_as(dp), case.z.ctypes.data_as(dp)) == 0
    c = vx.Simulation(p, case, backend=oracle)
    c.init()
    c.step(20)
    for k, f in (("x", "X"), ("y", "Y"), ("u", "U"), ("v", "V"), ("p", "P")):
        a, b = getattr(out, k), c.get(f)
        assert np.max(np.abs(a - b)) <= 1e-6 * max(np.max(np.abs(b)), 1.0), k
    assert "Total time steps: 20" in r.stdout
    c.close()


@pytest.mark.parametrize("test", ["drop", "shear", "turb", "oil"])
def test_cli_runs_every_test(test, tmp_path):
    inputs = GOLDEN
    if test == "oil":
        inputs = tmp_path / "inputs"
        assert gen.main(["oil", str(inputs)]) == 0
    if test == "turb":                      # the Poiseuille input is kept compressed
        import gzip
        inputs = tmp_path / "inputs"
        os.makedirs(inputs, exist_ok=True)
        with gzip.open(os.path.join(GOLDEN, "input-turb.grd.gz"), "rb") as f:
            (inputs / "input-turb.grd").write_bytes(f.read())
    r = _voxmps([test, "novisc", "noturb", "nocomp", "nomatopt", "nosrcopt", "openmp", "2d", "no",
                 "printfluid", "1000.0", "0.000001", "500.0", "0.001", "1", "0.0005"], tmp_path,
                inputs=inputs, steps=10, every=10)
    assert r.returncode == 0
    assert "0 NaN, 0 outside the box" in r.stdout
    out = vx.read_vtu(tmp_path / "output_fluid" / "out10.vtu", 2, 0.01)
    assert out.tp > 0 and np.all(out.ptype > 0) and np.all(np.isfinite(out.p))
    if test == "oil":
        assert "WEAKLY COMPRESSIBLE" in r.stdout and set(np.unique(out.ptype)) == {1, 2}


def test_cli_refuses_bad_arguments(tmp_path):
    r = subprocess.run([CLI, "dam"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "expected 16" in r.stderr
    r = _voxmps(["dam", "no", "no", "no", "no", "no", "cuda", "2d", "no", "no", "1000", "1e-6",
                 "500", "1e-3", "1", "0.0005"], tmp_path)
    assert r.returncode == 1 and "code optimization" in r.stderr
