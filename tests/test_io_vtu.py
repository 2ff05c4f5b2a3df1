"""VTU restart input (getTP / getVTU of inOut.cpp:25-102) on the CPU: a file
in the writer's layout parses back exactly, the reference's own line-skipping
reader (restated below from inOut.cpp:43-102) reads the same values from it,
and malformed files are refused."""
import ctypes as C

import numpy as np
import pytest

import voxarmps_amd as vx


def vtu_text(x, y, z, u, v, w, p, t):
    """The layout vox_write_vtu writes (Float64, 17 digits)."""
    m = len(x)
    f = lambda a: " ".join("%.17g" % q for q in a)
    lines = [
        "<?xml version='1.0' encoding='UTF-8'?>",
        "<VTKFile xmlns='VTK' byte_order='LittleEndian' version='0.1' type='UnstructuredGrid'>",
        " <UnstructuredGrid>",
        f"  <Piece NumberOfCells='{m}' NumberOfPoints='{m}'>",
        "   <Points>",
        "    <DataArray NumberOfComponents='3' type='Float64' Name='Position' format='ascii'>",
        f(np.stack([x, y, z], 1).ravel()) + " ",
        "    </DataArray>", "   </Points>", "   <PointData>",
        "    <DataArray NumberOfComponents='3' type='Float64' Name='Velocity' format='ascii'>",
        f(np.stack([u, v, w], 1).ravel()) + " ",
        "    </DataArray>",
        "    <DataArray NumberOfComponents='1' type='Float64' Name='Pressure' format='ascii'>",
        f(p) + " ",
        "    </DataArray>",
        "    <DataArray NumberOfComponents='1' type='Int32' Name='partType' format='ascii'>",
        " ".join(str(int(q)) for q in t) + " ",
        "    </DataArray>", "   </PointData>",
    ]
    return "\n".join(lines) + "\n   <Cells>\n  </Cells>\n  </Piece>\n </UnstructuredGrid>\n</VTKFile>"


def reference_getvtu(text):
    """inOut.cpp:25-102 restated: skip 3 lines, read TP after the first quote,
    then skip 3 / 5 / 3 / 3 lines before the position, velocity, pressure and
    type blocks (istream >> reads whitespace-separated numbers)."""
    lines = text.split("\n")
    tp = int(lines[3].split("'")[1])
    pos = np.array(lines[6].split(), float).reshape(tp, 3)
    # after the position numbers: the rest of that line, then 4 more lines
    vel = np.array(lines[6 + 5].split(), float).reshape(tp, 3)
    pr = np.array(lines[11 + 3].split(), float)
    ty = np.array(lines[14 + 3].split(), int)
    return tp, pos, vel, pr, ty


def _read(path, dim, tp):
    be = vx.hip_backend()
    a = {k: np.zeros(tp) for k in ("x", "y", "z", "u", "v", "w", "p")}
    t = np.zeros(tp, np.int32)
    g, w, f = C.c_int(), C.c_int(), C.c_int()
    dp = C.POINTER(C.c_double)
    rc = be.lib.vox_vtu_read(str(path).encode(), dim, tp,
                             *(a[k].ctypes.data_as(dp) for k in ("x", "y", "z", "u", "v", "w", "p")),
                             t.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(g), C.byref(w), C.byref(f))
    return rc, a, t, (g.value, w.value, f.value)


@pytest.mark.parametrize("dim", [2, 3])
def test_vtu_reader_round_trip(tmp_path, dim):
    rng = np.random.default_rng(dim)
    n = 257
    arr = [rng.normal(0, 1, n) * s for s in (1, 1e-3, 1e3, 1, 1e-7, 1, 1e4)]
    if dim == 2:
        arr[2][:] = 0.0
        arr[5][:] = 0.0
    t = rng.integers(-1, 3, n).astype(np.int32)
    text = vtu_text(*arr, t)
    path = tmp_path / "in.vtu"
    path.write_text(text)
    be = vx.hip_backend()
    cnt = C.c_int(0)
    assert be.lib.vox_vtu_count(str(path).encode(), C.byref(cnt)) == 0 and cnt.value == n
    rc, a, tt, counts = _read(path, dim, n)
    assert rc == 0, be.lib.vox_last_error()
    for k, ref in zip(("x", "y", "z", "u", "v", "w", "p"), arr):
        assert np.array_equal(a[k], ref), k                  # 17 digits: exact
    assert np.array_equal(tt, t)
    assert counts == (int((t < 0).sum()), int((t == 0).sum()), int((t > 0).sum()))
    # the reference's reader finds the same numbers in the same file
    tp, pos, vel, pr, ty = reference_getvtu(text)
    assert tp == n
    assert np.array_equal(pos[:, 0], arr[0]) and np.array_equal(vel[:, 2], arr[5])
    assert np.array_equal(pr, arr[6]) and np.array_equal(ty, t)


def test_vtu_reader_refuses_malformed_files(tmp_path):
    n = 10
    z = np.zeros(n)
    text = vtu_text(z, z, z, z, z, z, z, np.ones(n, np.int32))
    be = vx.hip_backend()
    bad = tmp_path / "short.vtu"
    bad.write_text(text.replace(" ".join(["0"] * n) + " ", " ".join(["0"] * (n - 3)) + " ", 1))
    rc, *_ = _read(bad, 2, n)
    assert rc == vx._abi.VOX_ERR_ARG
    rc, *_ = _read(tmp_path / "in.vtu", 2, n + 1)           # missing file
    assert rc == vx._abi.VOX_ERR_ARG
    good = tmp_path / "good.vtu"
    good.write_text(text)
    rc, *_ = _read(good, 2, n + 1)                           # count differs
    assert rc == vx._abi.VOX_ERR_ARG
    assert b"differs" in be.lib.vox_last_error()
