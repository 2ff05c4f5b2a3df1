"""Scenario input generators: the reference's case generator
(generateSimulationCases/generateSimulationCases/genParticlesFunctions.cpp)
restated, writing the same .grd text.

Each function returns the file text the reference's generator writes for the
same (DL, mult): a particle-count line, then one row per particle, x y [z]
PTYPE u v [w] p, coordinates printed with C's "%lf" (6 decimals), in the
generator's order (dummy, wall, fluid). The reference's own inputs of the dam
break, the shear-driven cavity, the water drop and the 3D dam break are these
generators' output at DL = 0.01, mult = 1; the tests compare them byte for
byte (the shear input differs in one lid layer's velocity, see there). The oil-spill input is not in the reference (it is listed as a missing
blob), so `oil_spill_2d` is the only way to obtain it.

`case_from_text` parses any of them into a Case (the reader of main.cu:640-668).
"""
import numpy as np


def _f(v):
    """C printf "%lf" (= "%f"): fixed notation, 6 decimals."""
    return "%f" % v


def _text(count, rows):
    return "%d \n" % count + "".join(r + " \n" for r in rows)


def oil_spill_2d(dl=0.004, mult=1.0):
    """createParticlesOilSpill2D (genParticlesFunctions.cpp:3-122): a long tank
    with a partition wall near its right end, open below a gap; water
    (PTYPE 1) left of the partition and oil (PTYPE 2, two layers deeper)
    right of it. The generator prints its count line from a double with "%d",
    which is undefined in C; here the count line is the number of rows."""
    mp = int(mult * 200)
    npart = int(mult * 25)
    xl = mp * dl
    dx = xl / mp
    dy = (mp * dl) / mp
    mp2 = float(mp)
    offset_wall = xl * 2 - 28 * dx * mult
    size_x_water = int(mp2 * 465 * dx)
    rows = []
    for i in range(2 * mp + 5):                                   # dummies: floor
        rows.append(f"{_f(i * dx)} 0 -1 0 0 0")
    for i in range(mp // 4 + 25):                                 # left
        rows.append(f"0 {_f(i * dy + dy)} -1 0 0 0")
    for i in range(mp // 4 + 26):                                 # right
        rows.append(f"{_f(2 * xl + 5 * dx)} {_f(i * dy)} -1 0 0 0")
    for i in range(2 * mp + 3):                                   # outer wall layer
        rows.append(f"{_f(i * dx + dx)} {_f(dy)} 0 0 0 0")
    for i in range(mp // 4 + 24):
        rows.append(f"{_f(dx)} {_f(i * dy + 2 * dy)} 0 0 0 0")
    for i in range(mp // 4 + 25):
        rows.append(f"{_f(2 * xl + 4 * dx)} {_f(i * dy + dy)} 0 0 0 0")
    for i in range(2 * mp + 1):                                   # inner wall layer
        rows.append(f"{_f(i * dx + 2 * dx)} {_f(2 * dy)} 0 0 0 0")
    for i in range(mp // 4 + 23):
        rows.append(f"{_f(2 * dx)} {_f(i * dy + 3 * dy)} 0 0 0 0")
    for i in range(mp // 4 + 24):
        rows.append(f"{_f(2 * xl + 3 * dx)} {_f(i * dy + 2 * dy)} 0 0 0 0")
    for i in range(3):                                            # partition, below the gap
        for j in range(mp // 20):
            rows.append(f"{_f(i * dx + 2 * dx + offset_wall)} {_f(j * dy + 3 * dy)} 0 0 0 0")
    upper = (mp // 20 + 3) + (mp // 10) * 3 - 20 * (mult - 1)     # a double bound, j < upper
    for i in range(3):                                            # partition, above the gap
        j = mp // 20 + 3
        while j < upper:
            rows.append(f"{_f(i * dx + 2 * dx + offset_wall)} {_f(j * dy + 3 * dy)} 0 0 0 0")
            j += 1
    for i in range(size_x_water - 1):                             # water
        for j in range(npart):
            rows.append(f"{_f(i * dx + 3 * dx)} {_f(j * dy + 3 * dy)} 1 0 0 0")
    i = size_x_water + 2
    while i < size_x_water + 28 * mult:                           # oil
        for j in range(2 * npart + 13):
            rows.append(f"{_f(i * dx + 3 * dx)} {_f(j * dy + 3 * dy)} 2 0 0 0")
        i += 1
    return _text(len(rows), rows)


def _tank_2d(mp, dx, dy, xl, left, right):
    """The three dummy / wall layers shared by the 2D dam-break generators
    (genParticlesFunctions.cpp:142-182, 339-379)."""
    rows = []
    for i in range(2 * mp + 5):
        rows.append(f"{_f(i * dx)} 0 -1 0 0 0")
    for i in range(mp + 6 + left):
        rows.append(f"0 {_f(i * dy + dy)} -1 0 0 0")
    for i in range(mp + 7 + right):
        rows.append(f"{_f(2 * xl + 5 * dx)} {_f(i * dy)} -1 0 0 0")
    for i in range(2 * mp + 3):
        rows.append(f"{_f(i * dx + dx)} {_f(dy)} 0 0 0 0")
    for i in range(mp + 5 + left):
        rows.append(f"{_f(dx)} {_f(i * dy + 2 * dy)} 0 0 0 0")
    for i in range(mp + 6 + right):
        rows.append(f"{_f(2 * xl + 4 * dx)} {_f(i * dy + dy)} 0 0 0 0")
    for i in range(2 * mp + 1):
        rows.append(f"{_f(i * dx + 2 * dx)} {_f(2 * dy)} 0 0 0 0")
    for i in range(mp + 4 + left):
        rows.append(f"{_f(2 * dx)} {_f(i * dy + 3 * dy)} 0 0 0 0")
    for i in range(mp + 5 + right):
        rows.append(f"{_f(2 * xl + 3 * dx)} {_f(i * dy + 2 * dy)} 0 0 0 0")
    return rows


def dam_break_2d(dl=0.01, mult=1.0):
    """createParticlesDamBreak2D (genParticlesFunctions.cpp:124-219): a water
    column of mp/2 x np particles in the corner of a tank twice mp wide.
    (DL, mult) = (0.01, 1) gives the reference's inputs/input-dam.grd."""
    mp = int(36 * mult)
    npart = int(36 * mult)
    xl = mp * dl
    dx = xl / mp
    dy = (npart * dl) / npart
    count = (mp // 2) * npart + (4 * mp + 3 * 1 + 7) + (4 * mp + 3 * 3 + 5) + (4 * mp + 3 * 5 + 3)
    rows = _tank_2d(mp, dx, dy, xl, 0, 0)
    for i in range(mp // 2):
        for j in range(npart):
            rows.append(f"{_f(i * dx + 3 * dx)} {_f(j * dy + 3 * dy)} 1 0 0 0")
    return _text(count, rows)


def shear_cavity_2d(dl=0.01, mult=1.0):
    """createParticlesShearCavity2D (genParticlesFunctions.cpp:221-287): a
    closed square cavity whose two top wall layers move at u = 1.
    (0.01, 1) gives the reference's inputs/input-shear.grd, except that the
    input moves its inner lid layer (y = yl - 2 dy) at u = 0.5 (an earlier
    generator's value; the source as it stands writes 1.0 for both layers)."""
    mp = int(36 * mult)
    npart = int(36 * mult)
    xl = mp * dl
    yl = npart * dl
    dx = xl / mp
    dy = yl / npart
    count = ((mp - 7) * (npart - 6)) + 2 * (mp - 4) + 2 * (npart - 4) + 2 * (mp - 2) + \
        2 * (npart - 2) + 4 * mp
    rows = []
    for i in range(mp):
        rows.append(f"{_f(i * dx)} {_f(0.0)} 0 0 0 0")
    for i in range(npart):
        rows.append(f"{_f(0.0 * dx)} {_f(i * dy + dy)} 0 0 0 0")
    for i in range(npart):
        rows.append(f"{_f(xl)} {_f(i * dy)} 0 0 0 0")
    for i in range(mp):
        rows.append(f"{_f(i * dx + dx)} {_f(yl)} 0 0 0 0")
    for i in range(mp - 2):
        rows.append(f"{_f(i * dx + dx)} {_f(dy)} 0 0 0 0")
    for i in range(npart - 2):
        rows.append(f"{_f(dx)} {_f(i * dy + 2 * dy)} 0 0 0 0")
    for i in range(npart - 2):
        rows.append(f"{_f(xl - dx)} {_f(i * dy + dy)} 0 0 0 0")
    for i in range(mp - 2):
        rows.append(f"{_f(i * dx + 2 * dx)} {_f(yl - dy)} 0 1.0 0 0")
    for i in range(mp - 4):
        rows.append(f"{_f(i * dx + 2 * dx)} {_f(2 * dy)} 0 0 0 0")
    for i in range(npart - 4):
        rows.append(f"{_f(2 * dx)} {_f(i * dy + 3 * dy)} 0 0 0 0")
    for i in range(npart - 4):
        rows.append(f"{_f(xl - 2 * dx)} {_f(i * dy + 2 * dy)} 0 0 0 0")
    for i in range(mp - 4):
        rows.append(f"{_f(i * dx + 3 * dx)} {_f(yl - 2 * dy)} 0 1.0 0 0")
    for i in range(mp - 7):
        for j in range(npart - 6):
            rows.append(f"{_f(i * dx + 4 * dx)} {_f(4 * dy + dy * j)} 1 0 0 0")
    return _text(count, rows)


def water_drop_2d(dl=0.01, mult=1.0):
    """createParticlesWaterDrop2D (genParticlesFunctions.cpp:289-319): a disc
    of fluid particles, i^2 + j^2 < mp^2 on the lattice. The generator's
    count line is the constant 1005 (its mult = 1 count). (0.01, 1) gives
    the reference's inputs/input-drop.grd."""
    mp = int(18 * mult)
    dx = (mp * dl) / mp
    rows = []
    for i in range(-mp, mp):
        for j in range(-mp, mp):
            if i * i + j * j < mp * mp:
                rows.append(f"{_f(i * dx + 2 * mp * dx)} {_f(j * dx + 2 * mp * dx)} 1 0 0 0")
    return _text(1005, rows)


def dam_break_3d(dl=0.01, mult=1.0):
    """createParticlesDamBreak3D (genParticlesFunctions.cpp:455-558): the 3D
    dam break, a water block of mp/2 x 2 np x op in a tank with three
    dummy / wall layers. (0.01, 1) gives the reference's
    inputs/input-dam3d.grd."""
    mp = int(16 * mult)
    npart = int(8 * mult)
    op = int(8 * mult)
    xl = mp * dl
    dx = xl / mp
    dy = (npart * dl) / npart
    dz = (op * dl) / op
    count = op * (mp * npart) + op * (4 * mp + 10) + (4 * mp + 4) * (mp + 5) + \
        (op + 4) * (4 * mp + 14) + (4 * mp + 4) * (mp + 5) + (op + 6) * (4 * mp + 18) + \
        (4 * mp + 8) * (mp + 6)
    z0 = " 0.0 0.0 0.0 0.0"
    rows = []
    for k in range(op + 6):
        zk = dz * k - 3 * dz
        for i in range(2 * mp + 6):
            rows.append(f"{_f(i * dx)} 0.0 {_f(zk)} -1" + z0)
        for i in range(mp + 6):
            rows.append(f"0.0 {_f(i * dy + dy)} {_f(zk)} -1" + z0)
        for i in range(mp + 6):
            rows.append(f"{_f(2 * xl + 5 * dx)} {_f(i * dy + dy)} {_f(zk)} -1" + z0)
    for zf in (-3 * dz, (op + 2) * dz):
        for i in range(2 * mp + 4):
            for j in range(mp + 6):
                rows.append(f"{_f(i * dx + dx)} {_f(j * dy + dy)} {_f(zf)} -1" + z0)
    for k in range(op + 4):
        zk = dz * k - 2 * dz
        for i in range(2 * mp + 4):
            rows.append(f"{_f(i * dx + dx)} {_f(dy)} {_f(zk)} 0" + z0)
        for i in range(mp + 5):
            rows.append(f"{_f(dx)} {_f(i * dy + 2 * dy)} {_f(zk)} 0" + z0)
        for i in range(mp + 5):
            rows.append(f"{_f(2 * xl + 4 * dx)} {_f(i * dy + 2 * dy)} {_f(zk)} 0" + z0)
    for zf in (-2 * dz, (op + 1) * dz):
        for i in range(2 * mp + 2):
            for j in range(mp + 5):
                rows.append(f"{_f(i * dx + 2 * dx)} {_f(2 * dy + j * dy)} {_f(zf)} 0" + z0)
    for k in range(op):
        for i in range(2 * mp + 1):
            rows.append(f"{_f(i * dx + 2 * dx)} {_f(2 * dy)} {_f(dz * k)} 0" + z0)
        for i in range(mp + 4):
            rows.append(f"{_f(2 * dx)} {_f(i * dy + 3 * dy)} {_f(dz * k)} 0" + z0)
        for i in range(mp + 5):
            rows.append(f"{_f(2 * xl + 3 * dx)} {_f(i * dy + 2 * dy)} {_f(dz * k)} 0" + z0)
    for zf in (-dz, op * dz):
        for i in range(2 * mp + 2):
            for j in range(mp + 5):
                rows.append(f"{_f(i * dx + 2 * dx)} {_f(2 * dy + j * dy)} {_f(zf)} 0" + z0)
    for k in range(op):
        for i in range(mp // 2):
            for j in range(2 * npart):
                rows.append(f"{_f(i * dx + 3 * dx)} {_f(j * dy + 3 * dy)} {_f(dz * k)} 1" + z0)
    return _text(count, rows)


def case_from_text(text, dim, l0):
    """A Case from .grd text: the first `count` rows (main.cu:640-668 reads
    TP rows after the count line)."""
    from .cases import Case
    lines = text.split("\n")
    tp = int(lines[0].split()[0])
    rows = np.array([ln.split() for ln in lines[1:1 + tp]], dtype=np.float64)
    if rows.shape[0] != tp:
        raise ValueError("fewer rows than the count line says")
    if dim == 3:
        x, y, z, t, u, v, w, p = (rows[:, k] for k in range(8))
    else:
        x, y, t, u, v, p = (rows[:, k] for k in range(6))
        z = np.zeros(tp)
        w = np.zeros(tp)
    return Case(dim, l0, x.copy(), y.copy(), z.copy(), u.copy(), v.copy(), w.copy(), p.copy(),
                t.astype(np.int32))


INPUTS = {
    "dam": ("input-dam.grd", lambda: dam_break_2d(0.01, 1.0)),
    "dam3d": ("input-dam3d.grd", lambda: dam_break_3d(0.01, 1.0)),
    "shear": ("input-shear.grd", lambda: shear_cavity_2d(0.01, 1.0)),
    "drop": ("input-drop.grd", lambda: water_drop_2d(0.01, 1.0)),
    "oil": ("input-oil.grd", lambda: oil_spill_2d(0.004, 1.0)),
}


def main(argv=None):
    """python -m voxarmps_amd.generators NAME|all DIR: write the named inputs
    (the reference's file names) into DIR, at the spacing main.cu runs them
    with (0.01 m; 0.004 m for the compressible oil spill)."""
    import os
    import sys
    args = sys.argv[1:] if argv is None else argv
    if len(args) != 2 or (args[0] not in INPUTS and args[0] != "all"):
        print("usage: python -m voxarmps_amd.generators {%s|all} DIR" % "|".join(INPUTS))
        return 2
    os.makedirs(args[1], exist_ok=True)
    for key in (INPUTS if args[0] == "all" else [args[0]]):
        name, make = INPUTS[key]
        with open(os.path.join(args[1], name), "w") as f:
            f.write(make())
        print(os.path.join(args[1], name))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
