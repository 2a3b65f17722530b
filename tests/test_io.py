"""Mesh and restart files of UCNS3D: GRID.cel / GRID.vrt / GRID.bnd, text and
Fortran unformatted (BINIO), and RESTART.dat (src/io.f90:4157-4301,
13457-13639; src/grid_p.f90:1655-1760; src/bc_p.f90:30-140).

Pinned by: the files' layout read back with numpy and the reference's shape
rule (the repeated node ids of a GRID.cel line), boundary codes and triangle
rows of GRID.bnd, record markers of the unformatted files; round trips of
every mesh construction (periodic, bounded, bumped, mixed) that rebuild the
same PRESTORE bit for bit; a mesh written independently here (shuffled node
and cell numbering) that reads back with its volumes, face pairing and a
uniform state at rest as a steady state; RESTART.dat's byte layout."""
import dataclasses
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from ucns3d_amd.case import HEX, HYBRID, TET, Case, CaseSpec, read_restart, reduced, write_restart

ARRAYS = ["cell_shape", "ineigh", "ineighn", "bcode", "volume", "surf", "anglex", "angley", "ihexl", "invmat",
          "qpoints", "indicator", "integ_basis", "admis", "full"]

SPECS = {
    "hybrid-walls": lambda: CaseSpec("io-hyb", HYBRID, 6, "weno4", jitter=0.1, seed=21, bc=(1, 2, 4, 4, 3, 3)),
    "tet-periodic": lambda: CaseSpec("io-tet", TET, 4, "weno4", jitter=0.1, seed=3),
    "channel-bump": lambda: reduced("c4", 3),
    "blast": lambda: dataclasses.replace(reduced("c5", 8), scheme="weno3", ees=0),
    "hex-muscl-periodic": lambda: CaseSpec("io-hex", HEX, 6, "muscl", seed=2),
}


def shape_of_line(I):
    """element shape of a GRID.cel line's 8 node ids (src/grid_p.f90:1670-1684):
    1 hexahedron, 2 tetrahedron, 3 pyramid, 4 prism"""
    i1, i2, i3, i4, i5, i6, i7, i8 = I
    if i5 != i6 and i6 != i7 and i7 != i8 and i3 != i4:
        return 1
    if i3 == i4 and i5 == i6 == i7 == i8:
        return 2
    if i3 != i4 and i5 == i6 == i7 == i8:
        return 3
    if i3 == i4 and i5 != i6 and i7 == i8:
        return 4
    return 0


def records(path, size):
    """payloads of a Fortran unformatted sequential file whose records all
    have `size` bytes"""
    raw = np.fromfile(path, dtype=np.uint8)
    n = len(raw) // (size + 8)
    assert len(raw) == n * (size + 8)
    r = raw.reshape(n, size + 8)
    lead = r[:, :4].copy().view(np.int32).ravel()
    tail = r[:, -4:].copy().view(np.int32).ravel()
    assert np.all(lead == size) and np.all(tail == size)
    return r[:, 4:-4]


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("binary", [False, True])
def test_round_trip_rebuilds_the_same_prestore(name, binary, tmp_path):
    spec = SPECS[name]()
    a = Case(spec)
    a.write_grid(str(tmp_path), binary)
    b = Case(spec, grid_dir=str(tmp_path), binary=binary)
    assert (a.ncell, a.nnode) == (b.ncell, b.nnode)
    for n in ARRAYS:
        assert np.array_equal(a.array(n), b.array(n)), n


@pytest.mark.parametrize("name", ["hybrid-walls", "blast", "tet-periodic"])
def test_written_files_follow_the_format(name, tmp_path):
    case = Case(SPECS[name]())
    case.write_grid(str(tmp_path), False)
    cel = np.loadtxt(tmp_path / "GRID.cel", dtype=np.int64)
    vrt = np.loadtxt(tmp_path / "GRID.vrt")
    bnd = np.loadtxt(tmp_path / "GRID.bnd", dtype=np.int64)
    assert np.array_equal(cel[:, 0], np.arange(1, case.ncell + 1))
    assert np.array_equal(vrt[:, 0], np.arange(1, len(vrt) + 1))
    np.testing.assert_array_equal(vrt[:, 1:], case.array("xyz").reshape(-1, 3))
    # the reference's shape rule on the written node ids gives the mesh's shapes
    shapes = np.array([shape_of_line(r) for r in cel[:, 1:]])
    ours = case.array("cell_shape")
    code = {HEX: 1, TET: 2, 4: 3, 3: 4}  # preprocessor PYRAMID = 4, PRISM = 3
    assert np.array_equal(shapes, np.array([code[s] for s in ours]))
    # boundary faces: one row per boundary (or periodic) face record, with its code
    bcode = case.array("bcode").reshape(case.ncell, case.nface_max)
    live = bcode > 0
    assert len(bnd) == live.sum()
    for c in np.unique(bcode[live]):
        assert (bnd[:, 5] == c).sum() == (bcode == c).sum()
    # triangles repeat their third node; quadrilaterals have four distinct ones
    tri = bnd[:, 3] == bnd[:, 4]
    nv = (case.array("face_nodes").reshape(case.ncell, case.nface_max, 4) >= 0).sum(2)
    assert tri.sum() == ((nv == 3) & live).sum()
    assert np.all(np.diff(np.sort(bnd[~tri, 1:5], axis=1), axis=1) > 0)
    # the binary files hold the same numbers in records of 36, 28 and 24 bytes
    bdir = tmp_path / "bin"
    bdir.mkdir()
    case.write_grid(str(bdir), True)
    assert np.array_equal(records(bdir / "GRID.cel", 36).copy().view(np.int32), cel)
    vb = records(bdir / "GRID.vrt", 28)
    assert np.array_equal(vb[:, :4].copy().view(np.int32).ravel(), vrt[:, 0].astype(np.int32))
    assert np.array_equal(vb[:, 4:].copy().view(np.float64), vrt[:, 1:])
    assert np.array_equal(records(bdir / "GRID.bnd", 24).copy().view(np.int32), bnd)


def box_files(d, n, rng):
    """an n^3 hexahedral unit box written here, with node and cell ids in
    random order, slip walls (code 4) on every side"""
    g = np.linspace(0.0, 1.0, n + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.c_[X.ravel(), Y.ravel(), Z.ravel()]
    perm = rng.permutation(len(pts))          # new id of grid node k: perm[k] + 1
    nid = lambda i, j, k: perm[(i * (n + 1) + j) * (n + 1) + k] + 1  # noqa: E731
    inv = np.argsort(perm)
    with open(os.path.join(d, "GRID.vrt"), "w") as f:
        for m in range(len(pts)):
            x = pts[inv[m]]
            f.write(f"{m + 1} {float(x[0])!r} {float(x[1])!r} {float(x[2])!r}\n")
    cells = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                c = [nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                     nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)]
                cells.append(c)
    order = rng.permutation(len(cells))
    with open(os.path.join(d, "GRID.cel"), "w") as f:
        for m, ci in enumerate(order):
            f.write(f"{m + 1} " + " ".join(map(str, cells[ci])) + "\n")
    rows = []
    for a in range(n):
        for b in range(n):
            rows += [(nid(0, a, b), nid(0, a + 1, b), nid(0, a + 1, b + 1), nid(0, a, b + 1)),
                     (nid(n, a, b), nid(n, a + 1, b), nid(n, a + 1, b + 1), nid(n, a, b + 1)),
                     (nid(a, 0, b), nid(a + 1, 0, b), nid(a + 1, 0, b + 1), nid(a, 0, b + 1)),
                     (nid(a, n, b), nid(a + 1, n, b), nid(a + 1, n, b + 1), nid(a, n, b + 1)),
                     (nid(a, b, 0), nid(a + 1, b, 0), nid(a + 1, b + 1, 0), nid(a, b + 1, 0)),
                     (nid(a, b, n), nid(a + 1, b, n), nid(a + 1, b + 1, n), nid(a, b + 1, n))]
    with open(os.path.join(d, "GRID.bnd"), "w") as f:
        for m, r in enumerate(rows):
            f.write(f"{m + 1} {r[0]} {r[1]} {r[2]} {r[3]} 4\n")


def test_a_mesh_written_elsewhere_reads_back(tmp_path):
    n = 6
    box_files(str(tmp_path), n, np.random.default_rng(4))
    case = Case(CaseSpec("io-box", HEX, n, "weno3", bc=(4,) * 6, length=1.0), grid_dir=str(tmp_path))
    assert case.ncell == n ** 3
    vol = case.array("volume")
    np.testing.assert_allclose(vol, 1.0 / n ** 3, rtol=1e-12)
    ineigh = case.array("ineigh").reshape(case.ncell, case.nface_max)
    ineighn = case.array("ineighn").reshape(case.ncell, case.nface_max)
    bcode = case.array("bcode").reshape(case.ncell, case.nface_max)
    inner = ineigh >= 0
    i, f = np.nonzero(inner)
    assert np.all(ineigh[ineigh[i, f], ineighn[i, f]] == i)
    assert (bcode == 4).sum() == 6 * n * n and inner.sum() == 6 * n ** 3 - 6 * n * n
    # a state at rest is steady between slip walls
    u = np.tile([1.2, 0.0, 0.0, 0.0, 2.5], (case.ncell, 1))
    rhs, _ = Oracle(case).residual(u)
    assert np.abs(rhs).max() <= 1e-13


def test_restart_file_layout(tmp_path):
    rng = np.random.default_rng(9)
    u = rng.standard_normal((37, 5))
    p = str(tmp_path / "RESTART.dat")
    write_restart(p, 123, 0.625, u)
    raw = open(p, "rb").read()
    assert len(raw) == 12 + 40 * 37
    assert np.frombuffer(raw[:4], np.int32)[0] == 123
    assert np.frombuffer(raw[4:12], np.float64)[0] == 0.625
    assert np.array_equal(np.frombuffer(raw[12:], np.float64).reshape(37, 5), u)
    it, t, v = read_restart(p, 37)
    assert (it, t) == (123, 0.625) and np.array_equal(v, u)
    with pytest.raises(RuntimeError, match="bytes"):
        read_restart(p, 36)


def test_reader_errors(tmp_path):
    spec = CaseSpec("io-err", TET, 4, "weno4", jitter=0.1, seed=3)
    with pytest.raises(RuntimeError, match="cannot open"):
        Case(spec, grid_dir=str(tmp_path))
    Case(spec).write_grid(str(tmp_path), False)
    lines = open(tmp_path / "GRID.cel").read().splitlines()
    lines[5] = "6 1 2 3 3 4 5 6 7"  # I3 = I4 but I5..I8 distinct: no element shape
    open(tmp_path / "GRID.cel", "w").write("\n".join(lines) + "\n")
    with pytest.raises(RuntimeError, match="element shape"):
        Case(spec, grid_dir=str(tmp_path))
