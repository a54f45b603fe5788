"""External meshes: the legacy VTK reader (GridIn::read_vtk in
CG/cells_batched/application.cc:1699-1756) and operators on meshes in an
unstructured input order, with and without the hierarchical reordering
(the reference's reorder_grid).  CPU only."""
import os

import numpy as np
import pytest

from conftest import rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_vtk_fixture_mixed_cells(smf):
    """Hand-written file: a FIELD block, float points, two tetrahedra (the
    second negatively oriented), a boundary triangle and cell data."""
    m = smf.Mesh.read_vtk(os.path.join(GOLDEN, "two_tets_mixed.vtk"))
    assert m.vertices.shape == (5, 3) and m.cells.shape == (2, 4)
    assert np.array_equal(m.vertices[4], [1, 1, 1])
    assert m.cells.tolist() == [[0, 1, 2, 3], [1, 2, 3, 4]]
    assert m.n_flipped == 1
    op = smf.Operator(m, 1, host_only=True, reorder=False)
    assert op.m() == 5
    vol = []
    for c in op.cell_vertices():
        J = (m.vertices[c[1:]] - m.vertices[c[0]]).T
        vol.append(np.linalg.det(J) / 6)
    assert sorted(vol) == pytest.approx([1 / 6, 1 / 3])


@pytest.mark.parametrize("layout", ["classic", "vtk5"])
def test_vtk_round_trip(smf, tmp_path, layout):
    m = smf.Mesh.kuhn_cube(3, 0.2, 4).scrambled(seed=2)
    path = str(tmp_path / f"mesh_{layout}.vtk")
    m.write_vtk(path, layout=layout)
    r = smf.Mesh.read_vtk(path)
    assert np.array_equal(r.vertices, m.vertices)
    assert np.array_equal(r.cells, m.cells)
    assert r.n_flipped == 0


def test_vtk_metadata_block_is_skipped(smf, tmp_path):
    m = smf.Mesh.kuhn_cube(1, 0.0, 0)
    path = str(tmp_path / "meta.vtk")
    m.write_vtk(path, layout="vtk5")
    text = open(path).read().replace("CELLS", "METADATA\nINFORMATION 0\n\nCELLS", 1)
    open(path, "w").write(text)
    r = smf.Mesh.read_vtk(path)
    assert np.array_equal(r.cells, m.cells)


def test_vtk_errors(smf, tmp_path):
    with pytest.raises(smf.MFError, match="cannot open"):
        smf.Mesh.read_vtk(str(tmp_path / "missing.vtk"))
    hexa = ("# vtk DataFile Version 3.0\nhex\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS 8 double\n"
            + "\n".join(f"{i & 1} {(i >> 1) & 1} {i >> 2}" for i in range(8))
            + "\nCELLS 1 9\n8 0 1 3 2 4 5 7 6\nCELL_TYPES 1\n12\n")
    p = tmp_path / "hex.vtk"
    p.write_text(hexa)
    with pytest.raises(smf.MFError, match="only tetrahedra"):
        smf.Mesh.read_vtk(str(p))
    p.write_text("# vtk DataFile Version 3.0\nb\nBINARY\nDATASET UNSTRUCTURED_GRID\n")
    with pytest.raises(smf.MFError, match="ASCII"):
        smf.Mesh.read_vtk(str(p))
    p.write_text("# vtk DataFile Version 3.0\nt\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS 4 double\n"
                 "0 0 0 1 0 0 0 1 0 0 0 1\nCELLS 1 5\n4 0 1 2 9\nCELL_TYPES 1\n10\n")
    with pytest.raises(smf.MFError, match="missing point"):
        smf.Mesh.read_vtk(str(p))


_VTK5_HEAD = ("# vtk DataFile Version 5.1\nt\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS 5 double\n"
              "0 0 0 1 0 0 0 1 0 0 0 1 1 1 1\n")


@pytest.mark.parametrize("offsets,conn,types", [
    # non-monotonic: skips a point cell, then a tetrahedron whose range would run past CONNECTIVITY
    ("0 50 54 4", "0 1 2 3", "1 10 10"),
    # an offset below zero
    ("0 -4 4", "0 1 2 3", "10 10"),
    # an offset beyond CONNECTIVITY in the middle
    ("0 9 4", "0 1 2 3", "10 10"),
])
def test_vtk5_offsets_validated(smf, tmp_path, offsets, conn, types):
    no = len(offsets.split())
    text = (_VTK5_HEAD + f"CELLS {no} {len(conn.split())}\nOFFSETS vtktypeint64\n{offsets}\n"
            f"CONNECTIVITY vtktypeint64\n{conn}\nCELL_TYPES {no - 1}\n{types}\n")
    p = tmp_path / "bad_offsets.vtk"
    p.write_text(text)
    with pytest.raises(smf.MFError, match="OFFSETS"):
        smf.Mesh.read_vtk(str(p))


def test_vtk_negative_counts_rejected(smf, tmp_path):
    p = tmp_path / "neg.vtk"
    p.write_text(_VTK5_HEAD + "CELLS 2 4\nOFFSETS vtktypeint64\n0 4\nCONNECTIVITY vtktypeint64\n0 1 2 3\n"
                 "CELL_TYPES -1\n")
    with pytest.raises(smf.MFError, match="negative CELL_TYPES"):
        smf.Mesh.read_vtk(str(p))
    p.write_text(_VTK5_HEAD + "FIELD f -2\n")
    with pytest.raises(smf.MFError, match="negative FIELD"):
        smf.Mesh.read_vtk(str(p))
    p.write_text(_VTK5_HEAD + "FIELD f 1\nx -1 3 double\n")
    with pytest.raises(smf.MFError, match="negative FIELD"):
        smf.Mesh.read_vtk(str(p))


def _by_point(op, y):
    """values of a scalar CG vector keyed by the rounded support points"""
    X = op.dof_points()
    key = np.round(X / 1e-9).astype(np.int64)
    order = np.lexsort(key.T[::-1])
    return key[order], y[order]


@pytest.mark.parametrize("kind", ["cg", "dg"])
@pytest.mark.parametrize("reorder", [True, False])
def test_scrambled_mesh_gives_the_same_operator(smf, orc, kind, reorder):
    """The operator of a mesh in an unstructured order (random vertex and
    cell order, rotated local vertices) equals that of the original mesh up
    to the DoF numbering; host execution of the kernel schedule against the
    oracle on the scrambled mesh."""
    base = smf.Mesh.kuhn_cube(4, 0.2, 3)
    scr = base.scrambled(seed=5)
    f = lambda X: np.sin(3 * X[:, 0]) + X[:, 1] * X[:, 2] - 0.5 * X[:, 2]
    ys = []
    for mesh in (base, scr):
        op = smf.Operator(mesh, 2, kind=kind, host_only=True, reorder=reorder)
        x = f(op.dof_points())
        pr = (orc.OracleProblem if kind == "cg" else orc.DGOracleProblem).from_operator(op)
        y = pr.vmult(x)
        assert rel_err(op.host_apply(x), y) < 1e-12
        ys.append((op, y))
    if kind == "cg":
        (k0, y0), (k1, y1) = _by_point(*ys[0]), _by_point(*ys[1])
        assert np.array_equal(k0, k1)
        assert rel_err(y1, y0) < 1e-12
    else:
        # DG nodes are per cell: compare the multiset of (point, value) pairs
        # through the sum over coincident nodes
        sums = []
        for op, y in ys:
            key = np.round(op.dof_points() / 1e-9).astype(np.int64)
            u, inv = np.unique(key, axis=0, return_inverse=True)
            s = np.zeros(len(u))
            np.add.at(s, inv.ravel(), y)
            sums.append((u, s))
        assert np.array_equal(sums[0][0], sums[1][0])
        assert rel_err(sums[1][1], sums[0][1]) < 1e-11
