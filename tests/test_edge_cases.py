"""Edge cases of the operator setup and application: a single cell, unused
vertices, an empty mesh, ragged mesh sizes (CPU: host execution against the
oracles; GPU: the device path through the C ABI)."""
import numpy as np
import pytest

from conftest import rel_err

V1 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)


def one_cell(smf, extra_vertex=False):
    v = np.vstack([V1, [[5.0, 5.0, 5.0]]]) if extra_vertex else V1
    return smf.Mesh(v, np.array([[0, 1, 2, 3]], np.int32))


@pytest.mark.parametrize("p", [1, 2, 3])
def test_single_cell_cg_is_identity(smf, p):
    """Every node of a lone tetrahedron is on the boundary: the CG operator is
    the identity (Dirichlet rows)."""
    op = smf.Operator(one_cell(smf), p, kind="cg", host_only=True)
    assert op.info["n_constrained"] == op.m()
    x = np.random.default_rng(p).uniform(-1, 1, op.m())
    np.testing.assert_array_equal(op.host_apply(x), x)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_single_cell_dg_and_nitsche_vs_oracle(smf, orc, p):
    dg = smf.Operator(one_cell(smf), p, kind="dg", host_only=True)
    x = np.random.default_rng(p).uniform(-1, 1, dg.m())
    assert rel_err(dg.host_apply(x), orc.DGOracleProblem.from_operator(dg).vmult(x)) < 1e-13
    # one cell: the Nitsche CG operator has the same matrix as SIPG up to the node numbering
    cg = smf.Operator(one_cell(smf), p, kind="cg", boundary="nitsche", host_only=True)
    A_dg = np.column_stack([dg.host_apply(e) for e in np.eye(dg.m())])
    A_cg = np.column_stack([cg.host_apply(e) for e in np.eye(cg.m())])
    assert np.allclose(np.sort(np.linalg.eigvalsh(A_dg)), np.sort(np.linalg.eigvalsh(A_cg)), rtol=1e-12, atol=1e-12)


def test_unused_vertex_is_ignored(smf, orc):
    for kind in ("cg", "dg"):
        a = smf.Operator(one_cell(smf), 2, kind=kind, host_only=True)
        b = smf.Operator(one_cell(smf, extra_vertex=True), 2, kind=kind, host_only=True)
        assert a.m() == b.m() == 10


@pytest.mark.parametrize("kind", ["cg", "dg", "cg_vector", "dg_vector", "dg_helmholtz"])
def test_empty_mesh(smf, kind):
    op = smf.Operator(smf.Mesh(V1, np.zeros((0, 4), np.int32)), 2, kind=kind, host_only=True)
    assert op.m() == 0 and op.host_apply(np.zeros(0)).shape == (0,)


def test_empty_mesh_matrix_curved_and_vtk(smf, tmp_path):
    empty = smf.Mesh(V1, np.zeros((0, 4), np.int32))
    M = smf.SystemMatrix(smf.Operator(empty, 2, host_only=True))
    ptr, col, val = M.to_host()
    assert M.info["n_rows"] == 0 and M.info["nnz"] == 0 and M.info["n_row_blocks"] == 0 and list(ptr) == [0]
    cur = smf.Operator(empty, 2, host_only=True, manifold="spherical")
    assert cur.m() == 0 and cur.host_apply(np.zeros(0)).shape == (0,)
    p = tmp_path / "tri.vtk"
    p.write_text("# vtk DataFile Version 3.0\nt\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS 3 double\n"
                 "0 0 0 1 0 0 0 1 0\nCELLS 1 4\n3 0 1 2\nCELL_TYPES 1\n5\n")
    m = smf.Mesh.read_vtk(str(p))
    assert m.cells.shape == (0, 4) and m.vertices.shape == (3, 3)


@pytest.mark.parametrize("n", [1, 3, 5, 7])
def test_ragged_sizes_host(smf, orc, n):
    op = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, n), 2, kind="cg", host_only=True)
    x = np.random.default_rng(n).uniform(-1, 1, op.m())
    assert rel_err(op.host_apply(x), orc.OracleProblem.from_operator(op).vmult(x)) < 1e-13


@pytest.mark.gpu
def test_gpu_single_cell_and_empty(gpu, orc):
    smf = gpu
    for kind in ("cg", "dg"):
        op = smf.Operator(one_cell(smf), 2, kind=kind)
        x = np.random.default_rng(1).uniform(-1, 1, op.m())
        y = op.initialize_dof_vector()
        op.vmult(y, smf.DeviceVector.from_numpy(x))
        assert rel_err(y.numpy(), op.host_apply(x)) < 1e-13
    for kind in ("cg", "dg"):
        op = smf.Operator(smf.Mesh(V1, np.zeros((0, 4), np.int32)), 2, kind=kind)
        assert op.m() == 0
        op.vmult(op.initialize_dof_vector(), op.initialize_dof_vector())
        op.compute_inverse_diagonal(op.initialize_dof_vector())
    # the assembled matrix and curved cells on one cell and on no cell
    for mesh in (one_cell(smf), smf.Mesh(V1, np.zeros((0, 4), np.int32))):
        for manifold in ("none", "spherical"):
            op = smf.Operator(mesh, 2, manifold=manifold)
            x = np.random.default_rng(2).uniform(-1, 1, op.m())
            y = op.initialize_dof_vector()
            op.vmult(y, smf.DeviceVector.from_numpy(x))
            assert op.m() == 0 or rel_err(y.numpy(), op.host_apply(x)) < 1e-13
            M = smf.SystemMatrix(op)
            z = op.initialize_dof_vector()
            M.vmult(z, smf.DeviceVector.from_numpy(x))
            assert op.m() == 0 or rel_err(z.numpy(), op.host_apply(x)) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 3, 5, 7, 13])
def test_gpu_ragged_sizes(gpu, orc, n):
    """Mesh sizes that fill no patch, wave or workgroup evenly."""
    smf = gpu
    print(f"\ncase: ragged n={n}", flush=True)
    for kind, pr in (("cg", orc.OracleProblem), ("dg", orc.DGOracleProblem)):
        op = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, n), 2, kind=kind)
        x = np.random.default_rng(n).uniform(-1, 1, op.m())
        y = op.initialize_dof_vector()
        op.vmult(y, smf.DeviceVector.from_numpy(x))
        assert rel_err(y.numpy(), pr.from_operator(op).vmult(x)) < 1e-12
    # curved cells and the assembled matrix at the same ragged sizes
    mesh = smf.Mesh.kuhn_cube(n, 0.2, n)
    op = smf.Operator(mesh, 2, manifold="spherical")
    pts, w = op.quadrature()
    x = np.random.default_rng(n + 1).uniform(-1, 1, op.m())
    src = smf.DeviceVector.from_numpy(x)
    ref = orc.curved_vmult(2, mesh.vertices, op.cell_vertices(), op.cell_dofs(), op.constrained(), pts, w, x)
    y = op.initialize_dof_vector()
    op.vmult(y, src)
    assert rel_err(y.numpy(), ref) < 1e-12
    z = op.initialize_dof_vector()
    smf.SystemMatrix(op).vmult(z, src)
    assert rel_err(z.numpy(), ref) < 1e-12
