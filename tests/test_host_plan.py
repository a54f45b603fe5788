"""The library's host plan (DoF numbering, patches, colourings, patch-local
tables, geometry) against the oracle, on the CPU: mf_debug_host_apply runs
the kernels' patch/colour schedule on the host from the very tables the GPU
reads."""
import numpy as np
import pytest

from conftest import rel_err

TOL = {"f64": 1e-13, "f32": 2e-6}


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("reorder", [True, False])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_host_schedule_matches_oracle(smf, orc, p, reorder, precision):
    op = smf.Operator(smf.Mesh.kuhn_cube(6, 0.2, 1), p, precision=precision, reorder=reorder,
                      host_only=True)
    pr = orc.OracleProblem.from_operator(op)
    x = np.random.default_rng(2).uniform(-1, 1, op.m())
    assert rel_err(op.host_apply(x).astype(float), pr.vmult(x)) < TOL[precision]
    assert rel_err(op.host_apply(None, mode=1).astype(float), pr.diagonal()) < TOL[precision]


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("n", [1, 3, 7])
def test_dof_counts_and_constraints(smf, p, n):
    op = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, 4), p, host_only=True)
    assert op.m() == (p * n + 1) ** 3
    assert op.info["n_constrained"] == (p * n + 1) ** 3 - max(p * n - 1, 0) ** 3
    X = op.dof_points()
    con = op.constrained().astype(bool)
    on_bdry = np.any((np.abs(X) < 1e-14) | (np.abs(X - 1) < 1e-14), axis=1)
    assert np.array_equal(con, on_bdry)
    # constrained DoFs are numbered last
    assert np.all(np.diff(con.astype(int)) >= 0)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_support_points_consistent(smf, orc, p):
    """Every cell maps its reference nodes onto the support points of its DoFs:
    a shared DoF has one position in all its cells (numbering conformity)."""
    mesh = smf.Mesh.kuhn_cube(3, 0.2, 8)
    op = smf.Operator(mesh, p, host_only=True)
    perm = op.cell_order()
    cd = op.cell_dofs()
    X = op.dof_points()
    from test_oracle_cg import nodes
    ref = nodes(p)
    for k, c in enumerate(perm):
        v = mesh.vertices[mesh.cells[c]]
        J = np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], axis=1)
        phys = v[0] + ref @ J.T
        assert np.allclose(X[cd[k]], phys, atol=1e-14)
    assert len(np.unique(cd)) == op.m()


def test_cg_p2_16cube_sizes(smf):
    op = smf.Operator(smf.Mesh.kuhn_cube(16, 0.2, 7), 2, host_only=True)
    assert op.info["n_cells"] == 24576 and op.m() == 35937
    assert op.info["n_constrained"] == 6146


def test_perturbation_is_seeded_and_bounded(smf):
    a = smf.Mesh.kuhn_cube(5, 0.2, 11)
    b = smf.Mesh.kuhn_cube(5, 0.2, 11)
    c = smf.Mesh.kuhn_cube(5, 0.2, 12)
    assert np.array_equal(a.vertices, b.vertices) and not np.array_equal(a.vertices, c.vertices)
    g = smf.Mesh.kuhn_cube(5, 0.0, 11).vertices
    d = np.abs(a.vertices - g)
    assert d.max() <= 0.2 / 5 + 1e-15 and d.max() > 0.15 / 5
    bd = np.any((g == 0) | (g == 1), axis=1)
    assert np.all(d[bd] == 0)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_cg_vector_host_matches_oracle(smf, orc, p, precision):
    """Components-batched vector Laplacian = the scalar operator on each
    component block (row c*n + d) (CG/components_batched/application.cc:1024-1216)."""
    op = smf.Operator(smf.Mesh.kuhn_cube(4, 0.2, 5), p, kind="cg_vector", precision=precision, host_only=True)
    n = op.n_scalar_dofs()
    assert op.m() == 3 * n
    pr = orc.OracleProblem.from_operator(op)
    x = np.random.default_rng(9).uniform(-1, 1, op.m())
    ref = np.concatenate([pr.vmult(x[c * n:(c + 1) * n]) for c in range(3)])
    assert rel_err(op.host_apply(x).astype(float), ref) < TOL[precision]
    d = np.tile(pr.diagonal(), 3)
    assert rel_err(op.host_apply(None, mode=1).astype(float), d) < TOL[precision]
