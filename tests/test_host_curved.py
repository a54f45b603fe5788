"""Curved cells: CG operators on a P2 geometry map whose edge nodes lie on a
spherical manifold (MappingFE(FE_SimplexP(2)) + SphericalManifold,
CG/cells_batched/application.cc:1613, 1668-1677) with the per-point metric
of the general D_e branch (:1361-1376).  CPU only (host execution of the
kernel schedule from the device tables).

What pins it without deal.II:
  * the operator's rule integrates monomials up to degree 2p exactly;
  * straight edge nodes reproduce the affine operator;
  * A (linear) = 0 on rows whose cells avoid the boundary (the integrand
    cof(J)^T grad phi is a polynomial of degree <= 2p there, so the rule is
    exact and the cell sums telescope);
  * the oracle's numpy restatement (its own bases and geodesic midpoints,
    the operator's rule) agrees to rounding;
  * symmetry, and the assembled matrix equals the operator."""
from math import factorial

import numpy as np
import pytest

from conftest import rel_err

CENTER = np.array([5.0, 0.0, 0.0])


def interior_rows(op):
    con = op.constrained().astype(bool)
    cd = op.cell_dofs()
    touch = np.zeros(op.m(), bool)
    touch[np.unique(cd[con[cd].any(axis=1)])] = True
    return ~touch


@pytest.mark.parametrize("p", [1, 2, 3])
def test_curved_rule_is_exact_to_degree_2p(smf, p):
    op = smf.Operator(smf.Mesh.kuhn_cube(1), p, host_only=True, manifold="spherical")
    pts, w = op.quadrature()
    assert abs(w.sum() - 1 / 6) < 1e-15
    for a in range(2 * p + 1):
        for b in range(2 * p + 1 - a):
            for c in range(2 * p + 1 - a - b):
                exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
                got = w @ (pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
                assert abs(got - exact) < 1e-15, (a, b, c)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_straight_edge_nodes_reproduce_affine(smf, p):
    mesh = smf.Mesh.kuhn_cube(4, 0.2, 3)
    a = smf.Operator(mesh, p, host_only=True)
    s = smf.Operator(mesh, p, host_only=True, manifold="straight")
    x = np.random.default_rng(p).uniform(-1, 1, a.m())
    assert rel_err(s.host_apply(x), a.host_apply(x)) < 1e-14
    assert rel_err(s.host_apply(mode=1), a.host_apply(mode=1)) < 1e-14
    assert np.allclose(s.dof_points(), a.dof_points(), atol=1e-15)


@pytest.mark.parametrize("p", [2, 3])
def test_curved_consistency(smf, p):
    op = smf.Operator(smf.Mesh.kuhn_cube(5, 0.2, 3), p, host_only=True, manifold="spherical")
    rows = interior_rows(op)
    assert rows.sum() > 100
    X = op.dof_points()
    lin = 0.3 + 1.5 * X[:, 0] - 0.7 * X[:, 1] + 2.0 * X[:, 2]
    assert np.max(np.abs(op.host_apply(lin)[rows])) < 1e-13
    # the cells are really curved: the affine support points differ
    aff = smf.Operator(smf.Mesh.kuhn_cube(5, 0.2, 3), p, host_only=True)
    assert np.max(np.abs(X - aff.dof_points())) > 1e-4


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("manifold", ["spherical", "straight"])
def test_curved_matches_oracle(smf, orc, p, manifold):
    mesh = smf.Mesh.kuhn_cube(4, 0.2, 3)
    op = smf.Operator(mesh, p, host_only=True, manifold=manifold)
    pts, w = op.quadrature()
    x = np.random.default_rng(p + 7).uniform(-1, 1, op.m())
    ref = orc.curved_vmult(p, mesh.vertices, op.cell_vertices(), op.cell_dofs(), op.constrained(), pts, w, x,
                           straight=manifold == "straight")
    assert rel_err(op.host_apply(x), ref) < 1e-13


def test_curved_edge_nodes_on_the_sphere(smf):
    """P2 edge DoFs sit at the geodesic midpoints: radius = mean of the end radii"""
    mesh = smf.Mesh.kuhn_cube(3, 0.2, 1)
    op = smf.Operator(mesh, 2, host_only=True, manifold="spherical", reorder=False)
    X = op.dof_points()
    cv, cd = op.cell_vertices(), op.cell_dofs()
    for k in range(0, len(cv), 7):
        for e, (a, b) in enumerate([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]):
            ra = np.linalg.norm(mesh.vertices[cv[k, a]] - CENTER)
            rb = np.linalg.norm(mesh.vertices[cv[k, b]] - CENTER)
            assert abs(np.linalg.norm(X[cd[k, 4 + e]] - CENTER) - 0.5 * (ra + rb)) < 1e-14


def test_curved_symmetry_and_matrix(smf):
    op = smf.Operator(smf.Mesh.kuhn_cube(2, 0.2, 5), 2, host_only=True, manifold="spherical")
    n = op.m()
    A = np.stack([op.host_apply(np.eye(n)[j]) for j in range(n)], axis=1)
    assert np.max(np.abs(A - A.T)) < 1e-14 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0
    ptr, col, val = smf.SystemMatrix(op).to_host()
    M = np.zeros((n, n))
    M[np.repeat(np.arange(n), np.diff(ptr)), col] = val
    assert np.max(np.abs(M - A)) < 1e-14 * np.abs(A).max()


def test_curved_options_rejected(smf):
    mesh = smf.Mesh.kuhn_cube(2)
    # DG operators take curved cells (tests/test_host_dg_curved.py), the Nitsche CG boundary does not
    with pytest.raises(smf.MFError, match="manifold"):
        smf.Operator(mesh, 2, boundary="nitsche", host_only=True, manifold="spherical")
    with pytest.raises(smf.MFError, match="manifold"):
        smf.Operator(mesh, 2, host_only=True, manifold=3)
