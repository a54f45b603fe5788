"""Curved DG cells (CPU): the plan's rules and the oracle's numpy restatement
oracle_py.dg_curved_matrix, pinned before it serves as the checker of the
curved-cell DG kernel (tests/test_gpu_dg_curved.py).

  * the cell rule is exact to degree 2p, the face rule to degree 2p;
  * with straight edge nodes (manifold 2) the restatement equals the affine
    DG oracle (oracle_dg.c, its own quadrature, traces by mapping points back)
    -- this pins the face parametrisation, normals, area elements, traces and
    penalty of the restatement;
  * on the sphere: symmetric, positive definite; A 1 = 0 on rows of cells
    without a boundary face; for P2/P3 (the FE space contains the physical
    linear functions: isoparametric P2 geometry) A x = 0 on those rows too;
  * the support points lie on the P2 map (edge midpoints on the sphere).
"""
import math

import numpy as np
import pytest

from conftest import rel_err


def _tet_moment(i, j, k):
    return math.factorial(i) * math.factorial(j) * math.factorial(k) / math.factorial(i + j + k + 3)


def _tri_moment(i, j):
    return math.factorial(i) * math.factorial(j) / math.factorial(i + j + 2)


def _op(smf, n, p, manifold, kind="dg", seed=3, center=(1.5, 0.5, 0.5)):
    return smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, seed), p, kind=kind, host_only=True, manifold=manifold,
                        manifold_center=center)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_curved_rules_exact(smf, p):
    op = _op(smf, 1, p, "spherical")
    pts, w = op.quadrature()
    for deg in range(2 * p + 1):
        for i in range(deg + 1):
            for j in range(deg + 1 - i):
                k = deg - i - j
                got = np.sum(w * pts[:, 0] ** i * pts[:, 1] ** j * pts[:, 2] ** k)
                assert abs(got - _tet_moment(i, j, k)) < 1e-14
    fp, fw = op.face_quadrature()
    for deg in range(2 * p + 1):
        for i in range(deg + 1):
            got = np.sum(fw * fp[:, 0] ** i * fp[:, 1] ** (deg - i))
            assert abs(got - _tri_moment(i, deg - i)) < 1e-14


def _curved(orc, op, p, straight, center=(1.5, 0.5, 0.5), **kw):
    pts, w = op.quadrature()
    fp, fw = op.face_quadrature()
    return orc.dg_curved_matrix(p, op.mesh.vertices, op.cell_vertices(), pts, w, fp, fw, straight=straight,
                                center=center, **kw)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_straight_restatement_equals_affine_oracle(smf, orc, p):
    op = _op(smf, 2 if p < 3 else 1, p, "straight")
    A = _curved(orc, op, p, straight=True)
    aff = orc.DGOracleProblem.from_operator(op)
    x = np.random.default_rng(p).uniform(-1, 1, op.m())
    assert rel_err(A @ x, aff.vmult(x)) < 1e-12
    # Helmholtz-like vector operator (mass 2.5, three components)
    opv = _op(smf, 1, p, "straight", kind="dg_helmholtz")
    Av = _curved(orc, opv, p, straight=True, ncomp=3, mass=2.5)
    affv = orc.DGOracleProblem.from_operator(opv, mass=2.5)
    xv = np.random.default_rng(p + 7).uniform(-1, 1, opv.m())
    assert rel_err(Av @ xv, affv.vmult(xv)) < 1e-12


@pytest.mark.parametrize("p", [1, 2, 3])
def test_curved_restatement_properties(smf, orc, p):
    op = _op(smf, 2, p, "spherical")
    A = _curved(orc, op, p, straight=False).toarray()
    assert np.max(np.abs(A - A.T)) < 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    # the sphere bends the cells: the operator differs from the straight one
    S = _curved(orc, op, p, straight=True).toarray()
    assert np.max(np.abs(A - S)) > 1e-3 * np.abs(S).max()
    C = op.cell_vertices()
    nl = op.info["dofs_per_cell"]
    keys = {}
    for k in range(C.shape[0]):
        for f in ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)):
            keys.setdefault(tuple(sorted(C[k, list(f)])), []).append(k)
    bdry = np.zeros(C.shape[0], bool)
    for cells in keys.values():
        if len(cells) == 1:
            bdry[cells[0]] = True
    rows = ~np.repeat(bdry, nl)
    assert rows.any()
    assert np.max(np.abs((A @ np.ones(op.m()))[rows])) < 1e-11 * np.abs(A).max()
    if p >= 2:
        X = op.dof_points()
        lin = 0.3 - 0.7 * X[:, 0] + 1.3 * X[:, 1] + 0.4 * X[:, 2]
        assert np.max(np.abs((A @ lin)[rows])) < 1e-10 * np.abs(A).max()


def test_support_points_on_the_map(smf, orc):
    op = _op(smf, 2, 2, "spherical")
    nodes = orc.curved_nodes(op.mesh.vertices, op.cell_vertices(), center=(1.5, 0.5, 0.5))
    X = op.dof_points().reshape(-1, 10, 3)
    # P2: the support points are the geometry nodes (vertices, then edge midpoints on the sphere)
    assert np.max(np.abs(X - nodes)) < 1e-13


def test_curved_dg_rejections(smf):
    m = smf.Mesh.kuhn_cube(2, 0.2, 1)
    with pytest.raises(smf.MFError, match="penalty_kind"):
        smf.Operator(m, 2, kind="dg", host_only=True, manifold="spherical", penalty_kind=1)
