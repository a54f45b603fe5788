"""Pinning the CPU oracle (CPU only).

Every check here is fixed by mathematics or by an independent computation,
never by re-typing the oracle's formulas:
  * basis: Kronecker property at the lattice nodes, partition of unity, and the
    closed barycentric forms of P1/P2/P3;
  * quadrature: exact monomial moments a!b!c!/(a+b+c+3)!;
  * integrals of the P2 basis: -|K|/20 (vertex) and |K|/5 (edge);
  * CG operator: golden fixtures from exact rational element matrices
    (tests/golden/make_golden.py), P1 closed form, symmetry and definiteness,
    linear consistency, A x^2 = -2 int phi_i, the energy identity
    u^T A u = int |grad u|^2, and the CSR matrix product.
"""
import math
import os

import numpy as np
import pytest

from conftest import rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
VX = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)


def nodes(p):
    """Node positions from the stated convention."""
    pts = [v for v in VX]
    for a, b in EDGES:
        for s in range(p - 1):
            pts.append(VX[a] + (s + 1) / p * (VX[b] - VX[a]))
    if p == 3:
        for f in range(4):
            pts.append(np.mean([VX[k] for k in range(4) if k != f], axis=0))
    return np.array(pts)


def bary(x):
    return np.array([1 - x.sum(), x[0], x[1], x[2]])


def closed_form(p, x):
    l = bary(x)
    if p == 1:
        return l
    if p == 2:
        return np.array([l[a] * (2 * l[a] - 1) for a in range(4)] + [4 * l[a] * l[b] for a, b in EDGES])
    vals = [0.5 * l[a] * (3 * l[a] - 1) * (3 * l[a] - 2) for a in range(4)]
    for a, b in EDGES:
        vals.append(4.5 * l[a] * l[b] * (3 * l[a] - 1))
        vals.append(4.5 * l[a] * l[b] * (3 * l[b] - 1))
    for f in range(4):
        o = [k for k in range(4) if k != f]
        vals.append(27 * l[o[0]] * l[o[1]] * l[o[2]])
    return np.array(vals)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_basis_kronecker_and_closed_form(orc, p):
    X = nodes(p)
    for j, x in enumerate(X):
        v, g = orc.basis(p, x)
        e = np.zeros(len(X))
        e[j] = 1
        assert np.allclose(v, e, atol=1e-13)
    rng = np.random.default_rng(p)
    for _ in range(20):
        x = rng.dirichlet(np.ones(4))[1:]
        v, g = orc.basis(p, x)
        assert abs(v.sum() - 1) < 1e-13
        assert np.allclose(g.sum(axis=0), 0, atol=1e-12)
        assert np.allclose(v, closed_form(p, x), atol=2e-15)
        # gradient against a central difference of the closed form
        h = 1e-6
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            fd = (closed_form(p, x + e) - closed_form(p, x - e)) / (2 * h)
            assert np.allclose(g[:, d], fd, atol=1e-7)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_quadrature_moments(orc, n):
    pts, w = orc.tet_rule(n)
    deg = 2 * n - 3
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                exact = math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
                val = np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
                assert abs(val - exact) < 1e-15 + 1e-13 * exact


def test_p2_basis_integrals(orc):
    pts, w = orc.tet_rule(4)
    tot = np.zeros(10)
    for x, wk in zip(pts, w):
        v, _ = orc.basis(2, x)
        tot += wk * v
    vol = 1 / 6
    assert np.allclose(tot[:4], -vol / 20, atol=1e-15)
    assert np.allclose(tot[4:], vol / 5, atol=1e-15)


def _match(points_a, points_b):
    """Permutation pi with points_a[pi[j]] == points_b[j]."""
    from scipy.spatial import cKDTree

    d, idx = cKDTree(points_a).query(points_b)
    assert np.max(d) < 1e-12
    assert len(set(idx.tolist())) == len(idx)
    return idx


def _oracle_from_fixture(smf, orc, fx, p, constrained=False):
    mesh = smf.Mesh(fx["vertices"], fx["cells"])
    op = smf.Operator(mesh, p, host_only=True)
    pr = orc.OracleProblem.from_operator(op)
    if not constrained:
        pr = orc.OracleProblem(pr.vertices, pr.cells, pr.cell_dofs, np.zeros(pr.n_dofs, np.uint8), p)
    perm = _match(op.dof_points(), fx["support_points"])  # fixture dof j -> op dof perm[j]
    return op, pr, perm


@pytest.mark.parametrize("p", [1, 2])
def test_golden_fixture(smf, orc, p):
    fx = np.load(os.path.join(GOLDEN, f"cg_p{p}_2x2x2.npz"))
    op, pr, perm = _oracle_from_fixture(smf, orc, fx, p)
    A = pr.dense()
    G = fx["matrix"]
    assert np.max(np.abs(A[np.ix_(perm, perm)] - G)) < 1e-13 * np.abs(G).max()
    # Dirichlet operator = free block + identity rows, matrix-free path
    op_pr = orc.OracleProblem.from_operator(op)
    con = op.constrained().astype(bool)
    x = np.random.default_rng(0).uniform(-1, 1, op.m())
    xf = np.where(con, 0.0, x)
    Gop = np.zeros_like(G)
    Gop[np.ix_(perm, perm)] = G
    ref = np.where(con, x, Gop @ xf)
    assert rel_err(op_pr.vmult(x), ref) < 1e-13


def test_p1_closed_form_single_tet(smf, orc):
    v = np.array([[0.1, 0.0, 0.0], [1.2, 0.1, 0.0], [0.2, 0.9, 0.1], [0.0, 0.3, 1.1]])
    mesh = smf.Mesh(v, np.array([[0, 1, 2, 3]], np.int32))
    op = smf.Operator(mesh, 1, host_only=True)
    pr = orc.OracleProblem.from_operator(op)
    pr = orc.OracleProblem(pr.vertices, pr.cells, pr.cell_dofs, np.zeros(4, np.uint8), 1)
    A = pr.dense()
    J = np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], axis=1)
    vol = abs(np.linalg.det(J)) / 6
    gl = np.zeros((4, 3))
    gl[1:] = np.linalg.inv(J)
    gl[0] = -gl[1:].sum(axis=0)
    K = vol * gl @ gl.T
    dofs = op.cell_dofs()[0]
    assert np.allclose(A[np.ix_(dofs, dofs)], K, atol=1e-14)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_symmetry_definiteness_and_csr(smf, orc, p):
    op = smf.Operator(smf.Mesh.kuhn_cube(3, 0.2, 3), p, host_only=True)
    pr = orc.OracleProblem.from_operator(op)
    A = pr.dense()
    assert np.max(np.abs(A - A.T)) < 1e-13 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0
    x = np.random.default_rng(1).uniform(-1, 1, op.m())
    assert rel_err(pr.vmult(x), A @ x) < 1e-13
    assert rel_err(pr.matrix_vmult(x), A @ x) < 1e-13
    # diagonal
    assert rel_err(pr.diagonal(), np.diag(A).copy()) < 1e-14


@pytest.mark.parametrize("p", [1, 2, 3])
def test_consistency_and_energy(smf, orc, p):
    """Neumann matrix (no constraints): A (linear) = 0 and A (x^2) = -2 int phi_i
    on rows of interior DoFs; u^T A u = int |grad u|^2 = 4/3 for u = x^2 (p >= 2)."""
    op = smf.Operator(smf.Mesh.kuhn_cube(4, 0.2, 9), p, host_only=True)
    pr0 = orc.OracleProblem.from_operator(op)
    pr = orc.OracleProblem(pr0.vertices, pr0.cells, pr0.cell_dofs, np.zeros(pr0.n_dofs, np.uint8), p)
    X = op.dof_points()
    interior = op.constrained() == 0
    lin = 0.3 + 1.5 * X[:, 0] - 0.7 * X[:, 1] + 2.0 * X[:, 2]
    assert np.max(np.abs(pr.vmult(lin)[interior])) < 1e-12
    if p >= 2:
        # int phi_i from the oracle's own rule (checked above against closed forms)
        pts, w = orc.tet_rule(4)
        vals = np.array([orc.basis(p, x)[0] for x in pts])
        ref_int = w @ vals  # per local node, reference cell
        intphi = np.zeros(op.m())
        V, C = pr.vertices, pr.cells
        for c in range(C.shape[0]):
            J = np.stack([V[C[c, k]] - V[C[c, 0]] for k in (1, 2, 3)], axis=1)
            intphi[pr.cell_dofs[c]] += abs(np.linalg.det(J)) * ref_int
        u = X[:, 0] ** 2
        assert np.max(np.abs(pr.vmult(u)[interior] + 2 * intphi[interior])) < 1e-12
        assert abs(u @ pr.vmult(u) - 4.0 / 3.0) < 1e-12


def test_planted_mistake_is_caught(smf, orc):
    """A transposed metric (J^-1 in place of J^-T) must fail the golden check:
    emulated by feeding the oracle a mirrored mesh of the fixture."""
    fx = np.load(os.path.join(GOLDEN, "cg_p2_2x2x2.npz"))
    op, pr, perm = _oracle_from_fixture(smf, orc, fx, 2)
    A = pr.dense()
    bad = orc.OracleProblem(pr.vertices * np.array([1.0, 1.0, 1.3]), pr.cells, pr.cell_dofs,
                            pr.constrained, 2).dense()
    G = fx["matrix"]
    assert np.max(np.abs(A[np.ix_(perm, perm)] - G)) < 1e-12
    assert np.max(np.abs(bad[np.ix_(perm, perm)] - G)) > 1e-3
