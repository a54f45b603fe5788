"""Pinning the DG SIPG oracle (CPU only).

Checks fixed by mathematics, independent of the oracle's code path:
  * symmetry and positive definiteness (SIPG with the reference penalty);
  * consistency: A 1 = 0 and A (linear) = 0 on rows of cells without a
    boundary face;
  * a(1,1) = sum over boundary faces of 2 sigma_B |F| (mirror principle),
    with sigma_B = (p+1)(p+3)/3 tau_K computed here from numpy geometry;
  * for continuous P1 functions vanishing on the boundary the SIPG form equals
    the CG Poisson form (jumps vanish), cross-checked against the CG oracle;
  * tau = (sum_bdry |F| + 1/2 sum_int |F|)/|K| from numpy geometry;
  * the Helmholtz mass term: a_mass(1, v) rows sum to mass * |Omega|.
"""
import numpy as np
import pytest

from conftest import rel_err

FACES = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]


def numpy_tau(V, C):
    keys = {}
    for c in range(C.shape[0]):
        for f in FACES:
            k = tuple(sorted(C[c, list(f)]))
            keys.setdefault(k, []).append(c)
    tau = np.zeros(C.shape[0])
    bfaces = []
    for k, cells in keys.items():
        a, b, c3 = V[list(k)]
        area = 0.5 * np.linalg.norm(np.cross(b - a, c3 - a))
        for c in cells:
            tau[c] += area * (1.0 if len(cells) == 1 else 0.5)
        if len(cells) == 1:
            bfaces.append((cells[0], area))
    vol = np.array([abs(np.linalg.det(np.stack([V[C[c, t]] - V[C[c, 0]] for t in (1, 2, 3)], 1))) / 6
                    for c in range(C.shape[0])])
    return tau / vol, bfaces


def _dg(smf, orc, n, p, kind="dg", mass=0.0, seed=3):
    op = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, seed), p, kind=kind, host_only=True,
                      mass_coefficient=1.0)
    return op, orc.DGOracleProblem.from_operator(op, mass=mass)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_symmetric_positive_definite(smf, orc, p):
    op, pr = _dg(smf, orc, 1 if p == 3 else 2, p)
    A = pr.dense()
    assert np.max(np.abs(A - A.T)) < 1e-13 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0


@pytest.mark.parametrize("p", [1, 2, 3])
def test_consistency_away_from_boundary(smf, orc, p):
    op, pr = _dg(smf, orc, 4, p)
    V, C = pr.vertices, pr.cells
    tau, bfaces = numpy_tau(V, C)
    has_bdry = np.zeros(C.shape[0], bool)
    for c, _ in bfaces:
        has_bdry[c] = True
    X = op.dof_points()
    rows = ~np.repeat(has_bdry, pr.n_loc)
    ones = np.ones(op.m())
    lin = 0.4 - 1.1 * X[:, 0] + 0.8 * X[:, 1] + 1.7 * X[:, 2]
    assert np.max(np.abs(pr.vmult(ones)[rows])) < 1e-11
    assert np.max(np.abs(pr.vmult(lin)[rows])) < 1e-11
    # a(1,1) = sum_B 2 sigma_B |F|
    factor = (p + 1) * (p + 3) / 3
    expect = sum(2 * factor * tau[c] * area for c, area in bfaces)
    assert abs(ones @ pr.vmult(ones) - expect) < 1e-10 * expect


def test_tau_matches_numpy_geometry(smf, orc):
    op, pr = _dg(smf, orc, 3, 1)
    tau, _ = numpy_tau(pr.vertices, pr.cells)
    assert rel_err(pr.tau(), tau) < 1e-13


def test_sipg_equals_cg_on_continuous_p1(smf, orc):
    mesh = smf.Mesh.kuhn_cube(4, 0.2, 5)
    cg = smf.Operator(mesh, 1, kind="cg", host_only=True)
    dg = smf.Operator(mesh, 1, kind="dg", host_only=True)
    cpr = orc.OracleProblem.from_operator(cg)
    dpr = orc.DGOracleProblem.from_operator(dg)
    from scipy.spatial import cKDTree

    d, idx = cKDTree(cg.dof_points()).query(dg.dof_points())  # DG node -> CG DoF
    assert d.max() < 1e-12
    rng = np.random.default_rng(0)
    con = cg.constrained().astype(bool)
    u = np.where(con, 0.0, rng.uniform(-1, 1, cg.m()))
    v = np.where(con, 0.0, rng.uniform(-1, 1, cg.m()))
    a_cg = v @ cpr.vmult(u)
    a_dg = v[idx] @ dpr.vmult(u[idx])
    assert abs(a_cg - a_dg) < 1e-12 * abs(a_cg)


def test_helmholtz_mass_rows(smf, orc):
    op, pr = _dg(smf, orc, 3, 2, kind="dg_helmholtz", mass=2.5)
    lap0 = orc.DGOracleProblem(pr.vertices, pr.cells, 2, 3, mass=2.5, laplace=0.0)
    ones = np.ones(op.m())
    # sum of all rows of the mass part = mass * |Omega| per component
    assert abs(ones @ lap0.vmult(ones) - 3 * 2.5) < 1e-12
