"""CG operator with a Nitsche (weak Dirichlet) boundary, the CG level under a
DG operator in the hybrid multigrid (application_phc.cc:1424-1456,
1798-1832).  For continuous u, v the interior-face terms of SIPG vanish, so
the operator equals P^T A_DG P with P the injection of the CG space into the
DG space; A_DG comes from the DG oracle (pinned separately).  CPU only: the
host execution of the kernel schedule plus the boundary rows."""
import numpy as np
import pytest


def cg_to_dg(cg_op, dg_op):
    """Injection P (rows: DG rows in the operator's layout, one 1 per row)
    matching support points; returns the CG column of every DG row."""
    pts_cg = cg_op.dof_points()
    h = 1e-7
    key = {tuple(np.round(x / h).astype(np.int64)): i for i, x in enumerate(pts_cg)}
    pts_dg = dg_op.dof_points()
    cols = np.empty(len(pts_dg), np.int64)
    for r, x in enumerate(pts_dg):
        k = tuple(np.round(x / h).astype(np.int64))
        if k not in key:
            raise AssertionError("DG node without a CG node at the same point")
        cols[r] = key[k]
    return cols


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_nitsche_equals_restricted_sipg(smf, orc, degree):
    mesh = smf.Mesh.kuhn_cube(3, perturb=0.2, seed=5)
    cg = smf.Operator(mesh, degree, kind="cg", boundary="nitsche", host_only=True)
    dg = smf.Operator(mesh, degree, kind="dg", host_only=True)
    assert not cg.constrained().any()
    cols = cg_to_dg(cg, dg)
    dgo = orc.DGOracleProblem.from_operator(dg)
    rng = np.random.default_rng(degree)
    x = rng.uniform(-1, 1, cg.m())
    y = cg.host_apply(x)
    want = np.zeros(cg.m())
    np.add.at(want, cols, dgo.vmult(x[cols]))
    assert np.max(np.abs(y - want)) <= 1e-12 * np.max(np.abs(want))


def test_nitsche_diagonal_and_spd(smf, orc):
    mesh = smf.Mesh.kuhn_cube(2, perturb=0.2, seed=3)
    cg = smf.Operator(mesh, 2, kind="cg", boundary="nitsche", host_only=True)
    n = cg.m()
    A = np.column_stack([cg.host_apply(e) for e in np.eye(n)])
    np.testing.assert_allclose(A, A.T, atol=1e-12 * np.abs(A).max())
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    np.testing.assert_allclose(cg.host_apply(mode=1), np.diag(A), rtol=1e-13)
    # a constant is not in the kernel (boundary terms see it), unlike pure Neumann
    assert np.linalg.norm(cg.host_apply(np.ones(n))) > 1e-3


def test_nitsche_rejected_for_other_kinds(smf):
    mesh = smf.Mesh.kuhn_cube(2)
    with pytest.raises(smf.MFError, match="Nitsche"):
        smf.Operator(mesh, 2, kind="dg", boundary="nitsche", host_only=True)
    with pytest.raises(smf.MFError, match="Nitsche"):
        smf.Operator(mesh, 2, kind="cg_vector", boundary="nitsche", host_only=True)


def test_nitsche_penalty_variant(smf, orc):
    mesh = smf.Mesh.kuhn_cube(3, perturb=0.2, seed=5)
    cg = smf.Operator(mesh, 2, kind="cg", boundary="nitsche", penalty_kind=1, host_only=True)
    dg = smf.Operator(mesh, 2, kind="dg", host_only=True)
    cols = cg_to_dg(cg, dg)
    dgo = orc.DGOracleProblem.from_operator(dg, penalty_kind=1)
    x = np.random.default_rng(9).uniform(-1, 1, cg.m())
    want = np.zeros(cg.m())
    np.add.at(want, cols, dgo.vmult(x[cols]))
    y = cg.host_apply(x)
    assert np.max(np.abs(y - want)) <= 1e-12 * np.max(np.abs(want))
