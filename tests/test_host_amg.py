"""The oracle's numpy restatement of the smoothed-aggregation AMG coarse
solver (oracle_py.amg_hierarchy / amg_vcycle), pinned on the CPU before it
serves as the checker of the device AMG (tests/test_gpu_amg.py):
  * every non-isolated row belongs to exactly one aggregate (the tentative
    prolongator's columns are disjoint, unit-norm) and identity rows of
    constrained DoFs to none;
  * Galerkin coarse matrices are symmetric and positive definite;
  * the V-cycle is a symmetric preconditioner (x^T V y = y^T V x) under
    which CG reaches 1e-10 in few iterations, far fewer than unpreconditioned
    CG; with max_dense above the size it is the exact inverse."""
import numpy as np
import pytest


def p1_matrix(smf, n, seed=3):
    op = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, seed).scrambled(seed=1), 1, host_only=True)
    return op, smf.SystemMatrix(op).to_host()


def pcg(A, b, prec, tol=1e-10, maxit=500):
    x = np.zeros_like(b)
    r = b.copy()
    z = prec(r)
    p = z.copy()
    rz = r @ z
    for it in range(1, maxit + 1):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) <= tol * np.linalg.norm(b):
            return x, it
        z = prec(r)
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
    return x, maxit


@pytest.mark.parametrize("theta", [0.0, 0.05])
def test_amg_restatement_structure(smf, orc, theta):
    op, (ptr, col, val) = p1_matrix(smf, 12)
    levels = orc.amg_hierarchy(ptr, col, val, max_dense=200, theta=theta)
    assert len(levels) >= 2 and levels[-1]["A"].shape[0] <= 200
    con = op.constrained().astype(bool)
    # constrained rows get no coarse correction, every free row some
    support = np.asarray((abs(levels[0]["P"]) > 0).sum(axis=1)).ravel()
    assert np.all(support[con] == 0) and np.all(support[~con] > 0)
    for L in levels:
        A = L["A"].toarray()
        assert np.max(np.abs(A - A.T)) <= 1e-12 * np.abs(A).max()
        free = np.abs(A).sum(axis=1) > 0
        assert np.linalg.eigvalsh(A[np.ix_(free, free)]).min() > 0


@pytest.mark.parametrize("theta", [0.0, 0.05])
def test_amg_restatement_preconditions_cg(smf, orc, theta):
    op, (ptr, col, val) = p1_matrix(smf, 14)
    levels = orc.amg_hierarchy(ptr, col, val, max_dense=300, theta=theta)
    A = levels[0]["A"]
    b = np.random.default_rng(2).uniform(-1, 1, A.shape[0])
    V = lambda r: orc.amg_vcycle(levels, r)
    y, z = np.random.default_rng(3).uniform(-1, 1, (2, A.shape[0]))
    assert abs(z @ V(y) - y @ V(z)) <= 1e-10 * abs(z @ V(y))
    x, it = pcg(A, b, V)
    _, it_plain = pcg(A, b, lambda r: r, maxit=2000)
    assert np.linalg.norm(b - A @ x) <= 1.01e-10 * np.linalg.norm(b)
    assert it <= 25 and it < it_plain / 3
    # no coarsening: the V-cycle is the dense inverse
    exact = orc.amg_hierarchy(ptr, col, val, max_dense=10 ** 6)
    assert len(exact) == 1
    assert np.max(np.abs(A @ orc.amg_vcycle(exact, b) - b)) <= 1e-10 * np.abs(b).max()


def test_amg_restatement_stall_retries_then_refuses(smf, orc):
    # a threshold no coupling meets isolates every row: the level retries with
    # theta / 4 until it coarsens (0.9 -> 0.225 here), as if called with that theta
    op, (ptr, col, val) = p1_matrix(smf, 10)
    weak = orc.amg_hierarchy(ptr, col, val, max_dense=200, theta=0.9)
    half = orc.amg_hierarchy(ptr, col, val, max_dense=200, theta=0.225)
    assert len(weak) >= 2 and weak[-1]["A"].shape[0] <= 200
    assert [L["A"].shape[0] for L in weak] == [L["A"].shape[0] for L in half]
    A = weak[0]["A"]
    b = np.random.default_rng(5).uniform(-1, 1, A.shape[0])
    _, it = pcg(A, b, lambda r: orc.amg_vcycle(weak, r))
    assert it <= 30
    # a diagonal matrix cannot coarsen at any threshold: above the dense limit it is refused
    n = 300
    dptr, dcol, dval = np.arange(n + 1), np.arange(n), np.full(n, 2.0)
    with pytest.raises(RuntimeError, match="stalled"):
        orc.amg_hierarchy(dptr, dcol, dval, max_dense=100)
