"""Assembled system matrix of the CG operators (the matrix-based comparison
path: get_system_matrix, CG/cells_batched/application_hp.cc:1411-1438, and
the SpMV timing of CG/cells_batched/application.cc:1944-1988).  CPU only:
host_only operators give host-only matrices, compared entry by entry with
the oracle's dense matrices."""
import numpy as np
import pytest

from test_host_nitsche import cg_to_dg


def dense_of(M):
    ptr, col, val = M.to_host()
    n = M.m()
    A = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(ptr))
    A[rows, col] = val
    return A, ptr, col


def check_pattern(ptr, col, constrained):
    n = len(ptr) - 1
    assert ptr[0] == 0 and np.all(np.diff(ptr) >= 1)
    for r in range(n):
        c = col[ptr[r]:ptr[r + 1]]
        assert np.all(np.diff(c) > 0), f"row {r}: columns not strictly ascending"
        if constrained[r]:
            assert list(c) == [r]
        else:
            assert not constrained[c].any(), f"free row {r} holds a constrained column"
            assert r in c


@pytest.mark.parametrize("p", [1, 2, 3])
def test_matrix_equals_oracle_dense(smf, orc, p):
    op = smf.Operator(smf.Mesh.kuhn_cube(3, 0.2, 4), p, host_only=True)
    M = smf.SystemMatrix(op)
    A, ptr, col = dense_of(M)
    check_pattern(ptr, col, op.constrained().astype(bool))
    ref = orc.OracleProblem.from_operator(op).dense()
    assert np.max(np.abs(A - ref)) <= 1e-13 * np.abs(ref).max()
    # the pattern is the cell-coupling graph of the free DoFs: every
    # structurally coupled free pair is present, no other column is
    free = ~op.constrained().astype(bool)
    cd = op.cell_dofs()
    S = np.zeros_like(A, dtype=bool)
    for dofs in cd:
        S[np.ix_(dofs, dofs)] = True
    S[~free, :] = False
    S[:, ~free] = False
    S[np.arange(op.m()), np.arange(op.m())] = True
    nnz = np.zeros_like(S)
    nnz[np.repeat(np.arange(op.m()), np.diff(ptr)), col] = True
    assert np.array_equal(nnz, S)
    info = M.info
    assert info["nnz"] == S.sum() and info["n_rows"] == op.m()
    # row blocks of the product kernel: greedy runs of at most block_entries
    # entries and 256 rows (every row fits one block)
    cap, nb, e, rows = info["block_entries"], 1, 0, 0
    for ln in np.diff(ptr):
        assert 1 <= ln <= cap
        if e + ln > cap or rows == 256:
            nb, e, rows = nb + 1, 0, 0
        e, rows = e + ln, rows + 1
    assert info["n_row_blocks"] == nb
    w = 8
    assert info["algorithmic_bytes"] == pytest.approx(info["nnz"] * (4 + w) + op.m() * (8 + 2 * w))


def test_matrix_fp32_rounds_the_fp64_values(smf):
    mesh = smf.Mesh.kuhn_cube(3, 0.2, 6)
    A64, _, _ = dense_of(smf.SystemMatrix(smf.Operator(mesh, 2, host_only=True)))
    A32, _, _ = dense_of(smf.SystemMatrix(smf.Operator(mesh, 2, precision="f32", host_only=True)))
    assert np.array_equal(A32, A64.astype(np.float32).astype(np.float64))


def test_vector_matrix_is_block_diagonal(smf):
    mesh = smf.Mesh.kuhn_cube(2, 0.2, 8)
    s = smf.Operator(mesh, 2, host_only=True)
    v = smf.Operator(mesh, 2, kind="cg_vector", host_only=True)
    As, _, _ = dense_of(smf.SystemMatrix(s))
    Av, ptr, col = dense_of(smf.SystemMatrix(v))
    n = s.m()
    assert v.m() == 3 * n
    want = np.zeros((3 * n, 3 * n))
    for c in range(3):
        want[c * n:(c + 1) * n, c * n:(c + 1) * n] = As
    assert np.array_equal(Av, want)


@pytest.mark.parametrize("p", [1, 2])
def test_nitsche_matrix_equals_restricted_sipg(smf, orc, p):
    """Nitsche CG matrix = P^T A_DG P with P the injection of the CG space
    into the DG space and A_DG the DG oracle's matrix."""
    mesh = smf.Mesh.kuhn_cube(2, perturb=0.2, seed=5)
    cg = smf.Operator(mesh, p, kind="cg", boundary="nitsche", host_only=True)
    dg = smf.Operator(mesh, p, kind="dg", host_only=True)
    cols = cg_to_dg(cg, dg)
    Adg = orc.DGOracleProblem.from_operator(dg).dense()
    P = np.zeros((dg.m(), cg.m()))
    P[np.arange(dg.m()), cols] = 1.0
    want = P.T @ Adg @ P
    A, ptr, col = dense_of(smf.SystemMatrix(cg))
    check_pattern(ptr, col, np.zeros(cg.m(), bool))
    assert np.max(np.abs(A - want)) <= 1e-12 * np.abs(want).max()


def test_matrix_rejects_dg_and_host_product(smf):
    mesh = smf.Mesh.kuhn_cube(2)
    with pytest.raises(smf.MFError, match="CG operators only"):
        smf.SystemMatrix(smf.Operator(mesh, 1, kind="dg", host_only=True))
    M = smf.SystemMatrix(smf.Operator(mesh, 1, host_only=True))
    rc = smf.lib().mf_csr_vmult(M.handle, 16, 32, None)
    assert rc != 0 and b"host_only" in smf.lib().mf_last_error()
