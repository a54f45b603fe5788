"""GPU: mf_solve_cg_noalloc (work memory in the library's static device array,
no allocation) against mf_solve_cg (work memory allocated per call).  Same
kernels in the same order: the solutions and statistics agree bitwise.  Two
calls in a row, the second on a smaller operator (the array keeps the first
call's contents), then the first again; an operator above the array's limit
is refused with a clear error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(smf, A, x, b, mg, rel_tol=1e-10, max_iter=100, flexible=False):
    # through the C ABI (mf_solve_cg_noalloc)
    return smf.solve_cg_noalloc(A, x, b, mg, rel_tol=rel_tol, max_iter=max_iter, flexible=flexible)


def _case(smf, n, seed):
    A = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, seed), 2, precision="f64")
    mg = smf.Multigrid(A)
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(A, "grf", seed=seed))
    return A, mg, b


def test_noalloc_solve_matches_allocating_solve(gpu):
    smf = gpu
    print("\ncase: mf_solve_cg_noalloc vs mf_solve_cg, CG P2 n=12 then n=6, then n=12 unpreconditioned",
          flush=True)
    big = _case(smf, 12, 3)
    small = _case(smf, 6, 4)
    for (A, mg, b), flexible in ((big, False), (small, True), (big, True)):
        x1, x2 = A.initialize_dof_vector(), A.initialize_dof_vector()
        s1 = smf.solve_cg(A, x1, b, mg, rel_tol=1e-10, flexible=flexible)
        s2 = _noalloc(smf, A, x2, b, mg, flexible=flexible)
        assert s1["converged"] and s2 == s1
        assert np.array_equal(x1.numpy(), x2.numpy())
    A, _, b = small
    x1, x2 = A.initialize_dof_vector(), A.initialize_dof_vector()
    s1 = smf.solve_cg(A, x1, b, None, rel_tol=1e-6, max_iter=500)
    s2 = _noalloc(smf, A, x2, b, None, rel_tol=1e-6, max_iter=500)
    assert s2 == s1 and np.array_equal(x1.numpy(), x2.numpy())


def test_noalloc_refuses_large_operator(gpu):
    smf = gpu
    A = smf.Operator(smf.Mesh.kuhn_cube(32, 0.2, 1), 2, kind="dg", precision="f64")
    assert A.m() > 1 << 20
    x = A.initialize_dof_vector()
    with pytest.raises(smf.MFError, match="at most 1048576"):
        _noalloc(smf, A, x, x, None)
