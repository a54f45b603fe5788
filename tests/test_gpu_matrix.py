"""GPU parity of the matrix-based comparison path: the device CSR product of
the assembled system matrix (get_system_matrix,
CG/cells_batched/application_hp.cc:1411-1438; SpMV timing,
CG/cells_batched/application.cc:1944-1988) against the CPU oracle and
against the matrix-free vmult of the same operator.

Bar: FP64 1e-12 and FP32 1e-5 relative max-norm error against the FP64
oracle; bitwise-identical repeated products.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}


def products(smf, op, M, x):
    src = smf.DeviceVector.from_numpy(x.astype(op.dtype))
    y_mf = op.initialize_dof_vector()
    y_mb = op.initialize_dof_vector()
    op.vmult(y_mf, src)
    M.vmult(y_mb, src)
    a = y_mb.numpy()
    M.vmult(y_mb, src)
    b = y_mb.numpy()
    assert np.array_equal(a, b), "repeated CSR products differ"
    return a.astype(np.float64), y_mf.numpy().astype(np.float64)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_csr_vmult_parity(gpu, orc, p, precision):
    smf = gpu
    print(f"\ncase: CSR vmult p={p} {precision} n=8", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(8, 0.2, 7), p, precision=precision)
    M = smf.SystemMatrix(op)
    x = np.random.default_rng(5).uniform(-1, 1, op.m())
    got, mf = products(smf, op, M, x)
    ref = orc.OracleProblem.from_operator(op).vmult(x)
    assert rel_err(got, ref) < TOL[precision]
    assert rel_err(got, mf) < TOL[precision]


def test_csr_vmult_vector_and_nitsche(gpu):
    smf = gpu
    mesh = smf.Mesh.kuhn_cube(6, 0.2, 9)
    for kind, boundary in (("cg_vector", "dirichlet"), ("cg", "nitsche")):
        print(f"\ncase: CSR vmult {kind} {boundary} p=2 n=6", flush=True)
        op = smf.Operator(mesh, 2, kind=kind, boundary=boundary)
        M = smf.SystemMatrix(op)
        x = np.random.default_rng(2).uniform(-1, 1, op.m())
        got, mf = products(smf, op, M, x)
        assert rel_err(got, mf) < 1e-12


def test_csr_vmult_full_size_matches_matrix_free(gpu):
    """64^3 x 6 P2 FP64 (2.1 M DoFs): the matrix-based and matrix-free
    products agree to rounding (size-independent property)."""
    smf = gpu
    print("\ncase: CSR vmult p=2 f64 n=64", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(64, 0.2, 1), 2)
    M = smf.SystemMatrix(op)
    assert M.m() == 129 ** 3
    x = np.random.default_rng(8).uniform(-1, 1, op.m())
    got, mf = products(smf, op, M, x)
    assert rel_err(got, mf) < 1e-12
    c = op.constrained().astype(bool)
    assert np.array_equal(got[c], x[c])
