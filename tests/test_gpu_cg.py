"""GPU parity of the CG Poisson operator against the CPU oracle.

Bar: FP64 1e-12 and FP32 1e-5 relative max-norm error against the FP64
oracle (restatement of CG/cells_batched/application.cc:1236-1416).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}


def _setup(smf, orc, n, p, precision, reorder=True, seed=7, patch_cells=0):
    mesh = smf.Mesh.kuhn_cube(n, 0.2, seed)
    op = smf.Operator(mesh, p, kind="cg", precision=precision, reorder=reorder,
                      patch_cells=patch_cells)
    pr = orc.OracleProblem.from_operator(op)
    return op, pr


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("reorder", [True, False])
def test_vmult_parity(gpu, orc, p, precision, reorder):
    smf = gpu
    print(f"\ncase: CG vmult p={p} {precision} reorder={reorder} n=8", flush=True)
    op, pr = _setup(smf, orc, 8, p, precision, reorder)
    x = np.random.default_rng(11).uniform(-1, 1, op.m())
    ref = pr.vmult(x)
    src = smf.DeviceVector.from_numpy(x.astype(op.dtype))
    dst = op.initialize_dof_vector()
    op.vmult(dst, src)
    got = dst.numpy().astype(np.float64)
    assert rel_err(got, ref) < TOL[precision]


@pytest.mark.parametrize("p", [1, 2, 3])
def test_vmult_parity_16cube(gpu, orc, p):
    """16^3 x 6 perturbed tets (24,576 cells), x ~ U[-1,1] on all DoFs."""
    smf = gpu
    print(f"\ncase: CG vmult p={p} f64 n=16", flush=True)
    op, pr = _setup(smf, orc, 16, p, "f64")
    if p == 2:
        assert op.m() == 35937 and op.info["n_constrained"] == 35937 - 31 ** 3
    x = np.random.default_rng(3).uniform(-1, 1, op.m())
    ref = pr.vmult(x)
    src = smf.DeviceVector.from_numpy(x)
    dst = op.initialize_dof_vector()
    op.vmult(dst, src)
    got = dst.numpy()
    assert rel_err(got, ref) < 1e-12
    # Dirichlet rows are identity rows
    con = op.constrained().astype(bool)
    assert np.array_equal(got[con], x[con])


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_diagonal(gpu, orc, p, precision):
    smf = gpu
    print(f"\ncase: CG diagonal p={p} {precision}", flush=True)
    op, pr = _setup(smf, orc, 6, p, precision)
    ref = pr.diagonal()
    d = op.initialize_dof_vector()
    op.compute_diagonal(d)
    assert rel_err(d.numpy().astype(np.float64), ref) < TOL[precision]
    inv = op.initialize_dof_vector()
    op.compute_inverse_diagonal(inv)
    refinv = np.where(np.abs(ref) > 1e-10, 1.0 / ref, 1.0)
    assert rel_err(inv.numpy().astype(np.float64), refinv) < TOL[precision]


def test_deterministic_and_tvmult(gpu):
    smf = gpu
    print("\ncase: CG determinism p=2 n=24", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(24, 0.2, 5), 2)
    x = np.random.default_rng(5).uniform(-1, 1, op.m())
    src = smf.DeviceVector.from_numpy(x)
    a, b, c = (op.initialize_dof_vector() for _ in range(3))
    op.vmult(a, src)
    op.vmult(b, src)
    op.Tvmult(c, src)
    assert np.array_equal(a.numpy(), b.numpy())
    assert np.array_equal(a.numpy(), c.numpy())


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("accumulate", [1, 2, 3])
@pytest.mark.parametrize("patch_cells", [64, 1024])
def test_accumulation_variants(gpu, orc, precision, accumulate, patch_cells):
    """The LDS accumulation schemes (atomics or read-add-write per colour, atomics per wave) and extreme
    patch sizes give the oracle's result, bitwise repeatably."""
    smf = gpu
    print(f"\ncase: CG p=2 {precision} accumulate={accumulate} patch={patch_cells}", flush=True)
    mesh = smf.Mesh.kuhn_cube(10, 0.2, 3)
    op = smf.Operator(mesh, 2, precision=precision, accumulate=accumulate, patch_cells=patch_cells)
    pr = orc.OracleProblem.from_operator(op)
    x = np.random.default_rng(7).uniform(-1, 1, op.m())
    src = smf.DeviceVector.from_numpy(x.astype(op.dtype))
    a, b = op.initialize_dof_vector(), op.initialize_dof_vector()
    op.vmult(a, src)
    op.vmult(b, src)
    assert np.array_equal(a.numpy(), b.numpy())
    assert rel_err(a.numpy().astype(np.float64), pr.vmult(x)) < TOL[precision]


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_cg_vector_parity(gpu, orc, p, precision):
    smf = gpu
    print(f"\ncase: CG vector p={p} {precision}", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(6, 0.2, 5), p, kind="cg_vector", precision=precision)
    n = op.n_scalar_dofs()
    pr = orc.OracleProblem.from_operator(op)
    x = np.random.default_rng(4).uniform(-1, 1, op.m())
    ref = np.concatenate([pr.vmult(x[c * n:(c + 1) * n]) for c in range(3)])
    src = smf.DeviceVector.from_numpy(x.astype(op.dtype))
    dst = op.initialize_dof_vector()
    op.vmult(dst, src)
    assert rel_err(dst.numpy().astype(np.float64), ref) < TOL[precision]
    d = op.initialize_dof_vector()
    op.compute_diagonal(d)
    assert rel_err(d.numpy().astype(np.float64), np.tile(pr.diagonal(), 3)) < TOL[precision]
