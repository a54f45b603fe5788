"""GPU parity of the DG SIPG operators (scalar, components-batched vector,
Helmholtz-like) against the CPU oracle: FP64 1e-12, FP32 1e-5."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-12, "f32": 1e-5}


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("kind", ["dg", "dg_vector", "dg_helmholtz"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_dg_vmult_and_diagonal(gpu, orc, p, kind, precision):
    smf = gpu
    print(f"\ncase: {kind} p={p} {precision} n=4", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(4, 0.2, 6), p, kind=kind, precision=precision,
                      mass_coefficient=1.0)
    pr = orc.DGOracleProblem.from_operator(op, mass=1.0 if kind == "dg_helmholtz" else 0.0)
    x = np.random.default_rng(8).uniform(-1, 1, op.m())
    src = smf.DeviceVector.from_numpy(x.astype(op.dtype))
    dst = op.initialize_dof_vector()
    op.vmult(dst, src)
    assert rel_err(dst.numpy().astype(float), pr.vmult(x)) < TOL[precision]
    d = op.initialize_dof_vector()
    op.compute_diagonal(d)
    assert rel_err(d.numpy().astype(float), pr.diagonal()) < TOL[precision]


@pytest.mark.parametrize("p", [1, 2, 3])
def test_dg_reorder_off_and_determinism(gpu, orc, p):
    smf = gpu
    print(f"\ncase: dg p={p} reorder off n=6", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(6, 0.2, 2), p, kind="dg", reorder=False)
    pr = orc.DGOracleProblem.from_operator(op)
    x = np.random.default_rng(1).uniform(-1, 1, op.m())
    src = smf.DeviceVector.from_numpy(x)
    a, b = op.initialize_dof_vector(), op.initialize_dof_vector()
    op.vmult(a, src)
    op.Tvmult(b, src)
    assert np.array_equal(a.numpy(), b.numpy())
    assert rel_err(a.numpy(), pr.vmult(x)) < 1e-12


@pytest.mark.parametrize("dg_kernel", [1, 2, 3, 4])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_dg_kernel_variants(gpu, orc, p, dg_kernel):
    smf = gpu
    print(f"\ncase: dg p={p} dg_kernel={dg_kernel}", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(5, 0.2, 3), p, kind="dg", dg_kernel=dg_kernel)
    pr = orc.DGOracleProblem.from_operator(op)
    x = np.random.default_rng(2).uniform(-1, 1, op.m())
    src = smf.DeviceVector.from_numpy(x)
    dst = op.initialize_dof_vector()
    op.vmult(dst, src)
    assert rel_err(dst.numpy(), pr.vmult(x)) < 1e-12
