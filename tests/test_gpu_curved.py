"""GPU parity of the curved-cell CG path (P2 geometry map on the spherical
manifold, per-point metric: CG/cells_batched/application.cc:1361-1376,
1613, 1668-1677) against the oracle's numpy restatement, plus the
size-independent consistency property A (linear) = 0 on rows whose cells
avoid the boundary, and a multigrid solve with curved levels.

Bar: FP64 1e-12, FP32 1e-5 relative max-norm error against the FP64 oracle.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}


def device_apply(smf, op, x, mode=0):
    dst = op.initialize_dof_vector()
    if mode == 0:
        op.vmult(dst, smf.DeviceVector.from_numpy(x.astype(op.dtype)))
    else:
        op.compute_diagonal(dst)
    return dst.numpy().astype(np.float64)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_curved_vmult_parity(gpu, orc, p, precision):
    smf = gpu
    print(f"\ncase: curved CG vmult p={p} {precision} n=8", flush=True)
    mesh = smf.Mesh.kuhn_cube(8, 0.2, 7)
    op = smf.Operator(mesh, p, precision=precision, manifold="spherical")
    pts, w = op.quadrature()
    x = np.random.default_rng(p).uniform(-1, 1, op.m())
    ref = orc.curved_vmult(p, mesh.vertices, op.cell_vertices(), op.cell_dofs(), op.constrained(), pts, w, x)
    assert rel_err(device_apply(smf, op, x), ref) < TOL[precision]


@pytest.mark.parametrize("p", [1, 2, 3])
def test_curved_diagonal_and_vector(gpu, p):
    smf = gpu
    print(f"\ncase: curved CG diagonal + vector p={p} n=6", flush=True)
    mesh = smf.Mesh.kuhn_cube(6, 0.2, 2)
    op = smf.Operator(mesh, p, manifold="spherical")
    host = smf.Operator(mesh, p, manifold="spherical", host_only=True)
    assert rel_err(device_apply(smf, op, None, mode=1), host.host_apply(mode=1)) < 1e-12
    vec = smf.Operator(mesh, p, kind="cg_vector", manifold="spherical")
    x = np.random.default_rng(3).uniform(-1, 1, vec.m())
    n = op.m()
    y = device_apply(smf, vec, x)
    for c in range(3):
        assert rel_err(y[c * n:(c + 1) * n], device_apply(smf, op, x[c * n:(c + 1) * n])) < 1e-12


def test_curved_consistency_full_size(gpu):
    """48^3 x 6 P2 FP64 (0.9 M DoFs): A (linear) = 0 on rows whose cells avoid the boundary"""
    smf = gpu
    print("\ncase: curved CG consistency p=2 n=48", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(48, 0.2, 1), 2, manifold="spherical")
    con = op.constrained().astype(bool)
    cd = op.cell_dofs()
    touch = np.zeros(op.m(), bool)
    touch[np.unique(cd[con[cd].any(axis=1)])] = True
    X = op.dof_points()
    lin = 0.3 + 1.5 * X[:, 0] - 0.7 * X[:, 1] + 2.0 * X[:, 2]
    y = device_apply(smf, op, lin)
    assert np.max(np.abs(y[~touch])) < 1e-12


def test_curved_mg_solve(gpu):
    smf = gpu
    print("\ncase: curved MG-CG p=2 n=16", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(16, 0.2, 4), 2, manifold="spherical")
    mg = smf.Multigrid(op)
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(op, "grf", seed=2))
    x = op.initialize_dof_vector()
    st = smf.solve_cg(op, x, b, mg, rel_tol=1e-10)
    assert st["converged"] and st["iterations"] < 30
    r = op.initialize_dof_vector()
    op.vmult(r, x)
    bb = b.numpy()
    assert np.linalg.norm(bb - r.numpy()) <= 2e-10 * np.linalg.norm(bb)
