"""Parity at the benchmark workloads' full sizes, through the C ABI:
CG P2 16^3, DG SIPG 64^3 (P1, P2, P3; FP64 and FP32, reordering on and off;
and size-independent properties),
Helmholtz-like vector DG P2 32^3, MG-preconditioned CG on CG P2 48^3.  The oracle runs with OpenMP on the host (seconds to tens
of seconds per case)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def gpu_vmult(smf, op, x):
    y = op.initialize_dof_vector()
    op.vmult(y, smf.DeviceVector.from_numpy(x.astype(op.dtype)))
    return y.numpy().astype(np.float64)


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_cg_p2_16cube(gpu, orc, precision, tol):
    smf = gpu
    op = smf.Operator(smf.Mesh.kuhn_cube(16, 0.2, 7), 2, kind="cg", precision=precision)
    assert op.m() == 35937 and op.info["n_constrained"] == 6146
    x = np.random.default_rng(0).uniform(-1, 1, op.m())
    want = orc.OracleProblem.from_operator(op).vmult(x.astype(op.dtype).astype(np.float64))
    assert rel_err(gpu_vmult(smf, op, x), want) < tol


def _cell_order(op):
    """internal cell -> rank of its (ascending) vertex tuple: matches the
    cells of two operators on one mesh whatever their internal orders"""
    cv = np.sort(op.cell_vertices(), axis=1)
    return np.lexsort(cv.T[::-1])


@pytest.mark.parametrize("p", [1, 2, 3])
def test_dg_64cube_vs_oracle(gpu, orc, p):
    """DG SIPG at 64^3 x 6 (the DG benchmark mesh) against the OpenMP oracle,
    for FP64 and FP32 with the hierarchical reordering and FP64 without it.
    The input is FP32-representable, so one oracle product serves both
    precisions; the reordering-off operator numbers the same cells in another
    order (cell-contiguous DoFs, local nodes by ascending vertex ids), so its
    result is the oracle's with the cell blocks permuted."""
    smf = gpu
    mesh = smf.Mesh.kuhn_cube(64, 0.2, 7)
    op = smf.Operator(mesh, p, kind="dg", precision="f64")
    x = np.random.default_rng(p).uniform(-1, 1, op.m()).astype(np.float32).astype(np.float64)
    want = orc.DGOracleProblem.from_operator(op).vmult_omp(x)
    print(f"\ncase: dg 64^3 p={p} f64 reorder on", flush=True)
    assert rel_err(gpu_vmult(smf, op, x), want) < 1e-12
    nl = op.info["dofs_per_cell"]
    o_on = _cell_order(op)
    del op
    print(f"case: dg 64^3 p={p} f32 reorder on", flush=True)
    op32 = smf.Operator(mesh, p, kind="dg", precision="f32")
    assert rel_err(gpu_vmult(smf, op32, x), want) < 1e-5
    del op32
    print(f"case: dg 64^3 p={p} f64 reorder off", flush=True)
    off = smf.Operator(mesh, p, kind="dg", precision="f64", reorder=False)
    o_off = _cell_order(off)
    # cell o_off[r] of `off` is cell o_on[r] of `op`
    to_on = np.empty_like(o_off)
    to_on[o_off] = o_on
    xo = x.reshape(-1, nl)[to_on].ravel()
    wo = want.reshape(-1, nl)[to_on].ravel()
    assert rel_err(gpu_vmult(smf, off, xo), wo) < 1e-12


def test_dg_64cube_properties(gpu):
    """Size-independent checks at 64^3 for P2 (FP64, FP32) and P3 (FP64):
    bitwise determinism, symmetry u.Av = v.Au, zero rows for a globally
    linear function on cells without a boundary face, and agreement of the
    one-lane and four-lane kernels."""
    smf = gpu
    mesh = smf.Mesh.kuhn_cube(64, 0.2, 7)
    for p, precision, tol in ((2, "f64", 1e-12), (2, "f32", 1e-5), (3, "f64", 1e-12)):
        print(f"\ncase: dg 64^3 p={p} {precision} properties", flush=True)
        op = smf.Operator(mesh, p, kind="dg", precision=precision)
        rng = np.random.default_rng(p)
        u = rng.uniform(-1, 1, op.m())
        v = rng.uniform(-1, 1, op.m())
        au, au2, av = gpu_vmult(smf, op, u), gpu_vmult(smf, op, u), gpu_vmult(smf, op, v)
        assert np.array_equal(au, au2)
        uu, vv = u.astype(op.dtype).astype(np.float64), v.astype(op.dtype).astype(np.float64)
        assert abs(vv @ au - uu @ av) <= tol * np.linalg.norm(vv) * np.linalg.norm(au)
        pts = op.dof_points()
        lin = 0.3 + pts @ np.array([1.0, -2.0, 0.5])
        al = gpu_vmult(smf, op, lin)
        nl = op.info["dofs_per_cell"]
        assert op.info["n_boundary_faces"] == 12 * 64 ** 2
        # cells with no vertex on the cube's boundary have no boundary face
        vb = (np.isclose(mesh.vertices, 0) | np.isclose(mesh.vertices, 1)).any(axis=1)
        inner = ~vb[op.cell_vertices()].any(axis=1)
        rows = (np.nonzero(inner)[0][:, None] * nl + np.arange(nl)).ravel()
        assert inner.sum() > 0.9 * len(inner)
        assert np.max(np.abs(al[rows])) <= 100 * tol * np.max(np.abs(al))
        if precision == "f64":
            other = smf.Operator(mesh, p, kind="dg", precision="f64", dg_kernel=2)
            assert rel_err(gpu_vmult(smf, other, u), au) < 1e-13


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_helmholtz_32cube(gpu, orc, precision, tol):
    smf = gpu
    print(f"\ncase: helmholtz 32^3 {precision}", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(32, 0.2, 7), 2, kind="dg_helmholtz", precision=precision,
                      mass_coefficient=1.0, laplace_coefficient=1.0)
    assert op.m() == 5898240
    x = np.random.default_rng(3).uniform(-1, 1, op.m())
    want = orc.DGOracleProblem.from_operator(op, mass=1.0).vmult_omp(x.astype(op.dtype).astype(np.float64))
    assert rel_err(gpu_vmult(smf, op, x), want) < tol


def test_mg_cg_48cube(gpu, orc):
    """Mixed-precision hybrid-MG CG on CG P2 48^3 (912,673 DoFs): the
    computed solution's residual with the oracle's operator."""
    smf = gpu
    op = smf.Operator(smf.Mesh.kuhn_cube(48, 0.2, 7), 2, kind="cg", precision="f64")
    assert op.m() == 912673 and op.info["n_constrained"] == 55298
    mg = smf.Multigrid(op)
    assert mg.info()["n_dofs"] == [2197, 15625, 117649, 912673]
    bh = smf.assemble_rhs(op, "grf", seed=7, n_features=256, corr_length=0.1)
    x = op.initialize_dof_vector()
    st = smf.solve_cg(op, x, smf.DeviceVector.from_numpy(bh), mg, rel_tol=1e-10)
    assert st["converged"] and st["iterations"] <= 12
    r = bh - orc.OracleProblem.from_operator(op).vmult_omp(x.numpy())
    assert np.linalg.norm(r) <= 2e-10 * np.linalg.norm(bh)
