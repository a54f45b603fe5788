"""GPU: the curved-cell DG kernel (dg_curved_kernel) against the oracle's
numpy restatement oracle_py.dg_curved_matrix (pinned in
tests/test_host_dg_curved.py), through the C ABI: vmult and diagonal, P1-P3,
FP64 (1e-12) and FP32 (1e-5), scalar SIPG, vector and Helmholtz-like
operators; and at a larger size the properties that do not need the oracle
(zero rows of constants and linear functions away from the boundary,
symmetry, determinism)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

CENTER = (1.5, 0.5, 0.5)


def _ref(orc, op, p, ncomp=1, mass=0.0):
    pts, w = op.quadrature()
    fp, fw = op.face_quadrature()
    return orc.dg_curved_matrix(p, op.mesh.vertices, op.cell_vertices(), pts, w, fp, fw, ncomp=ncomp, mass=mass,
                                center=CENTER)


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("kind,ncomp,mass", [("dg", 1, 0.0), ("dg_vector", 3, 0.0), ("dg_helmholtz", 3, 2.5)])
@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_curved_dg_vs_restatement(gpu, orc, p, kind, ncomp, mass, precision, tol):
    smf = gpu
    print(f"\ncase: curved DG {kind} P{p} {precision}", flush=True)
    n = 3 if p < 3 else 2
    op = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, 5), p, kind=kind, precision=precision, manifold="spherical",
                      manifold_center=CENTER, mass_coefficient=max(mass, 1.0))
    A = _ref(orc, op, p, ncomp, mass)
    x = np.random.default_rng(p).uniform(-1, 1, op.m()).astype(op.dtype)
    y = op.initialize_dof_vector()
    op.vmult(y, smf.DeviceVector.from_numpy(x))
    want = A @ x.astype(np.float64)
    assert rel_err(y.numpy().astype(np.float64), want) < tol
    d = op.initialize_dof_vector()
    op.compute_diagonal(d)
    assert rel_err(d.numpy().astype(np.float64), A.diagonal()) < tol


def test_curved_dg_properties_at_size(gpu):
    smf = gpu
    print("\ncase: curved DG P2 f64 n=12 properties", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(12, 0.2, 8), 2, kind="dg", manifold="spherical", manifold_center=CENTER)
    C = op.cell_vertices()
    keys = {}
    for k in range(C.shape[0]):
        for f in ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)):
            keys.setdefault(tuple(sorted(C[k, list(f)])), []).append(k)
    bdry = np.zeros(C.shape[0], bool)
    for cells in keys.values():
        if len(cells) == 1:
            bdry[cells[0]] = True
    rows = ~np.repeat(bdry, 10)

    def apply(v):
        y = op.initialize_dof_vector()
        op.vmult(y, smf.DeviceVector.from_numpy(v))
        return y.numpy()

    X = op.dof_points()
    one = apply(np.ones(op.m()))
    lin = apply(0.3 - 0.7 * X[:, 0] + 1.3 * X[:, 1] + 0.4 * X[:, 2])
    rng = np.random.default_rng(3)
    u, v = rng.uniform(-1, 1, (2, op.m()))
    au, av = apply(u), apply(v)
    scale = np.max(np.abs(au))
    assert np.max(np.abs(one[rows])) < 1e-11 * scale
    assert np.max(np.abs(lin[rows])) < 1e-10 * scale
    assert abs(v @ au - u @ av) < 1e-11 * abs(v @ au)
    assert np.array_equal(au, apply(u))


def test_curved_dg_rejects_multigrid(gpu):
    smf = gpu
    op = smf.Operator(smf.Mesh.kuhn_cube(4, 0.2, 1), 2, kind="dg", manifold="spherical")
    with pytest.raises(smf.MFError, match="curved DG"):
        smf.Multigrid(op)
