"""GPU: multigrid components and the random-field right-hand side against
the oracle's restatements.

  * b_i = int f phi_i for the seeded Gaussian random field (the right-hand
    side of the multigrid solves, the pattern of
    CG/cells_batched/application_hp.cc:1441-1465) against
    oracle_py.grf_features + the oracle's collapsed Gauss quadrature of high
    order (exact to degree 25).  The library integrates with a symmetric rule
    of degree 6.  With l = 100 the field is a polynomial to 1e-18 on every
    cell, so the two agree to rounding (1e-12): the frequencies (their linear
    term is 1e-2 of the value), phases, amplitude, Jacobian and basis order are
    pinned there.  At l = 1 and l = 0.1 the difference is the degree-6 rule's
    quadrature error: 2.5e-7 and 1.3e-3 relative as measured (tolerances 1e-6
    and 3e-3); a wrong feature is an O(1) error.
  * the Chebyshev interval: lambda_max of every level against
    oracle_py.lanczos_lambda_max on the oracle's level matrix (10 CG steps,
    application_ph.cc:1916-1918), FP64 levels;
  * one whole V-cycle (pre-smoothing, residual, restriction, coarse solve,
    prolongation, post-smoothing) against oracle_py.mg_vcycle on the oracle's
    level matrices and the library's transfer matrices (pinned separately by
    the lattice rule in test_gpu_mg.py), FP64 levels, 1e-10.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["cg", "dg"])
@pytest.mark.parametrize("n,nf,ell,tol", [(4, 1, 100.0, 1e-12), (4, 256, 100.0, 1e-12), (4, 1, 1.0, 1e-6),
                                          (4, 256, 1.0, 1e-6), (8, 1, 0.1, 3e-3), (8, 256, 0.1, 3e-3)])
def test_grf_rhs_matches_oracle_quadrature(gpu, orc, kind, n, nf, ell, tol):
    smf = gpu
    print(f"\ncase: GRF rhs {kind} P2 n={n} features={nf} l={ell}", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, 3), 2, kind=kind, precision="f64")
    seed = 17
    b = smf.assemble_rhs(op, "grf", seed=seed, n_features=nf, corr_length=ell)
    om, ph = orc.grf_features(seed, nf, ell)
    f = lambda X: orc.grf_eval(om, ph, X)  # noqa: E731
    if kind == "cg":
        pr = orc.OracleProblem.from_operator(op)
        ref = orc.rhs_quadrature(2, pr.vertices, pr.cells, pr.cell_dofs, op.m(), f, n_rule=14)
        ref[pr.constrained.astype(bool)] = 0.0
    else:
        V, Cv = op.mesh.vertices, op.cell_vertices()
        # DG DoF (k, i) is row k*10 + i; nodes in the cell's ascending vertex order
        ref = orc.rhs_quadrature(2, V, Cv, np.arange(op.m()).reshape(-1, 10), op.m(), f, n_rule=14)
    assert rel_err(b, ref) < tol


def _level_mesh(smf, fine_mesh, n_level):
    """Kuhn mesh of n_level^3 hexes whose vertices are the fine mesh's vertices
    at the matching lattice points (the h-coarsening of the hierarchy)"""
    nf = fine_mesh.kuhn_n
    s = nf // n_level
    base = smf.Mesh.kuhn_cube(n_level, 0.0, 0)
    n1, f1 = n_level + 1, nf + 1
    ids = [s * i + f1 * (s * j + f1 * s * k) for k in range(n1) for j in range(n1) for i in range(n1)]
    return smf.Mesh(fine_mesh.vertices[ids], base.cells, kuhn_n=n_level)


def _oracle_levels(smf, orc, mesh, mg, fine_degree):
    info = mg.info()
    As, Ps, cons = [], [], []
    for l in range(info["n_levels"]):
        xyz, con = mg.level_dofs(l)
        deg = info["degree"][l]
        lm = mesh if info["kuhn_n"][l] == mesh.kuhn_n else _level_mesh(smf, mesh, info["kuhn_n"][l])
        op = smf.Operator(lm, deg, host_only=True)
        A = orc.OracleProblem.from_operator(op).dense()
        # map the oracle operator's numbering onto the level's by support point
        X = op.dof_points()
        key = {tuple(np.round(p * 1e8).astype(np.int64)): i for i, p in enumerate(X)}
        perm = np.array([key[tuple(np.round(p * 1e8).astype(np.int64))] for p in xyz])
        As.append(A[np.ix_(perm, perm)])
        cons.append(con.astype(bool))
        assert np.array_equal(op.constrained()[perm].astype(bool), cons[-1])
        if l == 0:
            Ps.append(None)
            continue
        nc = info["n_dofs"][l - 1]
        P = np.zeros((info["n_dofs"][l], nc))
        e = np.zeros(nc)
        dst = smf.DeviceVector(info["n_dofs"][l])
        for j in range(nc):
            e[j] = 1.0
            mg.transfer(l, 0, dst, smf.DeviceVector.from_numpy(e))
            P[:, j] = dst.numpy()
            e[j] = 0.0
        Ps.append(P)
    return As, Ps, cons


@pytest.fixture(scope="module")
def mg_case(gpu, orc):
    smf = gpu
    print("\ncase: oracle V-cycle restatement, CG P2 n=4, f64 levels", flush=True)
    mesh = smf.Mesh.kuhn_cube(4, 0.2, 3)
    fine = smf.Operator(mesh, 2, precision="f64")
    mg = smf.Multigrid(fine, level_precision="f64", max_coarse_dofs=64)
    info = mg.info()
    assert info["degree"] == [1, 1, 2] and info["kuhn_n"] == [2, 4, 4]
    As, Ps, cons = _oracle_levels(smf, orc, mesh, mg, 2)
    return smf, fine, mg, orc.mg_levels(As, Ps, cons), cons


def test_lambda_max_matches_restatement(mg_case):
    smf, fine, mg, levels, _ = mg_case
    lam = mg.info()["lambda_max"]
    for l in range(1, len(levels)):
        assert abs(lam[l] - levels[l]["lmax"]) < 1e-9 * levels[l]["lmax"], (l, lam[l], levels[l]["lmax"])
        # the estimate lies below the true lambda_max of D^-1 A, the interval above it
        A, dinv = levels[l]["A"], levels[l]["dinv"]
        s = np.sqrt(dinv)
        exact = np.linalg.eigvalsh(s[:, None] * A * s[None, :]).max()
        assert levels[l]["lam"] <= exact * (1 + 1e-12) < lam[l]


def test_vcycle_matches_restatement(mg_case, orc):
    smf, fine, mg, levels, cons = mg_case
    rng = np.random.default_rng(5)
    for free_only in (True, False):
        b = rng.uniform(-1, 1, fine.m())
        if free_only:
            b[cons[-1]] = 0.0
        z = fine.initialize_dof_vector()
        mg.vmult(z, smf.DeviceVector.from_numpy(b))
        want = orc.mg_vcycle(levels, b)
        assert rel_err(z.numpy(), want) < 1e-10, free_only
