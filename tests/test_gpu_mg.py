"""GPU: hybrid multigrid transfers, V-cycle and the (F)CG solve.

Transfers are checked against the Kuhn-lattice interpolation rule derived
independently here (a fine node at integer lattice position Y, coarse
vertices at even positions, takes the mean of the coarse values at floor(Y/2)
and ceil(Y/2)); R = P^T by duality; solves against a dense FP64 solve of the
oracle's matrix and against the analytic solution prod sin(3 pi x_d)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _keys(x, m):
    y = np.rint(x * m).astype(np.int64)
    assert np.max(np.abs(y - x * m)) < 1e-9
    return y


def test_transfers_reproduce_lattice_interpolation(gpu):
    smf = gpu
    print("\ncase: MG transfers P2, n=8, unperturbed, f64 levels", flush=True)
    fine = smf.Operator(smf.Mesh.kuhn_cube(8, 0.0, 0), 2)
    mg = smf.Multigrid(fine, level_precision="f64", max_coarse_dofs=64)  # keep the h-levels
    info = mg.info()
    assert info["kuhn_n"][0] == 2
    assert info["degree"][-1] == 2 and info["degree"][-2] == 1
    rng = np.random.default_rng(0)
    for l in range(1, info["n_levels"]):
        xf, cf = mg.level_dofs(l)
        xc, cc = mg.level_dofs(l - 1)
        nc_lat = info["kuhn_n"][l - 1]
        kc = _keys(xc, nc_lat)
        lookup = {tuple(k): i for i, k in enumerate(kc)}
        vc = np.where(cc == 1, 0.0, rng.uniform(-1, 1, len(xc)))
        yf = _keys(xf, 2 * nc_lat)
        ref = np.zeros(len(xf))
        for i, y in enumerate(yf):
            if cf[i]:
                continue
            lo, hi = tuple(y // 2), tuple(-((-y) // 2))
            ref[i] = 0.5 * (vc[lookup[lo]] + vc[lookup[hi]])
        dvc = smf.DeviceVector.from_numpy(vc)
        dvf = smf.DeviceVector(len(xf))
        mg.transfer(l, 0, dvf, dvc)
        got = dvf.numpy()
        assert np.max(np.abs(got - ref)) < 1e-14, l
        # restriction is the transpose
        uf = np.where(cf == 1, 0.0, rng.uniform(-1, 1, len(xf)))
        duf = smf.DeviceVector.from_numpy(uf)
        drc = smf.DeviceVector(len(xc))
        mg.transfer(l, 1, drc, duf)
        assert abs(drc.numpy() @ vc - uf @ got) < 1e-12 * (1 + abs(uf @ got))


def test_solve_matches_dense_oracle(gpu, orc):
    smf = gpu
    print("\ncase: MG-CG P2 n=6 vs dense oracle solve", flush=True)
    fine = smf.Operator(smf.Mesh.kuhn_cube(6, 0.2, 4), 2)
    mg = smf.Multigrid(fine)
    b = smf.assemble_rhs(fine, "grf", seed=3, n_features=64, corr_length=0.3)
    A = orc.OracleProblem.from_operator(fine).dense()
    x_ref = np.linalg.solve(A, b)
    for flexible in (False, True):
        x = fine.initialize_dof_vector()
        st = smf.solve_cg(fine, x, smf.DeviceVector.from_numpy(b), mg, rel_tol=1e-10, flexible=flexible)
        assert st["converged"] and st["final_residual"] <= 1e-10 * st["rhs_norm"]
        assert st["iterations"] < 20
        assert rel_err(x.numpy(), x_ref) < 1e-8


def test_vcycle_deterministic_and_h_independent(gpu):
    smf = gpu
    its = []
    for n in (8, 16, 32):
        print(f"\ncase: MG-CG P2 sine n={n}", flush=True)
        fine = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, 7), 2)
        mg = smf.Multigrid(fine)
        b = smf.DeviceVector.from_numpy(smf.assemble_rhs(fine, "sine"))
        x = fine.initialize_dof_vector()
        st = smf.solve_cg(fine, x, b, mg, rel_tol=1e-10)
        assert st["converged"]
        its.append(st["iterations"])
        # discretisation error against the analytic solution decreases
        u = smf.exact_solution(fine)
        err = np.linalg.norm(x.numpy() - u) / np.linalg.norm(u)
        assert err < {8: 0.2, 16: 0.03, 32: 0.005}[n]
        if n == 8:
            z1, z2 = fine.initialize_dof_vector(), fine.initialize_dof_vector()
            mg.vmult(z1, b)
            mg.vmult(z2, b)
            assert np.array_equal(z1.numpy(), z2.numpy())
    assert max(its) <= min(its) + 4, its


@pytest.mark.parametrize("degree,prec,smoother", [(2, "f32", 5), (2, "f64", 5), (1, "f32", 3), (3, "f64", 2)])
def test_cheby_run_matches_per_step_fixup(gpu, degree, prec, smoother):
    # The Chebyshev run (interface rows completed by the next step's kernel,
    # Dirichlet rows by the closing kernel) against one operator application
    # and fix-up per step: the same recurrences and sums in the same order, so
    # the V-cycles agree to rounding and the solves take the same iterations.
    smf = gpu
    print(f"\ncase: Chebyshev run vs per-step fix-up, P{degree} {prec} degree {smoother}", flush=True)
    fine = smf.Operator(smf.Mesh.kuhn_cube(12, 0.2, 5), degree)
    mg_run = smf.Multigrid(fine, smoother_degree=smoother, level_precision=prec, fuse_steps=True)
    mg_ref = smf.Multigrid(fine, smoother_degree=smoother, level_precision=prec, fuse_steps=False)
    b = smf.DeviceVector.from_numpy(np.random.default_rng(2).uniform(-1, 1, fine.m()))
    for _ in range(2):  # the run's buffers are reused from one V-cycle to the next
        z1, z2 = fine.initialize_dof_vector(), fine.initialize_dof_vector()
        mg_run.vmult(z1, b)
        mg_ref.vmult(z2, b)
        tol = 1e-5 if prec == "f32" else 1e-12
        assert rel_err(z1.numpy(), z2.numpy()) < tol
    z3 = fine.initialize_dof_vector()
    mg_run.vmult(z3, b)
    assert np.array_equal(z1.numpy(), z3.numpy())  # deterministic
    x1, x2 = fine.initialize_dof_vector(), fine.initialize_dof_vector()
    s1 = smf.solve_cg(fine, x1, b, mg_run, rel_tol=1e-10)
    s2 = smf.solve_cg(fine, x2, b, mg_ref, rel_tol=1e-10)
    assert s1["converged"] and s2["converged"]
    assert abs(s1["iterations"] - s2["iterations"]) <= 1
    assert rel_err(x1.numpy(), x2.numpy()) < 1e-8


def test_rhs_against_oracle_quadrature(gpu, orc):
    smf = gpu
    print("\ncase: RHS sine P2 n=4", flush=True)
    fine = smf.Operator(smf.Mesh.kuhn_cube(4, 0.2, 2), 2)
    b = smf.assemble_rhs(fine, "sine")
    pr = orc.OracleProblem.from_operator(fine)
    pts, w = orc.tet_rule(7)
    vals = np.array([orc.basis(2, x)[0] for x in pts])
    ref = np.zeros(fine.m())
    V, C = pr.vertices, pr.cells
    a = 3 * np.pi
    for c in range(C.shape[0]):
        J = np.stack([V[C[c, k]] - V[C[c, 0]] for k in (1, 2, 3)], axis=1)
        X = V[C[c, 0]] + pts @ J.T
        f = 3 * a * a * np.sin(a * X[:, 0]) * np.sin(a * X[:, 1]) * np.sin(a * X[:, 2])
        ref[pr.cell_dofs[c]] += abs(np.linalg.det(J)) * ((w * f) @ vals)
    ref[pr.constrained.astype(bool)] = 0
    assert rel_err(b, ref) < 2e-3
