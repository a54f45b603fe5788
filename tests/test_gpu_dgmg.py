"""DG hybrid multigrid on the GPU: the Nitsche CG operator (c-level) against
the DG oracle restricted to continuous functions, the c-transfer, and DG SIPG
P2 solves with the DG -> CG -> P1 -> h hierarchy (application_phc.cc:
3416-3552), including the learned smoother on the DG operator."""
import numpy as np
import pytest

from test_host_nitsche import cg_to_dg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_nitsche_vmult_vs_restricted_dg_oracle(gpu, orc, degree, precision, tol):
    smf = gpu
    mesh = smf.Mesh.kuhn_cube(4, perturb=0.2, seed=11)
    cg = smf.Operator(mesh, degree, kind="cg", precision=precision, boundary="nitsche")
    dg = smf.Operator(mesh, degree, kind="dg", host_only=True)
    cols = cg_to_dg(cg, dg)
    dgo = orc.DGOracleProblem.from_operator(dg)
    rng = np.random.default_rng(degree)
    x = rng.uniform(-1, 1, cg.m())
    y = cg.initialize_dof_vector()
    cg.vmult(y, smf.DeviceVector.from_numpy(x.astype(cg.dtype)))
    want = np.zeros(cg.m())
    np.add.at(want, cols, dgo.vmult(x.astype(cg.dtype).astype(np.float64)[cols]))
    assert np.max(np.abs(y.numpy() - want)) <= tol * np.max(np.abs(want))
    d = cg.initialize_dof_vector()
    cg.compute_diagonal(d)
    np.testing.assert_allclose(d.numpy(), cg.host_apply(mode=1), rtol=tol)


@pytest.fixture(scope="module")
def dgmg(gpu):
    smf = gpu
    mesh = smf.Mesh.kuhn_cube(8, perturb=0.2, seed=7)
    A = smf.Operator(mesh, 2, kind="dg", precision="f64")
    mg = smf.Multigrid(A)
    return smf, mesh, A, mg


def test_dg_hierarchy(dgmg):
    smf, _, A, mg = dgmg
    inf = mg.info()
    assert inf["level_kind"][-1] == smf.MF_DG_SIPG
    assert all(k == smf.MF_CG_POISSON for k in inf["level_kind"][:-1])
    assert all(inf["nitsche"][:-1]) and not inf["nitsche"][-1]
    assert inf["degree"][-2:] == [2, 2] and inf["degree"][-3] == 1
    assert inf["n_dofs"][-1] == A.m() == 6 * 8 ** 3 * 10
    assert inf["n_dofs"][-2] == 17 ** 3  # CG P2 on 8^3 with every boundary node free
    # h-coarsening stops at the first level the dense coarse solver takes
    assert inf["kuhn_n"][0] == 8 and inf["n_dofs"][0] == 9 ** 3
    deep = smf.Multigrid(A, max_coarse_dofs=64).info()
    assert deep["kuhn_n"][0] == 2 and deep["n_dofs"][-3:] == inf["n_dofs"][-3:]


def test_c_transfer_is_injection(dgmg):
    smf, _, A, mg = dgmg
    L = mg.info()["n_levels"] - 1
    xyz_cg, con_cg = mg.level_dofs(L - 1)
    xyz_dg, _ = mg.level_dofs(L)
    assert not con_cg.any()
    rng = np.random.default_rng(3)
    xc = rng.uniform(-1, 1, len(xyz_cg)).astype(np.float32)
    y = smf.DeviceVector(len(xyz_dg), np.float32)
    mg.transfer(L, 0, y, smf.DeviceVector.from_numpy(xc))
    # each DG node takes the CG value at the same point
    key = {tuple(np.round(p / 1e-7).astype(np.int64)): i for i, p in enumerate(xyz_cg)}
    cols = np.array([key[tuple(np.round(p / 1e-7).astype(np.int64))] for p in xyz_dg])
    np.testing.assert_array_equal(y.numpy(), xc[cols])
    # restriction is the transpose: sums over the DG copies of each CG node
    yd = rng.uniform(-1, 1, len(xyz_dg)).astype(np.float32)
    r = smf.DeviceVector(len(xyz_cg), np.float32)
    mg.transfer(L, 1, r, smf.DeviceVector.from_numpy(yd))
    want = np.zeros(len(xyz_cg))
    np.add.at(want, cols, yd.astype(np.float64))
    np.testing.assert_allclose(r.numpy(), want, rtol=1e-6, atol=1e-6)


def test_dg_mg_solve_vs_oracle(dgmg, orc):
    smf, _, A, mg = dgmg
    bh = smf.assemble_rhs(A, "grf", seed=5)
    b = smf.DeviceVector.from_numpy(bh)
    x = A.initialize_dof_vector()
    st = smf.solve_cg(A, x, b, mg, rel_tol=1e-10, max_iter=100)
    assert st["converged"] and st["iterations"] <= 25
    # residual of the computed solution with the oracle's operator
    r = bh - orc.DGOracleProblem.from_operator(A).vmult(x.numpy())
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(bh)


def test_dg_mg_h_independent(gpu):
    smf = gpu
    its = []
    for n in (8, 16):
        A = smf.Operator(smf.Mesh.kuhn_cube(n, perturb=0.2, seed=7), 2, kind="dg")
        mg = smf.Multigrid(A)
        b = smf.DeviceVector.from_numpy(smf.assemble_rhs(A, "grf", seed=6))
        st = smf.solve_cg(A, A.initialize_dof_vector(), b, mg, rel_tol=1e-10, max_iter=100)
        assert st["converged"]
        its.append(st["iterations"])
    assert its[1] <= its[0] + 3, its


def test_dg_sine_discretisation_error(gpu):
    """SIPG P2 with the sine right-hand side converges at order ~3 in the
    max norm at the nodes (exact solution prod sin(3 pi x_d))."""
    smf = gpu
    errs = []
    for n in (6, 12):
        A = smf.Operator(smf.Mesh.kuhn_cube(n, perturb=0.0), 2, kind="dg")
        mg = smf.Multigrid(A)
        b = smf.DeviceVector.from_numpy(smf.assemble_rhs(A, "sine"))
        x = A.initialize_dof_vector()
        assert smf.solve_cg(A, x, b, mg, rel_tol=1e-12, max_iter=100)["converged"]
        errs.append(np.max(np.abs(x.numpy() - smf.exact_solution(A))))
    rate = np.log2(errs[0] / errs[1])
    assert rate > 2.5, (errs, rate)


def test_dg_learned_smoother(dgmg):
    smf, _, A, _ = dgmg
    mg2 = smf.Multigrid(A, smoother_degree=2)
    ls = smf.LearnedSmoother(A, seed=3)
    st = ls.train(mg2, n_rhs=12, pairs_per_rhs=8, steps_per_pair=2, loss=0)
    assert st["last_loss"] < 0.8 * st["first_loss"]
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(A, "grf", seed=99))
    xp, xl = A.initialize_dof_vector(), A.initialize_dof_vector()
    sp = smf.solve_fcg_learned(A, xp, b, mg2, None, 1e-10, 200)
    sl = smf.solve_fcg_learned(A, xl, b, mg2, ls, 1e-10, 200)
    assert sp["converged"] and sl["converged"] and sl["fallback"] == 0
    assert sl["iterations"] <= sp["iterations"]
    ls.set_params(np.zeros(smf.LS_N_PARAMS, np.float32))
    s0 = smf.solve_fcg_learned(A, xl, b, mg2, ls, 1e-10, 200)
    assert s0["converged"] and s0["fallback"] == 1


@pytest.mark.parametrize("dg_kernel", [0, 2])
def test_dg_penalty_variant_gpu(gpu, orc, dg_kernel):
    smf = gpu
    mesh = smf.Mesh.kuhn_cube(6, perturb=0.2, seed=4)
    A = smf.Operator(mesh, 2, kind="dg", penalty_kind=1, dg_kernel=dg_kernel)
    x = np.random.default_rng(1).uniform(-1, 1, A.m())
    y = A.initialize_dof_vector()
    A.vmult(y, smf.DeviceVector.from_numpy(x))
    want = orc.DGOracleProblem.from_operator(A, penalty_kind=1).vmult(x)
    assert np.max(np.abs(y.numpy() - want)) <= 1e-12 * np.max(np.abs(want))
    mg = smf.Multigrid(A)  # the CG (Nitsche) levels inherit the penalty variant
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(A, "grf", seed=3))
    st = smf.solve_cg(A, A.initialize_dof_vector(), b, mg, rel_tol=1e-10, max_iter=100)
    assert st["converged"] and st["iterations"] <= 25


def test_penalty_variant_rejected_by_table_kernel(gpu):
    smf = gpu
    A = smf.Operator(smf.Mesh.kuhn_cube(2), 2, kind="dg", penalty_kind=1, dg_kernel=1)
    with pytest.raises(smf.MFError, match="penalty_kind"):
        A.vmult(A.initialize_dof_vector(), A.initialize_dof_vector())


def test_dg_fused_smoother_matches_unfused(gpu):
    """The four-lane DG kernel applies the Chebyshev step and the residual in
    its epilogue (double-buffered iterate); the one-lane kernel takes the
    separate vmult + vector-update path.  One V-cycle of each agrees to FP32
    rounding, and repeated V-cycles are bitwise reproducible."""
    smf = gpu
    mesh = smf.Mesh.kuhn_cube(6, perturb=0.2, seed=4)
    rng = np.random.default_rng(8)
    r = None
    zs = []
    for k in (2, 3):
        A = smf.Operator(mesh, 2, kind="dg", dg_kernel=k)
        mg = smf.Multigrid(A)
        if r is None:
            r = smf.DeviceVector.from_numpy(rng.uniform(-1, 1, A.m()))
        z1, z2 = A.initialize_dof_vector(), A.initialize_dof_vector()
        mg.vmult(z1, r)
        mg.vmult(z2, r)
        assert np.array_equal(z1.numpy(), z2.numpy())
        zs.append(z1.numpy())
    assert np.max(np.abs(zs[1] - zs[0])) <= 1e-5 * np.max(np.abs(zs[0]))


@pytest.mark.parametrize("degree", [1, 2])
def test_nitsche_cheby_run_matches_per_step(gpu, degree):
    """The CG (Nitsche) levels of the DG hierarchy in Chebyshev runs: the
    boundary terms are added per cell inside the kernel from the staged (and
    pending-updated) values, against the per-step path with the boundary-row
    kernel.  V-cycles agree to FP32 rounding, runs are reproducible and the
    solves take the same number of iterations."""
    smf = gpu
    print(f"\ncase: Nitsche levels, Chebyshev run vs per-step, DG P{degree} n=8", flush=True)
    A = smf.Operator(smf.Mesh.kuhn_cube(8, perturb=0.2, seed=6), degree, kind="dg")
    mg_run = smf.Multigrid(A, fuse_steps=True)
    mg_ref = smf.Multigrid(A, fuse_steps=False)
    b = smf.DeviceVector.from_numpy(np.random.default_rng(9).uniform(-1, 1, A.m()))
    for _ in range(2):
        z1, z2 = A.initialize_dof_vector(), A.initialize_dof_vector()
        mg_run.vmult(z1, b)
        mg_ref.vmult(z2, b)
        assert np.max(np.abs(z1.numpy() - z2.numpy())) <= 1e-5 * np.max(np.abs(z2.numpy()))
    z3 = A.initialize_dof_vector()
    mg_run.vmult(z3, b)
    assert np.array_equal(z1.numpy(), z3.numpy())
    x1, x2 = A.initialize_dof_vector(), A.initialize_dof_vector()
    s1 = smf.solve_cg(A, x1, b, mg_run, rel_tol=1e-10)
    s2 = smf.solve_cg(A, x2, b, mg_ref, rel_tol=1e-10)
    assert s1["converged"] and s2["converged"]
    assert abs(s1["iterations"] - s2["iterations"]) <= 1
    assert np.max(np.abs(x1.numpy() - x2.numpy())) <= 1e-8 * np.max(np.abs(x2.numpy()))
