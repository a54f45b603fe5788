"""Smoothed-aggregation AMG coarse solver (the reference's MGCoarseAMG, one
ML V-cycle, CG/cells_batched/application_hp.cc:1684-1731, 1918-1930): the
coarse solve of the hybrid multigrid on meshes without a Kuhn lattice,
whose coarsest geometric level (P1 on the fine mesh) is too large for the
dense solver.

Bar: the preconditioned CG reaches 1e-10 relative residual (true residual
checked with the operator), in a bounded number of iterations that does not
grow much with the mesh size; results are bitwise repeatable.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def solve(smf, op, mg):
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(op, "grf", seed=3))
    x = op.initialize_dof_vector()
    st = smf.solve_cg(op, x, b, mg, rel_tol=1e-10)
    r = op.initialize_dof_vector()
    op.vmult(r, x)
    bb = b.numpy()
    res = np.linalg.norm(bb - r.numpy()) / np.linalg.norm(bb)
    return st, res, x.numpy()


@pytest.mark.parametrize("n", [16, 24])
def test_amg_coarse_on_external_mesh(gpu, n):
    smf = gpu
    print(f"\ncase: AMG coarse solve, scrambled {n}^3 mesh, P2", flush=True)
    mesh = smf.Mesh.kuhn_cube(n, 0.2, 5).scrambled(seed=1)
    op = smf.Operator(mesh, 2)
    mg = smf.Multigrid(op)
    info = mg.info()
    rows = info["amg_rows"]
    assert info["n_dofs"][0] == (n + 1) ** 3 > 4096
    assert len(rows) >= 2 and rows[0] == info["n_dofs"][0] and rows[-1] <= 4096
    assert all(a > b for a, b in zip(rows, rows[1:]))
    st, res, x1 = solve(smf, op, mg)
    assert st["converged"] and st["iterations"] <= 20
    assert res <= 2e-10
    st2, _, x2 = solve(smf, op, mg)
    assert st2["iterations"] == st["iterations"] and np.array_equal(x1, x2)


def test_amg_alone_preconditions_p1(gpu):
    """P1 on a scrambled 24^3 mesh: the hierarchy is the P1 level only and
    every V-cycle is one AMG V-cycle on its matrix."""
    smf = gpu
    print("\ncase: AMG-only preconditioner, scrambled 24^3 mesh, P1", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(24, 0.2, 2).scrambled(seed=4), 1)
    mg = smf.Multigrid(op)
    assert mg.info()["n_levels"] == 1 and len(mg.info()["amg_rows"]) >= 2
    st, res, _ = solve(smf, op, mg)
    assert st["converged"] and st["iterations"] <= 30
    assert res <= 2e-10


def test_coarse_solver_options(gpu):
    smf = gpu
    print("\ncase: coarse solver options on a Kuhn 16^3 mesh", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(16, 0.2, 6), 2)
    dense = smf.Multigrid(op, coarse_solver="dense")
    amg = smf.Multigrid(op, coarse_solver="amg")
    assert dense.info()["amg_rows"] == [] and len(amg.info()["amg_rows"]) >= 1
    sd, rd, _ = solve(smf, op, dense)
    sa, ra, _ = solve(smf, op, amg)
    assert sd["converged"] and sa["converged"] and rd <= 2e-10 and ra <= 2e-10
    assert abs(sa["iterations"] - sd["iterations"]) <= 3
    with pytest.raises(smf.MFError, match="dense coarse solver"):
        smf.Multigrid(smf.Operator(smf.Mesh.kuhn_cube(16, 0.2, 5).scrambled(seed=1), 2), coarse_solver="dense")


def test_amg_coarse_under_dg(gpu):
    """DG SIPG P2 on a scrambled mesh: DG -> CG (Nitsche) P2 -> P1, whose
    Nitsche matrix (no identity rows) the AMG coarsens."""
    smf = gpu
    print("\ncase: AMG coarse solve under DG, scrambled 16^3 mesh, P2", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(16, 0.2, 7).scrambled(seed=2), 2, kind="dg")
    mg = smf.Multigrid(op)
    info = mg.info()
    assert info["nitsche"][0] == 1 and len(info["amg_rows"]) >= 2
    st, res, _ = solve(smf, op, mg)
    assert st["converged"] and st["iterations"] <= 30
    assert res <= 2e-10


@pytest.mark.parametrize("theta", [0.05, 0.9])
def test_amg_vcycle_matches_oracle_restatement(gpu, orc, theta):
    """P1 on a scrambled 20^3 mesh (9,261 rows): the hierarchy is the P1
    level only, so one FP64 multigrid application is one AMG V-cycle; it
    matches the oracle's numpy restatement built from the same assembled
    matrix (itself checked against the oracle's dense matrix).  theta = 0.05
    is the default threshold; theta = 0.9 isolates every row, so every level
    takes the stall retries (theta / 4 until it coarsens)."""
    smf = gpu
    print(f"\ncase: AMG V-cycle vs oracle restatement, scrambled 20^3, P1, theta {theta}", flush=True)
    op = smf.Operator(smf.Mesh.kuhn_cube(20, 0.2, 4).scrambled(seed=6), 1)
    mg = smf.Multigrid(op, level_precision="f64", amg_threshold=theta)
    assert mg.info()["n_levels"] == 1
    ptr, col, val = smf.SystemMatrix(op).to_host()
    levels = orc.amg_hierarchy(ptr, col, val, max_dense=4096, theta=theta)
    assert [L["A"].shape[0] for L in levels] == mg.info()["amg_rows"]
    r = np.random.default_rng(9).uniform(-1, 1, op.m())
    z = op.initialize_dof_vector()
    mg.vmult(z, smf.DeviceVector.from_numpy(r))
    ref = orc.amg_vcycle(levels, r)
    assert np.max(np.abs(z.numpy() - ref)) <= 1e-10 * np.abs(ref).max()
