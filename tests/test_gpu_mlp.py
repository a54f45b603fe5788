"""Learned per-cell smoother on the GPU, through the C ABI: BF16-MFMA forward
and FP32 backward against the CPU oracle, Adam, state blob, features and
assembly, GPU training, and flexible CG with the learned step, its fallback
and its gain over the same solver without it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP = 6154


@pytest.fixture(scope="module")
def setup(gpu):
    smf = gpu
    mesh = smf.Mesh.kuhn_cube(8, perturb=0.2, seed=7)
    A = smf.Operator(mesh, 2, kind="cg", precision="f64")
    return smf, mesh, A


def rows(rng, n):
    return rng.uniform(-1, 1, (n, 20)).astype(np.float32)


def test_forward_mfma_vs_oracle(setup, orc):
    smf, _, A = setup
    ls = smf.LearnedSmoother(A, seed=11)
    p = ls.params().astype(np.float64)
    rng = np.random.default_rng(0)
    for n in (1, 15, 16, 17, 64, 1000):  # ragged tiles and groups
        x = rows(rng, n)
        y = smf.DeviceVector(n * 10, np.float32)
        ls.forward(smf.DeviceVector.from_numpy(x.ravel()), n, y)
        got = y.numpy().reshape(n, 10)
        want_bf = orc.mlp_forward(p, x, bf16=True)
        want = orc.mlp_forward(p, x)
        scale = np.max(np.abs(want))
        # BF16 inputs/weights, FP32 accumulation: the BF16 oracle differs only
        # by accumulation order, the erf approximation and rare ties
        assert np.max(np.abs(got - want_bf)) <= 2e-3 * scale, n
        assert np.max(np.abs(got - want)) <= 3e-2 * scale, n


def test_gradient_vs_oracle(setup, orc):
    smf, _, A = setup
    ls = smf.LearnedSmoother(A, seed=12)
    p = ls.params().astype(np.float64)
    # larger output layer so every layer's gradient is well above FP32 noise
    p[1280 + 64 + 4096 + 64:1280 + 64 + 4096 + 64 + 640] *= 10
    ls.set_params(p.astype(np.float32))
    p = ls.params().astype(np.float64)
    rng = np.random.default_rng(1)
    for n in (1, 63, 64, 65, 3000):
        x = rows(rng, n)
        t = rng.normal(size=(n, 10)).astype(np.float32)
        loss, g = ls.gradient(smf.DeviceVector.from_numpy(x.ravel()), smf.DeviceVector.from_numpy(t.ravel()), n)
        lo, go = orc.mlp_loss_grad(p, x, t)
        assert loss == pytest.approx(lo, rel=1e-5)
        assert np.max(np.abs(g - go)) <= 1e-5 * np.max(np.abs(go)), n


def test_gradient_deterministic(setup):
    smf, _, A = setup
    ls = smf.LearnedSmoother(A, seed=13)
    rng = np.random.default_rng(2)
    x = smf.DeviceVector.from_numpy(rows(rng, 5000).ravel())
    t = smf.DeviceVector.from_numpy(rng.normal(size=50000).astype(np.float32))
    l1, g1 = ls.gradient(x, t, 5000)
    l2, g2 = ls.gradient(x, t, 5000)
    assert l1 == l2 and np.array_equal(g1, g2)


def test_adam_step_vs_oracle(setup, orc):
    smf, _, A = setup
    ls = smf.LearnedSmoother(A, seed=14)
    rng = np.random.default_rng(3)
    xh, th = rows(rng, 500), rng.normal(size=(500, 10)).astype(np.float32)
    x, t = smf.DeviceVector.from_numpy(xh.ravel()), smf.DeviceVector.from_numpy(th.ravel())
    p = ls.params().astype(np.float64)
    m, v = np.zeros(NP), np.zeros(NP)
    for step in (1, 2, 3):
        _, g = ls.gradient(x, t, 500)
        ls.train_step(x, t, 500, lr=1e-3)
        orc.adam(p, g.astype(np.float64), m, v, 1e-3, step)
        got = ls.params()
        assert np.max(np.abs(got - p)) <= 1e-6 * np.max(np.abs(p)), step
        p = got.astype(np.float64)  # keep both on the device's FP32 iterate


def test_state_blob_round_trip(setup):
    smf, _, A = setup
    ls = smf.LearnedSmoother(A, seed=15)
    rng = np.random.default_rng(4)
    x = smf.DeviceVector.from_numpy(rows(rng, 100).ravel())
    t = smf.DeviceVector.from_numpy(rng.normal(size=1000).astype(np.float32))
    ls.train_step(x, t, 100)
    blob = ls.save()
    assert blob[:8] == b"SMFMLP01" and len(blob) == 24 + 14 * NP
    other = smf.LearnedSmoother(A, seed=99)
    other.load(blob)
    assert np.array_equal(other.params(), ls.params())
    assert other.save() == blob
    bad = bytearray(blob)
    bad[-1] ^= 0x01  # BF16 copy no longer matches the FP32 master
    with pytest.raises(smf.MFError, match="BF16"):
        other.load(bytes(bad))
    with pytest.raises(smf.MFError):
        other.load(blob[:-2])


def test_features_scale_and_geometry(gpu):
    smf = gpu
    A = smf.Operator(smf.Mesh.kuhn_cube(4, perturb=0.0), 2, kind="cg", precision="f64")
    ls = smf.LearnedSmoother(A, seed=1)
    rng = np.random.default_rng(5)
    r = rng.uniform(-1, 1, A.m())
    x, _ = ls.features(smf.DeviceVector.from_numpy(r))
    x = x.numpy().reshape(-1, 20)
    cd = A.cell_dofs()
    con = A.constrained().astype(bool)
    rk = np.where(con[cd], 0.0, r[cd])
    s = np.linalg.norm(rk, axis=1)
    np.testing.assert_allclose(x[:, :10] * s[:, None], rk, rtol=1e-6, atol=1e-7)
    # unperturbed Kuhn cells: h J^-T has entries in {0, +-1}, |det J| = h^3
    np.testing.assert_allclose(np.abs(x[:, 10:19]) * (1 - np.abs(x[:, 10:19])), 0, atol=1e-6)
    np.testing.assert_allclose(x[:, 19], 1.0, rtol=1e-6)


def test_apply_constant_network(setup):
    """W3 = 0, b3 = v: the correction of cell K is v s_K kappa_K, averaged
    over the cells sharing each DoF (Dirichlet rows 0)."""
    smf, _, A = setup
    ls = smf.LearnedSmoother(A, seed=2)
    p = ls.params()
    v = np.linspace(-1, 1, 10).astype(np.float32)
    p[5504:6144] = 0
    p[6144:] = v
    ls.set_params(p)
    rng = np.random.default_rng(6)
    r = rng.uniform(-1, 1, A.m())
    out = A.initialize_dof_vector()
    ls.apply(out, smf.DeviceVector.from_numpy(r))
    got = out.numpy()
    inv = A.initialize_dof_vector()
    A.compute_inverse_diagonal(inv)
    dinv = inv.numpy()
    cd = A.cell_dofs()
    con = A.constrained().astype(bool)
    free = ~con[cd]
    rk = np.where(free, r[cd], 0.0)
    s = np.linalg.norm(rk, axis=1)
    kappa = np.array([dinv[c][f].mean() for c, f in zip(cd, free)])
    acc = np.zeros(A.m())
    cnt = np.zeros(A.m())
    corr = (v[None, :] * (s * kappa)[:, None]).astype(np.float64)
    np.add.at(acc, cd[free], corr[free])
    np.add.at(cnt, cd[free], 1)
    want = np.where(cnt > 0, acc / np.maximum(cnt, 1), 0.0)
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))
    assert np.all(got[con] == 0)


@pytest.fixture(scope="module")
def trained(setup):
    """Network trained on the GPU with the energy loss behind a degree-2
    Chebyshev V-cycle (pairs from plain multigrid CG solves)."""
    smf, _, A = setup
    mg = smf.Multigrid(A, smoother_degree=2)
    ls = smf.LearnedSmoother(A, seed=3)
    st = ls.train(mg, n_rhs=24, pairs_per_rhs=10, steps_per_pair=2, lr=1e-3, loss=0)
    return ls, mg, st


def test_training_lowers_energy_loss(trained):
    _, _, st = trained
    assert st["pairs"] > 24 and st["steps"] == 2 * st["pairs"]
    assert st["zero_loss"] == 1.0
    assert 0.9 < st["first_loss"] <= 1.0
    assert st["last_loss"] < 0.5 * st["first_loss"]


def test_training_mse_targets(setup):
    smf, _, A = setup
    mg = smf.Multigrid(A, smoother_degree=2)
    ls = smf.LearnedSmoother(A, seed=4)
    st = ls.train(mg, n_rhs=8, pairs_per_rhs=4, steps_per_pair=4, lr=1e-3, loss=1)
    assert st["last_loss"] < st["first_loss"]
    assert st["last_loss"] < st["zero_loss"]


def test_plain_fcg_matches_flexible_cg(setup):
    smf, _, A = setup
    mg = smf.Multigrid(A)
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(A, "grf", seed=21))
    x1, x2 = A.initialize_dof_vector(), A.initialize_dof_vector()
    s1 = smf.solve_cg(A, x1, b, mg, 1e-10, 100, flexible=True)
    s2 = smf.solve_fcg_learned(A, x2, b, mg, None, 1e-10, 100)
    assert s2["converged"] and s2["learned_steps"] == 0 and s2["fallback"] == 0
    assert abs(s1["iterations"] - s2["iterations"]) <= 1
    assert np.max(np.abs(x1.numpy() - x2.numpy())) <= 1e-8 * np.max(np.abs(x1.numpy()))


def test_learned_fcg_gain_and_error(setup, trained):
    smf, _, A = setup
    ls, mg, _ = trained
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(A, "grf", seed=777))  # unseen right-hand side
    xs = A.initialize_dof_vector()
    smf.solve_cg(A, xs, b, smf.Multigrid(A), 1e-13, 200)
    xref = xs.numpy()
    xp, xl = A.initialize_dof_vector(), A.initialize_dof_vector()
    sp = smf.solve_fcg_learned(A, xp, b, mg, None, 1e-10, 100)
    sl = smf.solve_fcg_learned(A, xl, b, mg, ls, 1e-10, 100)
    assert sp["converged"] and sl["converged"]
    assert sl["fallback"] == 0 and sl["learned_steps"] == sl["iterations"]
    # the trained step saves iterations over the same solver without it
    assert sl["iterations"] < sp["iterations"]
    for k in (3, 5):
        rp = smf.solve_fcg_learned(A, xp, b, mg, None, 1e-16, k)["final_residual"]
        rl = smf.solve_fcg_learned(A, xl, b, mg, ls, 1e-16, k)["final_residual"]
        assert rl < 0.5 * rp, k
    # error against the FP64 reference solution, energy norm |e|_A = sqrt(e.A e)
    xp, xl = A.initialize_dof_vector(), A.initialize_dof_vector()
    sp = smf.solve_fcg_learned(A, xp, b, mg, None, 1e-10, 100)
    sl = smf.solve_fcg_learned(A, xl, b, mg, ls, 1e-10, 100)
    ep, el = energy_error(smf, A, xp, xref), energy_error(smf, A, xl, xref)
    mp, ml = np.max(np.abs(xp.numpy() - xref)), np.max(np.abs(xl.numpy() - xref))
    print(f"energy error plain {ep:.3e} learned {el:.3e}; max error plain {mp:.3e} learned {ml:.3e}; "
          f"final residual plain {sp['final_residual']:.3e} learned {sl['final_residual']:.3e}")
    assert el <= 2 * ep


def energy_error(smf, A, x, xref):
    e = smf.DeviceVector.from_numpy(x.numpy() - xref)
    ae = A.initialize_dof_vector()
    A.vmult(ae, e)
    return float(np.sqrt(max(np.dot(e.numpy(), ae.numpy()), 0.0)))


def test_untrained_network_gains_less(setup, trained):
    smf, _, A = setup
    ls, mg, _ = trained
    raw = smf.LearnedSmoother(A, seed=3)
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(A, "grf", seed=778))
    x = A.initialize_dof_vector()
    rt = smf.solve_fcg_learned(A, x, b, mg, ls, 1e-16, 4)["final_residual"]
    ru = smf.solve_fcg_learned(A, x, b, mg, raw, 1e-16, 4)["final_residual"]
    assert rt < 0.5 * ru


def _solve(smf, A, mg, ls, seed=31, **kw):
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(A, "grf", seed=seed))
    x = A.initialize_dof_vector()
    return smf.solve_fcg_learned(A, x, b, mg, ls, 1e-10, 200, **kw), x


def test_fallback_zero_weights(setup):
    smf, _, A = setup
    mg = smf.Multigrid(A)
    ls = smf.LearnedSmoother(A, seed=5)
    ls.set_params(np.zeros(NP, np.float32))
    st, _ = _solve(smf, A, mg, ls)
    assert st["converged"] and st["fallback"] == 1 and st["fallback_iteration"] == 0
    assert st["learned_steps"] == 0


def test_fallback_nan_weights(setup):
    smf, _, A = setup
    mg = smf.Multigrid(A)
    ls = smf.LearnedSmoother(A, seed=6)
    p = ls.params()
    p[7] = np.nan
    ls.set_params(p)
    st, x = _solve(smf, A, mg, ls)
    assert st["converged"] and st["fallback"] == 1 and st["fallback_iteration"] == 0
    assert np.all(np.isfinite(x.numpy()))


def test_fallback_forced_and_stall(setup, trained):
    smf, _, A = setup
    ls, mg, _ = trained
    st, _ = _solve(smf, A, mg, ls, force_fallback_at=2)
    assert st["converged"] and st["fallback"] == 1 and st["fallback_iteration"] == 2
    assert st["learned_steps"] == 2
    # a stall threshold no iteration can meet: the detector fires after two iterations
    st, _ = _solve(smf, A, mg, ls, stall_ratio=1e-30)
    assert st["converged"] and st["fallback"] == 1 and st["fallback_iteration"] == 2
    st, _ = _solve(smf, A, mg, ls, use_learned=0)
    assert st["converged"] and st["learned_steps"] == 0 and st["fallback"] == 0


def test_iteration_cap(setup):
    smf, _, A = setup
    st, _ = _solve(smf, A, smf.Multigrid(A), None)
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(A, "grf", seed=31))
    x = A.initialize_dof_vector()
    s2 = smf.solve_fcg_learned(A, x, b, smf.Multigrid(A), None, 1e-10, 2)
    assert s2["iterations"] == 2 and not s2["converged"]
    assert st["iterations"] > 2
