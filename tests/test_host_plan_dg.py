"""DG host plan (canonical vertex order, neighbours, geometry, penalty,
tables) against the DG oracle, via the host execution of the kernel's per-cell
algorithm (CPU only)."""
import numpy as np
import pytest

from conftest import rel_err

TOL = {"f64": 1e-13, "f32": 3e-6}


@pytest.mark.parametrize("dg_kernel", [2, 3, 4])
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("kind", ["dg", "dg_vector", "dg_helmholtz"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_host_dg_matches_oracle(smf, orc, p, kind, precision, dg_kernel):
    """dg_kernel 2: one lane per cell; 3: four lanes per cell (one face each,
    the lane's share of the cell rule), summed in lane order."""
    op = smf.Operator(smf.Mesh.kuhn_cube(3, 0.2, 2), p, kind=kind, precision=precision,
                      host_only=True, mass_coefficient=1.0, dg_kernel=dg_kernel)
    pr = orc.DGOracleProblem.from_operator(op, mass=1.0 if kind == "dg_helmholtz" else 0.0)
    x = np.random.default_rng(4).uniform(-1, 1, op.m())
    assert rel_err(op.host_apply(x).astype(float), pr.vmult(x)) < TOL[precision]
    assert rel_err(op.host_apply(None, mode=1).astype(float), pr.diagonal()) < TOL[precision]


def test_dg_sizes(smf):
    op = smf.Operator(smf.Mesh.kuhn_cube(4, 0.2, 1), 2, kind="dg", host_only=True)
    n = 4
    assert op.info["n_cells"] == 6 * n ** 3 and op.m() == 10 * 6 * n ** 3
    assert op.info["n_boundary_faces"] == 12 * n ** 2
    assert op.info["n_interior_faces"] == (4 * 6 * n ** 3 - 12 * n ** 2) // 2
    assert op.info["n_constrained"] == 0


@pytest.mark.parametrize("dg_kernel", [2, 3, 4])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_host_dg_penalty_variant(smf, orc, p, dg_kernel):
    """penalty_kind 1: sigma = (p+1)(p+2)/h_F with h_F = |K|/|F| per side (the
    config's variant) against the oracle with the same penalty; it differs
    from the reference penalty."""
    m = smf.Mesh.kuhn_cube(3, 0.2, 2)
    op = smf.Operator(m, p, kind="dg", host_only=True, penalty_kind=1, dg_kernel=dg_kernel)
    ref = smf.Operator(m, p, kind="dg", host_only=True, dg_kernel=dg_kernel)
    pr = orc.DGOracleProblem.from_operator(op, penalty_kind=1)
    x = np.random.default_rng(5).uniform(-1, 1, op.m())
    y = op.host_apply(x)
    assert rel_err(y, pr.vmult(x)) < 1e-13
    assert rel_err(op.host_apply(None, mode=1), pr.diagonal()) < 1e-13
    assert rel_err(y, ref.host_apply(x)) > 1e-3
