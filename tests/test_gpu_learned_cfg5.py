"""GPU: the learned-smoother flexible CG as configured -- CG P2 and DG P2 on
48^3 x 6 perturbed Kuhn tetrahedra, random-field right-hand sides, with the
committed networks in simplex-benchmarks_amd/weights/ (trained on the GPU by
train_smoother.py on 4096 right-hand sides).

Checks, for an unseen right-hand side:
  * the learned step stays on (no fallback) and the solve converges to 1e-10;
  * CG: it takes fewer iterations than the same V-cycle without the step;
    DG: no more (the DG network gains nothing at 48^3: 15 against 15, see
    its training metadata and the design notes);
  * its error against an FP64 reference solution is within 2x of the plain
    solver's, in the energy norm.  The reference solution is a 1e-13 plain
    multigrid solve whose residual is checked with the oracle's operator (one
    OpenMP oracle product at 48^3).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "simplex-benchmarks_amd",
                       "weights")


def _energy(smf, A, e):
    de = smf.DeviceVector.from_numpy(e)
    ae = A.initialize_dof_vector()
    A.vmult(ae, de)
    return float(np.sqrt(max(e @ ae.numpy(), 0.0)))


@pytest.mark.parametrize("kind,name", [("cg", "learned_smoother_p2_cheb3"), ("dg", "learned_smoother_p2_dg_cheb3")])
def test_committed_network_at_48cube(gpu, orc, kind, name):
    smf = gpu
    meta = json.load(open(os.path.join(WEIGHTS, name + ".json")))
    assert meta["rhs"] == 4096 and meta.get("kind", "cg") == kind
    degree = meta["chebyshev_degree"]
    print(f"\ncase: committed {name} ({kind} P2 48^3, Chebyshev({degree}))", flush=True)
    A = smf.Operator(smf.Mesh.kuhn_cube(48, 0.2, 7), 2, kind=kind, precision="f64")
    ls = smf.LearnedSmoother(A, seed=0)
    with open(os.path.join(WEIGHTS, name + ".bin"), "rb") as f:
        ls.load(f.read())
    mg = smf.Multigrid(A, smoother_degree=degree)
    bh = smf.assemble_rhs(A, "grf", seed=90001)
    b = smf.DeviceVector.from_numpy(bh)
    xr = A.initialize_dof_vector()
    sr = smf.solve_cg(A, xr, b, smf.Multigrid(A), rel_tol=1e-13, max_iter=200)
    assert sr["converged"]
    xref = xr.numpy()
    pr = orc.OracleProblem.from_operator(A) if kind == "cg" else orc.DGOracleProblem.from_operator(A)
    res = bh - pr.vmult_omp(xref)
    assert np.linalg.norm(res) <= 1e-11 * np.linalg.norm(bh)
    xp, xl = A.initialize_dof_vector(), A.initialize_dof_vector()
    sp = smf.solve_fcg_learned(A, xp, b, mg, None, 1e-10, 200)
    sl = smf.solve_fcg_learned(A, xl, b, mg, ls, 1e-10, 200)
    ep, el = _energy(smf, A, xp.numpy() - xref), _energy(smf, A, xl.numpy() - xref)
    print(f"{kind}: plain {sp['iterations']} it, learned {sl['iterations']} it, mean omega "
          f"{sl['mean_omega']:.3f}, energy error plain {ep:.3e} learned {el:.3e}", flush=True)
    assert sp["converged"] and sl["converged"]
    assert sl["fallback"] == 0 and sl["learned_steps"] == sl["iterations"]
    if kind == "cg":
        assert sl["iterations"] < sp["iterations"]
    else:
        assert sl["iterations"] <= sp["iterations"]
    assert el <= 2.0 * ep
