"""The CPU baseline (oracle/cpu_baseline.cpp), the reference's optimised DG
SIPG vmult restated for CPUs, against the oracle (CPU only): its quadrature
rules are exact to the degrees the operator needs (stiffness 2p-2, face
terms 2p), and its product equals the oracle's to 1e-12 for P1-P3 on
perturbed meshes, with any thread count."""
import math

import numpy as np
import pytest

from conftest import rel_err


def _tri_moment(i, j):
    return math.factorial(i) * math.factorial(j) / math.factorial(i + j + 2) * 2  # normalised to area 1


def _tet_moment(i, j, k):
    return math.factorial(i) * math.factorial(j) * math.factorial(k) / math.factorial(i + j + k + 3) * 6


@pytest.mark.parametrize("p", [1, 2, 3])
def test_rules_exact(orc, p):
    bc, w = orc.dgb_rule(0, p)
    assert abs(w.sum() - 1) < 1e-14
    for deg in range(2 * p - 1):
        for i in range(deg + 1):
            for j in range(deg + 1 - i):
                k = deg - i - j
                got = np.sum(w * bc[:, 1] ** i * bc[:, 2] ** j * bc[:, 3] ** k)
                assert abs(got - _tet_moment(i, j, k)) < 1e-14, (p, i, j, k)
    bc, w = orc.dgb_rule(1, p)
    assert abs(w.sum() - 1) < 1e-14
    for deg in range(2 * p + 1):
        for i in range(deg + 1):
            j = deg - i
            got = np.sum(w * bc[:, 1] ** i * bc[:, 2] ** j)
            assert abs(got - _tri_moment(i, j)) < 1e-14, (p, i, j)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_baseline_matches_oracle(smf, orc, p):
    op = smf.Operator(smf.Mesh.kuhn_cube(3 if p < 3 else 2, 0.25, 5), p, kind="dg", host_only=True)
    pr = orc.DGOracleProblem.from_operator(op)
    cpu = orc.DGBaseline(op.mesh.vertices, op.cell_vertices(), p, nthreads=4)
    x = np.random.default_rng(p).uniform(-1, 1, op.m())
    want = pr.vmult(x)
    assert rel_err(cpu.vmult(x), want) < 1e-12
    one = orc.DGBaseline(op.mesh.vertices, op.cell_vertices(), p, nthreads=1)
    assert rel_err(one.vmult(x), want) < 1e-12


def test_baseline_rejects_unsorted_cells(smf, orc):
    m = smf.Mesh.kuhn_cube(2, 0.2, 1)
    with pytest.raises(ValueError):
        orc.DGBaseline(m.vertices, m.cells, 2)
