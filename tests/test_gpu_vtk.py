"""GPU parity on external meshes: a mesh written to and read back from a
legacy VTK file in an unstructured (scrambled) order, with and without the
hierarchical reordering (CG/cells_batched/application.cc:1699-1756), against
the oracle; and a multigrid solve on it (p-levels only, no lattice).

Bar: FP64 1e-12 relative max-norm error against the FP64 oracle; the solve
reaches 1e-10 relative residual.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext_mesh(gpu, tmp_path_factory):
    path = str(tmp_path_factory.mktemp("vtk") / "scrambled.vtk")
    gpu.Mesh.kuhn_cube(10, 0.2, 11).scrambled(seed=3).write_vtk(path, layout="vtk5")
    return gpu.Mesh.read_vtk(path)


@pytest.mark.parametrize("kind", ["cg", "dg"])
@pytest.mark.parametrize("reorder", [True, False])
@pytest.mark.parametrize("p", [1, 2])
def test_external_mesh_vmult_parity(gpu, orc, ext_mesh, kind, reorder, p):
    smf = gpu
    print(f"\ncase: external mesh {kind} p={p} reorder={reorder}", flush=True)
    op = smf.Operator(ext_mesh, p, kind=kind, reorder=reorder)
    pr = (orc.OracleProblem if kind == "cg" else orc.DGOracleProblem).from_operator(op)
    x = np.random.default_rng(p).uniform(-1, 1, op.m())
    src = smf.DeviceVector.from_numpy(x)
    dst = op.initialize_dof_vector()
    op.vmult(dst, src)
    assert rel_err(dst.numpy(), pr.vmult(x)) < 1e-12


def test_external_mesh_mg_solve(gpu, ext_mesh):
    smf = gpu
    print("\ncase: external mesh MG-CG p=2", flush=True)
    op = smf.Operator(ext_mesh, 2)
    mg = smf.Multigrid(op)
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(op, "grf", seed=1))
    x = op.initialize_dof_vector()
    st = smf.solve_cg(op, x, b, mg, rel_tol=1e-10)
    assert st["converged"] and st["iterations"] < 40
    r = op.initialize_dof_vector()
    op.vmult(r, x)
    bb = b.numpy()
    assert np.linalg.norm(bb - r.numpy()) <= 2e-10 * np.linalg.norm(bb)
