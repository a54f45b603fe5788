"""The drop-in adaptor's one-time renumbering (CPU): a caller that numbers
its DoFs its own way (here as a DoFHandler does: input cells in order, its
own local node order, a new id at the first encounter of a DoF) builds the
permutation caller -> library from mf_get_cell_order, mf_get_cell_dofs and
the support points (mf_get_dof_points), the renumbering that
DoFRenumbering / DoFHandler::renumber_dofs would apply.  Checks: it is a
bijection, and the library's operator, conjugated by it, equals the oracle's
operator assembled directly in the caller's numbering (CG with Dirichlet
rows, and DG)."""
import numpy as np
import pytest

from conftest import rel_err


def caller_numbering(op, node_order, points_of_cell):
    """caller cell -> caller DoF ids (first-encounter numbering over input cells)"""
    ids, table = {}, []
    for c in range(op.mesh.cells.shape[0]):
        row = []
        for x in points_of_cell(c)[node_order]:
            key = tuple(np.round(x * 1e9).astype(np.int64))
            if key not in ids:
                ids[key] = len(ids)
            row.append(ids[key])
        table.append(row)
    return np.array(table), ids


def renumbering(op, caller_table, caller_points):
    """new_numbers[caller dof] = library dof, from the library's per-cell DoFs
    (internal cell k is input cell perm[k]) matched by support point"""
    perm = op.cell_order()
    nl = op.info["dofs_per_cell"]
    lib_dofs = op.cell_dofs().reshape(-1, nl)
    X = op.dof_points()
    new = -np.ones(caller_table.max() + 1, np.int64)
    for k, c in enumerate(perm):
        lp = {tuple(np.round(X[d] * 1e9).astype(np.int64)): d for d in lib_dofs[k]}
        for j, cd in enumerate(caller_table[c]):
            d = lp[tuple(np.round(caller_points[c][j] * 1e9).astype(np.int64))]
            assert new[cd] in (-1, d)
            new[cd] = d
    return new


def cell_points(op, degree):
    """support points of input cell c in a caller's own local node order:
    vertices in input order, then the lattice points of higher degree"""
    import itertools
    V, Cc = op.mesh.vertices, op.mesh.cells
    bary = [np.array(a) / degree for a in itertools.product(range(degree + 1), repeat=4) if sum(a) == degree]

    def pts(c):
        return np.array([b @ V[Cc[c]] for b in bary])
    return pts


@pytest.mark.parametrize("kind,degree", [("cg", 2), ("cg", 3), ("dg", 2)])
def test_caller_numbering_round_trips(smf, orc, kind, degree):
    op = smf.Operator(smf.Mesh.kuhn_cube(3, 0.2, 4).scrambled(seed=2), degree, kind=kind, host_only=True)
    pts = cell_points(op, degree)
    nl = op.info["dofs_per_cell"]
    order = np.random.default_rng(1).permutation(nl)  # the caller's local node order
    table, _ = caller_numbering(op, order, pts) if kind == "cg" else (None, None)
    if kind == "dg":
        # DG: every cell owns its DoFs, numbered cell by cell in the caller's node order
        table = np.arange(op.mesh.cells.shape[0] * nl).reshape(-1, nl)
    caller_pts = [pts(c)[order] for c in range(op.mesh.cells.shape[0])]
    new = renumbering(op, table, caller_pts)
    assert np.array_equal(np.sort(new), np.arange(op.m()))
    x = np.random.default_rng(3).uniform(-1, 1, op.m())  # a caller vector
    xl = np.empty(op.m())
    xl[new] = x                                          # to the library numbering
    yl = op.host_apply(xl)                               # the library's vmult (host schedule)
    y_caller = yl[new]                                   # back to the caller numbering
    # the oracle assembled directly in the caller's numbering
    if kind == "cg":
        con = op.constrained().astype(bool)[new]
        want = _oracle_in_caller_numbering(orc, op, degree, table, caller_pts, con, x)
    else:
        want = _dg_oracle_in_caller_numbering(orc, op, degree, table, caller_pts, x)
    assert rel_err(y_caller, want) < 1e-12


def _oracle_in_caller_numbering(orc, op, degree, table, caller_pts, con, x):
    # the oracle's local node i of input cell c sits at its own support point;
    # give it the caller's DoF id of that point
    nl = table.shape[1]
    V, Cc = op.mesh.vertices, op.mesh.cells
    ocd = np.empty_like(table)
    for c in range(Cc.shape[0]):
        J = np.stack([V[Cc[c, t]] - V[Cc[c, 0]] for t in (1, 2, 3)], axis=1)
        own = {tuple(np.round(p * 1e9).astype(np.int64)): table[c][j] for j, p in enumerate(caller_pts[c])}
        for i in range(nl):
            xr = _oracle_node(degree, i)
            ocd[c, i] = own[tuple(np.round((V[Cc[c, 0]] + J @ xr) * 1e9).astype(np.int64))]
    pr = orc.OracleProblem(V, Cc, ocd, con.astype(np.uint8), degree)
    return pr.vmult(x)


def _oracle_node(p, i):
    """reference position of the oracle's local node i: where its basis function is 1"""
    import itertools
    cand = [np.array(a[1:], float) / p for a in itertools.product(range(p + 1), repeat=4) if sum(a) == p]
    import oracle_py
    for x in cand:
        v, _ = oracle_py.basis(p, x)
        if abs(v[i] - 1) < 1e-12:
            return x
    raise AssertionError("node not found")


def _dg_oracle_in_caller_numbering(orc, op, degree, table, caller_pts, x):
    nl = table.shape[1]
    V, Cc = op.mesh.vertices, op.mesh.cells
    ocd = np.empty_like(table)
    for c in range(Cc.shape[0]):
        J = np.stack([V[Cc[c, t]] - V[Cc[c, 0]] for t in (1, 2, 3)], axis=1)
        own = {tuple(np.round(p * 1e9).astype(np.int64)): table[c][j] for j, p in enumerate(caller_pts[c])}
        for i in range(nl):
            ocd[c, i] = own[tuple(np.round((V[Cc[c, 0]] + J @ _oracle_node(degree, i)) * 1e9).astype(np.int64))]
    pr = orc.DGOracleProblem(V, Cc, degree)
    # DGOracleProblem numbers DoF (k, i) as k*nl + i over the given cells: permute into the caller's ids
    xo = x[ocd.ravel()]
    yo = pr.vmult(xo)
    y = np.empty(op.m())
    y[ocd.ravel()] = yo
    return y
