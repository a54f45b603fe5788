"""ctypes loader of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")


class OrcMesh(C.Structure):
    _fields_ = [("n_vertices", C.c_int64), ("vertices", C.c_void_p),
                ("n_cells", C.c_int64), ("cells", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        for name in ["orc_cg_vmult", "orc_cg_diagonal", "orc_cg_matrix_vmult", "orc_cg_dense",
                     "orc_dg_vmult", "orc_dg_diagonal", "orc_dg_tau", "orc_dg_vmult_omp", "orc_dg_set_penalty_kind"]:
            getattr(_lib, name).restype = C.c_int
        _lib.orc_basis.argtypes = [C.c_int, vp, vp, vp]
        _lib.orc_tet_rule.argtypes = [C.c_int, vp, vp]
        _lib.orc_gauss01.argtypes = [C.c_int, vp, vp]
    return _lib


class OracleProblem:
    """A mesh in the operator's internal cell order plus its DoF table."""

    def __init__(self, vertices, cells, cell_dofs, constrained, degree):
        self.vertices = np.ascontiguousarray(vertices, np.float64)
        self.cells = np.ascontiguousarray(cells, np.int32)
        self.cell_dofs = np.ascontiguousarray(cell_dofs, np.int64)
        self.constrained = np.ascontiguousarray(constrained, np.uint8)
        self.degree = int(degree)
        self.n_dofs = int(self.constrained.size)
        self._m = OrcMesh(self.vertices.shape[0], self.vertices.ctypes.data,
                          self.cells.shape[0], self.cells.ctypes.data)

    @staticmethod
    def from_operator(op):
        return OracleProblem(op.mesh.vertices, op.cell_vertices(), op.cell_dofs(),
                             op.constrained(), op.info["degree"])

    def _call(self, fn, *args):
        rc = fn(C.byref(self._m), self.degree, C.c_void_p(self.cell_dofs.ctypes.data),
                C.c_void_p(self.constrained.ctypes.data), C.c_int64(self.n_dofs), *args)
        if rc < 0:
            raise RuntimeError(f"oracle call failed ({rc})")
        return rc

    def vmult(self, src):
        src = np.ascontiguousarray(src, np.float64)
        dst = np.empty(self.n_dofs)
        self._call(lib().orc_cg_vmult, C.c_void_p(src.ctypes.data), C.c_void_p(dst.ctypes.data))
        return dst

    def matrix_vmult(self, src):
        src = np.ascontiguousarray(src, np.float64)
        dst = np.empty(self.n_dofs)
        self._call(lib().orc_cg_matrix_vmult, C.c_void_p(src.ctypes.data), C.c_void_p(dst.ctypes.data))
        return dst

    def diagonal(self):
        d = np.empty(self.n_dofs)
        self._call(lib().orc_cg_diagonal, C.c_void_p(d.ctypes.data))
        return d

    def dense(self):
        A = np.empty((self.n_dofs, self.n_dofs))
        self._call(lib().orc_cg_dense, C.c_void_p(A.ctypes.data))
        return A

    def colors(self):
        cptr = np.zeros(258, np.int32)
        ccells = np.empty(self.cells.shape[0], np.int64)
        L = lib()
        n = L.orc_cg_colors(C.byref(self._m), self.degree, C.c_void_p(self.cell_dofs.ctypes.data),
                            C.c_int64(self.n_dofs), C.c_void_p(cptr.ctypes.data),
                            C.c_void_p(ccells.ctypes.data))
        if n < 0:
            raise RuntimeError("oracle colouring failed")
        return cptr[: n + 1].copy(), ccells, n

    def vmult_omp(self, src, colors=None, dst=None):
        if colors is None:
            colors = self.colors()
        cptr, ccells, n = colors
        src = np.ascontiguousarray(src, np.float64)
        if dst is None:
            dst = np.empty(self.n_dofs)
        self._call(lib().orc_cg_vmult_omp, C.c_void_p(cptr.ctypes.data),
                   C.c_void_p(ccells.ctypes.data), C.c_int(n), C.c_void_p(src.ctypes.data),
                   C.c_void_p(dst.ctypes.data))
        return dst


class DGOracleProblem:
    """DG SIPG operator (optionally vector / Helmholtz) on the operator's
    internal cells; rows (k*ncomp + c)*n_loc + i."""

    def __init__(self, vertices, cells, degree, ncomp=1, mass=0.0, laplace=1.0, penalty_kind=0):
        self.penalty_kind = int(penalty_kind)
        self.vertices = np.ascontiguousarray(vertices, np.float64)
        self.cells = np.ascontiguousarray(cells, np.int32)
        self.degree = int(degree)
        self.ncomp = int(ncomp)
        self.mass = float(mass)
        self.laplace = float(laplace)
        nl = (self.degree + 1) * (self.degree + 2) * (self.degree + 3) // 6
        nc = self.cells.shape[0]
        self.n_loc = nl
        self.n_scalar = nc * nl
        # oracle layout: component c of scalar DoF d is row c*n_scalar + d, with
        # d = k*nl + i; map to the operator layout (k*ncomp + c)*nl + i
        k = np.arange(nc)[:, None, None]
        c = np.arange(self.ncomp)[None, :, None]
        i = np.arange(nl)[None, None, :]
        self.op_rows = ((k * self.ncomp + c) * nl + i)            # [k, c, i]
        self.orc_rows = (c * self.n_scalar + k * nl + i)          # same entries, oracle layout
        self.cell_dofs = np.ascontiguousarray(np.arange(self.n_scalar, dtype=np.int64))
        self._m = OrcMesh(self.vertices.shape[0], self.vertices.ctypes.data, nc, self.cells.ctypes.data)

    @staticmethod
    def from_operator(op, mass=0.0, laplace=1.0, penalty_kind=0):
        return DGOracleProblem(op.mesh.vertices, op.cell_vertices(), op.info["degree"],
                               op.info["n_components"], mass, laplace, penalty_kind)

    def _to_orc(self, x):
        y = np.empty(x.size)
        y[self.orc_rows.ravel()] = x[self.op_rows.ravel()]
        return y

    def _from_orc(self, y):
        x = np.empty(y.size)
        x[self.op_rows.ravel()] = y[self.orc_rows.ravel()]
        return x

    def vmult(self, src):
        s = self._to_orc(np.asarray(src, np.float64))
        d = np.empty_like(s)
        lib().orc_dg_set_penalty_kind(self.penalty_kind)
        rc = lib().orc_dg_vmult(C.byref(self._m), self.degree, C.c_void_p(self.cell_dofs.ctypes.data),
                                C.c_int64(self.n_scalar), self.ncomp, C.c_double(self.mass),
                                C.c_double(self.laplace), C.c_void_p(s.ctypes.data), C.c_void_p(d.ctypes.data))
        if rc:
            raise RuntimeError("orc_dg_vmult failed")
        return self._from_orc(d)

    def vmult_omp(self, src):
        s = self._to_orc(np.asarray(src, np.float64))
        d = np.empty_like(s)
        lib().orc_dg_set_penalty_kind(self.penalty_kind)
        rc = lib().orc_dg_vmult_omp(C.byref(self._m), self.degree, C.c_void_p(self.cell_dofs.ctypes.data),
                                    C.c_int64(self.n_scalar), self.ncomp, C.c_double(self.mass),
                                    C.c_double(self.laplace), C.c_void_p(s.ctypes.data),
                                    C.c_void_p(d.ctypes.data))
        if rc:
            raise RuntimeError("orc_dg_vmult_omp failed")
        return self._from_orc(d)

    def diagonal(self):
        d = np.empty(self.n_scalar * self.ncomp)
        lib().orc_dg_set_penalty_kind(self.penalty_kind)
        rc = lib().orc_dg_diagonal(C.byref(self._m), self.degree, C.c_void_p(self.cell_dofs.ctypes.data),
                                   C.c_int64(self.n_scalar), self.ncomp, C.c_double(self.mass),
                                   C.c_double(self.laplace), C.c_void_p(d.ctypes.data))
        if rc:
            raise RuntimeError("orc_dg_diagonal failed")
        return self._from_orc(d)

    def tau(self):
        t = np.empty(self.cells.shape[0])
        lib().orc_dg_tau(C.byref(self._m), C.c_void_p(t.ctypes.data))
        return t

    def dense(self):
        n = self.n_scalar * self.ncomp
        A = np.empty((n, n))
        e = np.zeros(n)
        for j in range(n):
            e[j] = 1.0
            A[:, j] = self.vmult(e)
            e[j] = 0.0
        return A


def basis(p, x):
    nl = (p + 1) * (p + 2) * (p + 3) // 6
    x = np.ascontiguousarray(x, np.float64)
    v = np.empty(nl)
    g = np.empty((nl, 3))
    if lib().orc_basis(p, x.ctypes.data, v.ctypes.data, g.ctypes.data):
        raise RuntimeError("orc_basis failed")
    return v, g


def tet_rule(n):
    pts = np.empty((n ** 3, 3))
    w = np.empty(n ** 3)
    if lib().orc_tet_rule(n, pts.ctypes.data, w.ctypes.data):
        raise RuntimeError("orc_tet_rule failed")
    return pts, w


# ---- curved cells (numpy restatement) ----------------------------------------
# Follows CG/cells_batched/application.cc:1361-1376 (curvilinear D_e:
# g_q <- JxW_q jac_q^T (jac_q g_q), jac = J^-T) with the geometry of
# MappingFE(FE_SimplexP(2)) + SphericalManifold (:1613, 1668-1677): a P2 map
# per cell through its vertices and, per edge, the manifold's midpoint.
# Both bases are the oracle's own (Vandermonde inversion, orc_basis); the
# quadrature points and weights are an input (the operator's rule, checked
# separately for exactness), so the comparison is exact up to rounding.

EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def sphere_midpoint(xa, xb, center):
    """Geodesic midpoint on the spherical manifold around center: direction
    the normalised mean of the two unit directions, radius the mean radius
    (deal.II SphericalManifold intermediate point at weight 1/2)."""
    va, vb = xa - center, xb - center
    ra = np.linalg.norm(va, axis=-1, keepdims=True)
    rb = np.linalg.norm(vb, axis=-1, keepdims=True)
    e = va / ra + vb / rb
    return center + 0.5 * (ra + rb) * e / np.linalg.norm(e, axis=-1, keepdims=True)


def curved_nodes(vertices, cells, center=(5.0, 0.0, 0.0), straight=False):
    """(n_cells, 10, 3) P2 geometry nodes: 4 vertices, then the 6 edge nodes"""
    X = np.asarray(vertices, np.float64)[np.asarray(cells)]
    c = np.asarray(center, np.float64)
    mids = [0.5 * (X[:, a] + X[:, b]) if straight else sphere_midpoint(X[:, a], X[:, b], c) for a, b in EDGES]
    return np.concatenate([X, np.stack(mids, axis=1)], axis=1)


def curved_element_matrices(p, nodes, pts, w):
    """(n_cells, n_loc, n_loc): sum_q w_q |det J_q| (J_q^-T grad phi_i).(J_q^-T grad phi_j)"""
    gg = np.stack([basis(2, x)[1] for x in pts])          # (nq, 10, 3) geometry gradients
    gf = np.stack([basis(p, x)[1] for x in pts])          # (nq, nl, 3) FE gradients
    J = np.einsum("kid,qie->kqde", nodes, gg)              # J[k, q][d][e] = dx_d / dxi_e
    det = np.linalg.det(J)
    if np.any(det <= 0):
        raise ValueError("curved cell with non-positive Jacobian")
    Jit = np.linalg.inv(J).transpose(0, 1, 3, 2)           # J^-T
    grads = np.einsum("kqde,qie->kqid", Jit, gf)            # physical gradients
    return np.einsum("q,kq,kqid,kqjd->kij", np.asarray(w), np.abs(det), grads, grads)


def curved_vmult(p, vertices, cells, cell_dofs, constrained, pts, w, x, center=(5.0, 0.0, 0.0), straight=False):
    """A x with identity rows on constrained DoFs, constrained columns read as 0
    (CG/cells_batched/application.cc:1173-1175, 1287-1292)"""
    K = curved_element_matrices(p, curved_nodes(vertices, cells, center, straight), pts, w)
    con = np.asarray(constrained).astype(bool)
    xf = np.where(con, 0.0, x)
    ye = np.einsum("kij,kj->ki", K, xf[cell_dofs])
    y = np.zeros(len(x))
    np.add.at(y, cell_dofs.ravel(), ye.ravel())
    y[con] = x[con]
    return y


# ---- smoothed-aggregation AMG (numpy restatement) ----------------------------
# The coarse solver in place of the reference's ML V-cycle (MGCoarseAMG,
# CG/cells_batched/application_hp.cc:1684-1731): smoothed aggregation
# (Vanek, Mandel, Brezina 1996) with the build's stated choices -- strong
# couplings |a_ij| >= theta sqrt(|a_ii a_jj|), three aggregation passes in row order, unit-norm
# aggregate indicators, P = (I - 4/(3 lambda) D^-1 A) T with lambda from 30
# power iterations on D^-1 A from v_i = 1 + sin(0.7 i)/2, Galerkin A_c, a
# dense inverse at <= max_dense rows, Chebyshev-Jacobi smoothing (degree,
# range 30, upper bound 1.1 lambda) before and after the coarse correction.


def _csr_rows(ptr, col, val):
    return [(col[ptr[i]:ptr[i + 1]], val[ptr[i]:ptr[i + 1]]) for i in range(len(ptr) - 1)]


def amg_hierarchy(ptr, col, val, max_dense=4096, theta=0.0):
    """levels: list of dicts with A (dense-free scipy CSR), P, lam; last level has inv"""
    import scipy.sparse as sp
    n = len(ptr) - 1
    A = sp.csr_matrix((np.asarray(val, float), np.asarray(col), np.asarray(ptr)), shape=(n, n))
    levels = []
    while A.shape[0] > max_dense:
        n = A.shape[0]
        d = A.diagonal().copy()
        v = 1.0 + 0.5 * np.sin(0.7 * np.arange(n))
        lam = 0.0
        for _ in range(30):
            nv = np.sqrt(v @ v)
            v = v / nv
            w = (A @ v) / d
            lam = v @ w
            v = w
        th = theta
        while True:
            agg, na, rows, iso = _aggregate(A, d, th)
            if (na > 0 and na <= 0.9 * n) or th <= 0.0:
                break
            th = 0.25 * th if th > 1e-3 else 0.0  # stalled: weaker strength threshold
        if na == 0 or na > 0.9 * n:
            break
        member = agg >= 0
        size = np.bincount(agg[member], minlength=na)
        T = sp.csr_matrix((1.0 / np.sqrt(size[agg[member]]), (np.nonzero(member)[0], agg[member])), shape=(n, na))
        P = T - sp.diags((4.0 / (3.0 * lam)) / d) @ (A @ T)
        levels.append({"A": A, "P": P, "lam": 1.1 * lam, "dinv": 1.0 / d})
        A = (P.T @ (A @ P)).tocsr()
    if A.shape[0] > max_dense:
        raise RuntimeError(f"AMG: coarsening stalled at {A.shape[0]} rows")
    levels.append({"A": A, "inv": np.linalg.inv(A.toarray()), "dinv": 1.0 / A.diagonal()})
    return levels


def _aggregate(A, d, theta):
    """three aggregation passes in row order over the strong couplings"""
    n = A.shape[0]
    rows = _csr_rows(A.indptr, A.indices, A.data)
    # strong couplings: |a_ij| >= theta sqrt(|a_ii a_jj|), off the diagonal, nonzero
    rows = [(c[(c != i) & (x != 0) & (np.abs(x) >= theta * np.sqrt(np.abs(d[i] * d[c])))],
             x[(c != i) & (x != 0) & (np.abs(x) >= theta * np.sqrt(np.abs(d[i] * d[c])))])
            for i, (c, x) in enumerate(rows)]
    iso = np.array([len(c) == 0 for c, _ in rows])
    agg = -np.ones(n, np.int64)
    na = 0
    for i, (c, x) in enumerate(rows):  # pass 1
        if agg[i] >= 0 or iso[i]:
            continue
        nb = c
        if np.any(agg[nb] >= 0):
            continue
        agg[i] = na
        agg[nb] = na
        na += 1
    agg1 = agg.copy()
    for i, (c, x) in enumerate(rows):  # pass 2
        if agg[i] >= 0 or iso[i]:
            continue
        best, bv = -1, 0.0
        for j, a in zip(c, x):
            if agg1[j] >= 0 and abs(a) > bv:
                best, bv = agg1[j], abs(a)
        agg[i] = best
    for i, (c, x) in enumerate(rows):  # pass 3
        if agg[i] >= 0 or iso[i]:
            continue
        agg[i] = na
        for j, a in zip(c, x):
            if agg[j] < 0 and not iso[j]:
                agg[j] = na
        na += 1
    return agg, na, rows, iso


def amg_vcycle(levels, b, degree=5, smoothing_range=30.0, l=0):
    """x = one V-cycle applied to b from x = 0"""
    L = levels[l]
    if "inv" in L:
        return L["inv"] @ b
    A, dinv = L["A"], L["dinv"]
    lmax = L["lam"]
    lmin = lmax / smoothing_range
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta

    def cheby(x, zero):
        rho = 1.0 / sigma
        d = np.zeros_like(b)
        for k in range(degree):
            if k == 0:
                c1, c2 = 0.0, 1.0 / theta
                if zero:
                    d = c2 * dinv * b
                    x = d.copy()
                    continue
            else:
                rn = 1.0 / (2.0 * sigma - rho)
                c1, c2 = rn * rho, 2.0 * rn / delta
                rho = rn
            r = b - A @ x
            d = c2 * dinv * r if k == 0 else c1 * d + c2 * dinv * r
            x = x + d
        return x

    x = cheby(None, True)
    r = b - A @ x
    xc = amg_vcycle(levels, L["P"].T @ r, degree, smoothing_range, l + 1)
    x = x + L["P"] @ xc
    return cheby(x, False)


# ---- learned-smoother network (oracle_mlp.c) --------------------------------
def _mlp_lib():
    L = lib()
    if not getattr(L, "_mlp_ready", False):
        vp = C.c_void_p
        L.orc_mlp_n_params.restype = C.c_int
        L.orc_bf16.argtypes = [C.c_float]
        L.orc_bf16.restype = C.c_uint16
        L.orc_mlp_forward.argtypes = [vp, vp, C.c_int64, C.c_int, vp]
        L.orc_mlp_loss.argtypes = [vp, vp, vp, C.c_int64, vp]
        L.orc_mlp_loss_grad.argtypes = [vp, vp, vp, C.c_int64, vp, vp]
        L.orc_adam.argtypes = [vp, vp, vp, vp, C.c_int64, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_int64]
        L._mlp_ready = True
    return L


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def mlp_n_params():
    return _mlp_lib().orc_mlp_n_params()


def bf16_bits(f):
    return int(_mlp_lib().orc_bf16(float(f)))


def mlp_forward(params, x, bf16=False):
    p, x = _f64(params), _f64(x).reshape(-1, 20)
    y = np.empty((x.shape[0], 10))
    _mlp_lib().orc_mlp_forward(p.ctypes.data, x.ctypes.data, x.shape[0], int(bf16), y.ctypes.data)
    return y


def mlp_loss(params, x, t):
    p, x, t = _f64(params), _f64(x).reshape(-1, 20), _f64(t).reshape(-1, 10)
    l = np.zeros(1)
    _mlp_lib().orc_mlp_loss(p.ctypes.data, x.ctypes.data, t.ctypes.data, x.shape[0], l.ctypes.data)
    return float(l[0])


def mlp_loss_grad(params, x, t):
    p, x, t = _f64(params), _f64(x).reshape(-1, 20), _f64(t).reshape(-1, 10)
    g = np.empty(p.size)
    l = np.zeros(1)
    _mlp_lib().orc_mlp_loss_grad(p.ctypes.data, x.ctypes.data, t.ctypes.data, x.shape[0], g.ctypes.data,
                                 l.ctypes.data)
    return float(l[0]), g


def adam(p, g, m, v, lr, step, b1=0.9, b2=0.999, eps=1e-8):
    """In place on float64 arrays p, m, v."""
    g = _f64(g)
    _mlp_lib().orc_adam(p.ctypes.data, g.ctypes.data, m.ctypes.data, v.ctypes.data, p.size, lr, b1, b2, eps,
                        step)


# ---- CPU baseline (cpu_baseline.cpp) ------------------------------------------
class DGBaseline:
    """The reference's optimised DG SIPG vmult restated for CPUs (vector lanes
    of cells or faces, batches of four vectors per shape matrix, OpenMP).
    cells: (n, 4) in ascending vertex order; DoF (k, i) is row k*n_loc + i."""

    def __init__(self, vertices, cells, degree, nthreads=0):
        L = lib()
        L.orc_dgb_create.restype = C.c_void_p
        L.orc_dgb_create.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_dgb_vmult.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_dgb_time.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_dgb_info.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_dgb_destroy.argtypes = [C.c_void_p]
        self.vertices = np.ascontiguousarray(vertices, np.float64)
        self.cells = np.ascontiguousarray(cells, np.int32)
        self.degree = int(degree)
        self.n_loc = (degree + 1) * (degree + 2) * (degree + 3) // 6
        self.n = self.cells.shape[0] * self.n_loc
        self._m = OrcMesh(self.vertices.shape[0], self.vertices.ctypes.data, self.cells.shape[0],
                          self.cells.ctypes.data)
        self.plan = L.orc_dgb_create(C.byref(self._m), self.degree, int(nthreads))
        if not self.plan:
            raise ValueError("CPU baseline: degree 1..3 and ascending vertex order per cell required")

    def info(self):
        out = np.zeros(6, np.int64)
        lib().orc_dgb_info(self.plan, out.ctypes.data)
        return dict(zip(["n_q", "n_qf", "face_groups", "colours", "threads", "vector_bits"], map(int, out)))

    def vmult(self, src, dst=None):
        src = np.ascontiguousarray(src, np.float64)
        if dst is None:
            dst = np.empty(self.n)
        lib().orc_dgb_vmult(self.plan, src.ctypes.data, dst.ctypes.data)
        return dst

    def time(self, src, reps=5, n=100):
        """seconds of each of `reps` repetitions of `n` products"""
        src = np.ascontiguousarray(src, np.float64)
        dst = np.empty(self.n)
        sec = np.zeros(reps)
        lib().orc_dgb_time(self.plan, src.ctypes.data, dst.ctypes.data, reps, n, sec.ctypes.data)
        return sec

    def __del__(self):
        if getattr(self, "plan", None) and _lib is not None:
            _lib.orc_dgb_destroy(self.plan)
            self.plan = None


def dgb_rule(kind, p):
    L = lib()
    L.orc_dgb_rule.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    n = L.orc_dgb_rule(kind, p, None, None)
    nb = 4 if kind == 0 else 3
    bc = np.empty((n, nb))
    w = np.empty(n)
    L.orc_dgb_rule(kind, p, bc.ctypes.data, w.ctypes.data)
    return bc, w


# ---- planted mistakes (tests only) -------------------------------------------
class planted:
    """Context manager that switches a planted mistake into the DG oracle
    (1 min for max penalty, 2 average without 1/2, 3 lumped mass), so a test
    can show that its check fails on it."""

    def __init__(self, code):
        self.code = int(code)

    def __enter__(self):
        L = lib()
        L.orc_dg_set_plant.argtypes = [C.c_int]
        if L.orc_dg_set_plant(self.code):
            raise ValueError("unknown plant")
        return self

    def __exit__(self, *exc):
        lib().orc_dg_set_plant(0)
        return False


# ---- Gaussian random field right-hand side (numpy restatement) ---------------
# f(x) = sqrt(2/M) sum_m cos(omega_m . x + phase_m), omega_m ~ N(0, I/l^2)
# (Box-Muller), phase_m ~ U[0, 2 pi), drawn from splitmix64(seed) in the order
# (u1, u2) per omega component, then the phase (Rahimi & Recht's random
# Fourier features of the squared-exponential covariance exp(-|d|^2/(2 l^2))).
# The reference's right-hand side pattern b_i = int f phi_i is
# CG/cells_batched/application_hp.cc:1441-1465; its f (:77-93) is replaced by
# this field.

_M64 = (1 << 64) - 1


def splitmix64(seed, n):
    """first n outputs of splitmix64 (Steele, Lea, Flood 2014) as Python ints"""
    s = seed & _M64
    out = []
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & _M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        out.append(z ^ (z >> 31))
    return out


def grf_features(seed, n_features, corr_length):
    """(M, 3) frequencies and (M,) phases of the seeded field"""
    r = splitmix64(seed, 7 * n_features)
    u = [((v >> 11) + 0.5) * 2.0 ** -53 for v in r]
    om = np.empty((n_features, 3))
    ph = np.empty(n_features)
    k = 0
    for f in range(n_features):
        for a in range(3):
            u1, u2 = u[k], u[k + 1]
            k += 2
            om[f, a] = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2) / corr_length
        ph[f] = 2.0 * np.pi * u[k]
        k += 1
    return om, ph


def grf_eval(om, ph, X):
    """field at points X (n, 3)"""
    X = np.atleast_2d(X)
    return np.sqrt(2.0 / len(ph)) * np.cos(X @ om.T + ph).sum(axis=1)


def rhs_quadrature(p, vertices, cells, cell_dofs, n_dofs, f, n_rule=10):
    """b_i = sum_K int_K f phi_i with the oracle's collapsed Gauss rule of
    n_rule points per direction (exact to degree 2 n_rule - 3) and its own
    basis; f maps physical points (n, 3) to values"""
    pts, w = tet_rule(n_rule)
    vals = np.array([basis(p, x)[0] for x in pts])          # (nq, nl)
    V = np.asarray(vertices, np.float64)
    Cc = np.asarray(cells)
    x0 = V[Cc[:, 0]]
    J = np.stack([V[Cc[:, k]] - x0 for k in (1, 2, 3)], axis=2)   # (nc, 3, 3), columns = edges
    det = np.abs(np.linalg.det(J))
    X = x0[:, None, :] + np.einsum("kde,qe->kqd", J, pts)          # (nc, nq, 3)
    fx = f(X.reshape(-1, 3)).reshape(X.shape[:2])
    loc = det[:, None] * np.einsum("kq,q,qi->ki", fx, w, vals)
    b = np.zeros(n_dofs)
    np.add.at(b, np.asarray(cell_dofs).ravel(), loc.ravel())
    return b


# ---- Chebyshev-Jacobi smoother, its eigenvalue estimate and the V-cycle -------
# Restates the multigrid of CG/cells_batched/application_ph.cc:1754-2026 /
# application_hp.cc:1893-1948 with the build's stated choices (deal.II's
# PreconditionChebyshev internals are not in the reference):
#  * lambda_max of D^-1 A from `iters` CG iterations on A x = D 1 ... started
#    from r = 1, z = D^-1 r (Lanczos tridiagonal from the CG coefficients:
#    T_kk = 1/alpha_k + beta_{k-1}/alpha_{k-1}, T_k,k+1 = sqrt(beta_k)/alpha_k);
#  * Chebyshev interval [1.2 lambda / range, 1.2 lambda], Saad's Algorithm
#    12.1 on the Jacobi-preconditioned system, `degree` residual polynomial
#    degree, the first step from x = 0 skipping the product;
#  * V-cycle: pre-smooth from 0, residual, restrict R = P^T, recurse, prolongate
#    and add, post-smooth; the coarsest level is solved exactly on its free
#    rows (constrained rows of the coarse solution are 0).


def lanczos_lambda_max(A, dinv, iters):
    n = A.shape[0]
    r = np.ones(n)
    z = dinv * r
    p = z.copy()
    rz = r @ z
    al, be = [], []
    for _ in range(iters):
        if not rz > 0:
            break
        q = A @ p
        a = rz / (p @ q)
        r = r - a * q
        z = dinv * r
        rzn = r @ z
        b = rzn / rz
        al.append(a)
        be.append(b)
        p = z + b * p
        rz = rzn
    m = len(al)
    if m == 0:
        return 1.0
    T = np.zeros((m, m))
    for k in range(m):
        T[k, k] = 1.0 / al[k] + (be[k - 1] / al[k - 1] if k > 0 else 0.0)
        if k + 1 < m:
            T[k, k + 1] = T[k + 1, k] = np.sqrt(be[k]) / al[k]
    return float(np.linalg.eigvalsh(T).max())


def inverse_diagonal(d):
    """1/d where |d| > 1e-10, else 1 (CG/cells_batched/application_hp.cc:1404-1407)"""
    d = np.asarray(d, np.float64)
    big = np.abs(d) > 1e-10
    return np.where(big, 1.0 / np.where(big, d, 1.0), 1.0)


def chebyshev(A, dinv, b, x, degree, lmax, lmin):
    """`degree` Chebyshev-Jacobi steps on A x = b; x None = start from zero"""
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta
    rho = 1.0 / sigma
    d = None
    for k in range(degree):
        if k == 0:
            c1, c2 = 0.0, 1.0 / theta
            if x is None:
                d = c2 * dinv * b
                x = d.copy()
                continue
        else:
            rn = 1.0 / (2.0 * sigma - rho)
            c1, c2 = rn * rho, 2.0 * rn / delta
            rho = rn
        r = b - A @ x
        d = c2 * dinv * r if k == 0 else c1 * d + c2 * dinv * r
        x = x + d
    return x


def chebyshev_error_factor(lam, degree, lmax, lmin):
    """error reduction of `degree` steps from x = 0 on an eigenvector of
    D^-1 A with eigenvalue lam: T_d((theta - lam)/delta) / T_d(theta/delta)"""
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    T = np.polynomial.chebyshev.Chebyshev.basis(degree)
    return T((theta - lam) / delta) / T(theta / delta)


def mg_levels(As, Ps, constrained, degree=5, smoothing_range=20.0, eig_iters=10):
    """level dicts, [0] = coarsest: As[l] dense level matrices, Ps[l] (l >= 1)
    the prolongation from level l-1 to l, constrained[l] boolean masks"""
    levels = []
    for l, A in enumerate(As):
        L = {"A": A, "dinv": inverse_diagonal(np.diag(A)), "con": np.asarray(constrained[l], bool)}
        if l == 0:
            free = ~L["con"]
            L["inv"] = np.linalg.inv(A[np.ix_(free, free)])
        else:
            L["P"] = Ps[l]
            lam = lanczos_lambda_max(A, L["dinv"], eig_iters)
            L["lmax"] = 1.2 * lam
            L["lmin"] = L["lmax"] / smoothing_range
            L["lam"] = lam
        L["degree"] = degree
        levels.append(L)
    return levels


def mg_vcycle(levels, b, l=None):
    if l is None:
        l = len(levels) - 1
    L = levels[l]
    if l == 0:
        x = np.zeros(len(b))
        free = ~L["con"]
        x[free] = L["inv"] @ b[free]
        return x
    A, dinv = L["A"], L["dinv"]
    x = chebyshev(A, dinv, b, None, L["degree"], L["lmax"], L["lmin"])
    r = b - A @ x
    xc = mg_vcycle(levels, L["P"].T @ r, l - 1)
    x = x + L["P"] @ xc
    return chebyshev(A, dinv, b, x, L["degree"], L["lmax"], L["lmin"])


# ---- curved DG cells (numpy restatement) ---------------------------------------
# Follows DG/cells_batched/application.cc:1614-1666 (general D_e:
# g <- JxW_q jac_q^T jac_q g) and :2191-2238 (curvilinear SIPG face flux with
# per-point normal x jacobian on both sides), with the penalty of :1193-1245,
# 2120-2123 evaluated on the curved geometry.  Geometry: each cell's P2 map
# through its vertices (ascending global order) and the manifold's edge
# midpoints (curved_nodes).  Independent of the product's code: the oracle's
# own P2 and P_p bases (orc_basis), the Jacobians from the basis gradients,
# the face area element as |dx/ds x dx/dt| of the face parametrisation, and
# an explicit check that both sides' points coincide.  The rules are inputs
# (the operator's, exactness checked separately).

_REF_V = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
_FACE_V = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]  # face f: opposite vertex f, ascending


def _p2_jac(nodes_k, xi):
    """J (3x3) of the P2 map with nodes (10, 3) at reference point xi"""
    g = basis(2, xi)[1]                    # (10, 3) gradients of the geometry basis
    return np.einsum("id,ie->de", nodes_k, g)


def _p2_map(nodes_k, xi):
    return basis(2, xi)[0] @ nodes_k


def dg_curved_matrix(p, vertices, cells, pts, w, fpts, fw, ncomp=1, mass=0.0, laplace=1.0,
                     center=(5.0, 0.0, 0.0), straight=False):
    """sparse matrix of the curved-cell DG SIPG operator; rows
    (k*ncomp + c)*n_loc + i; cells with vertices in ascending order"""
    import scipy.sparse as sp
    cells = np.asarray(cells)
    nc = cells.shape[0]
    nl = (p + 1) * (p + 2) * (p + 3) // 6
    nodes = curved_nodes(vertices, cells, center, straight)
    factor = (p + 1) * (p + 3) / 3.0
    # faces
    faces = {}
    for k in range(nc):
        for f, fv in enumerate(_FACE_V):
            faces.setdefault(tuple(sorted(cells[k, list(fv)])), []).append((k, f))

    def face_point(f, s, t):
        a, b, c = _FACE_V[f]
        return _REF_V[a] + s * (_REF_V[b] - _REF_V[a]) + t * (_REF_V[c] - _REF_V[a])

    def face_geom(k, f, s, t):
        """physical point, area element, outward unit normal at face point (s, t)"""
        a, b, c = _FACE_V[f]
        xi = face_point(f, s, t)
        J = _p2_jac(nodes[k], xi)
        ts, tt = J @ (_REF_V[b] - _REF_V[a]), J @ (_REF_V[c] - _REF_V[a])
        A = np.cross(ts, tt)
        # outward: n is normal to the face tangents J e1, J e2, so J^T n is
        # normal to the reference face; it must point out of the reference cell
        xq = _p2_map(nodes[k], xi)
        n = A / np.linalg.norm(A)
        out_ref = _REF_V[a] - _REF_V[f]     # from the opposite vertex to the face
        if (J.T @ n) @ out_ref < 0:
            n = -n
        return xq, np.linalg.norm(A), n, J, xi

    # volumes and surfaces for tau
    vol = np.zeros(nc)
    for k in range(nc):
        for x, wq in zip(pts, w):
            vol[k] += wq * abs(np.linalg.det(_p2_jac(nodes[k], x)))
    surf = np.zeros(nc)
    for key, sides in faces.items():
        for k, f in sides:
            area = sum(wq * face_geom(k, f, s, t)[1] for (s, t), wq in zip(fpts, fw))
            surf[k] += area * (1.0 if len(sides) == 1 else 0.5)
    tau = surf / vol
    n = nc * nl
    rows, cols, vals = [], [], []

    def add(block, ka, kb):
        ia, ib = np.arange(nl) + ka * nl, np.arange(nl) + kb * nl
        rows.append(np.repeat(ia, nl))
        cols.append(np.tile(ib, nl))
        vals.append(block.ravel())

    for k in range(nc):
        Ke = np.zeros((nl, nl))
        for x, wq in zip(pts, w):
            J = _p2_jac(nodes[k], x)
            v, g = basis(p, x)
            gp = g @ np.linalg.inv(J)          # physical gradients (nl, 3)
            dj = abs(np.linalg.det(J))
            Ke += wq * dj * (laplace * gp @ gp.T + mass * np.outer(v, v))
        add(Ke, k, k)
    for key, sides in faces.items():
        (km, fm) = sides[0]
        Bmm, Bmp, Bpm, Bpp = (np.zeros((nl, nl)) for _ in range(4))
        sig = factor * (max(tau[km], tau[sides[1][0]]) if len(sides) == 2 else tau[km])
        for (s, t), wq in zip(fpts, fw):
            xq, da, nrm, J, xi = face_geom(km, fm, s, t)
            vm, gm = basis(p, xi)
            dm = gm @ (np.linalg.inv(J) @ nrm)
            jw = wq * da * laplace
            if len(sides) == 2:
                kp, fp = sides[1]
                xqp, _, _, Jp, xip = face_geom(kp, fp, s, t)
                assert np.max(np.abs(xqp - xq)) < 1e-11, "the two sides' face points differ"
                vp, gpp = basis(p, xip)
                dp = gpp @ (np.linalg.inv(Jp) @ nrm)
                Jv = np.concatenate([vm, -vp])
                Av = 0.5 * np.concatenate([dm, dp])
                B = jw * (sig * np.outer(Jv, Jv) - np.outer(Jv, Av) - np.outer(Av, Jv))
                Bmm += B[:nl, :nl]
                Bmp += B[:nl, nl:]
                Bpm += B[nl:, :nl]
                Bpp += B[nl:, nl:]
            else:
                Bmm += jw * (2 * sig * np.outer(vm, vm) - np.outer(vm, dm) - np.outer(dm, vm))
        add(Bmm, km, km)
        if len(sides) == 2:
            kp = sides[1][0]
            add(Bmp, km, kp)
            add(Bpm, kp, km)
            add(Bpp, kp, kp)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    if ncomp == 1:
        return A
    # components share the geometry: row (k ncomp + c) nl + i
    A3 = sp.kron(A, sp.eye(1)).tocoo()
    r, c_, v = [], [], []
    for comp in range(ncomp):
        kr, ir = np.divmod(A3.row, nl)
        kc, ic = np.divmod(A3.col, nl)
        r.append((kr * ncomp + comp) * nl + ir)
        c_.append((kc * ncomp + comp) * nl + ic)
        v.append(A3.data)
    return sp.csr_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(c_))), shape=(n * ncomp, n * ncomp))
