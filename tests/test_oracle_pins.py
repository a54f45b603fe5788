"""Pins of oracle functions that the symmetric / consistency checks leave open
(CPU only).  Each check is fixed by mathematics computed here independently of
the oracle's code path, and each is shown to fail on a plausible mistake
planted into the oracle (oracle_py.planted) or into this file's restatement:

  * DG SIPG interior penalty and consistency terms: a(u, u) of a
    discontinuous piecewise-linear u in closed form,
      sum_K |K| |g_K|^2
      + sum_F int_F sigma_F [u]^2 - 2 {d_n u} [u]
      + sum_B int_F 2 sigma_B u^2 - 2 u d_n u,
    sigma_F = (p+1)(p+3)/3 max(tau-, tau+)  (DG/cells_batched/application.cc:
    2120-2123, 1193-1194), sigma_B = (p+1)(p+3)/3 tau_K (:2668-2669), face
    integrals by the edge-midpoint rule (exact for quadratics);
    planted min-for-max and a dropped 1/2 in the average must fail;
  * the DG diagonal equals the diagonal of the dense DG matrix
    (DG/cells_batched/application_phc.cc:1491-1526);
  * the Helmholtz mass term (DG/components_batched/application_helmholtz.cc:
    1025-1111) per cell equals mass |K| M_ref with M_ref the exact barycentric
    integrals of the Lagrange basis products; a lumped mass must fail;
  * the Gaussian-random-field restatement: splitmix64 stream, Box-Muller
    normals, unit variance and the covariance exp(-|d|^2 / (2 l^2));
  * the Chebyshev-Jacobi restatement (application_ph.cc:1912-1932): the
    Lanczos estimate equals the exact lambda_max of D^-1 A when it runs n
    steps and stays below it after 10; the smoother's error on an
    eigenvector equals the Chebyshev factor T_d((theta-lam)/delta)/
    T_d(theta/delta); the restated V-cycle is symmetric and contracting.
"""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

from test_oracle_dg import numpy_tau

FACES = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]


# ---------------------------------------------------------------------------
# DG SIPG energy of a discontinuous piecewise-linear function

def _dg_problem(smf, orc, n, p, seed=3, ncomp_kind="dg"):
    op = smf.Operator(smf.Mesh.kuhn_cube(n, 0.3, seed), p, kind=ncomp_kind, host_only=True)
    return op, orc.DGOracleProblem.from_operator(op)


def _faces(C):
    keys = {}
    for c in range(C.shape[0]):
        for lf, f in enumerate(FACES):
            keys.setdefault(tuple(sorted(C[c, list(f)])), []).append((c, lf))
    return keys


def _sipg_energy_closed_form(V, C, p, coef, grad, combine=max, avg_factor=0.5):
    """a(u, u) for u_K(x) = coef_K + grad_K . x (per cell)"""
    tau, _ = numpy_tau(V, C)
    factor = (p + 1) * (p + 3) / 3
    vol = np.array([abs(np.linalg.det(np.stack([V[C[c, t]] - V[C[c, 0]] for t in (1, 2, 3)], 1))) / 6
                    for c in range(C.shape[0])])
    a = float(np.sum(vol * np.sum(grad * grad, axis=1)))

    def u(c, x):
        return coef[c] + grad[c] @ x

    for key, sides in _faces(C).items():
        X = V[list(key)]
        area = 0.5 * np.linalg.norm(np.cross(X[1] - X[0], X[2] - X[0]))
        nrm = np.cross(X[1] - X[0], X[2] - X[0])
        nrm /= np.linalg.norm(nrm)
        cm, lfm = sides[0]
        opp = V[C[cm, lfm]]
        if nrm @ (X[0] - opp) < 0:
            nrm = -nrm  # outward of the first cell
        mids = [0.5 * (X[i] + X[j]) for i, j in ((0, 1), (0, 2), (1, 2))]
        if len(sides) == 2:
            cp = sides[1][0]
            sigma = factor * combine(tau[cm], tau[cp])
            jump = [u(cm, x) - u(cp, x) for x in mids]
            avg = avg_factor * (grad[cm] + grad[cp]) @ nrm
            int_j2 = area / 3 * sum(j * j for j in jump)
            int_j = area / 3 * sum(jump)
            a += sigma * int_j2 - 2 * avg * int_j
        else:
            sigma = factor * tau[cm]
            vals = [u(cm, x) for x in mids]
            a += 2 * sigma * area / 3 * sum(v * v for v in vals) - 2 * (grad[cm] @ nrm) * area / 3 * sum(vals)
    return a


@pytest.mark.parametrize("p", [1, 2, 3])
def test_dg_interior_penalty_closed_form(smf, orc, p):
    op, pr = _dg_problem(smf, orc, 2, p)
    V, C = pr.vertices, pr.cells
    rng = np.random.default_rng(10 + p)
    coef = rng.uniform(-1, 1, C.shape[0])
    grad = rng.uniform(-1, 1, (C.shape[0], 3))
    X = op.dof_points().reshape(C.shape[0], pr.n_loc, 3)
    uvec = (coef[:, None] + np.einsum("kd,kid->ki", grad, X)).ravel()
    got = uvec @ pr.vmult(uvec)
    want = _sipg_energy_closed_form(V, C, p, coef, grad)
    assert abs(got - want) < 1e-11 * abs(want)
    # the perturbed mesh has cells of different tau, so min(tau) differs from max(tau)
    alt = _sipg_energy_closed_form(V, C, p, coef, grad, combine=min)
    assert abs(alt - want) > 1e-3 * abs(want)
    # planted mistakes in the oracle fail the closed form
    with orc.planted(1):
        bad_min = uvec @ pr.vmult(uvec)
    with orc.planted(2):
        bad_avg = uvec @ pr.vmult(uvec)
    assert abs(bad_min - want) > 1e-4 * abs(want)
    assert abs(bad_min - alt) < 1e-11 * abs(alt)  # the plant is exactly the min variant
    assert abs(bad_avg - want) > 1e-4 * abs(want)


def test_dg_piecewise_constant_jump_energy(smf, orc):
    # u constant per cell: a(u,u) = sum_F sigma_F |F| [u]^2 + sum_B 2 sigma_B |F| u^2
    op, pr = _dg_problem(smf, orc, 2, 2, seed=8)
    V, C = pr.vertices, pr.cells
    tau, _ = numpy_tau(V, C)
    c = np.random.default_rng(4).uniform(-1, 1, C.shape[0])
    uvec = np.repeat(c, pr.n_loc)
    want = 0.0
    for key, sides in _faces(C).items():
        X = V[list(key)]
        area = 0.5 * np.linalg.norm(np.cross(X[1] - X[0], X[2] - X[0]))
        if len(sides) == 2:
            (a, _), (b, _) = sides
            want += 5.0 * max(tau[a], tau[b]) * area * (c[a] - c[b]) ** 2
        else:
            a = sides[0][0]
            want += 2 * 5.0 * tau[a] * area * c[a] ** 2
    got = uvec @ pr.vmult(uvec)
    assert abs(got - want) < 1e-12 * want


# ---------------------------------------------------------------------------
# DG diagonal

@pytest.mark.parametrize("p,kind,mass,penalty", [(1, "dg", 0.0, 0), (2, "dg", 0.0, 0), (3, "dg", 0.0, 0),
                                                   (2, "dg", 0.0, 1), (2, "dg_helmholtz", 2.5, 0),
                                                   (1, "dg_vector", 0.0, 0)])
def test_dg_diagonal_equals_dense_diagonal(smf, orc, p, kind, mass, penalty):
    op = smf.Operator(smf.Mesh.kuhn_cube(1 if p == 3 else 2, 0.3, 5), p, kind=kind, host_only=True,
                      mass_coefficient=max(mass, 1.0), penalty_kind=penalty)
    pr = orc.DGOracleProblem.from_operator(op, mass=mass, penalty_kind=penalty)
    A = pr.dense()
    d = pr.diagonal()
    assert np.max(np.abs(d - np.diag(A))) < 1e-12 * np.abs(np.diag(A)).max()


# ---------------------------------------------------------------------------
# Helmholtz mass term: exact barycentric element mass matrices

def _lam_integral(alpha):
    """int_K prod l_k^alpha_k / |K| = 3! alpha! / (|alpha| + 3)!"""
    num = math.factorial(3)
    for a in alpha:
        num *= math.factorial(a)
    return Fraction(num, math.factorial(sum(alpha) + 3))


def _poly_mul(a, b):
    out = {}
    for ea, ca in a.items():
        for eb, cb in b.items():
            e = tuple(x + y for x, y in zip(ea, eb))
            out[e] = out.get(e, 0) + ca * cb
    return out


def _lagrange_barycentric(alpha, p):
    """phi_alpha = prod_k prod_{j < alpha_k} (p l_k - j)/(j + 1) as {exponent: Fraction}"""
    poly = {(0, 0, 0, 0): Fraction(1)}
    for k in range(4):
        for j in range(alpha[k]):
            e = [0, 0, 0, 0]
            e[k] = 1
            poly = _poly_mul(poly, {tuple(e): Fraction(p, j + 1), (0, 0, 0, 0): Fraction(-j, j + 1)})
    return poly


def _exact_mass_ref(alphas, p):
    polys = [_lagrange_barycentric(a, p) for a in alphas]
    n = len(alphas)
    M = np.empty((n, n))
    for i in range(n):
        for j in range(i, n):
            prod = _poly_mul(polys[i], polys[j])
            M[i, j] = M[j, i] = float(sum(c * _lam_integral(e) for e, c in prod.items()))
    return M


def test_lagrange_barycentric_is_nodal():
    # the closed-form basis of this file is Lagrange: phi_a(node_b) = delta_ab
    p = 3
    alphas = [a for a in itertools.product(range(p + 1), repeat=4) if sum(a) == p]
    for a in alphas:
        poly = _lagrange_barycentric(a, p)
        for b in alphas:
            lam = [Fraction(x, p) for x in b]
            v = sum(c * math.prod(l ** e for l, e in zip(lam, ex)) for ex, c in poly.items())
            assert v == (1 if a == b else 0)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_helmholtz_mass_matches_exact_element_mass(smf, orc, p):
    mass = 2.5
    op = smf.Operator(smf.Mesh.kuhn_cube(1, 0.3, 2), p, kind="dg_helmholtz", host_only=True,
                      mass_coefficient=mass)
    pr0 = orc.DGOracleProblem.from_operator(op, mass=mass, laplace=0.0)
    A = pr0.dense()
    V = pr0.vertices
    C = pr0.cells
    nl, nc = pr0.n_loc, C.shape[0]
    X = op.dof_points().reshape(nc, nl, 3)
    for k in range(nc):
        P = V[C[k]]
        T = np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], axis=1)
        lam123 = np.linalg.solve(T, (X[k] - P[0]).T).T
        lam = np.hstack([1 - lam123.sum(1, keepdims=True), lam123])
        alphas = [tuple(int(v) for v in np.rint(p * l)) for l in lam]
        assert np.max(np.abs(np.array(alphas) - p * lam)) < 1e-10
        Mref = _exact_mass_ref(alphas, p)
        vol = abs(np.linalg.det(T)) / 6
        for c in range(3):
            rows = (k * 3 + c) * nl + np.arange(nl)
            blk = A[np.ix_(rows, rows)]
            assert np.max(np.abs(blk - mass * vol * Mref)) < 1e-13 * mass * vol
    # nothing couples cells or components in the mass part
    off = A.copy()
    for k in range(nc * 3):
        off[k * nl:(k + 1) * nl, k * nl:(k + 1) * nl] = 0
    assert np.max(np.abs(off)) == 0
    # a lumped mass (row sums on the diagonal) fails the same check
    x = np.random.default_rng(p).uniform(-1, 1, op.m())
    with orc.planted(3):
        lumped = pr0.vmult(x)
    assert np.max(np.abs(lumped - A @ x)) > 1e-3 * np.max(np.abs(A @ x))


# ---------------------------------------------------------------------------
# Gaussian random field restatement

def test_splitmix64_matches_the_generator_of_the_seeded_mesh(smf, orc):
    # the mesh perturbation draws perturb*h*(2 u - 1), u = (next >> 11) 2^-53,
    # per coordinate of the interior vertices in id order, from the same stream
    n, pert, seed = 3, 0.25, 12345
    m = smf.Mesh.kuhn_cube(n, pert, seed)
    base = smf.Mesh.kuhn_cube(n, 0.0, 0)
    ids = [i + (n + 1) * (j + (n + 1) * k) for k in range(n + 1) for j in range(n + 1) for i in range(n + 1)
           if 0 < i < n and 0 < j < n and 0 < k < n]
    disp = (m.vertices - base.vertices)[ids] / (pert / n)
    # only interior vertices draw, three values each, in vertex-id order
    r = orc.splitmix64(seed, 3 * len(ids))
    u = np.array([2.0 * ((v >> 11) * 2.0 ** -53) - 1.0 for v in r]).reshape(-1, 3)
    assert np.max(np.abs(disp - u)) < 1e-12
    # the widely quoted first outputs of splitmix64 from state 0
    assert orc.splitmix64(0, 2) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4]


def test_grf_restatement_statistics(orc):
    # unit variance and covariance exp(-|d|^2/(2 l^2)) over many seeds
    ell, M, nseeds = 0.1, 256, 400
    X = np.array([[0.3, 0.4, 0.5], [0.35, 0.4, 0.5], [0.3, 0.5, 0.5], [0.6, 0.4, 0.5]])
    F = np.array([orc.grf_eval(*orc.grf_features(s, M, ell), X) for s in range(nseeds)])
    assert abs(F.mean()) < 0.1
    assert abs(F[:, 0].var() - 1.0) < 0.2
    for j, d in ((1, 0.05), (2, 0.1), (3, 0.3)):
        cov = np.mean(F[:, 0] * F[:, j])
        assert abs(cov - np.exp(-d * d / (2 * ell * ell))) < 0.12, (j, cov)
    # a field without the 1/l scaling has a far longer correlation
    om, ph = orc.grf_features(0, M, 1.0)
    assert np.std(om) < 2.0 and np.std(orc.grf_features(0, M, ell)[0]) > 5.0


def test_rhs_quadrature_is_exact_for_polynomials(smf, orc):
    # b_i = int phi_i for f = 1 (P2 vertex -|K|/20, edge |K|/5), and linear f
    # reproduces the exact moments sum_i b_i = int f
    op = smf.Operator(smf.Mesh.kuhn_cube(2, 0.2, 1), 2, kind="dg", host_only=True)
    pr = orc.DGOracleProblem.from_operator(op)
    b = orc.rhs_quadrature(2, pr.vertices, pr.cells, np.arange(op.m()).reshape(-1, 10), op.m(),
                           lambda X: np.ones(len(X)))
    assert abs(b.sum() - 1.0) < 1e-13
    bl = orc.rhs_quadrature(2, pr.vertices, pr.cells, np.arange(op.m()).reshape(-1, 10), op.m(),
                            lambda X: X[:, 0] + 2 * X[:, 1])
    assert abs(bl.sum() - 1.5) < 1e-13


# ---------------------------------------------------------------------------
# Chebyshev-Jacobi restatement and its eigenvalue estimate

def _cg_matrix(smf, orc, n=2, p=1, seed=6):
    op = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, seed), p, host_only=True)
    pr = orc.OracleProblem.from_operator(op)
    return op, pr.dense()


def test_lanczos_estimate_exact_with_full_steps(smf, orc):
    _, A = _cg_matrix(smf, orc, 2, 2)
    dinv = orc.inverse_diagonal(np.diag(A))
    s = np.sqrt(dinv)
    exact = np.linalg.eigvalsh(s[:, None] * A * s[None, :]).max()
    # the starting vector 1 has components on every eigenvector of the free block
    full = orc.lanczos_lambda_max(A, dinv, 60)
    assert abs(full - exact) < 1e-9 * exact
    ten = orc.lanczos_lambda_max(A, dinv, 10)
    assert ten <= exact * (1 + 1e-12) and ten > 0.7 * exact
    # a tridiagonal built without the beta/alpha term overestimates or underestimates
    assert orc.lanczos_lambda_max(A, dinv, 1) <= ten


def test_chebyshev_error_factor_on_eigenvectors(smf, orc):
    _, A = _cg_matrix(smf, orc, 2, 1)
    dinv = orc.inverse_diagonal(np.diag(A))
    s = np.sqrt(dinv)
    lam, W = np.linalg.eigh(s[:, None] * A * s[None, :])
    lmax = 1.2 * lam.max()
    lmin = lmax / 20.0
    for k in (0, len(lam) // 2, len(lam) - 1):
        v = s * W[:, k]                       # A v = lam D v
        b = A @ v
        for degree in (1, 2, 5):
            x = orc.chebyshev(A, dinv, b, None, degree, lmax, lmin)
            f = orc.chebyshev_error_factor(lam[k], degree, lmax, lmin)
            assert np.max(np.abs((v - x) - f * v)) < 1e-11 * np.max(np.abs(v)), (k, degree)
    # the smoother damps the upper part of the spectrum below 1/T_5(theta/delta)
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    bound = 1.0 / np.cosh(5 * np.arccosh(theta / delta))
    for l in lam[lam >= lmin]:
        assert abs(orc.chebyshev_error_factor(l, 5, lmax, lmin)) <= bound * (1 + 1e-12)


def test_restated_vcycle_symmetric_and_contracting(smf, orc):
    # two-level P2 -> P1 on the same mesh; P1 -> P2 interpolation built here
    # from the P1 hat functions at the P2 nodes
    mesh = smf.Mesh.kuhn_cube(3, 0.2, 9)
    f2 = smf.Operator(mesh, 2, host_only=True)
    f1 = smf.Operator(mesh, 1, host_only=True)
    A2 = orc.OracleProblem.from_operator(f2).dense()
    A1 = orc.OracleProblem.from_operator(f1).dense()
    con2, con1 = f2.constrained().astype(bool), f1.constrained().astype(bool)
    P = np.zeros((f2.m(), f1.m()))
    cd2, cd1 = f2.cell_dofs().reshape(-1, 10), f1.cell_dofs().reshape(-1, 4)
    X2, X1 = f2.dof_points(), f1.dof_points()
    # interpolation by barycentric coordinates of every P2 node in its cell
    V, Cv = mesh.vertices, mesh.cells
    perm = f2.cell_order()
    for k in range(cd2.shape[0]):
        Pk = V[Cv[perm[k]]]
        T = np.stack([Pk[1] - Pk[0], Pk[2] - Pk[0], Pk[3] - Pk[0]], axis=1)
        for i in range(10):
            l3 = np.linalg.solve(T, X2[cd2[k, i]] - Pk[0])
            lam = np.concatenate([[1 - l3.sum()], l3])
            for a in range(4):
                j = np.argmin(np.linalg.norm(X1 - Pk[a], axis=1))
                P[cd2[k, i], j] = lam[a] if abs(lam[a]) > 1e-12 else 0.0
    P[con2, :] = 0
    P[:, con1] = 0
    levels = orc.mg_levels([A1, A2], [None, P], [con1, con2])
    n = f2.m()
    B = np.empty((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        B[:, j] = orc.mg_vcycle(levels, e)
        e[j] = 0.0
    free = ~con2
    Bf = B[np.ix_(free, free)]
    assert np.max(np.abs(Bf - Bf.T)) < 1e-10 * np.abs(Bf).max()
    Af = A2[np.ix_(free, free)]
    rho = np.max(np.abs(np.linalg.eigvals(np.eye(Af.shape[0]) - Bf @ Af)))
    assert rho < 0.3, rho
