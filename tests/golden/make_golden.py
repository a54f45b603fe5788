#!/usr/bin/env python3
"""Write golden fixtures for the CG Poisson operator (independent of both the
oracle and the library).

The element stiffness matrices of P1 and P2 are computed in exact rational
arithmetic from the barycentric form of the bases:
  P1: grad phi_k = grad l_k
  P2: vertex a: grad phi = (4 l_a - 1) grad l_a; edge (a,b): 4 (l_b grad l_a + l_a grad l_b)
and the closed-form simplex integral  int_K l^alpha = |K| 3! alpha! / (|alpha|+3)!.
The only floating-point step is the geometry (grad l_k from the vertex
coordinates).  Assembled on a 2x2x2-hex Kuhn mesh (48 tets) with a fixed
vertex perturbation, without boundary conditions (pure Neumann matrix), and
stored with the support point of every DoF so a test can map the fixture onto
any DoF numbering.

Run: python3 make_golden.py  (writes cg_p1_2x2x2.npz, cg_p2_2x2x2.npz)
"""
import itertools
import math
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def kuhn_mesh(n):
    """Kuhn split of n^3 hexes (same orientation rule as the library: odd
    permutations swap local vertices 2 and 3), unit cube."""
    n1 = n + 1
    verts = np.array([[i / n, j / n, k / n] for k in range(n1) for j in range(n1) for i in range(n1)])
    cells = []
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    odd = [False, True, True, False, False, True]
    for k in range(n):
        for j in range(n):
            for i in range(n):
                for s, pm in enumerate(perms):
                    p = [i, j, k]
                    vid = [p[0] + n1 * (p[1] + n1 * p[2])]
                    for t in range(3):
                        p[pm[t]] += 1
                        vid.append(p[0] + n1 * (p[1] + n1 * p[2]))
                    if odd[s]:
                        vid[2], vid[3] = vid[3], vid[2]
                    cells.append(vid)
    return verts, np.array(cells, dtype=np.int32)


def lam_integral(alpha):
    """int over the reference-volume-normalised simplex of prod l_k^alpha_k,
    divided by |K|:  3! alpha! / (|alpha|+3)!"""
    num = math.factorial(3)
    for a in alpha:
        num *= math.factorial(a)
    return Fraction(num, math.factorial(sum(alpha) + 3))


# gradient coefficients: grad phi_i = sum_k C_i[k](l) grad l_k, C linear in l
# represented as dict {alpha (4-tuple): coefficient}
def p1_coeffs():
    out = []
    for a in range(4):
        c = [dict() for _ in range(4)]
        c[a][(0, 0, 0, 0)] = Fraction(1)
        out.append(c)
    return out


def unit(k):
    e = [0, 0, 0, 0]
    e[k] = 1
    return tuple(e)


def p2_coeffs():
    out = []
    for a in range(4):
        c = [dict() for _ in range(4)]
        c[a][unit(a)] = Fraction(4)
        c[a][(0, 0, 0, 0)] = Fraction(-1)
        out.append(c)
    for a, b in EDGES:
        c = [dict() for _ in range(4)]
        c[a][unit(b)] = Fraction(4)
        c[b][unit(a)] = Fraction(4)
        out.append(c)
    return out


def poly_mul_int(p, q):
    s = Fraction(0)
    for ea, ca in p.items():
        for eb, cb in q.items():
            s += ca * cb * lam_integral(tuple(x + y for x, y in zip(ea, eb)))
    return s


def reference_tensor(coeffs):
    """T[i][j][k][l] = (1/|K|) int C_i[k] C_j[l]  (exact rationals)"""
    n = len(coeffs)
    T = np.zeros((n, n, 4, 4))
    for i, j, k, l in itertools.product(range(n), range(n), range(4), range(4)):
        T[i, j, k, l] = float(poly_mul_int(coeffs[i][k], coeffs[j][l]))
    return T


def element(xv, T):
    J = np.stack([xv[1] - xv[0], xv[2] - xv[0], xv[3] - xv[0]], axis=1)
    vol = abs(np.linalg.det(J)) / 6.0
    Jinv = np.linalg.inv(J)
    gl = np.zeros((4, 3))
    gl[1:] = Jinv
    gl[0] = -Jinv.sum(axis=0)
    G = gl @ gl.T
    return vol * np.einsum("ijkl,kl->ij", T, G)


def assemble(verts, cells, p):
    coeffs = p1_coeffs() if p == 1 else p2_coeffs()
    T = reference_tensor(coeffs)
    # DoFs: vertices, then edges (sorted vertex pairs); support points
    edge_id = {}
    if p == 2:
        for c in cells:
            for a, b in EDGES:
                key = tuple(sorted((c[a], c[b])))
                edge_id.setdefault(key, len(verts) + len(edge_id))
    nd = len(verts) + len(edge_id)
    pts = np.zeros((nd, 3))
    pts[: len(verts)] = verts
    for (a, b), d in edge_id.items():
        pts[d] = 0.5 * (verts[a] + verts[b])
    A = np.zeros((nd, nd))
    for c in cells:
        dofs = list(c)
        if p == 2:
            dofs += [edge_id[tuple(sorted((c[a], c[b])))] for a, b in EDGES]
        K = element(verts[c], T)
        A[np.ix_(dofs, dofs)] += K
    return A, pts


def main():
    verts, cells = kuhn_mesh(2)
    # fixed perturbation of the single interior vertex (1/2,1/2,1/2)
    center = 1 + 3 * (1 + 3 * 1)
    verts[center] += np.array([0.07, -0.05, 0.03])
    for p in (1, 2):
        A, pts = assemble(verts, cells, p)
        np.savez(os.path.join(HERE, f"cg_p{p}_2x2x2.npz"), vertices=verts, cells=cells,
                 support_points=pts, matrix=A)
        print(f"P{p}: {A.shape[0]} DoFs, |A| = {np.abs(A).max():.4f}")


if __name__ == "__main__":
    main()
