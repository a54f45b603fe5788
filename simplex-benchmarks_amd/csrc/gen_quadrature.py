#!/usr/bin/env python3
"""Derive the symmetric simplex quadrature rules used by the library and write
``quad_tables.h``.

Each rule is a union of symmetry orbits in barycentric coordinates.  The orbit
parameters and weights are found by Newton/least-squares iteration on the
moment equations

    sum_k w_k * x_k^a y_k^b z_k^c = a! b! c! / (a+b+c+3)!      (tetrahedron)
    sum_k w_k * x_k^a y_k^b       = a! b!   / (a+b+2)!          (triangle)

for every monomial of total degree <= d, refined by Gauss-Newton in 40-digit arithmetic,
and accepted only if all of those equations hold to 1e-30 before rounding to
double.  The reference element is the unit simplex
with vertices 0, e_x, e_y (, e_z).

Rules (degree of exactness, number of points):
  tetrahedron: (1,1) (2,4) (4,11) (5,14) (6,24)
  triangle:    (1,1) (2,3) (4,6)  (6,12)

Run:  python3 gen_quadrature.py > quad_tables.h
"""
import itertools
import math
import sys

import mpmath
import numpy as np
from scipy.optimize import least_squares


def tet_orbit(kind, p):
    if kind == "S4":
        b = [(0.25, 0.25, 0.25, 0.25)]
    elif kind == "S31":
        a = p[0]
        b = set(itertools.permutations((a, a, a, 1 - 3 * a)))
    elif kind == "S22":
        a = p[0]
        b = set(itertools.permutations((a, a, 0.5 - a, 0.5 - a)))
    elif kind == "S211":
        a, c = p
        b = set(itertools.permutations((a, a, c, 1 - 2 * a - c)))
    else:
        raise ValueError(kind)
    return sorted(b)


def tri_orbit(kind, p):
    if kind == "S3":
        b = [(1 / 3, 1 / 3, 1 / 3)]
    elif kind == "S21":
        a = p[0]
        b = set(itertools.permutations((a, a, 1 - 2 * a)))
    elif kind == "S111":
        a, c = p
        b = set(itertools.permutations((a, c, 1 - a - c)))
    else:
        raise ValueError(kind)
    return sorted(b)


NPAR = {"S4": 0, "S31": 1, "S22": 1, "S211": 2, "S3": 0, "S21": 1, "S111": 2}


def build(dim, orbits, params):
    """Return points (n, dim) and weights (n,) for the orbit list."""
    pts, wts = [], []
    k = 0
    for kind in orbits:
        npar = NPAR[kind]
        p = params[k : k + npar]
        w = params[k + npar]
        k += npar + 1
        orb = tet_orbit(kind, p) if dim == 3 else tri_orbit(kind, p)
        for bary in orb:
            # barycentric (l0, l1, l2[, l3]) -> cartesian (l1, l2[, l3])
            pts.append(bary[1:])
            wts.append(w)
    return np.array(pts), np.array(wts)


def monomials(dim, deg):
    for tot in range(deg + 1):
        for e in itertools.product(range(tot + 1), repeat=dim):
            if sum(e) == tot:
                yield e


def exact_moment(e):
    dim = len(e)
    num = 1.0
    for a in e:
        num *= math.factorial(a)
    return num / math.factorial(sum(e) + dim)


def residual(dim, deg, orbits, params):
    pts, wts = build(dim, orbits, params)
    res = []
    for e in monomials(dim, deg):
        val = np.sum(wts * np.prod(pts ** np.array(e), axis=1))
        res.append(val - exact_moment(e))
    return np.array(res)


def refine_mp(dim, deg, orbits, x0):
    """Gauss-Newton in 40-digit arithmetic (mpmath) on the moment equations,
    starting from the double-precision solution."""
    mpmath.mp.dps = 40
    x = [mpmath.mpf(float(v)) for v in x0]
    mons = list(monomials(dim, deg))

    def res(xv):
        pts, wts = build_mp(dim, orbits, xv)
        out = []
        for e in mons:
            val = mpmath.mpf(0)
            for pt, w in zip(pts, wts):
                term = w
                for a, xa in zip(e, pt):
                    term *= xa ** a
                val += term
            num = mpmath.mpf(1)
            for a in e:
                num *= mpmath.factorial(a)
            out.append(val - num / mpmath.factorial(sum(e) + dim))
        return mpmath.matrix(out)

    for _ in range(30):
        r = res(x)
        h = mpmath.mpf(10) ** -25
        Jm = mpmath.matrix(len(r), len(x))
        for j in range(len(x)):
            xp = list(x)
            xp[j] += h
            rp = res(xp)
            for i in range(len(r)):
                Jm[i, j] = (rp[i] - r[i]) / h
        dx = mpmath.lu_solve(Jm.T * Jm, Jm.T * r)
        x = [xi - dxi for xi, dxi in zip(x, dx)]
        if mpmath.norm(dx) < mpmath.mpf(10) ** -32:
            break
    return x, float(mpmath.norm(res(x), mpmath.inf))


def build_mp(dim, orbits, params):
    pts, wts = [], []
    k = 0
    one, half = mpmath.mpf(1), mpmath.mpf(1) / 2
    for kind in orbits:
        npar = NPAR[kind]
        p = params[k: k + npar]
        w = params[k + npar]
        k += npar + 1
        if kind == "S4":
            orb = [(one / 4,) * 4]
        elif kind == "S31":
            a = p[0]
            orb = sorted(set(itertools.permutations((a, a, a, one - 3 * a))))
        elif kind == "S22":
            a = p[0]
            orb = sorted(set(itertools.permutations((a, a, half - a, half - a))))
        elif kind == "S211":
            a, c = p
            orb = sorted(set(itertools.permutations((a, a, c, one - 2 * a - c))))
        elif kind == "S3":
            orb = [(one / 3,) * 3]
        elif kind == "S21":
            a = p[0]
            orb = sorted(set(itertools.permutations((a, a, one - 2 * a))))
        elif kind == "S111":
            a, c = p
            orb = sorted(set(itertools.permutations((a, c, one - a - c))))
        for bary in orb:
            pts.append(bary[1:])
            wts.append(w)
    return pts, wts


def solve(dim, deg, orbits, guess):
    sol = least_squares(lambda p: residual(dim, deg, orbits, p), guess,
                        xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=20000)
    xmp, rmp = refine_mp(dim, deg, orbits, sol.x)
    if rmp > 1e-30:
        raise RuntimeError(f"rule dim={dim} deg={deg} not exact in extended precision: {rmp}")
    r = np.max(np.abs(residual(dim, deg, orbits, np.array([float(v) for v in xmp]))))
    pts_mp, wts_mp = build_mp(dim, orbits, xmp)
    pts = np.array([[float(v) for v in pt] for pt in pts_mp])
    wts = np.array([float(w) for w in wts_mp])
    # all points must lie inside the simplex
    bary0 = 1 - pts.sum(axis=1)
    if np.min(pts) < -1e-14 or np.min(bary0) < -1e-14:
        raise RuntimeError(f"rule dim={dim} deg={deg} has points outside")
    return pts, wts, r


# (dim, degree, orbit list, initial guess: per orbit its parameters then weight)
RULES = [
    ("tet", 1, ["S4"], [1 / 6]),
    ("tet", 2, ["S31"], [0.138, 1 / 24]),
    ("tet", 4, ["S4", "S31", "S22"], [-0.0131, 0.0714, 0.0076, 0.1005, 0.0249]),
    ("tet", 5, ["S31", "S31", "S22"], [0.0927, 0.0122, 0.3109, 0.0188, 0.0455, 0.0071]),
    ("tet", 6, ["S31", "S31", "S31", "S211"],
     [0.2146, 0.0066, 0.0406, 0.0017, 0.3223, 0.0092, 0.0637, 0.2696, 0.0080]),
    ("tri", 1, ["S3"], [0.5]),
    ("tri", 2, ["S21"], [1 / 6, 1 / 6]),
    ("tri", 4, ["S21", "S21"], [0.4459, 0.1117, 0.0916, 0.0550]),
    ("tri", 6, ["S21", "S21", "S111"], [0.0631, 0.0254, 0.2493, 0.0584, 0.0531, 0.3104, 0.0414]),
]


def main():
    out = sys.stdout
    out.write("// Generated by gen_quadrature.py: symmetric simplex quadrature rules on the\n")
    out.write("// unit simplex, solved from the moment equations and verified exact to the\n")
    out.write("// stated degree (max moment residual printed per rule).  Do not edit.\n")
    out.write("#pragma once\n\nnamespace smf {\n\n")
    names = []
    for shape, deg, orbits, guess in RULES:
        dim = 3 if shape == "tet" else 2
        pts, wts, r = solve(dim, deg, orbits, np.array(guess, dtype=float))
        name = f"q_{shape}_d{deg}"
        names.append((shape, deg, len(wts), name))
        out.write(f"// {shape} degree {deg}: {len(wts)} points, orbits {'+'.join(orbits)}, "
                  f"max moment residual {r:.1e}\n")
        out.write(f"constexpr double {name}_pts[{len(wts)}][{dim}] = {{\n")
        for p in pts:
            out.write("  {" + ", ".join(f"{v:.17g}" for v in p) + "},\n")
        out.write("};\n")
        out.write(f"constexpr double {name}_w[{len(wts)}] = {{\n  ")
        out.write(", ".join(f"{w:.17g}" for w in wts))
        out.write("\n};\n\n")
    out.write("}  // namespace smf\n")


if __name__ == "__main__":
    main()
