"""Generate the reference-element golden fixtures with exact rational arithmetic.

This is an independent restatement of the mathematics behind
``src/reference.mojo:519-660`` and ``src/reference_2d.mojo:130-261``.  It does
not follow the reference's route (homogeneous barycentric monomials, FP64
Gauss-Jordan); instead it uses Cartesian monomials r^a s^b t^c of total
degree <= P and exact ``fractions.Fraction`` linear algebra:

    V[j, k]  = m_k(x_j)                  (Lagrange nodes x_j)
    G[k, l]  = int m_k m_l               (monomial Gram matrix)
    Gd[k,l]  = int (d m_k / d r_d) m_l
    M^-1     = V G^-1 V^T                (since phi = V^-T m)
    D_d      = M^-1 S_d = V G^-1 Gd V^-1
    Lift_f   = M^-1 L_raw_f,  L_raw_f[i, m] = |ref face|^-1 int_f phi_i psi_m
    w_i      = int phi_i / |ref element|

The node orderings (vertex, edge, face-interior, volume-interior, the VTK
edge order and the face-local ordering) are conventions of the reference and
are restated here as they must be alike.  Run from the repository root:

    python tests/golden/make_golden.py

It writes ``tests/golden/refelem_golden.npz``.
"""
from __future__ import annotations

import os
from fractions import Fraction as Fr
from math import factorial

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "refelem_golden.npz")


def tet_nodes(P):
    """Barycentric exponents (a0..a3) in the reference's canonical order."""
    ea, eb = [0, 1, 2, 0, 1, 2], [1, 2, 0, 3, 3, 3]
    out = [tuple(P if d == v else 0 for d in range(4)) for v in range(4)]
    if P >= 2:
        for e in range(6):
            for k in range(1, P):
                out.append(tuple(P - k if ea[e] == d else (k if eb[e] == d else 0) for d in range(4)))
    if P >= 3:
        for f in range(4):
            ni = [i for i in range(4) if i != f]
            for a0 in range(P - 2, 0, -1):
                for a1 in range(P - a0 - 1, 0, -1):
                    a2 = P - a0 - a1
                    x = [0, 0, 0, 0]
                    x[ni[0]], x[ni[1]], x[ni[2]] = a0, a1, a2
                    out.append(tuple(x))
    if P >= 4:
        for a0 in range(P - 2, 0, -1):
            for a1 in range(P - a0 - 1, 0, -1):
                for a2 in range(P - a0 - a1 - 1, 0, -1):
                    out.append((a0, a1, a2, P - a0 - a1 - a2))
    return out


def tri_nodes(P):
    ea, eb = [0, 1, 2], [1, 2, 0]
    out = [tuple(P if d == v else 0 for d in range(3)) for v in range(3)]
    if P >= 2:
        for e in range(3):
            for k in range(1, P):
                out.append(tuple(P - k if ea[e] == d else (k if eb[e] == d else 0) for d in range(3)))
    if P >= 3:
        for a0 in range(P - 2, 0, -1):
            for a1 in range(P - a0 - 1, 0, -1):
                out.append((a0, a1, P - a0 - a1))
    return out


def face_to_elem(P, nodes):
    """Face-local order: 3 vertices (ascending), edges vA-vB, vB-vC, vC-vA,
    then interior nodes lex-descending."""
    idx = {n: i for i, n in enumerate(nodes)}
    out = []
    for f in range(4):
        vA, vB, vC = [v for v in range(4) if v != f]
        row = []
        for v in (vA, vB, vC):
            row.append(idx[tuple(P if d == v else 0 for d in range(4))])
        for a, b in ((vA, vB), (vB, vC), (vC, vA)):
            for k in range(1, P):
                row.append(idx[tuple(P - k if d == a else (k if d == b else 0) for d in range(4))])
        for a0 in range(P - 2, 0, -1):
            for a1 in range(P - a0 - 1, 0, -1):
                x = [0, 0, 0, 0]
                x[vA], x[vB], x[vC] = a0, a1, P - a0 - a1
                row.append(idx[tuple(x)])
        out.append(row)
    return out


def edge_to_elem(P, nodes):
    idx = {n: i for i, n in enumerate(nodes)}
    out = []
    for a, b in ((0, 1), (1, 2), (2, 0)):
        out.append([idx[tuple(P - k if d == a else (k if d == b else 0) for d in range(3))]
                    for k in range(P + 1)])
    return out


def inv(A):
    n = len(A)
    M = [list(r) + [Fr(int(i == j)) for j in range(n)] for i, r in enumerate(A)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        pv = M[c][c]
        M[c] = [x / pv for x in M[c]]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [r[n:] for r in M]


def mm(A, B):
    Bt = list(zip(*B))
    return [[sum(a * b for a, b in zip(r, c)) for c in Bt] for r in A]


def T(A):
    return [list(r) for r in zip(*A)]


def simplex_monomial_integral(exps, dim):
    """int over the unit simplex of prod x_d^e_d = prod e_d! / (sum e + dim)!"""
    num = 1
    for e in exps:
        num *= factorial(e)
    return Fr(num, factorial(sum(exps) + dim))


def cart_monomials(P, dim):
    out = []
    if dim == 3:
        for s in range(P + 1):
            for a in range(s, -1, -1):
                for b in range(s - a, -1, -1):
                    out.append((a, b, s - a - b))
    else:
        for s in range(P + 1):
            for a in range(s, -1, -1):
                out.append((a, s - a))
    return out


def element_ops(P, dim):
    nodes = tet_nodes(P) if dim == 3 else tri_nodes(P)
    N = len(nodes)
    # Cartesian coordinates (r, s[, t]) = (l1, l2[, l3]).
    X = [tuple(Fr(n[d + 1], P) for d in range(dim)) for n in nodes]
    mons = cart_monomials(P, dim)
    assert len(mons) == N
    V = [[np.prod([x[d] ** m[d] for d in range(dim)], dtype=object) for m in mons] for x in X]
    V = [[Fr(v) for v in row] for row in V]
    G = [[simplex_monomial_integral([a + b for a, b in zip(mk, ml)], dim) for ml in mons] for mk in mons]
    Vi, Gi = inv(V), inv(G)
    Minv = mm(mm(V, Gi), T(V))
    D = []
    for d in range(dim):
        Gd = []
        for mk in mons:
            row = []
            for ml in mons:
                if mk[d] == 0:
                    row.append(Fr(0))
                    continue
                e = list(mk)
                e[d] -= 1
                row.append(mk[d] * simplex_monomial_integral([a + b for a, b in zip(e, ml)], dim))
            Gd.append(row)
        D.append(mm(mm(mm(V, Gi), Gd), Vi))
    vol = Fr(1, 6) if dim == 3 else Fr(1, 2)
    g = [simplex_monomial_integral(m, dim) for m in mons]
    # w_i = int phi_i / vol,  phi = V^-T m  ->  int phi = V^-T g
    w = [sum(Vi[k][i] * g[k] for k in range(N)) / vol for i in range(N)]
    if dim == 3:
        fmap = face_to_elem(P, nodes)
        Pf, Mf = element_face_mass(P, 2)
        scale = 2  # per unit area on the reference triangle
    else:
        fmap = edge_to_elem(P, nodes)
        Pf, Mf = element_face_mass(P, 1)
        scale = 1  # the reference edge [0, 1] has unit length
    NFP = len(fmap[0])
    Lift = []
    for f, row in enumerate(fmap):
        L = [[Fr(0)] * NFP for _ in range(N)]
        for l, i in enumerate(row):
            for m in range(NFP):
                L[i][m] = scale * Mf[l][m]
        Lift.append(mm(Minv, L))
    return dict(D=D, Lift=Lift, w=w, fmap=fmap)


def element_face_mass(P, fdim):
    """Mass matrix of the face-local Lagrange basis on the unit simplex of
    dimension fdim, in the face-local node order."""
    if fdim == 2:
        nodes = tri_nodes(P)
        X = [(Fr(n[1], P), Fr(n[2], P)) for n in nodes]
    else:
        X = [(Fr(k, P),) for k in range(P + 1)]
    mons = cart_monomials(P, fdim) if fdim == 2 else [(a,) for a in range(P + 1)]
    V = [[Fr(np.prod([x[d] ** m[d] for d in range(fdim)], dtype=object)) for m in mons] for x in X]
    if fdim == 2:
        G = [[simplex_monomial_integral([a + b for a, b in zip(mk, ml)], 2) for ml in mons] for mk in mons]
    else:
        G = [[Fr(1, mk[0] + ml[0] + 1) for ml in mons] for mk in mons]
    Vi = inv(V)
    return X, mm(mm(T(Vi), G), Vi)


def to_np(A):
    return np.array([[float(x) for x in r] for r in A], dtype=np.float64)


def main():
    out = {}
    for P in range(1, 6):
        for dim in (3, 2):
            ops = element_ops(P, dim)
            tag = f"{dim}d_p{P}"
            out[f"D_{tag}"] = np.stack([to_np(d) for d in ops["D"]])
            out[f"Lift_{tag}"] = np.stack([to_np(L) for L in ops["Lift"]])
            out[f"w_{tag}"] = np.array([float(x) for x in ops["w"]])
            out[f"fmap_{tag}"] = np.array(ops["fmap"], dtype=np.int32)
            print(tag, "done", flush=True)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
