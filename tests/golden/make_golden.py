"""Writes exact-rational golden fixtures for the assembly path.

Independent of the oracle and of the library: meshes, first-appearance face
numbering, P1/P2 basis gradients, quadrature and assembly are restated here
in exact rational arithmetic (fractions.Fraction), from the mathematics:

* Laplace stiffness  K_ij = kappa * sum_T |T| grad(phi_i).grad(phi_j)
  (integrand of degree <= 2 on P2, integrated exactly),
* load vector        b_i  = sum_T int_T phi_i,
* homogeneous Dirichlet penalty on all boundary dofs (diagonal = 1e60),
* P1 mass matrix     M_ij = sum_T |T|/12 (1 + delta_ij),
* P1 tetrahedra      K_ij = sum_T |T| grad(lambda_i).grad(lambda_j) on the
  Kuhn cube (six tetrahedra per cube along the paths 000 -> 111),
* Bernardi-Raugel    H1BR{2} x L2P0 Stokes blocks: nu (grad u, grad v) and
  -(div v, q) with its transposed copy, face bubbles 6 l_a l_b times the
  face normal (src/fedefs/h1v_br.jl:117-130,150-162), in exact arithmetic
  over Q(sqrt 2) (the diagonal faces have normals (1, -1)/sqrt 2).

An entry is stored iff some cell contributes an exactly nonzero local value
to it: the reference's rule (_addnz skips v == 0, src/fematrix.jl:54-65),
applied to exact values.  The reference decides it with floating-point zero
tests; on these meshes the two coincide, which the tests assert.

Run:  python tests/golden/make_golden.py   (writes tests/golden/*.json)
"""
import json
import os
from fractions import Fraction as F

HERE = os.path.dirname(os.path.abspath(__file__))


def unitsquare(N):
    n1 = N + 1
    coords = [(F(i, N), F(j, N)) for j in range(n1) for i in range(n1)]
    cells = []
    for j in range(N):
        for i in range(N):
            v00 = j * n1 + i
            v10, v01, v11 = v00 + 1, v00 + n1, v00 + n1 + 1
            cells.append((v00, v10, v11))
            cells.append((v00, v11, v01))
    return coords, cells


def faces(cells):
    """first appearance in cell order, local faces (1,2), (2,3), (3,1)"""
    ids, cellfaces, count = {}, [], {}
    for c in cells:
        cf = []
        for a, b in ((0, 1), (1, 2), (2, 0)):
            key = (min(c[a], c[b]), max(c[a], c[b]))
            if key not in ids:
                ids[key] = len(ids)
            count[key] = count.get(key, 0) + 1
            cf.append(ids[key])
        cellfaces.append(cf)
    bfaces = sorted(ids[k] for k, v in count.items() if v == 1)
    bnodes = sorted({n for k, v in count.items() if v == 1 for n in k})
    return cellfaces, len(ids), bfaces, bnodes


def lam_grads(X):
    (x1, y1), (x2, y2), (x3, y3) = X
    det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)
    g = [((y2 - y3) / det, (x3 - x2) / det), ((y3 - y1) / det, (x1 - x3) / det),
         ((y1 - y2) / det, (x2 - x1) / det)]
    return g, abs(det) / 2


def p2_grad(g, lam, a):
    """gradient of the P2 basis (nodes, then faces (1,2),(2,3),(3,1)) as a
    linear combination of barycentric gradients, at barycentric point lam"""
    if a < 3:
        s = 4 * lam[a] - 1
        return (s * g[a][0], s * g[a][1])
    i, j = ((0, 1), (1, 2), (2, 0))[a - 3]
    return (4 * (lam[i] * g[j][0] + lam[j] * g[i][0]), 4 * (lam[i] * g[j][1] + lam[j] * g[i][1]))


def assemble(N, fe, kappa):
    coords, cells = unitsquare(N)
    nn = len(coords)
    cellfaces, nf, bfaces, bnodes = faces(cells)
    ndofs = nn + (nf if fe == 2 else 0)
    A, b = {}, [F(0)] * ndofs
    # face-midpoint rule in barycentric coordinates, exact for degree 2
    qp = [(F(0), F(1, 2), F(1, 2)), (F(1, 2), F(0), F(1, 2)), (F(1, 2), F(1, 2), F(0))]
    for c, cf in zip(cells, cellfaces):
        X = [coords[v] for v in c]
        g, vol = lam_grads(X)
        dofs = list(c) + ([nn + f for f in cf] if fe == 2 else [])
        nd = len(dofs)
        for a in range(nd):
            for bb in range(nd):
                if fe == 1:
                    v = vol * (g[a][0] * g[bb][0] + g[a][1] * g[bb][1])
                else:
                    v = F(0)
                    for lam in qp:
                        ga, gb = p2_grad(g, lam, a), p2_grad(g, lam, bb)
                        v += vol / 3 * (ga[0] * gb[0] + ga[1] * gb[1])
                v *= kappa
                add(A, dofs[a], dofs[bb], v)
        # load vector, f = 1: P1 -> |T|/3 per vertex; P2 -> 0 vertex, |T|/3 per face
        for a in range(nd):
            if fe == 1:
                b[dofs[a]] += vol / 3
            elif a >= 3:
                b[dofs[a]] += vol / 3
    fixed = sorted(set(bnodes) | ({nn + f for f in bfaces} if fe == 2 else set()))
    return A, b, fixed, ndofs


def add(A, i, j, v):
    """_addnz: a local value enters (and creates) the entry only when nonzero"""
    if v != 0:
        A[(i, j)] = A.get((i, j), 0) + v


def to_csc(A, ndofs, fixed, penalty=1e60):
    ent = {k: float(v) for k, v in A.items()}
    for d in fixed:
        ent[(d, d)] = penalty
    cols = [[] for _ in range(ndofs)]
    for (i, j), v in ent.items():
        cols[j].append((i, v))
    colptr, rowval, nzval = [0], [], []
    for j in range(ndofs):
        for i, v in sorted(cols[j]):
            rowval.append(i)
            nzval.append(v)
        colptr.append(len(rowval))
    return colptr, rowval, nzval


# ------------------------------------------------------------- P1 mass

def assemble_mass(N):
    """consistent P1 mass matrix on the unit square: |T|/12 (1 + delta_ij)"""
    coords, cells = unitsquare(N)
    A = {}
    for c in cells:
        _, vol = lam_grads([coords[v] for v in c])
        for a in range(3):
            for bb in range(3):
                add(A, c[a], c[bb], vol / 12 * (2 if a == bb else 1))
    return A, len(coords)


# ----------------------------------------------------- Kuhn tetrahedra

KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def unitcube(n):
    """n^3 cubes, node (i, j, k) -> (k (n+1) + j)(n+1) + i; cube (i, j, k) ->
    six tetrahedra along the paths 000 -> e_a -> e_a + e_b -> 111 for the
    permutations (a, b, c) of KUHN, cell 6 (i + n (j + n k)) + p"""
    n1 = n + 1
    node = lambda p: (p[2] * n1 + p[1]) * n1 + p[0]
    coords = [(F(i, n), F(j, n), F(k, n)) for k in range(n1) for j in range(n1) for i in range(n1)]
    cells = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                for perm in KUHN:
                    pos = [i, j, k]
                    vs = [node(pos)]
                    for ax in perm:
                        pos[ax] += 1
                        vs.append(node(pos))
                    cells.append(tuple(vs))
    return coords, cells


def inv3(A):
    """inverse of a 3x3 rational matrix (adjugate / det) and det"""
    (a, b, c), (d, e, f), (g, h, i) = A
    det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g)
    adj = [[e * i - f * h, c * h - b * i, b * f - c * e],
           [f * g - d * i, a * i - c * g, c * d - a * f],
           [d * h - e * g, b * g - a * h, a * e - b * d]]
    return [[x / det for x in row] for row in adj], det


def assemble_tet(n):
    coords, cells = unitcube(n)
    A = {}
    for c in cells:
        X = [coords[v] for v in c]
        J = [[X[k + 1][r] - X[0][r] for k in range(3)] for r in range(3)]  # columns x_k - x_0
        Ji, det = inv3(J)
        # grad lambda_{k+1} = row k of J^{-1}; grad lambda_0 = -sum
        g = [None] + [Ji[k] for k in range(3)]
        g[0] = [-(g[1][r] + g[2][r] + g[3][r]) for r in range(3)]
        vol = abs(det) / 6
        for a in range(4):
            for bb in range(4):
                add(A, c[a], c[bb], vol * sum(g[a][r] * g[bb][r] for r in range(3)))
    return A, coords, cells


# ------------------------------------------------- Q(sqrt 2) arithmetic

class Q2:
    """a + b sqrt(2) with rational a, b"""
    __slots__ = ("a", "b")

    def __init__(self, a, b=0):
        self.a, self.b = F(a), F(b)

    def _c(self, o):
        return o if isinstance(o, Q2) else Q2(o)

    def __add__(self, o):
        o = self._c(o)
        return Q2(self.a + o.a, self.b + o.b)
    __radd__ = __add__

    def __sub__(self, o):
        o = self._c(o)
        return Q2(self.a - o.a, self.b - o.b)

    def __rsub__(self, o):
        return self._c(o) - self

    def __neg__(self):
        return Q2(-self.a, -self.b)

    def __mul__(self, o):
        o = self._c(o)
        return Q2(self.a * o.a + 2 * self.b * o.b, self.a * o.b + self.b * o.a)
    __rmul__ = __mul__

    def inv(self):
        d = self.a * self.a - 2 * self.b * self.b
        return Q2(self.a / d, -self.b / d)

    def __truediv__(self, o):
        return self * self._c(o).inv()

    def __eq__(self, o):
        o = self._c(o)
        return self.a == o.a and self.b == o.b

    def __ne__(self, o):
        return not self == o

    def __hash__(self):
        return hash((self.a, self.b))

    def __float__(self):
        # correctly rounded: |b| sqrt 2 via an exact integer square root
        from math import isqrt
        if self.b == 0:
            return float(self.a)
        scale = 1 << 200
        s2 = F(isqrt(2 * scale * scale), scale)  # sqrt 2 to 2^-200
        return float(self.a + self.b * s2)


def sqrt_sq(x):
    """sqrt of a rational that is q^2 or 2 q^2, in Q(sqrt 2)"""
    from math import isqrt
    x = F(x)
    for k, root in ((1, None), (2, None)):
        y = x / k
        nu, de = y.numerator, y.denominator
        rn, rd = isqrt(nu), isqrt(de)
        if rn * rn == nu and rd * rd == de:
            return Q2(F(rn, rd)) if k == 1 else Q2(0, F(rn, rd))
    raise ValueError(x)


def assemble_br_stokes(N, nu):
    """H1BR{2} x L2P0 on the N x N square: [1,1] = nu (grad u, grad v)
    (symmetric), [1,2] = -(div v, q), [2,1] its transposed copy (+), the
    no-slip fixed dofs and the pinned first pressure dof."""
    coords, cells = unitsquare(N)
    nn = len(coords)
    ids, orient, cellfaces, count = {}, {}, [], {}
    for c in cells:
        cf = []
        for a, b in ((0, 1), (1, 2), (2, 0)):
            key = (min(c[a], c[b]), max(c[a], c[b]))
            if key not in ids:
                ids[key] = len(ids)
                orient[ids[key]] = (c[a], c[b])   # global orientation: first appearance
            count[key] = count.get(key, 0) + 1
            cf.append(ids[key])
        cellfaces.append(cf)
    nf = len(ids)
    normal = {}
    for f, (a, b) in orient.items():
        tx, ty = coords[b][0] - coords[a][0], coords[b][1] - coords[a][1]
        L = sqrt_sq(tx * tx + ty * ty)
        normal[f] = (Q2(ty) / L, Q2(-tx) / L)
    nU = 2 * nn + nf
    A = {}
    mids = [(F(1, 2), F(0)), (F(1, 2), F(1, 2)), (F(0), F(1, 2))]  # reference face midpoints
    for e, (c, cf) in enumerate(zip(cells, cellfaces)):
        X = [coords[v] for v in c]
        g, vol = lam_grads(X)
        dofs = [c[0], c[1], c[2], nn + c[0], nn + c[1], nn + c[2]] + [2 * nn + f for f in cf]

        def grad(dof, lam):
            """2x2 gradient [component][direction] of BR basis function dof"""
            if dof < 6:
                comp, v = divmod(dof, 3)
                G = [[Q2(0), Q2(0)], [Q2(0), Q2(0)]]
                G[comp] = [Q2(g[v][0]), Q2(g[v][1])]
                return G
            k = dof - 6
            i, j = k, (k + 1) % 3
            db = [6 * (lam[i] * g[j][r] + lam[j] * g[i][r]) for r in range(2)]
            n = normal[cf[k]]
            return [[n[0] * db[0], n[0] * db[1]], [n[1] * db[0], n[1] * db[1]]]

        pts = []
        for x0, x1 in mids:
            pts.append((1 - x0 - x1, x0, x1))
        for a in range(9):
            for bb in range(9):
                v = Q2(0)
                for lam in pts:
                    Ga, Gb = grad(a, lam), grad(bb, lam)
                    v = v + vol / 3 * sum((Ga[p][r] * Gb[p][r] for p in range(2) for r in range(2)), Q2(0))
                add(A, dofs[a], dofs[bb], v * nu)
        lamc = (F(1, 3), F(1, 3), F(1, 3))
        for a in range(9):
            G = grad(a, lamc)
            div = (G[0][0] + G[1][1]) * vol
            add(A, dofs[a], nU + e, -div)
            add(A, nU + e, dofs[a], div)
    bn = sorted({n for (a, b), v in count.items() if v == 1 for n in (a, b)})
    bf = sorted(ids[k] for k, v in count.items() if v == 1)
    fixed = sorted(set(bn) | {n + nn for n in bn} | {2 * nn + f for f in bf} | {nU})
    return A, nU + len(cells), fixed


def main():
    cases = [(2, 1, F(1)), (4, 1, F(1)), (2, 2, F(1)), (4, 2, F(1, 1000)), (8, 2, F(1, 1000))]
    for N, fe, kappa in cases:
        A, b, fixed, ndofs = assemble(N, fe, kappa)
        colptr, rowval, nzval = to_csc(A, ndofs, fixed)
        name = f"laplace_p{fe}_N{N}.json"
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump({"N": N, "fe": fe, "kappa": str(kappa), "ndofs": ndofs, "fixed": fixed,
                       "colptr": colptr, "rowval": rowval, "nzval": nzval,
                       "b": [float(v) for v in b]}, fh)
        print(name, ndofs, len(rowval))
    for N in (2, 4):
        A, nn = assemble_mass(N)
        colptr, rowval, nzval = to_csc(A, nn, [])
        name = f"mass_p1_N{N}.json"
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump({"N": N, "ndofs": nn, "colptr": colptr, "rowval": rowval, "nzval": nzval}, fh)
        print(name, nn, len(rowval))
    for n in (2, 3):
        A, coords, cells = assemble_tet(n)
        colptr, rowval, nzval = to_csc(A, len(coords), [])
        name = f"tet_p1_n{n}.json"
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump({"n": n, "ndofs": len(coords), "coords": [[float(x) for x in p] for p in coords],
                       "cells": [list(c) for c in cells], "colptr": colptr, "rowval": rowval,
                       "nzval": nzval}, fh)
        print(name, len(coords), len(rowval))
    for N in (2, 4):
        nu = F(1, 100)
        A, n, fixed = assemble_br_stokes(N, nu)
        colptr, rowval, nzval = to_csc(A, n, fixed)
        name = f"stokes_br_N{N}.json"
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump({"N": N, "nu": str(nu), "n": n, "fixed": fixed, "colptr": colptr, "rowval": rowval,
                       "nzval": nzval}, fh)
        print(name, n, len(rowval))


if __name__ == "__main__":
    main()
