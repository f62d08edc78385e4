"""Writes exact-rational golden fixtures for the assembly path.

Independent of the oracle and of the library: meshes, first-appearance face
numbering, P1/P2 basis gradients, quadrature and assembly are restated here
in exact rational arithmetic (fractions.Fraction), from the mathematics:

* Laplace stiffness  K_ij = kappa * sum_T |T| grad(phi_i).grad(phi_j)
  (integrand of degree <= 2 on P2, integrated exactly),
* load vector        b_i  = sum_T int_T phi_i,
* homogeneous Dirichlet penalty on all boundary dofs (diagonal = 1e60).

An entry is stored iff its exact value is nonzero.  The reference decides
its pattern with floating-point zero tests (src/fematrix.jl:54-65); on these
dyadic meshes the two coincide, which the tests assert.

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
                A[(dofs[a], dofs[bb])] = A.get((dofs[a], dofs[bb]), F(0)) + v
        # load vector, f = 1: P1 -> |T|/3 per vertex; P2 -> 0 vertex, |T|/3 per face
        for a in range(nd):
            if fe == 1:
                b[dofs[a]] += vol / 3
            elif a >= 3:
                b[dofs[a]] += vol / 3
    fixed = sorted(set(bnodes) | ({nn + f for f in bfaces} if fe == 2 else set()))
    return A, b, fixed, ndofs


def to_csc(A, ndofs, fixed, penalty=1e60):
    ent = {k: float(v) for k, v in A.items() if v != 0}
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


if __name__ == "__main__":
    main()
