"""Pins of the oracle's ItemIntegrator (src/assemblypatterns/itemintegrator.jl:
169-360) and L2ErrorIntegrator (:40-89) by exact integrals: a function in
the discrete space is integrated exactly by the element's rule, and the
result is compared with an independent conical Gauss rule of high degree
(Gauss-Legendre x Gauss-Jacobi(1, 0) from scipy) on jittered meshes, cell
by cell (the reference checks integrals of interpolants the same way,
test/runtests.jl:224-264)."""
import numpy as np
import pytest
from scipy.special import roots_jacobi, roots_legendre


def _tri_rule(n):
    """points (barycentric l1, l2 = x, y on the reference triangle) and
    weights (sum 1/2) exact to degree 2n - 1"""
    xu, wu = roots_legendre(n)
    xv, wv = roots_jacobi(n, 1.0, 0.0)
    u, v = 0.5 * (xu + 1.0), 0.5 * (xv + 1.0)
    P, W = [], []
    for i in range(n):
        for j in range(n):
            P.append((u[i] * (1.0 - v[j]), v[j]))
            W.append(wu[i] * 0.5 * wv[j] * 0.25)
    return np.array(P), np.array(W)


def _jittered_square(O, N, seed, amp=0.2):
    coords, cells = O.unitsquare(N)
    rng = np.random.default_rng(seed)
    inner = np.all((coords > 0) & (coords < 1), axis=1)
    coords = coords.copy()
    coords[inner] += amp / N * (2 * rng.random((inner.sum(), 2)) - 1)
    cellfaces, facenodes, facecells = O.tri_faces(cells, coords.shape[0])
    return coords, cells, cellfaces, facenodes


def _poly(x, y):
    return 1.0 + 2.0 * x - 0.5 * y + 3.0 * x * x - 1.5 * x * y + 0.75 * y * y


def _exact_cell_integrals(coords, cells, f):
    P, W = _tri_rule(6)
    X = coords[cells]                      # [nc, 3, 2]
    det = ((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1])
           - (X[:, 2, 0] - X[:, 0, 0]) * (X[:, 1, 1] - X[:, 0, 1]))
    out = np.zeros(cells.shape[0])
    for (a, b), w in zip(P, W):
        p = X[:, 0] + a * (X[:, 1] - X[:, 0]) + b * (X[:, 2] - X[:, 0])
        out += w * f(p[:, 0], p[:, 1])
    return out * np.abs(det)


@pytest.mark.parametrize("fe", [1, 2])
def test_integrate_identity_exact(O, fe):
    """ItemIntegrator([Identity]) of the P2 interpolant of a quadratic
    (P1: of a linear function) equals its exact integral on every cell."""
    coords, cells, cellfaces, facenodes = _jittered_square(O, 7, 11)
    nn = coords.shape[0]
    cd, ndofs = O.celldofs(fe, 1, cells, nn, cellfaces, facenodes.shape[0])
    f = _poly if fe == 2 else (lambda x, y: 1.0 + 2.0 * x - 0.5 * y)
    pts = coords if fe == 1 else np.vstack([coords, 0.5 * (coords[facenodes[:, 0]] + coords[facenodes[:, 1]])])
    u = f(pts[:, 0], pts[:, 1])
    got = O.integrate_identity(u, fe, 1, fe, coords, cells, cd)[:, 0]
    ref = _exact_cell_integrals(coords, cells, f)
    assert np.max(np.abs(got - ref)) <= 1e-15 * np.max(np.abs(ref)) * 8


def test_integrate_identity_tet_exact(O):
    """P1 on jittered tetrahedra: the integral of a linear function is the
    volume times its mean over the vertices."""
    rng = np.random.default_rng(5)
    n = 3
    n1 = n + 1
    g = np.stack(np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij"), -1)
    coords = (g.transpose(2, 1, 0, 3).reshape(-1, 3) / n).astype(np.float64)  # node (k n1 + j) n1 + i
    inner = np.all((coords > 0) & (coords < 1), axis=1)
    coords[inner] += 0.1 / n * (2 * rng.random((inner.sum(), 3)) - 1)
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    cells = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                for p in perms:
                    pos = [i, j, k]
                    vs = [(pos[2] * n1 + pos[1]) * n1 + pos[0]]
                    for ax in p:
                        pos[ax] += 1
                        vs.append((pos[2] * n1 + pos[1]) * n1 + pos[0])
                    cells.append(vs)
    cells = np.array(cells, dtype=np.int32)
    u = 0.5 + coords[:, 0] - 2.0 * coords[:, 1] + 0.25 * coords[:, 2]
    got = O.integrate_identity(u, O.FE_P1, 1, 1, coords, cells, cells, geom=3)[:, 0]
    X = coords[cells]
    vol = np.abs(np.linalg.det(X[:, 1:] - X[:, :1])) / 6.0
    ref = vol * u[cells].mean(axis=1)
    assert abs(vol.sum() - 1.0) < 1e-13
    assert np.max(np.abs(got - ref)) <= 1e-15 * np.max(np.abs(ref)) * 8
