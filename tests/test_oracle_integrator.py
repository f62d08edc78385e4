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


def _cell_integrals(coords, cells, fvals):
    """sum over an exact-to-degree-11 conical rule of fvals(l1, l2, x, y)"""
    P, W = _tri_rule(6)
    X = coords[cells]
    det = ((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1])
           - (X[:, 2, 0] - X[:, 0, 0]) * (X[:, 1, 1] - X[:, 0, 1]))
    out = np.zeros(cells.shape[0])
    for (a, b), w in zip(P, W):
        p = X[:, 0] + a * (X[:, 1] - X[:, 0]) + b * (X[:, 2] - X[:, 0])
        out += w * fvals(a, b, p[:, 0], p[:, 1])
    return out * np.abs(det)


def test_l2error_p1_interpolant(O):
    """L2ErrorIntegrator of the P1 interpolant of a quadratic g, quadorder
    3 (+1 for P1 = the Stroud rule of order 4, exact for the degree-4
    integrand): equal to the exact per-cell integral of (g - I g)^2; zero
    for linear data."""
    coords, cells, cellfaces, facenodes = _jittered_square(O, 6, 3)
    nn = coords.shape[0]
    cd, ndofs = O.celldofs(O.FE_P1, 1, cells, nn)
    coef = [1.0, 2.0, -0.5, 3.0, -1.5, 0.75]
    u = _poly(coords[:, 0], coords[:, 1])
    got = O.l2error_cells(O.FE_P1, 1, u, [coef], 4, 1.0, coords, cells, cd)

    def err2(l1, l2, x, y):
        uh = (1 - l1 - l2) * u[cells[:, 0]] + l1 * u[cells[:, 1]] + l2 * u[cells[:, 2]]
        return (_poly(x, y) - uh) ** 2
    ref = _cell_integrals(coords, cells, err2)
    assert np.all(ref > 0)
    assert np.max(np.abs(got - ref) / ref) < 1e-12
    lin = [0.5, -1.0, 2.0]
    ul = 0.5 - coords[:, 0] + 2.0 * coords[:, 1]
    z = O.l2error_cells(O.FE_P1, 1, ul, [lin], 2, 1.0, coords, cells, cd)
    assert np.max(z) < 1e-28


def test_l2norm_br_identity(O):
    """L2NormIntegrator of a Bernardi-Raugel function on a jittered mesh:
    P1 parts per component plus face bubbles 6 l_a l_b times the face
    normal, against the conical rule (degree 4 integrand, Stroud order 4)."""
    coords, cells, cellfaces, facenodes = _jittered_square(O, 5, 9)
    S = O.StokesSpace(5, coords, cells)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(S.nU)
    nrm = S.fnormal.reshape(-1, 2)[S.cellfaces]           # (nc, 3, 2)
    got = O.l2error_cells(O.FE_BR, 2, u, None, 4, -1.0, S.coords, S.cells, S.celldofs, normals=nrm)
    cdf = S.celldofs

    def sq(l1, l2, x, y):
        lam = [1 - l1 - l2, l1, l2]
        val = np.zeros((cells.shape[0], 2))
        for k in range(2):
            for a in range(3):
                val[:, k] += u[cdf[:, 3 * k + a]] * lam[a]
        for f in range(3):
            b = 6 * lam[f] * lam[(f + 1) % 3]
            val += (u[cdf[:, 6 + f]] * b)[:, None] * nrm[:, f, :]
        return (val ** 2).sum(axis=1)
    ref = _cell_integrals(S.coords, S.cells, sq)
    assert np.max(np.abs(got - ref) / ref) < 1e-12
