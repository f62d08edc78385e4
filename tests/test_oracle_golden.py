"""The oracle against exact-rational golden fixtures (tests/golden) and
closed forms.  CPU only."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as fh:
        return json.load(fh)


@pytest.mark.parametrize("name", ["laplace_p1_N2.json", "laplace_p1_N4.json", "laplace_p2_N2.json",
                                  "laplace_p2_N4.json", "laplace_p2_N8.json"])
def test_laplace_matrix_and_load_vector(O, name):
    g = _load(name)
    fe = O.FE_P1 if g["fe"] == 1 else O.FE_P2
    S = O.Space(g["N"], fe)
    assert S.ndofs == g["ndofs"]
    assert np.array_equal(S.fixed, np.array(g["fixed"]))
    kappa = float(Fraction(g["kappa"]))
    q = 0 if fe == O.FE_P1 else 2
    A = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_LAPLACE, fe=fe, quadorder=q, factor=kappa,
                                        coords=S.coords, cells=S.cells, celldofs=S.celldofs)], fixed=S.fixed)
    # pattern: the FP zero test gives the exact-arithmetic pattern on dyadic meshes
    assert np.array_equal(A.colptr, np.array(g["colptr"]))
    assert np.array_equal(A.rowval, np.array(g["rowval"]))
    ref = np.array(g["nzval"])
    assert np.max(np.abs(A.nzval - ref) / np.abs(ref)) < 4e-16
    b = np.zeros(S.ndofs)
    O.assemble_lf_const(b, fe, 1, fe, [1.0], S.coords, S.cells, S.celldofs)
    bref = np.array(g["b"])
    # P1 uses the midpoint rule whose point 1/3 is not a dyadic: a few ulps
    assert np.all(np.abs(b - bref) <= 1e-15 * np.abs(bref))


def test_p1_five_point_stencil(O):
    """P1 Laplace on the right-isosceles mesh is the 5-point stencil."""
    N = 8
    S = O.Space(N, O.FE_P1)
    A = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_LAPLACE, fe=O.FE_P1, quadorder=0, factor=1.0,
                                        coords=S.coords, cells=S.cells, celldofs=S.celldofs)])
    M = A.to_scipy().toarray()
    n1 = N + 1
    for j in range(1, N):
        for i in range(1, N):
            r = j * n1 + i
            row = {c: M[r, c] for c in np.nonzero(M[r])[0]}
            assert row == {r: 4.0, r - 1: -1.0, r + 1: -1.0, r - n1: -1.0, r + n1: -1.0}
    # K 1 = 0 before boundary conditions
    assert np.max(np.abs(M.sum(axis=1))) < 1e-14


def test_quadrature_exactness(O):
    """Triangle rules integrate monomials of their order exactly
    (the reference's own check, test/runtests.jl:81-137)."""
    import ctypes
    from math import factorial
    L = O.lib()
    for order, deg in [(1, 1), (2, 2)]:
        xr = np.zeros(32)
        w = np.zeros(16)
        f = L.or_quad_tri
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int, np.ctypeslib.ndpointer(np.float64), np.ctypeslib.ndpointer(np.float64)]
        nq = f(order, xr, w)
        for a in range(deg + 1):
            for b in range(deg + 1 - a):
                exact = factorial(a) * factorial(b) / factorial(a + b + 2)
                got = sum(w[q] * xr[2 * q] ** a * xr[2 * q + 1] ** b for q in range(nq)) * 0.5
                assert abs(got - exact) < 1e-15


def test_faces_first_appearance(O):
    """Face numbering by first appearance: count and boundary of the
    structured square (3N^2+2N faces, 4N boundary faces)."""
    for N in [1, 3, 8]:
        S = O.Space(N, O.FE_P2)
        assert S.nfaces == 3 * N * N + 2 * N
        assert int(np.sum(S.facecells[:, 1] < 0)) == 4 * N
        # every interior face is shared by exactly two cells
        cnt = np.bincount(S.cellfaces.ravel(), minlength=S.nfaces)
        assert set(cnt.tolist()) <= {1, 2}


def _csc_close(A, g, rtol):
    """pattern identical; every entry within rtol of the golden's (exact) value"""
    assert np.array_equal(A.colptr, np.array(g["colptr"]))
    assert np.array_equal(A.rowval, np.array(g["rowval"]))
    ref = np.array(g["nzval"])
    err = np.abs(A.nzval - ref)
    assert np.all(err <= rtol * np.abs(ref)), float(np.max(err / np.maximum(np.abs(ref), 1e-300)))


@pytest.mark.parametrize("N", [2, 4])
def test_p1_mass_matrix(O, N):
    """P1 mass |T|/12 (1 + delta_ij) (the 3-point rule is exact on P1 x P1;
    a one-point or lumped rule fails the off-diagonal values)."""
    g = _load(f"mass_p1_N{N}.json")
    S = O.Space(N, O.FE_P1)
    M = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_MASS, fe=O.FE_P1, quadorder=2, factor=1.0,
                                        coords=S.coords, cells=S.cells, celldofs=S.celldofs)])
    # the rule's weight 1/3 is not dyadic: a few ulps
    _csc_close(M, g, 4e-16)


@pytest.mark.parametrize("n", [2, 3])
def test_tet_p1_stiffness(O, n):
    """P1 stiffness on the Kuhn cube: the 7-point stencil with its exact
    values (off-path couplings of each tetrahedron are exactly zero, so the
    reference's zero test drops them)."""
    g = _load(f"tet_p1_n{n}.json")
    coords = np.array(g["coords"])
    cells = np.array(g["cells"], dtype=np.int32)
    A = O.assemble_matrix(g["ndofs"], [dict(kind=O.BLF_LAPLACE, geom=3, fe=O.FE_P1, quadorder=0, factor=1.0,
                                           coords=coords, cells=cells, celldofs=cells)])
    # |det|/6 is not dyadic: a few ulps
    _csc_close(A, g, 8e-16)
    # 7-point stencil at the interior node: 6/n, -1/n on the axis neighbours
    if n == 3:
        n1 = n + 1
        r = (1 * n1 + 1) * n1 + 1
        M = A.to_scipy().tocsr()
        row = M.getrow(r)
        assert sorted(row.indices.tolist()) == sorted([r, r - 1, r + 1, r - n1, r + n1, r - n1 * n1, r + n1 * n1])
        assert abs(M[r, r] - 6.0 / n) < 1e-14 and all(abs(M[r, c] + 1.0 / n) < 1e-14 for c in row.indices if c != r)


@pytest.mark.parametrize("N", [2, 4])
def test_br_stokes_blocks(O, N):
    """H1BR{2} x L2P0: the viscous block nu (grad u, grad v), the divergence
    block -(div v, q) and its transposed copy, with the no-slip penalty,
    against exact arithmetic over Q(sqrt 2)."""
    g = _load(f"stokes_br_N{N}.json")
    S = O.StokesSpace(N)
    assert S.n == g["n"] and np.array_equal(S.fixed, np.array(g["fixed"]))
    A = S.matrix(float(Fraction(g["nu"])))
    _csc_close(A, g, 2e-15)
