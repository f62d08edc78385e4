"""Pins of the oracle's inhomogeneous Dirichlet data (boundarydata!,
src/boundarydata.jl:198-352) on the unit square and on shuffled meshes.

* The boundary mass matrix sums to the perimeter (4).
* Data in the trace space (linear for P1, quadratic for P2) is reproduced
  exactly by the best approximation, and equals the interpolation.
* For data outside it, the error is L2-orthogonal to every trace basis
  function, checked with an independent 8-point rule per face.
* End to end: P2 Poisson with the quadratic solution u = 1 + x - 2y + x^2/2 + xy
  + 3y^2/2 (-Delta u = -4) and best-approximated Dirichlet data gives
  u_h = u at every dof (P2 contains u)."""
import numpy as np

from meshes import shuffled_square

Q2 = [1.0, 1.0, -2.0, 0.5, 1.0, 1.5]
CUBIC = [0.3, -0.2, 0.5, 0.1, 0.7, -0.4, 1.1, -0.6, 0.9, 0.25]
_MONO = ((0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3))


def _mesh(O, fe, N, seed=None):
    if seed is None:
        coords, cells = O.unitsquare(N)
    else:
        coords, cells = shuffled_square(O, N, seed)
    cf, fn, fc = O.tri_faces(cells, coords.shape[0])
    cd, nd = O.celldofs(fe, 1, cells, coords.shape[0], cf, fn.shape[0])
    return coords, cells, cd, nd, cf, fn, fc


def _points(coords, fn, fe):
    pts = coords
    if fe == 2:
        pts = np.vstack([coords, 0.5 * (coords[fn[:, 0]] + coords[fn[:, 1]])])
    return pts


def test_mass_sums_to_perimeter(O):
    for fe in (O.FE_P1, O.FE_P2):
        coords, cells, cd, nd, cf, fn, fc = _mesh(O, fe, 6, seed=2)
        t, Mb, bd = O.boundary_bestapprox(fe, [1.0], 0, coords, cells, cd, cf, fc, nd)
        assert abs(Mb.sum() - 4.0) < 1e-13
        assert np.allclose(t[bd], 1.0, rtol=0, atol=1e-14)


def test_trace_space_data_reproduced(O):
    for fe, coef in ((O.FE_P1, [0.5, 2.0, -1.0]), (O.FE_P2, Q2)):
        for seed in (None, 3):
            coords, cells, cd, nd, cf, fn, fc = _mesh(O, fe, 7, seed)
            deg = 1 if fe == O.FE_P1 else 2
            t, Mb, bd = O.boundary_bestapprox(fe, coef, fe + deg, coords, cells, cd, cf, fc, nd)
            ti = O.boundary_interpolate(fe, coef, coords, cells, cd, cf, fc, nd)
            P = _points(coords, fn, fe)
            exact = O.poly_eval(coef, P[:, 0], P[:, 1])
            assert np.max(np.abs(t[bd] - exact[bd])) < 1e-13
            assert np.array_equal(np.isnan(ti), np.isnan(t))
            assert np.max(np.abs(ti[bd] - exact[bd])) < 1e-15


def test_galerkin_orthogonality(O):
    """(g - g_h, phi_i)_boundary = 0 for every boundary trace basis function,
    integrated with an 8-point Gauss rule (exact for the degree-5 integrand)."""
    xg, wg = np.polynomial.legendre.leggauss(8)
    s8, w8 = 0.5 * (xg + 1.0), 0.5 * wg
    for fe in (O.FE_P1, O.FE_P2):
        coords, cells, cd, nd, cf, fn, fc = _mesh(O, fe, 5, seed=4)
        t, Mb, bd = O.boundary_bestapprox(fe, CUBIC, fe + 3, coords, cells, cd, cf, fc, nd)
        res = {d: 0.0 for d in bd.tolist()}
        for c, k in O.boundary_faces(cells, cf, fc):
            a, b = cells[c, k], cells[c, (k + 1) % 3]
            p0, p1 = coords[a], coords[b]
            L = np.hypot(*(p1 - p0))
            dofs = [cd[c, k], cd[c, (k + 1) % 3]] + ([cd[c, 3 + k]] if fe == O.FE_P2 else [])
            for s, w in zip(s8, w8):
                x, y = p0 + s * (p1 - p0)
                if fe == O.FE_P1:
                    phi = [1 - s, s]
                else:
                    phi = [(1 - s) * (1 - 2 * s), s * (2 * s - 1), 4 * s * (1 - s)]
                gh = sum(t[d] * ph for d, ph in zip(dofs, phi))
                g = sum(ck * x ** i * y ** j for ck, (i, j) in zip(CUBIC, _MONO))
                for d, ph in zip(dofs, phi):
                    res[d] += L * w * (g - gh) * ph
        assert max(abs(v) for v in res.values()) < 1e-14
        # and the projection differs from the interpolation for cubic data
        ti = O.boundary_interpolate(fe, CUBIC, coords, cells, cd, cf, fc, nd)
        assert np.max(np.abs(ti[bd] - t[bd])) > 1e-6


def test_p2_poisson_with_quadratic_data(O):
    N = 6
    coords, cells, cd, nd, cf, fn, fc = _mesh(O, O.FE_P2, N, seed=5)
    fixed = O.boundary_dofs(O.FE_P2, 1, coords.shape[0], fn, fc, fn.shape[0])
    A = O.assemble_matrix(nd, [dict(kind=O.BLF_LAPLACE, fe=O.FE_P2, quadorder=2, factor=1.0, coords=coords,
                                    cells=cells, celldofs=cd)], fixed=fixed)
    b = np.zeros(nd)
    O.assemble_lf_const(b, O.FE_P2, 1, 2, [-4.0], coords, cells, cd)
    t, Mb, bd = O.boundary_bestapprox(O.FE_P2, Q2, 4, coords, cells, cd, cf, fc, nd)
    assert np.array_equal(np.sort(bd), np.sort(fixed.astype(np.int64)))
    b = O.penalized_rhs(b, fixed, values=t[fixed])
    x = O.direct_solve(A, b)
    P = _points(coords, fn, 2)
    u = O.poly_eval(Q2, P[:, 0], P[:, 1])
    assert np.max(np.abs(x - u)) < 1e-11


def _br_space(O, N, seed):
    if seed is None:
        return O.StokesSpace(N)
    coords, cells = shuffled_square(O, N, seed)
    return O.StokesSpace(N, coords, cells)


def test_br_linear_data_reproduced(O):
    """H1BR{2}: linear data lies in the trace space (P1 part, zero bubble):
    the best approximation and the interpolation both reproduce it."""
    lin = [[1.0, 0.0, 1.0], [-1.0, 1.0, 0.0]]     # (y + 1, x - 1): the reference's Stokes test data
    for seed in (None, 6):
        S = _br_space(O, 5, seed)
        t, bd = O.boundary_bestapprox_br(S, lin, 3)
        ti = O.boundary_interpolate_br(S, lin)
        nn = S.nnodes
        assert np.array_equal(np.isnan(ti), np.isnan(t))
        nodes = bd[bd < nn]
        X, Y = S.coords[nodes, 0], S.coords[nodes, 1]
        assert np.max(np.abs(t[nodes] - (Y + 1))) < 1e-14
        assert np.max(np.abs(t[nodes + nn] - (X - 1))) < 1e-14
        bub = bd[bd >= 2 * nn]
        assert bub.size == 4 * S.N and np.max(np.abs(t[bub])) < 1e-14
        assert np.max(np.abs(ti[bd] - t[bd])) < 1e-14


def test_br_galerkin_orthogonality(O):
    """Cubic data: (g - g_h, v_i) = 0 on the boundary for all five face
    functions (8-point Gauss per face), and the interpolation keeps each
    face's normal flux int_F g.n."""
    xg, wg = np.polynomial.legendre.leggauss(8)
    s8, w8 = 0.5 * (xg + 1.0), 0.5 * wg
    g2 = [CUBIC, [0.1, 0.4, -0.3, 0.2, -0.5, 0.6, -0.7, 0.3, 0.2, -0.1]]
    S = _br_space(O, 4, 7)
    t, bd = O.boundary_bestapprox_br(S, g2, 5)
    ti = O.boundary_interpolate_br(S, g2)
    res = {d: 0.0 for d in bd.tolist()}
    for c, k, dofs, n in O._br_face_units(S):
        a, b = S.cells[c, k], S.cells[c, (k + 1) % 3]
        p0, p1 = S.coords[a], S.coords[b]
        L = np.hypot(*(p1 - p0))
        flux_g = flux_i = 0.0
        for s, w in zip(s8, w8):
            x, y = p0 + s * (p1 - p0)
            bub = 6 * s * (1 - s)
            V = [np.array([1 - s, 0]), np.array([s, 0]), np.array([0, 1 - s]), np.array([0, s]), bub * n]
            g = np.array([sum(ck * x ** i * y ** j for ck, (i, j) in zip(gc, _MONO)) for gc in g2])
            gh = sum(t[d] * v for d, v in zip(dofs, V))
            gi = sum(ti[d] * v for d, v in zip(dofs, V))
            for d, v in zip(dofs, V):
                res[d] += L * w * float((g - gh) @ v)
            flux_g += L * w * float(g @ n)
            flux_i += L * w * float(gi @ n)
        assert abs(flux_g - flux_i) < 1e-14
    assert max(abs(v) for v in res.values()) < 1e-14
