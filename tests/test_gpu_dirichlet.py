"""Inhomogeneous Dirichlet data on the device (boundarydata!,
src/boundarydata.jl:198-352; penalty of solve_direct!, src/solvers.jl:647-652)
against the oracle.

Bar: best-approximated boundary values within 1e-12 of max|target| of the
oracle's (sparse LU of the same boundary system; the device PCG stops at
relative residual 1e-14); interpolated values within 1e-14; entries off the
boundary untouched; scalar and two-component spaces; P2 (P1) Poisson and P1{2}
elasticity solutions with quadratic (linear) data and a quadratic (linear)
exact solution reproduce it at every dof to 1e-9."""
import ctypes

import numpy as np
import pytest

from meshes import shuffled_square

pytestmark = pytest.mark.gpu

Q2 = [1.0, 1.0, -2.0, 0.5, 1.0, 1.5]
CUBIC = [0.3, -0.2, 0.5, 0.1, 0.7, -0.4, 1.1, -0.6, 0.9, 0.25]


def _grid(grmp, O, N, seed):
    if seed is None:
        grid = grmp.grid_unitsquare(N)
        coords, cells = grid.download()
    else:
        coords, cells = shuffled_square(O, N, seed)
        grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    return grid, coords, cells


@pytest.mark.parametrize("fe", [1, 2])
@pytest.mark.parametrize("N,seed", [(9, None), (7, 3)])
def test_boundary_values_match_oracle(grmp, O, fe, N, seed):
    print(f"[case] boundary values fe={fe} N={N} seed={seed}", flush=True)
    grid, coords, cells = _grid(grmp, O, N, seed)
    fes = grmp.FESpace(grmp.H1P1(1) if fe == 1 else grmp.H1P2(1), grid)
    cf, fn, fc = O.tri_faces(cells, coords.shape[0])
    cd = fes.celldofs()
    nd = fes.ndofs
    data = grmp.PolynomialData(CUBIC)
    q = fe + data.bonus_quadorder
    ref, Mb, bd = O.boundary_bestapprox(fe, CUBIC, q, coords, cells, cd, cf, fc, nd)
    refi = O.boundary_interpolate(fe, CUBIC, coords, cells, cd, cf, fc, nd)
    assert np.array_equal(bd, fes.boundary_dofs())
    sentinel = np.full(nd + 5, 7.0)
    for kind in ("bestapprox", "interpolate"):
        t = grmp.DeviceVector(grid.ctx, nd + 5, sentinel)
        if kind == "bestapprox":
            it, rr = ctypes.c_int64(), ctypes.c_double()
            rc = grmp.lib().grmp_boundary_bestapprox(fes.h, grmp._abi.dptr(data.c), data.ncoef, q, 1e-14, t.h, 5,
                                                     ctypes.byref(it), ctypes.byref(rr))
        else:
            rc = grmp.lib().grmp_boundary_interpolate(fes.h, grmp._abi.dptr(data.c), data.ncoef, t.h, 5)
        grid.ctx.check(rc)
        v = t.download()
        assert np.all(v[:5] == 7.0)
        v = v[5:]
        off = np.setdiff1d(np.arange(nd), bd)
        assert np.all(v[off] == 7.0)
        r = ref if kind == "bestapprox" else refi
        err = np.max(np.abs(v[bd] - r[bd])) / np.max(np.abs(r[bd]))
        extra = f", {it.value} iterations, relres {rr.value:.1e}" if kind == "bestapprox" else ""
        print(f"  {kind}: max rel {err:.2e}{extra}")
        assert err <= (1e-12 if kind == "bestapprox" else 1e-14)


def test_multigrid_solve_with_dirichlet_data(grmp, O):
    """solve(solver="mg") lifts best-approximated data: P2 on a generated
    square reproduces the quadratic solution."""
    N = 16
    print(f"[case] multigrid solve with dirichlet data N={N}", flush=True)
    grid = grmp.grid_unitsquare(N)
    coords, cells = grid.download()
    fes = grmp.FESpace(grmp.H1P2(1), grid)
    pde = grmp.PoissonProblem(1.0)
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([-4.0])))
    pde.add_boundarydata(0, grmp.BestapproxDirichletBoundary, data=grmp.PolynomialData(Q2))
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, solver="mg", rtol=1e-12)
    cf, fn, fc = O.tri_faces(cells, coords.shape[0])
    P = np.vstack([coords, 0.5 * (coords[fn[:, 0]] + coords[fn[:, 1]])])
    u = O.poly_eval(Q2, P[:, 0], P[:, 1])
    err = np.max(np.abs(sol.entries() - u)) / np.max(np.abs(u))
    print(f"  {info.iterations} iterations, max rel err vs exact {err:.2e}")
    assert info.status == 0 and err <= 1e-9


@pytest.mark.parametrize("fe,kind", [(2, "bestapprox"), (1, "interpolate"), (1, "bestapprox")])
def test_poisson_with_dirichlet_data(grmp, O, fe, kind):
    """u = 1 + x - 2y + x^2/2 + xy + 3y^2/2 (P2, -Delta u = -4) or
    u = 0.5 + 2x - y (P1, -Delta u = 0) on a shuffled mesh."""
    N, seed = 8, 6
    print(f"[case] poisson with dirichlet data fe={fe} {kind}", flush=True)
    grid, coords, cells = _grid(grmp, O, N, seed)
    fes = grmp.FESpace(grmp.H1P1(1) if fe == 1 else grmp.H1P2(1), grid)
    coef, f = (Q2, -4.0) if fe == 2 else ([0.5, 2.0, -1.0], 0.0)
    pde = grmp.PoissonProblem(1.0)
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([f])))
    K = grmp.BestapproxDirichletBoundary if kind == "bestapprox" else grmp.InterpolateDirichletBoundary
    pde.add_boundarydata(0, K, data=grmp.PolynomialData(coef))
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, precision="fp64", rtol=1e-13)
    assert info.status == 0
    x = sol.entries()
    P = coords
    if fe == 2:
        cf, fn, fc = O.tri_faces(cells, coords.shape[0])
        P = np.vstack([coords, 0.5 * (coords[fn[:, 0]] + coords[fn[:, 1]])])
    u = O.poly_eval(coef, P[:, 0], P[:, 1])
    err = np.max(np.abs(x - u)) / np.max(np.abs(u))
    print(f"  {info.iterations} iterations, max rel err vs exact {err:.2e}")
    assert err <= 1e-9


@pytest.mark.parametrize("fe", [1, 2])
def test_vector_boundary_values(grmp, O, fe):
    """Two components with their own cubic polynomials on a shuffled mesh."""
    N, seed = 6, 8
    print(f"[case] vector boundary values fe={fe}", flush=True)
    grid, coords, cells = _grid(grmp, O, N, seed)
    fes = grmp.FESpace(grmp.H1P1(2) if fe == 1 else grmp.H1P2(2), grid)
    cf, fn, fc = O.tri_faces(cells, coords.shape[0])
    coef = np.array([CUBIC, CUBIC[::-1]])
    data = grmp.PolynomialData(coef)
    q = fe + data.bonus_quadorder
    nd = fes.ndofs
    for kind in ("bestapprox", "interpolate"):
        ref = O.boundary_values_vec(kind, fe, 2, coef, q, coords, cells, fes.celldofs(), cf, fc, nd)
        bd = np.nonzero(~np.isnan(ref))[0]
        assert np.array_equal(bd, fes.boundary_dofs())
        t = grmp.DeviceVector(grid.ctx, nd, np.zeros(nd))
        if kind == "bestapprox":
            rc = grmp.lib().grmp_boundary_bestapprox(fes.h, grmp._abi.dptr(data.c), data.ncoef, q, 1e-14, t.h, 0,
                                                     None, None)
        else:
            rc = grmp.lib().grmp_boundary_interpolate(fes.h, grmp._abi.dptr(data.c), data.ncoef, t.h, 0)
        grid.ctx.check(rc)
        v = t.download()
        err = np.max(np.abs(v[bd] - ref[bd])) / np.max(np.abs(ref[bd]))
        print(f"  {kind}: max rel {err:.2e}")
        assert err <= (1e-12 if kind == "bestapprox" else 1e-14)


def test_elasticity_with_linear_displacement(grmp, O):
    """A linear displacement has constant stress, so with f = 0 and its
    boundary values the P1{2} elasticity solution reproduces it."""
    N, seed = 8, 9
    print("[case] elasticity with linear Dirichlet displacement", flush=True)
    grid, coords, cells = _grid(grmp, O, N, seed)
    fes = grmp.FESpace(grmp.H1P1(2), grid)
    coef = np.array([[0.1, 0.5, -0.3], [-0.2, 0.25, 0.4]])
    pde = grmp.LinearElasticityProblem(1.0, 0.3)
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([0.0, 0.0])))
    pde.add_boundarydata(0, grmp.BestapproxDirichletBoundary, data=grmp.PolynomialData(coef))
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, precision="fp64", rtol=1e-13)
    assert info.status == 0
    x = sol.entries()
    nn = coords.shape[0]
    u = np.concatenate([O.poly_eval(coef[m], coords[:, 0], coords[:, 1]) for m in range(2)])
    err = np.max(np.abs(x - u)) / np.max(np.abs(u))
    print(f"  {info.iterations} iterations, max rel err vs exact {err:.2e}")
    assert x.size == 2 * nn and err <= 1e-9


def test_dirichlet_data_errors(grmp):
    print("[case] dirichlet data argument errors", flush=True)
    grid = grmp.grid_unitsquare(4)
    fesq = grmp.FESpace(grmp.L2P0(1), grid)  # no boundary dofs: not built
    t = grmp.DeviceVector(grid.ctx, fesq.ndofs)
    c = np.array([1.0, 1.0])
    with pytest.raises(grmp.GrmpError):
        grid.ctx.check(grmp.lib().grmp_boundary_interpolate(fesq.h, grmp._abi.dptr(c), 1, t.h, 0))
    fes = grmp.FESpace(grmp.H1P1(1), grid)
    short = grmp.DeviceVector(grid.ctx, fes.ndofs - 1)
    with pytest.raises(grmp.GrmpError):
        grid.ctx.check(grmp.lib().grmp_boundary_interpolate(fes.h, grmp._abi.dptr(c), 1, short.h, 0))
    with pytest.raises(grmp.GrmpError):
        grmp.PolynomialData(np.ones(16))
    with pytest.raises(grmp.GrmpError):
        grmp.PoissonProblem().add_boundarydata(0, grmp.BestapproxDirichletBoundary)
    pde = grmp.PoissonProblem()
    pde.add_boundarydata(0, grmp.BestapproxDirichletBoundary, data=grmp.PolynomialData([[1.0], [2.0]]))
    with pytest.raises(grmp.GrmpError):  # two-component data on a scalar space
        grmp.boundary_values(pde, grmp.FEVector([fes]))


@pytest.mark.parametrize("N,seed", [(6, None), (7, 5)])
def test_br_boundary_values_match_oracle(grmp, O, N, seed):
    """H1BR{2}: best approximation (5 dofs per face, components coupled by the
    normal bubble) within 1e-12 of the oracle's sparse-LU solve;
    interpolation (node values, flux-keeping bubble) within 1e-14."""
    print(f"[case] BR boundary values N={N} seed={seed}", flush=True)
    grid, coords, cells = _grid(grmp, O, N, seed)
    fes = grmp.FESpace(grmp.H1BR(2), grid)
    S = O.StokesSpace(N, coords, cells) if seed is not None else O.StokesSpace(N)
    assert np.array_equal(S.celldofs, fes.celldofs())
    g2 = np.array([CUBIC, [0.1, 0.4, -0.3, 0.2, -0.5, 0.6, -0.7, 0.3, 0.2, -0.1]])
    data = grmp.PolynomialData(g2)
    q = 2 + data.bonus_quadorder
    ref, bd = O.boundary_bestapprox_br(S, g2, q)
    refi = O.boundary_interpolate_br(S, g2)
    assert np.array_equal(bd, fes.boundary_dofs())
    nd = fes.ndofs
    for kind in ("bestapprox", "interpolate"):
        t = grmp.DeviceVector(grid.ctx, nd, np.full(nd, 7.0))
        if kind == "bestapprox":
            it, rr = ctypes.c_int64(), ctypes.c_double()
            rc = grmp.lib().grmp_boundary_bestapprox(fes.h, grmp._abi.dptr(data.c), data.ncoef, q, 1e-14, t.h, 0,
                                                     ctypes.byref(it), ctypes.byref(rr))
        else:
            rc = grmp.lib().grmp_boundary_interpolate(fes.h, grmp._abi.dptr(data.c), data.ncoef, t.h, 0)
        grid.ctx.check(rc)
        v = t.download()
        off = np.setdiff1d(np.arange(nd), bd)
        assert np.all(v[off] == 7.0)
        r = ref if kind == "bestapprox" else refi
        tol = 1e-12 if kind == "bestapprox" else 1e-14
        assert np.max(np.abs(v[bd] - r[bd])) <= tol * np.max(np.abs(r[bd])), kind
