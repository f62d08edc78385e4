"""Algebraic multigrid CG (grmp_amg_*) on meshes the geometric multigrid
does not take: uploaded (shuffled, jittered) triangle meshes with P1, P2 and
P1{2} elasticity, and jittered Kuhn tetrahedra, all with log-normal cell
coefficients.

Bar: solutions within 1e-8 of the sparse-LU solution of the same penalized
system (scipy SuperLU in place of UMFPACK, src/solvers.jl:656-660); a
quadratic solution reproduced by P2 with inhomogeneous Dirichlet data to
1e-9; two solves and two hierarchies give the same bits; iterations stay
bounded as the mesh is refined (the Jacobi CG's grow with the mesh)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from meshes import shuffled_square

pytestmark = pytest.mark.gpu


def _poisson(grmp, O, fe, N, seed, sigma=1.0, uploaded=True):
    if uploaded:
        coords, cells = shuffled_square(O, N, seed)
        grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    else:
        grid = grmp.grid_unitsquare(N)
    ft = {1: grmp.H1P1(1), 2: grmp.H1P2(1), 3: grmp.H1P1(2)}[fe]
    fes = grmp.FESpace(ft, grid)
    kap = grmp.DeviceVector(grid.ctx, grid.ncells)
    kap.lognormal(seed, 0.0, sigma)
    if fe == 3:
        pde = grmp.LinearElasticityProblem(1.0, 0.3, cell_coefficient=grmp.CellCoefficient(kap))
        pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([0.3, -1.0])))
    else:
        pde = grmp.PoissonProblem(cell_coefficient=grmp.CellCoefficient(kap))
        pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
    pde.add_boundarydata(0)
    return grid, fes, pde, kap


def _lu(A, b):
    return spla.spsolve(A.to_scipy().tocsc(), b.entries())


@pytest.mark.parametrize("fe,N", [(1, 40), (2, 24), (3, 24)])
def test_amg_matches_lu_uploaded(grmp, O, fe, N):
    print(f"[case] amg vs LU, fe {fe}, uploaded N={N}", flush=True)
    grid, fes, pde, kap = _poisson(grmp, O, fe, N, seed=5)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, solver="amg", rtol=1e-12)
    xref = _lu(A, b)
    err = np.linalg.norm(sol.entries() - xref) / np.linalg.norm(xref)
    amg = grmp.AlgebraicMultigridSolver(A)
    nl, coarse, cx = amg.info()
    print(f"  {info.iterations} iterations, rel err {err:.2e}, {nl} levels, coarsest {coarse} rows, "
          f"operator complexity {cx:.2f}")
    assert info.status == 0 and err <= 1e-8
    assert nl >= 2 and coarse <= 64 and cx < 2.5
    assert info.iterations < (120 if fe == 3 else 60)


def test_amg_tetrahedra(grmp, O):
    print("[case] amg on jittered Kuhn tetrahedra n=14", flush=True)
    grid = grmp.grid_unitcube(14, 0.1, 3)
    fes = grmp.FESpace(grmp.H1P1(1), grid)
    kap = grmp.DeviceVector(grid.ctx, grid.ncells)
    kap.lognormal(11, 0.0, 0.5)
    pde = grmp.PoissonProblem(cell_coefficient=grmp.CellCoefficient(kap))
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
    pde.add_boundarydata(0)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, solver="amg", rtol=1e-12)
    xref = _lu(A, b)
    err = np.linalg.norm(sol.entries() - xref) / np.linalg.norm(xref)
    print(f"  {info.iterations} iterations, rel err {err:.2e}")
    assert info.status == 0 and err <= 1e-8 and info.iterations < 40


def test_amg_inhomogeneous_data_exact(grmp, O):
    """P2 reproduces a quadratic solution with best-approximated data: the
    penalized rows carry the data into the solve (x_d = b_d / A_dd)."""
    print("[case] amg P2 quadratic solution, uploaded mesh", flush=True)
    coords, cells = shuffled_square(O, 14, 2)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    fes = grmp.FESpace(grmp.H1P2(1), grid)
    pde = grmp.PoissonProblem(1.0)
    # u = x^2 + x y - y^2 + 0.5: -lap u = -2 + 2 = 0
    pde.add_boundarydata(0, grmp.BestapproxDirichletBoundary, data=grmp.PolynomialData([0.5, 0, 0, 1.0, 1.0, -1.0]))
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, solver="amg", rtol=1e-13)
    err = float(grmp.evaluate(grmp.L2ErrorIntegrator(grmp.PolynomialData([0.5, 0, 0, 1.0, 1.0, -1.0]), quadorder=4),
                              sol[0])) ** 0.5
    print(f"  {info.iterations} iterations, L2 error {err:.2e}")
    assert info.status == 0 and err < 1e-9


def test_amg_deterministic_and_scaling(grmp, O):
    print("[case] amg determinism and iteration growth", flush=True)
    its = []
    for N in (32, 96):
        grid, fes, pde, kap = _poisson(grmp, O, 1, N, seed=9)
        sol = grmp.FEVector([fes])
        A, b, fixed = grmp.assemble(pde, sol)
        m1 = grmp.AlgebraicMultigridSolver(A)
        x1, x2, x3 = grmp.FEVector([fes]), grmp.FEVector([fes]), grmp.FEVector([fes])
        i1 = m1.solve(b, x1, 1e-10)
        m1.solve(b, x2, 1e-10)
        m2 = grmp.AlgebraicMultigridSolver(A)
        m2.solve(b, x3, 1e-10)
        assert np.array_equal(x1.entries(), x2.entries()) and np.array_equal(x1.entries(), x3.entries())
        ij = grmp.solve_system(A, b, grmp.FEVector([fes]), grmp.SolverConfig(rtol=1e-10))
        print(f"  N={N}: amg {i1.iterations} iterations, Jacobi CG {ij.iterations}")
        its.append((i1.iterations, ij.iterations))
    # the Jacobi count grows about 3x from N=32 to N=96, the multigrid's far less
    assert its[1][0] < 1.6 * its[0][0] + 5 and its[1][1] > 2 * its[0][1]


def test_amg_auto_on_uploaded_mesh(grmp, O):
    print("[case] solve(solver='auto') picks the algebraic multigrid off generated squares", flush=True)
    grid, fes, pde, kap = _poisson(grmp, O, 1, 30, seed=4)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, solver="auto", rtol=1e-11)
    assert info.status == 0 and info.iterations < 60
    assert not grmp.amg_eligible(grmp.IncompressibleNavierStokesProblem(2, 1e-2),
                                 grmp.FEVector([grmp.FESpace(grmp.H1BR(2), grid), grmp.FESpace(grmp.L2P0(1), grid)]))


@pytest.mark.parametrize("fe", [1, 2])
def test_amg_equal_strengths(grmp, O, fe):
    """A constant coefficient on generated squares: every pair of a row has the
    same strength, so the matching ranks ties by its symmetric pair key (ranking
    them by index stalled the coarsening of P1 and P2 at 256 x 256)."""
    print(f"[case] amg with equal strengths, fe {fe}, generated 256x256", flush=True)
    grid, fes, pde, kap = _poisson(grmp, O, fe, 256, seed=1, sigma=0.0, uploaded=False)
    sol = grmp.FEVector([fes])
    A, b, fixed = grmp.assemble(pde, sol)
    m = grmp.AlgebraicMultigridSolver(A)
    nl, coarse, cx = m.info()
    xa, xj = grmp.FEVector([fes]), grmp.FEVector([fes])
    ia = m.solve(b, xa, 1e-12)
    ij = grmp.solve_system(A, b, xj, grmp.SolverConfig(rtol=1e-12, maxiterations=20000))
    u, v = xa.entries(), xj.entries()
    err = np.abs(u - v).max() / np.abs(v).max()
    print(f"  {nl} levels, coarsest {coarse} rows, complexity {cx:.3f}; amg {ia.iterations} iterations, "
          f"Jacobi CG {ij.iterations}; max rel diff {err:.2e}")
    assert ia.status == 0 and ij.status == 0
    assert nl >= 4 and cx < 1.6 and ia.iterations < 80 and err < 1e-9
