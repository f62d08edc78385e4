"""BackwardEuler time stepping on the device (TimeControlSolver and
advance!, src/solvers_timedependent.jl:116-594) against the oracle's steps
on the P1 heat equation u_t - div(kappa grad u) = 1 with homogeneous
Dirichlet data, and with best-approximated linear data and f = 0.

Bar:
* mass matrix: pattern and values bitwise equal to the oracle's;
* step right-hand side b = f + M u / dt: within 1e-13 of max|b| on free rows,
  penalty rows b_d = 1e60 u_d bitwise;
* every state within 1e-9 (max norm, relative to the oracle state's) of the
  oracle's sparse-LU steps, the device solve stopping at relative residual
  1e-13 (Krylov) or 1e-12 (multigrid);
* the system matrix is reassembled on the first step and on each change of
  the step size only;
* both step solvers: the Krylov solver, and the multigrid V-cycle CG on the
  rediscretized unit-square hierarchy (odd N included);
* inhomogeneous data: boundary values kept to 1e-13 at every step, and the
  stationary limit equal to the harmonic (linear) data to 1e-9;
* CrankNicolson (src/solvers_timedependent.jl:376-433): every state within
  1e-9 of the oracle's CrankNicolson steps, both step solvers, and
  advance_until_time! stopping at the final time."""
import numpy as np
import pytest

from meshes import shuffled_square

pytestmark = pytest.mark.gpu


def _heat(grmp, grid, kap=None, diffusion=1.0):
    fes = grmp.FESpace(grmp.H1P1(1), grid)
    coef = grmp.CellCoefficient(kap) if kap is not None else None
    pde = grmp.PoissonProblem(diffusion=diffusion, cell_coefficient=coef)
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
    pde.add_boundarydata(0, grmp.HomogeneousDirichletBoundary)
    return fes, pde


def _oracle(O, coords, cells, kap=None, diffusion=1.0):
    nn = coords.shape[0]
    cf, fn, fc = O.tri_faces(cells, nn)
    cd, nd = O.celldofs(O.FE_P1, 1, cells, nn, cf, fn.shape[0])
    fixed = O.boundary_dofs(O.FE_P1, 1, nn, fn, fc, fn.shape[0])
    base = dict(fe=O.FE_P1, coords=coords, cells=cells, celldofs=cd)
    lap = dict(base, kind=O.BLF_LAPLACE, quadorder=0, factor=diffusion)
    if kap is not None:
        lap.update(factor=1.0, coef=O.COEF_CELL, kappa=kap)
    A = O.assemble_matrix(nd, [lap])
    M = O.assemble_matrix(nd, [dict(base, kind=O.BLF_MASS, quadorder=2, factor=1.0)])
    f = np.zeros(nd)
    O.assemble_lf_const(f, O.FE_P1, 1, 1, [1.0], coords, cells, cd)
    return A, M, f, fixed


def test_step_rhs_and_mass(grmp, O):
    N, dt = 12, 0.02
    print(f"[case] timestep rhs N={N}", flush=True)
    grid = grmp.grid_unitsquare(N)
    fes, pde = _heat(grmp, grid, diffusion=0.8)
    X = grmp.FEVector([fes])
    tcs = grmp.TimeControlSolver(pde, X, grmp.BackwardEuler, timedependent_equations=[0])
    coords, cells = grid.download()
    A, M, f, fixed = _oracle(O, coords, cells, diffusion=0.8)
    colptr, rowval, nzval = tcs.AM.csc()
    assert np.array_equal(colptr, M.colptr) and np.array_equal(rowval, M.rowval)
    assert np.array_equal(nzval, M.nzval)
    assert np.array_equal(tcs.f.entries(), f)
    rng = np.random.default_rng(11)
    u = rng.standard_normal(fes.ndofs)  # nonzero on the boundary too: exercises the penalty rows
    X.set_entries(u)
    tcs.S = tcs._system(dt)
    b = tcs.rhs(dt).entries()
    bref = O.backward_euler_rhs(M, f, u, dt, fixed)
    free = np.setdiff1d(np.arange(fes.ndofs), fixed)
    err = np.max(np.abs(b[free] - bref[free])) / np.max(np.abs(bref[free]))
    print(f"  rhs max rel {err:.2e}")
    assert err <= 1e-13
    assert np.array_equal(b[fixed], bref[fixed])


@pytest.mark.parametrize("solver", ["krylov", "mg"])
@pytest.mark.parametrize("N", [2, 24, 37])
def test_steps_match_oracle(grmp, O, N, solver):
    print(f"[case] backward euler steps N={N} {solver}", flush=True)
    grid = grmp.grid_unitsquare(N)
    kap = grmp.DeviceVector(grid.ctx, grid.ncells)
    kap.lognormal(7, 0.0, 1.0)
    fes, pde = _heat(grmp, grid, kap=kap)
    X = grmp.FEVector([fes])
    rng = np.random.default_rng(N)
    X.set_entries(rng.standard_normal(fes.ndofs))
    tcs = grmp.TimeControlSolver(pde, X, grmp.BackwardEuler, timedependent_equations=[0],
                                 rtol=1e-13 if solver == "krylov" else 1e-12, solver=solver)
    assert tcs.solver == solver
    u0 = X.entries()
    coords, cells = grid.download()
    A, M, f, fixed = _oracle(O, coords, cells, kap=kap.download())
    assert np.all(u0[fixed] == 0.0)
    dts = [0.01, 0.01, 0.01, 0.03, 0.03, 0.01]
    ref = O.backward_euler(A, M, f, u0, dts, fixed)
    for k, dt in enumerate(dts):
        info = tcs.advance(dt)
        assert info.status == 0
        u = X.entries()
        err = np.max(np.abs(u - ref[k])) / np.max(np.abs(ref[k]))
        print(f"  step {k} dt={dt} iterations {info.iterations} max rel {err:.2e}")
        assert err <= 1e-9, (k, err)
    assert tcs.assemblies == 3
    assert tcs.cstep == len(dts) and abs(tcs.ctime - sum(dts)) < 1e-15


def test_uploaded_mesh_until_stationarity(grmp, O):
    """A shuffled, partly clockwise mesh: steps with dt = 0.5 until the change
    drops below 1e-9, then the state equals the stationary solve."""
    N, seed = 10, 5
    print(f"[case] stationarity on shuffled mesh N={N} seed={seed}", flush=True)
    coords, cells = shuffled_square(O, N, seed)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    fes, pde = _heat(grmp, grid, diffusion=1.3)
    X = grmp.FEVector([fes])
    tcs = grmp.TimeControlSolver(pde, X, grmp.BackwardEuler, timedependent_equations=[0], rtol=1e-13)
    # an uploaded mesh: the default step solver is the algebraic multigrid CG
    assert tcs.solver == "amg"
    nsteps = tcs.advance_until_stationarity(0.5, stationarity_threshold=1e-9, maxtimesteps=200)
    assert nsteps < 200 and tcs.last_change < 1e-9
    A, M, f, fixed = _oracle(O, coords, cells, diffusion=1.3)
    ref = O.backward_euler(A, M, f, np.zeros(fes.ndofs), [0.5] * nsteps, fixed)[-1]
    u = X.entries()
    err = np.max(np.abs(u - ref)) / np.max(np.abs(ref))
    us = O.direct_solve(O.assemble_matrix(fes.ndofs, [dict(kind=O.BLF_LAPLACE, fe=O.FE_P1, quadorder=0,
                                                           factor=1.3, coords=coords, cells=cells,
                                                           celldofs=fes.celldofs())], fixed=fixed),
                        O.penalized_rhs(f, fixed))
    errs = np.max(np.abs(u - us)) / np.max(np.abs(us))
    print(f"  {nsteps} steps, vs oracle steps {err:.2e}, vs stationary {errs:.2e}")
    assert err <= 1e-9
    assert errs <= 1e-8


@pytest.mark.parametrize("solver", ["krylov", "mg"])
def test_dirichlet_data_kept_and_stationary_limit(grmp, O, solver):
    """Best-approximated linear Dirichlet data g = 0.5 + 2x - y with f = 0:
    the boundary values hold at every step and the state tends to the
    harmonic P1 function, which is g itself."""
    N = 12
    print(f"[case] heat with Dirichlet data N={N} {solver}", flush=True)
    grid = grmp.grid_unitsquare(N)
    fes = grmp.FESpace(grmp.H1P1(1), grid)
    coef = [0.5, 2.0, -1.0]
    pde = grmp.PoissonProblem(1.0)
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([0.0])))
    pde.add_boundarydata(0, grmp.BestapproxDirichletBoundary, data=grmp.PolynomialData(coef))
    X = grmp.FEVector([fes])
    X.set_entries(np.random.default_rng(3).standard_normal(fes.ndofs))
    tcs = grmp.TimeControlSolver(pde, X, grmp.BackwardEuler, timedependent_equations=[0],
                                 rtol=1e-13 if solver == "krylov" else 1e-12, solver=solver)
    assert tcs.solver == solver
    coords, _ = grid.download()
    g = O.poly_eval(coef, coords[:, 0], coords[:, 1])
    bd = fes.boundary_dofs()
    assert np.max(np.abs(X.entries()[bd] - g[bd])) < 1e-14
    tcs.advance(0.01)
    assert np.max(np.abs(X.entries()[bd] - g[bd])) < 1e-13
    n = tcs.advance_until_stationarity(1.0, stationarity_threshold=1e-11, maxtimesteps=100)
    err = np.max(np.abs(X.entries() - g)) / np.max(np.abs(g))
    print(f"  {n} steps to stationarity, max rel err vs g {err:.2e}")
    assert n < 100 and err <= 1e-9


def test_argument_errors(grmp):
    print("[case] time control argument errors", flush=True)
    grid = grmp.grid_unitsquare(3)
    fes, pde = _heat(grmp, grid)
    X = grmp.FEVector([fes])
    with pytest.raises(grmp.GrmpError):
        grmp.TimeControlSolver(pde, X, grmp.CrankNicolson, timedependent_equations=[])
    with pytest.raises(grmp.GrmpError):
        grmp.TimeControlSolver(pde, X, grmp.BackwardEuler)
    tcs = grmp.TimeControlSolver(pde, X, grmp.BackwardEuler, timedependent_equations=[0])
    with pytest.raises(grmp.GrmpError):
        tcs.advance(0.0)
    with pytest.raises(grmp.GrmpError):  # multigrid needs a generated unit square
        up = grmp.ExtendableGrid.from_arrays(*grid.download())
        fu, pu = _heat(grmp, up)
        grmp.TimeControlSolver(pu, grmp.FEVector([fu]), grmp.BackwardEuler, timedependent_equations=[0],
                               solver="mg")
    other = grmp.FEVector([fes])
    with pytest.raises(grmp.GrmpError):  # b aliasing u
        tcs.S = tcs._system(0.1)
        X.ctx.check(grmp.lib().grmp_timestep_rhs(tcs.S.h, tcs.AM.h, tcs.f.vec.h, other.vec.h, 10.0, other.vec.h))


@pytest.mark.parametrize("solver", ["krylov", "mg", "amg"])
def test_crank_nicolson_matches_oracle(grmp, O, solver):
    N = 24
    print(f"[case] crank nicolson steps N={N} {solver}", flush=True)
    grid = grmp.grid_unitsquare(N)
    kap = grmp.DeviceVector(grid.ctx, grid.ncells)
    kap.lognormal(9, 0.0, 1.0)
    fes, pde = _heat(grmp, grid, kap=kap)
    X = grmp.FEVector([fes])
    X.set_entries(np.random.default_rng(2).standard_normal(fes.ndofs))
    tcs = grmp.TimeControlSolver(pde, X, grmp.CrankNicolson, timedependent_equations=[0],
                                 rtol=1e-13 if solver == "krylov" else 1e-12, solver=solver)
    assert tcs.solver == solver
    u0 = X.entries()
    coords, cells = grid.download()
    A, M, f, fixed = _oracle(O, coords, cells, kap=kap.download())
    dts = [0.01, 0.01, 0.03, 0.03, 0.01]
    ref = O.crank_nicolson(A, M, f, u0, dts, fixed)
    for k, dt in enumerate(dts):
        info = tcs.advance(dt)
        assert info.status == 0
        err = np.max(np.abs(X.entries() - ref[k])) / np.max(np.abs(ref[k]))
        print(f"  step {k} dt={dt} iterations {info.iterations} max rel {err:.2e}")
        assert err <= 1e-9, (k, err)
    assert tcs.assemblies == 3


def test_advance_until_time_crank_nicolson_uploaded(grmp, O):
    """advance_until_time! on a shuffled mesh: 0.05 steps to t = 0.4, the
    states equal the oracle's, the time is met, and the callback runs every
    step."""
    coords, cells = shuffled_square(O, 9, 6)
    print("[case] advance until time (CN, uploaded mesh)", flush=True)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    fes, pde = _heat(grmp, grid, diffusion=0.9)
    X = grmp.FEVector([fes])
    tcs = grmp.TimeControlSolver(pde, X, grmp.CrankNicolson, timedependent_equations=[0], rtol=1e-13)
    seen = []
    n = tcs.advance_until_time(0.05, 0.4, finaltime_tolerance=1e-12, do_after_each_timestep=lambda t: seen.append(t.cstep))
    assert n == 8 and seen == list(range(1, 9)) and abs(tcs.ctime - 0.4) < 1e-12
    A, M, f, fixed = _oracle(O, coords, cells, diffusion=0.9)
    ref = O.crank_nicolson(A, M, f, np.zeros(fes.ndofs), [0.05] * 8, fixed)[-1]
    err = np.max(np.abs(X.entries() - ref)) / np.max(np.abs(ref))
    print(f"  vs oracle {err:.2e}")
    assert err <= 1e-9


def test_auto_solver_inhomogeneous_data_tight_tolerance(grmp, O):
    """solver="auto" with inhomogeneous Dirichlet data picks the lifted
    multigrid steps; at rtol 1e-13 its states match the oracle's to 1e-9 and
    keep the boundary values to 1e-13."""
    N = 16
    print(f"[case] auto solver, inhomogeneous data, rtol 1e-13, N={N}", flush=True)
    grid = grmp.grid_unitsquare(N)
    fes = grmp.FESpace(grmp.H1P1(1), grid)
    coef = [0.5, 2.0, -1.0, 0.3]
    pde = grmp.PoissonProblem(1.0)
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
    pde.add_boundarydata(0, grmp.InterpolateDirichletBoundary, data=grmp.PolynomialData(coef))
    X = grmp.FEVector([fes])
    tcs = grmp.TimeControlSolver(pde, X, grmp.BackwardEuler, timedependent_equations=[0], rtol=1e-13)
    assert tcs.solver == "mg"
    coords, cells = grid.download()
    A, M, f, fixed = _oracle(O, coords, cells)
    g = O.poly_eval(coef, coords[:, 0], coords[:, 1])
    u0 = X.entries()
    assert np.max(np.abs(u0[fixed] - g[fixed])) < 1e-15
    dts = [0.02, 0.02, 0.05]
    ref = O.backward_euler(A, M, f, u0, dts, fixed)
    for k, dt in enumerate(dts):
        info = tcs.advance(dt)
        assert info.status == 0, info.status
        u = X.entries()
        err = np.max(np.abs(u - ref[k])) / np.max(np.abs(ref[k]))
        print(f"  step {k} iterations {info.iterations} relres {info.relres:.2e} max rel {err:.2e}")
        assert err <= 1e-9 and np.max(np.abs(u[fixed] - g[fixed])) < 1e-13
