"""Device solves against the oracle's FP64 sparse LU (the direct-solve
stand-in for UMFPACK) on the same penalized systems.

Bar: relative error of the solution vector <= 1e-8 (stated tolerance of
the path; the solves stop at relative residual 1e-12)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _poisson(grmp, O, N, fe, seed):
    grid = grmp.grid_unitsquare(N)
    ft = grmp.H1P1(1) if fe == 1 else grmp.H1P2(1)
    fes = grmp.FESpace(ft, grid)
    kap = grmp.DeviceVector(grid.ctx, grid.ncells)
    kap.lognormal(seed, 0.0, 1.0)
    pde = grmp.PoissonProblem(cell_coefficient=grmp.CellCoefficient(kap))
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
    pde.add_boundarydata(0, grmp.HomogeneousDirichletBoundary)
    S = O.Space(N, fe)
    ref = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_LAPLACE, fe=fe, quadorder=2 * fe - 2, factor=1.0,
                                          coef=O.COEF_CELL, kappa=kap.download(), coords=S.coords, cells=S.cells,
                                          celldofs=S.celldofs)], fixed=S.fixed)
    b = np.zeros(S.ndofs)
    O.assemble_lf_const(b, fe, 1, fe, [1.0], S.coords, S.cells, S.celldofs)
    b = O.penalized_rhs(b, S.fixed)
    xref = O.direct_solve(ref, b)
    return grid, fes, pde, xref


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
@pytest.mark.parametrize("N,fe", [(24, 1), (24, 2), (2, 2)])
def test_poisson_solve(grmp, O, precision, N, fe):
    print(f"[case] poisson solve N={N} fe={fe} {precision}", flush=True)
    grid, fes, pde, xref = _poisson(grmp, O, N, fe, 17)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, precision=precision, rtol=1e-12)
    x = sol.entries()
    err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    assert info.status == 0
    assert info.relres <= 1e-12
    assert err <= 1e-8, err


def test_elasticity_solve(grmp, O):
    N = 20
    grid = grmp.grid_unitsquare(N)
    fes = grmp.FESpace(grmp.H1P1(2), grid)
    kap = grmp.DeviceVector(grid.ctx, grid.ncells)
    kap.lognormal(5, 0.0, 1.0)
    pde = grmp.LinearElasticityProblem(1.0, 0.3, grmp.CellCoefficient(kap))
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([0.0, -1.0])))
    pde.add_boundarydata(0, grmp.HomogeneousDirichletBoundary)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, precision="fp32_ir", rtol=1e-12)
    S = O.Space(N, O.FE_P1, ncomp=2)
    mu, lam = 1.0 / 2.6, 0.3 / (1.3 * 0.4)
    ref = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_HOOKE2D, fe=O.FE_P1, ncomp=2, quadorder=0, factor=1.0,
                                          coef=O.COEF_CELL, kappa=kap.download(), params=np.array([lam, mu]),
                                          coords=S.coords, cells=S.cells, celldofs=S.celldofs)], fixed=S.fixed)
    bb = np.zeros(S.ndofs)
    O.assemble_lf_const(bb, O.FE_P1, 2, 1, [0.0, -1.0], S.coords, S.cells, S.celldofs)
    xref = O.direct_solve(ref, O.penalized_rhs(bb, S.fixed))
    err = np.linalg.norm(sol.entries() - xref) / np.linalg.norm(xref)
    assert info.status == 0 and err <= 1e-8, err


@pytest.mark.parametrize("N,fe", [(24, 1), (37, 1), (24, 2), (3, 2)])
def test_poisson_solve_multigrid(grmp, O, N, fe):
    """solve(..., solver="mg"): the multigrid V-cycle CG on the rediscretized
    unit-square hierarchy, against sparse LU."""
    print(f"[case] multigrid solve N={N} fe={fe}", flush=True)
    grid, fes, pde, xref = _poisson(grmp, O, N, fe, 23)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, solver="mg", rtol=1e-12)
    err = np.linalg.norm(sol.entries() - xref) / np.linalg.norm(xref)
    print(f"  {info.iterations} iterations, rel err vs LU {err:.2e}")
    assert info.status == 0 and info.iterations <= 60
    assert err <= 1e-8, err


def test_elasticity_solve_multigrid(grmp, O):
    N = 20
    print(f"[case] multigrid elasticity N={N}", flush=True)
    grid = grmp.grid_unitsquare(N)
    fes = grmp.FESpace(grmp.H1P1(2), grid)
    kap = grmp.DeviceVector(grid.ctx, grid.ncells)
    kap.lognormal(5, 0.0, 1.0)
    pde = grmp.LinearElasticityProblem(1.0, 0.3, grmp.CellCoefficient(kap))
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([0.0, -1.0])))
    pde.add_boundarydata(0, grmp.HomogeneousDirichletBoundary)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, solver="mg", rtol=1e-12)
    S = O.Space(N, O.FE_P1, ncomp=2)
    mu, lam = 1.0 / 2.6, 0.3 / (1.3 * 0.4)
    ref = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_HOOKE2D, fe=O.FE_P1, ncomp=2, quadorder=0, factor=1.0,
                                          coef=O.COEF_CELL, kappa=kap.download(), params=np.array([lam, mu]),
                                          coords=S.coords, cells=S.cells, celldofs=S.celldofs)], fixed=S.fixed)
    bb = np.zeros(S.ndofs)
    O.assemble_lf_const(bb, O.FE_P1, 2, 1, [0.0, -1.0], S.coords, S.cells, S.celldofs)
    xref = O.direct_solve(ref, O.penalized_rhs(bb, S.fixed))
    err = np.linalg.norm(sol.entries() - xref) / np.linalg.norm(xref)
    print(f"  {info.iterations} iterations, rel err vs LU {err:.2e}")
    assert info.status == 0 and err <= 1e-8, err


def test_multigrid_handle_reuse(grmp, O):
    """One assembled hierarchy, several right-hand sides: each solve equals,
    bit for bit, a fresh one-shot solve of the same right-hand side."""
    print("[case] multigrid handle reuse", flush=True)
    grid, fes, pde, xref = _poisson(grmp, O, 30, 1, 11)
    sol = grmp.FEVector([fes])
    A, b, fixed = grmp.assemble(pde, sol)
    levels, kappa0 = grmp.mg_hierarchy(pde, fes, A)
    M = grmp.MultigridSolver(levels, kappa0)
    b2 = grmp.FEVector([fes])
    v = np.random.default_rng(1).standard_normal(fes.ndofs)
    v[fixed] = 0.0
    b2.set_entries(v)
    outs = []
    for rhs in (b, b2, b):
        x = grmp.FEVector([fes])
        info = M.solve(rhs, x, 1e-12)
        assert info.status == 0
        outs.append(x.entries())
    x1 = grmp.FEVector([fes])
    grmp.mg_solve(levels, kappa0, b, x1, 1e-12)
    assert np.array_equal(outs[0], outs[2])
    assert np.array_equal(outs[0], x1.entries())
    err = np.linalg.norm(outs[0] - xref) / np.linalg.norm(xref)
    assert err <= 1e-8 and not np.array_equal(outs[0], outs[1])


def test_multigrid_not_applicable(grmp, O):
    """solver="mg" on an uploaded mesh is refused; "auto" falls back."""
    print("[case] multigrid not applicable", flush=True)
    grid, fes, pde, xref = _poisson(grmp, O, 6, 1, 3)
    up = grmp.ExtendableGrid.from_arrays(*grid.download())
    fu = grmp.FESpace(grmp.H1P1(1), up)
    pu = grmp.PoissonProblem(1.0)
    pu.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
    pu.add_boundarydata(0, grmp.HomogeneousDirichletBoundary)
    assert not grmp.mg_eligible(pu, grmp.FEVector([fu]))
    with pytest.raises(grmp.GrmpError):
        grmp.solve(grmp.FEVector([fu]), pu, solver="mg")
    info, A, b = grmp.solve(grmp.FEVector([fu]), pu, solver="auto", rtol=1e-12)
    assert info.status == 0


def test_nonconvergence_reported(grmp, O):
    """An unreachable tolerance with few iterations returns GRMP_ERR_NOCONV
    (the reference warns, src/solvers.jl:1398-1400)."""
    grid, fes, pde, xref = _poisson(grmp, O, 16, 2, 1)
    sol = grmp.FEVector([fes])
    with pytest.raises(grmp.GrmpError) as ei:
        grmp.solve(sol, pde, precision="fp64", rtol=1e-30, maxiterations=5)
    assert ei.value.code == grmp._abi.ERR_NOCONV
    # the device stops after exactly maxit iterations (its scalars and the
    # stop test live on the device, the host reads them every 8 iterations)
    import ctypes
    A, b, fixed = grmp.assemble(pde, sol)
    for maxit in (5, 8, 13):
        o = grmp._abi.SolveOpts()
        grmp.lib().grmp_solve_opts_default(ctypes.byref(o))
        o.method, o.precision = grmp._abi.SOLVE_CG, grmp._abi.PRECISION_FP64
        o.rtol, o.maxit, o.check_every = 1e-30, maxit, 1
        info = grmp._abi.SolveInfo()
        rc = grmp.lib().grmp_solve(A.h, b.vec.h, sol.vec.h, ctypes.byref(o), ctypes.byref(info))
        assert rc == grmp._abi.ERR_NOCONV and info.iterations == maxit, (maxit, rc, info.iterations)


def test_plain_cg_iterations_pinned(grmp, O):
    """The device plain FP64 CG takes the iterations of the oracle's host
    Jacobi-PCG on the same penalized system (same recurrences; the dot
    products are summed in other orders, so the count may differ by one) and
    stops within the tolerance."""
    grid, fes, pde, xref = _poisson(grmp, O, 32, 1, 3)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, precision="fp64", rtol=1e-10)
    S = O.Space(32, O.FE_P1)
    colptr, rowval, nzval = A.csc()
    Ao = O.CSC(S.ndofs, colptr.astype(np.int64), rowval.astype(np.int32), nzval)
    x, it, rr = O.pcg(Ao, b.entries(), S.fixed, rtol=1e-10)
    assert info.status == 0 and info.relres <= 1e-10 and rr <= 1e-10
    assert abs(info.iterations - it) <= 1, (info.iterations, it)
    assert np.linalg.norm(sol.entries() - x) / np.linalg.norm(x) < 1e-8


@pytest.mark.parametrize("precision", [1, 2])
def test_solve_csc_plugin(grmp, O, precision):
    """The LinearSolve plug-in entry on a CSC assembled by the oracle."""
    import ctypes
    N = 20
    S = O.Space(N, O.FE_P2)
    rng = np.random.default_rng(4)
    kap = np.exp(rng.standard_normal(S.cells.shape[0]))
    A = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_LAPLACE, fe=O.FE_P2, quadorder=2, factor=1.0, coef=O.COEF_CELL,
                                        kappa=kap, coords=S.coords, cells=S.cells, celldofs=S.celldofs)],
                          fixed=S.fixed)
    b = np.zeros(S.ndofs)
    O.assemble_lf_const(b, O.FE_P2, 1, 2, [1.0], S.coords, S.cells, S.celldofs)
    b = O.penalized_rhs(b, S.fixed)
    xref = O.direct_solve(A, b)
    ab = grmp._abi
    colptr = np.ascontiguousarray(A.colptr + 1, dtype=np.int64)
    rowval = np.ascontiguousarray(A.rowval + 1, dtype=np.int64)
    fixed = np.ascontiguousarray(S.fixed + 1, dtype=np.int64)
    x = np.zeros(S.ndofs)
    ctx = grmp.default_context()
    ctx.check(grmp.lib().grmp_solve_csc(ctx.h, S.ndofs, ab.i64ptr(colptr), ab.i64ptr(rowval), ab.dptr(A.nzval),
                                        ab.dptr(b), ab.dptr(x), fixed.size, ab.i64ptr(fixed), precision, 1e-12))
    err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    assert err <= 1e-8, err
