"""NonlinearForm Newton assembly and the fixpoint loop on the device
(src/assemblypatterns/nonlinearform.jl:44-246, src/solvers.jl:812-1017)
against the oracle, on the nonlinear Poisson problem
-div((1 + u^2) grad u) = f of examples/Example205_NonlinearPoisson2D.jl.

Bar: Newton matrix within 1e-12 of the largest entry of its row and the
Newton right-hand side within 1e-12 of its largest entry, at a random
iterate, P1 and P2 on a shuffled mesh (the oracle sums in numpy's order, and
the reference's Jacobian comes from ForwardDiff and BLAS: not bitwise);
polynomial right-hand sides within 1e-13; the device fixpoint loop
(BiCGStab per Newton step) reproduces u = x y with P2 at every dof to 1e-10,
and on P1 its solution is within 1e-9 of the oracle's Newton solution."""
import numpy as np
import pytest

from meshes import shuffled_square

pytestmark = pytest.mark.gpu

C = (1.0, 0.0, 1.0)
F = [0.0] * 15
F[11], F[13] = -2.0, -2.0


def _setup(grmp, O, N, fe, seed=4):
    coords, cells = shuffled_square(O, N, seed)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    fes = grmp.FESpace(grmp.H1P1(1) if fe == 1 else grmp.H1P2(1), grid)
    cf, fn, fc = O.tri_faces(cells, coords.shape[0])
    pts = coords if fe == 1 else np.vstack([coords, 0.5 * (coords[fn[:, 0]] + coords[fn[:, 1]])])
    return grid, fes, coords, cells, pts


def _pde(grmp):
    pde = grmp.PDEDescription("nonlinear Poisson problem")
    pde.add_unknown("u")
    nlf = grmp.NonlinearForm(grmp.Gradient, [grmp.Identity, grmp.Gradient], [0, 0], grmp.NonlinearDiffusion(*C),
                             [2, 3], bonus_quadorder=2)
    pde.add_operator((0, 0), nlf)
    pde.add_boundarydata(0, grmp.InterpolateDirichletBoundary, data=grmp.PolynomialData([0, 0, 0, 0, 1.0]))
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.PolynomialData(F)))
    return pde, nlf


@pytest.mark.parametrize("fe", [1, 2])
def test_newton_system_matches_oracle(grmp, O, fe):
    print(f"[case] newton matrix / rhs fe={fe}", flush=True)
    grid, fes, coords, cells, pts = _setup(grmp, O, 7, fe)
    pde, nlf = _pde(grmp)
    X = grmp.FEVector([fes])
    u = np.random.default_rng(fe).standard_normal(fes.ndofs)
    X.set_entries(u)
    A = grmp.FEMatrix([fes])
    form, keep = nlf.form(fes, 0, 0, grid.ctx, current=X)
    A.add_form(form, keep)
    A.build_pattern()
    A.assemble()
    b = grmp.FEVector([fes])
    b.vec.fill(0.0)
    import ctypes
    lf = nlf.newton_lform(fes, X)
    grid.ctx.check(grmp.lib().grmp_assemble_linearform(grid.ctx.h, fes.h, ctypes.byref(lf), b.vec.h, 0))
    q = nlf.quadorder(fes)
    cd = fes.celldofs()
    Ao, bo = O.newton_assemble(fe, coords, cells, cd, fes.ndofs, u, C, q)
    Ad = A.to_scipy().tocsr()
    scale = np.maximum(abs(Ao).max(axis=1).toarray().ravel(), 1e-300)
    D = (Ad - Ao).tocoo()
    rel = np.abs(D.data) / scale[D.row] if D.nnz else np.zeros(1)
    print(f"  matrix max rel {rel.max():.2e}, rhs max rel "
          f"{np.max(np.abs(b.entries() - bo)) / np.max(np.abs(bo)):.2e}")
    assert rel.max() <= 1e-12
    assert np.max(np.abs(b.entries() - bo)) <= 1e-12 * np.max(np.abs(bo))
    # the structural pattern: every local pair of every cell
    assert A.info()[1] == (Ao != 0).nnz or A.info()[1] >= (Ao != 0).nnz


@pytest.mark.parametrize("fe", [1, 2])
def test_polynomial_linear_form(grmp, O, fe):
    print(f"[case] polynomial linear form fe={fe}", flush=True)
    grid, fes, coords, cells, pts = _setup(grmp, O, 8, fe, seed=2)
    coef = [0.5, -1.0, 2.0, 0.3, -0.7, 1.1, 0.2, -0.4, 0.9, 0.6, -1.2, 0.8, 0.25, -0.5, 0.7]
    b = grmp.FEVector([fes])
    grmp.assemble_linearform(b, 0, grmp.LinearForm(grmp.Identity, grmp.PolynomialData(coef)))
    ref = O.assemble_lf_poly(fe, coef, fe + 4, coords, cells, fes.celldofs(), fes.ndofs)
    assert np.max(np.abs(b.entries() - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_newton_p2_reproduces_xy(grmp, O):
    print("[case] newton P2 u = xy", flush=True)
    grid, fes, coords, cells, pts = _setup(grmp, O, 6, 2)
    pde, nlf = _pde(grmp)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, target_residual=1e-11, nonlinear_maxiterations=12, rtol=1e-13)
    u_ex = pts[:, 0] * pts[:, 1]
    err = np.max(np.abs(sol.entries() - u_ex))
    print(f"  {info.iterations} Newton steps, residuals {['%.1e' % r for r in info.residuals]}, "
          f"BiCGStab {info.linear_iterations}, max err {err:.2e}")
    assert info.status == 0 and err < 1e-10


@pytest.mark.parametrize("pc", ["amg", "jacobi"])
def test_newton_p1_matches_oracle(grmp, O, pc):
    print(f"[case] newton P1 vs oracle, {pc}", flush=True)
    grid, fes, coords, cells, pts = _setup(grmp, O, 12, 1, seed=9)
    pde, nlf = _pde(grmp)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, target_residual=1e-11, nonlinear_maxiterations=12, rtol=1e-13,
                            newton_preconditioner=pc)
    assert info.preconditioner == pc
    cd = fes.celldofs()
    fixed = fes.boundary_dofs()
    xo, res = O.newton_solve(1, coords, cells, cd, fes.ndofs, fixed, pts[:, 0] * pts[:, 1], C, F, quadorder=3,
                             lf_quadorder=5, maxit=12, target_residual=1e-11)
    err = np.max(np.abs(sol.entries() - xo)) / np.max(np.abs(xo))
    print(f"  device {info.iterations} steps {['%.1e' % r for r in info.residuals]}; oracle {len(res)} steps; "
          f"max rel diff {err:.2e}")
    assert info.status == 0 and err <= 1e-9
    assert abs(info.iterations - len(res)) <= 1


def test_bicgstab_initial_guess(grmp, O):
    """a solution passed back as the initial guess is accepted at once; CG
    and MINRES refuse an initial guess"""
    print("[case] bicgstab initial guess", flush=True)
    import ctypes
    grid, fes, coords, cells, pts = _setup(grmp, O, 9, 2, seed=5)
    pde, nlf = _pde(grmp)
    X = grmp.FEVector([fes])
    X.set_entries(np.random.default_rng(3).standard_normal(fes.ndofs))
    A = grmp.FEMatrix([fes])
    form, keep = nlf.form(fes, 0, 0, grid.ctx, current=X)
    A.add_form(form, keep)
    A.set_fixed(fes.boundary_dofs())
    A.build_pattern()
    A.assemble(grmp._abi.VALUES_FP64)
    b = grmp.FEVector([fes])
    b.set_entries(np.random.default_rng(4).standard_normal(fes.ndofs))
    cfg = grmp.SolverConfig(rtol=1e-12)
    x = grmp.FEVector([fes])
    i1 = grmp.solve_system(A, b, x, cfg, grmp._abi.SOLVE_BICGSTAB)
    x1 = x.entries()
    i2 = grmp.solve_system(A, b, x, cfg, grmp._abi.SOLVE_BICGSTAB, initial_guess=True)
    print(f"  cold {i1.iterations} iterations ({i1.outer - 1} restarts), warm {i2.iterations}")
    assert i1.status == 0 and i2.status == 0
    assert i2.iterations == 0 and np.array_equal(x.entries(), x1)
    # against a sparse LU of the free block
    K = A.to_scipy().tocsr()
    fixed = fes.boundary_dofs()
    free = np.setdiff1d(np.arange(fes.ndofs), fixed)
    import scipy.sparse.linalg as spla
    bb = b.entries()
    xr = np.zeros(fes.ndofs)
    xr[fixed] = bb[fixed] / K.diagonal()[fixed]
    xr[free] = spla.spsolve(K[free][:, free].tocsc(), bb[free] - K[free][:, fixed] @ xr[fixed])
    assert np.max(np.abs(x1 - xr)) <= 1e-9 * np.max(np.abs(xr))
    o = grmp._abi.SolveOpts()
    grmp.lib().grmp_solve_opts_default(ctypes.byref(o))
    o.initial_guess = 1
    for m in (grmp._abi.SOLVE_CG, grmp._abi.SOLVE_MINRES):
        o.method = m
        inf = grmp._abi.SolveInfo()
        assert grmp.lib().grmp_solve(A.h, b.vec.h, x.vec.h, ctypes.byref(o), ctypes.byref(inf)) == \
            grmp._abi.ERR_UNSUPPORTED


@pytest.mark.parametrize("pc", ["jacobi", "amg"])
def test_newton_p2_large(grmp, pc):
    """the fixpoint loop at 148,225 P2 dofs: with Jacobi, BiCGStab's
    recurrence meets near-breakdowns and restarts; with the algebraic
    multigrid V-cycle it needs far fewer iterations.  ||u_h - x y||_L2 below
    1e-9 either way"""
    print(f"[case] newton P2 192x192, {pc}", flush=True)
    N = 192
    grid = grmp.grid_unitsquare(N)
    fes = grmp.FESpace(grmp.H1P2(1), grid)
    pde, nlf = _pde(grmp)
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, target_residual=1e-10, nonlinear_maxiterations=12, rtol=1e-12,
                            newton_preconditioner=pc)
    assert info.preconditioner == pc
    if pc == "amg":
        assert max(info.linear_iterations) < 200, info.linear_iterations
    err = float(grmp.evaluate(grmp.L2ErrorIntegrator(grmp.PolynomialData([0, 0, 0, 0, 1.0]), quadorder=2),
                              sol[0])) ** 0.5
    print(f"  {info.iterations} Newton steps {['%.1e' % r for r in info.residuals]}, BiCGStab "
          f"{info.linear_iterations}, restarts {info.linear_restarts}, L2 error {err:.2e}")
    assert info.status == 0 and err < 1e-9
