"""Stokes on the device (H1BR{2} x L2P0, LaplaceOperator(nu) +
LagrangeMultiplier(Divergence), reconstructed BR -> RT0 right-hand side,
MINRES) against the oracle.

Stated tolerances: sparsity pattern identical; matrix values bitwise equal
(the summation order follows the reference's); right-hand side within 1e-12
relative to max|b| (the device's sin/cos differ from the host libm in the
last bits); solutions within 1e-8 relative of the FP64 sparse-LU solution
after the FixedIntegralMean shift; a gradient force gives |u_h| below
1e-12 max|p_h| when the device-assembled system is solved exactly (pressure
robustness, examples/Example222_PressureRobustness2D.jl), and below 1e-6
max|p_h| after MINRES to 1e-13."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NU = 1e-2
MODES = np.array([[0.3, 1, 0, 0.1], [-0.2, 1, 2, 1.3], [0.25, 0, 1, 0.7],   # phi
                  [0.5, 1, 1, 0.4], [0.1, 2, -1, 2.0], [-0.3, 1, 3, 0.9]])  # psi
POLY = [0.4, -0.3, 0.2]


def _stokes(grmp, N, modes=MODES, poly=POLY, values=None, precision="fp64"):
    grid = grmp.grid_unitsquare(N)
    V = grmp.FESpace(grmp.H1BR(2), grid)
    Q = grmp.FESpace(grmp.L2P0(1), grid)
    pde = grmp.IncompressibleNavierStokesProblem(2, viscosity=NU)
    K = len(modes) // 2
    f = grmp.GradientCurlData(modes[:K], modes[K:], poly, bonus_quadorder=2)
    pde.add_rhsdata(0, grmp.LinearForm(grmp.ReconstructionIdentity(grmp.HDIVRT0), f))
    sol = grmp.FEVector([V, Q])
    return grid, V, Q, pde, sol


def test_stokes_dofmaps(grmp, O):
    print("[case] stokes dofmaps N=6", flush=True)
    grid, V, Q, pde, sol = _stokes(grmp, 6)
    S = O.StokesSpace(6)
    assert V.ndofs == S.nU and Q.ndofs == S.nP
    assert np.array_equal(V.celldofs(), S.celldofs)
    assert np.array_equal(Q.celldofs()[:, 0], np.arange(S.nc))
    fixed = np.concatenate([V.boundary_dofs(), [S.nU]])
    assert np.array_equal(np.unique(fixed), S.fixed)


@pytest.mark.parametrize("N", [4, 13, 32])
def test_stokes_matrix_parity(grmp, O, N):
    print(f"[case] stokes matrix N={N}", flush=True)
    grid, V, Q, pde, sol = _stokes(grmp, N)
    A, b, fixed = grmp.assemble(pde, sol)
    S = O.StokesSpace(N)
    R = S.matrix(NU)
    colptr, rowval, nzval = A.csc()
    assert np.array_equal(colptr, R.colptr)
    assert np.array_equal(rowval, R.rowval)
    Ar = R.to_scipy().tocsr()
    scale = np.maximum(np.abs(Ar).max(axis=1).toarray().ravel(), 1e-300)
    diff = np.abs((A.to_scipy().tocsr() - Ar)).tocoo()
    rel = diff.data / scale[diff.row] if diff.nnz else np.zeros(1)
    print(f"  bitwise equal: {np.count_nonzero(nzval == R.nzval)} of {nzval.size}, max rel {rel.max():.2e}")
    # entries that differ: where they lie (velocity node / bubble / pressure rows and columns)
    nU = V.ndofs
    nn = (N + 1) * (N + 1)

    def cls(i):
        return "p" if i >= nU else ("node" if i < 2 * nn else "bubble")
    d = np.nonzero(nzval != R.nzval)[0]
    if d.size:
        rows = rowval[d] - 1
        cols = np.searchsorted(colptr, d + 1, side="right") - 1
        per = np.abs(nzval[d] - R.nzval[d]) / np.maximum(np.abs(R.nzval[d]), 1e-300)
        kinds = {}
        for r_, c_ in zip(rows, cols):
            k = cls(r_) + "/" + cls(c_)
            kinds[k] = kinds.get(k, 0) + 1
        k = int(np.argmax(per))
        print(f"  differing entries by (row, col) kind: {kinds}; max per-entry rel {per.max():.2e} at "
              f"({rows[k]}, {cols[k]}) value {R.nzval[d[k]]:.3e}, row max {scale[rows[k]]:.3e}")
    assert rel.max() <= 1e-12
    assert np.array_equal(nzval, R.nzval)


@pytest.mark.parametrize("N", [4, 16])
def test_stokes_rhs_parity(grmp, O, N):
    print(f"[case] stokes rhs N={N}", flush=True)
    grid, V, Q, pde, sol = _stokes(grmp, N)
    b = grmp.FEVector([V, Q])
    grmp.assemble_linearform(b, 0, pde.RHS[0][0])
    got = b.entries()
    S = O.StokesSpace(N)
    ref = S.rhs(2 + 2, MODES, POLY)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"  max rel {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_stokes_minres_vs_lu(grmp, O, precision):
    N = 24
    print(f"[case] stokes minres {precision} N={N}", flush=True)
    grid, V, Q, pde, sol = _stokes(grmp, N)
    info, A, b = grmp.solve(sol, pde, precision=precision, rtol=1e-11)
    assert info.status == 0, info.status
    x = sol.entries()
    S = O.StokesSpace(N)
    bref = O.penalized_rhs(S.rhs(4, MODES, POLY), S.fixed)
    xref = S.pressure_mean_shift(O.direct_solve(S.matrix(NU), bref))
    err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    print(f"  iterations {info.iterations} outer {info.outer} rel err {err:.2e}")
    assert err <= 1e-8
    # FixedIntegralMean: the pressure has zero mean
    p = x[S.nU:]
    assert abs(np.sum(p) / p.size) <= 1e-12 * max(1.0, np.max(np.abs(p)))


def test_stokes_pressure_robust(grmp, O):
    """A polynomial gradient force is balanced by the pressure alone: the device-assembled
    system solved exactly gives u_h = 0 to rounding, and MINRES leaves only
    its iteration error in u_h."""
    import scipy.sparse.linalg as spla
    N = 16
    print(f"[case] stokes pressure robustness N={N}", flush=True)
    # polynomial phi: (grad phi, RT0) is integrated exactly by the order-4 rule,
    # which the identity (grad phi, Pi v) = -(phi, div Pi v) = 0 needs
    grid, V, Q, pde, sol = _stokes(grmp, N, modes=np.zeros((0, 4)), poly=[1.0, 0.5, -2.0])
    A, b, fixed = grmp.assemble(pde, sol)
    x = spla.spsolve(A.to_scipy().tocsc(), b.entries())
    u, p = x[:V.ndofs], x[V.ndofs:]
    print(f"  LU on the device system: max|u| {np.max(np.abs(u)):.2e} max|p| {np.max(np.abs(p)):.2e}")
    assert np.max(np.abs(p)) > 1e-3
    assert np.max(np.abs(u)) <= 1e-12 * np.max(np.abs(p))
    info, A, b = grmp.solve(sol, pde, rtol=1e-13)
    assert info.status == 0
    x = sol.entries()
    u, p2 = x[:V.ndofs], x[V.ndofs:]
    print(f"  MINRES: max|u| {np.max(np.abs(u)):.2e}, iterations {info.iterations}")
    assert np.max(np.abs(u)) <= 1e-6 * np.max(np.abs(p2))


def test_stokes_csc_plugin_picks_minres(grmp, O):
    """The LinearSolve plug-in entry point recognises the saddle point."""
    import ctypes
    N = 8
    print(f"[case] stokes csc plug-in N={N}", flush=True)
    S = O.StokesSpace(N)
    R = S.matrix(NU)
    bref = O.penalized_rhs(S.rhs(4, MODES, POLY), S.fixed)
    xref = O.direct_solve(R, bref)
    ctx = grmp.default_context()
    x = np.zeros(S.n)
    cp = np.ascontiguousarray(R.colptr + 1, dtype=np.int64)
    rv = np.ascontiguousarray(R.rowval + 1, dtype=np.int64)
    nz = np.ascontiguousarray(R.nzval)
    fx = np.ascontiguousarray(S.fixed + 1, dtype=np.int64)
    lib = grmp.lib()
    rc = lib.grmp_solve_csc(ctx.h, S.n, grmp._abi.i64ptr(cp), grmp._abi.i64ptr(rv), grmp._abi.dptr(nz),
                            grmp._abi.dptr(bref), grmp._abi.dptr(x), fx.size, grmp._abi.i64ptr(fx),
                            grmp._abi.PRECISION_FP64, ctypes.c_double(1e-11))
    ctx.check(rc)
    err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    assert err <= 1e-8, err


@pytest.mark.parametrize("seed", [None, 8])
def test_stokes_exactness_br_p0(grmp, O, seed):
    """The reference's Stokes test for H1BR{2} x L2P0, orders (1, 0)
    (test/runtests.jl:521-540,581,606-640): u = (y + 1, x - 1), p = 0,
    f = 0, viscosity 1, BestapproxDirichletBoundary with u on all regions,
    FixedIntegralMean; L2 errors of velocity (quadorder 1) and pressure
    (quadorder 0) below the reference's tolerance 6e-12 after MINRES."""
    from meshes import shuffled_square
    print(f"[case] stokes exactness seed={seed}", flush=True)
    if seed is None:
        grid = grmp.grid_unitsquare(8)
    else:
        coords, cells = shuffled_square(O, 8, seed)
        grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    V = grmp.FESpace(grmp.H1BR(2), grid)
    Q = grmp.FESpace(grmp.L2P0(1), grid)
    pde = grmp.IncompressibleNavierStokesProblem(2, viscosity=1.0)
    lin = grmp.PolynomialData(np.array([[1.0, 0.0, 1.0], [-1.0, 1.0, 0.0]]), bonus_quadorder=1)
    pde.add_boundarydata(0, grmp.BestapproxDirichletBoundary, data=lin)
    sol = grmp.FEVector([V, Q])   # f = 0: no right-hand side operator
    info, A, b = grmp.solve(sol, pde, precision="fp64", rtol=1e-14)
    assert info.status == 0, info.status
    eV = np.sqrt(grmp.evaluate(grmp.L2ErrorIntegrator(lin, quadorder=1), sol[0]))
    eP = np.sqrt(grmp.evaluate(grmp.L2ErrorIntegrator(grmp.PolynomialData([0.0]), quadorder=0), sol[1]))
    print(f"  MINRES {info.iterations} it, errorV {eV:.2e}, errorP {eP:.2e}")
    assert max(eV, eP) < 6e-12


@pytest.mark.parametrize("N", [24, 65])
def test_stokes_minres_mg_vs_lu(grmp, O, N):
    """MINRES with the multigrid velocity block (grmp_solve_minres_mg): the
    same solution as sparse LU to 1e-8, in far fewer iterations than the
    diagonal preconditioner, and an iteration count that does not grow
    with N."""
    print(f"[case] stokes minres multigrid N={N}", flush=True)
    grid, V, Q, pde, sol = _stokes(grmp, N)
    assert grmp.stokes_mg_eligible(pde, sol)
    info, A, b = grmp.solve(sol, pde, solver="mg", rtol=1e-11)
    assert info.status == 0, info.status
    x = sol.entries()
    S = O.StokesSpace(N)
    bref = O.penalized_rhs(S.rhs(4, MODES, POLY), S.fixed)
    xref = S.pressure_mean_shift(O.direct_solve(S.matrix(NU), bref))
    err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    sol2 = grmp.FEVector([V, Q])
    info2, _, _ = grmp.solve(sol2, pde, solver="krylov", rtol=1e-11)
    print(f"  multigrid block: {info.iterations} iterations, rel err {err:.2e}; diagonal: {info2.iterations}")
    assert err <= 1e-8
    assert info.iterations < 400 and info.iterations < info2.iterations
    if N >= 65:
        assert info.iterations * 3 < info2.iterations
    # the two velocity components' V-cycles run at the same time on two
    # copies of the hierarchy; they reproduce one hierarchy applied to the
    # components in turn, which took these iterations, and a second solve
    # gives the same bits
    assert info.iterations == {24: 187, 65: 214}[N]
    sol3 = grmp.FEVector([V, Q])
    grmp.solve(sol3, pde, solver="mg", rtol=1e-11)
    assert np.array_equal(sol3.entries(), x)
    p = x[S.nU:]
    assert abs(np.sum(p) / p.size) <= 1e-12 * max(1.0, np.max(np.abs(p)))


def test_stokes_minres_mg_errors(grmp):
    print("[case] stokes minres multigrid errors", flush=True)
    import ctypes
    grid, V, Q, pde, sol = _stokes(grmp, 8)
    A, b, fixed = grmp.assemble(pde, sol)
    mg = grmp.stokes_velocity_mg(pde, V)
    o = grmp._abi.SolveOpts()
    grmp.lib().grmp_solve_opts_default(ctypes.byref(o))
    inf = grmp._abi.SolveInfo()
    L = grmp.lib()
    # more velocity rows than the system has
    assert L.grmp_solve_minres_mg(A.h, b.vec.h, sol.vec.h, ctypes.byref(o), mg.h, 40, ctypes.byref(inf)) == \
        grmp._abi.ERR_ARG
    o.precision = grmp._abi.PRECISION_FP32_IR
    assert L.grmp_solve_minres_mg(A.h, b.vec.h, sol.vec.h, ctypes.byref(o), mg.h, 2, ctypes.byref(inf)) == \
        grmp._abi.ERR_UNSUPPORTED
    assert L.grmp_solve_minres_mg(A.h, b.vec.h, sol.vec.h, ctypes.byref(o), None, 2, ctypes.byref(inf)) == \
        grmp._abi.ERR_ARG
