"""HIP assembly path against the oracle on the same inputs.

Bar: meshes, dofmaps and sparsity patterns bit-exact; FP64 values within
1e-12 relative (they are in fact bit-identical: same operation order, same
per-entry summation order); load vectors within 1e-12."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _values_close(got, ref, rtol=RTOL):
    den = np.maximum(np.abs(ref), 1e-300)
    return float(np.max(np.abs(got - ref) / den)) if len(ref) else 0.0


def _check_csc(A, ref):
    colptr, rowval, nzval = A.csc()
    assert np.array_equal(colptr, ref.colptr), "colptr differs"
    assert np.array_equal(rowval, ref.rowval), "rowval differs"
    err = _values_close(nzval, ref.nzval)
    assert err <= RTOL, f"max rel value error {err}"
    return float(np.mean(nzval == ref.nzval))


@pytest.mark.parametrize("N", [1, 2, 5, 16, 64])
def test_unitsquare_mesh_and_dofmaps(grmp, O, N):
    print(f"[case] unitsquare N={N}", flush=True)
    grid = grmp.grid_unitsquare(N)
    c, cn = grid.download()
    cref, cnref = O.unitsquare(N)
    assert np.array_equal(c, cref) and np.array_equal(cn, cnref)
    S = O.Space(N, O.FE_P2)
    fes = grmp.FESpace(grmp.H1P2(1), grid)
    assert fes.ndofs == S.ndofs
    assert np.array_equal(fes.celldofs(), S.celldofs)
    assert np.array_equal(fes.boundary_dofs(), S.fixed)
    S1 = O.Space(N, O.FE_P1, ncomp=2)
    fes1 = grmp.FESpace(grmp.H1P1(2), grid)
    assert np.array_equal(fes1.celldofs(), S1.celldofs)
    assert np.array_equal(fes1.boundary_dofs(), S1.fixed)
    assert np.allclose(grid.cellvolumes(), 0.5 / N / N, rtol=1e-15)


def _laplace_case(grmp, O, N, fe, mode, seed=7):
    print(f"[case] laplace N={N} fe={fe} mode={mode}", flush=True)
    grid = grmp.grid_unitsquare(N)
    ft = grmp.H1P1(1) if fe == O.FE_P1 else grmp.H1P2(1)
    fes = grmp.FESpace(ft, grid)
    S = O.Space(N, fe)
    if mode == "factor":
        op = grmp.LaplaceOperator(1e-3)
        oform = dict(kind=O.BLF_LAPLACE, fe=fe, quadorder=op.quadorder(fes), factor=1e-3)
    else:
        kap = grmp.DeviceVector(grid.ctx, grid.ncells)
        kap.lognormal(seed, 0.0, 1.0)
        op = grmp.BilinearForm([grmp.Gradient, grmp.Gradient], grmp.CellCoefficient(kap))
        oform = dict(kind=O.BLF_LAPLACE, fe=fe, quadorder=op.quadorder(fes), factor=1.0, coef=O.COEF_CELL,
                     kappa=kap.download())
    A = grmp.FEMatrix([fes])
    form, keep = op.form(fes, 0, 0, grid.ctx)
    A.add_form(form, keep)
    A.set_fixed(S.fixed)
    A.build_pattern()
    A.assemble(grmp._abi.VALUES_FP64 | grmp._abi.VALUES_FP32)
    ref = O.assemble_matrix(S.ndofs, [dict(oform, coords=S.coords, cells=S.cells, celldofs=S.celldofs)],
                            fixed=S.fixed)
    frac = _check_csc(A, ref)
    return A, S, ref, frac


def test_c1_p2_noaction_pattern_and_values(grmp, O):
    """Config 1: P2, 64x64, LaplaceOperator(1e-3): the reference pattern."""
    A, S, ref, frac = _laplace_case(grmp, O, 64, O.FE_P2, "factor")
    assert A.info()[1] == 99329 + 0  # exactly-nonzero couplings (+ penalized diagonals already present)
    assert frac == 1.0  # bit-identical values


@pytest.mark.parametrize("N,fe,mode", [(64, 1, "factor"), (48, 1, "cell"), (40, 2, "cell"), (3, 2, "cell"),
                                       (1, 1, "factor")])
def test_laplace_parity(grmp, O, N, fe, mode):
    _, _, _, frac = _laplace_case(grmp, O, N, fe, mode)
    assert frac == 1.0


def test_hooke_parity(grmp, O):
    N = 24
    print(f"[case] hooke N={N}", flush=True)
    grid = grmp.grid_unitsquare(N)
    fes = grmp.FESpace(grmp.H1P1(2), grid)
    S = O.Space(N, O.FE_P1, ncomp=2)
    kap = grmp.DeviceVector(grid.ctx, grid.ncells)
    kap.lognormal(11, 0.0, 1.0)
    E, nu = 1.0, 0.3
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    op = grmp.HookStiffnessOperator2D(mu, lam, grmp.CellCoefficient(kap))
    A = grmp.FEMatrix([fes])
    form, keep = op.form(fes, 0, 0, grid.ctx)
    A.add_form(form, keep)
    A.set_fixed(S.fixed)
    A.build_pattern()
    A.assemble()
    ref = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_HOOKE2D, fe=O.FE_P1, ncomp=2, quadorder=0, factor=1.0,
                                          coef=O.COEF_CELL, kappa=kap.download(), params=np.array([lam, mu]),
                                          coords=S.coords, cells=S.cells, celldofs=S.celldofs)], fixed=S.fixed)
    assert _check_csc(A, ref) == 1.0


def test_mass_parity(grmp, O):
    N = 16
    grid = grmp.grid_unitsquare(N)
    fes = grmp.FESpace(grmp.H1P1(1), grid)
    S = O.Space(N, O.FE_P1)
    op = grmp.ReactionOperator(2.0)
    A = grmp.FEMatrix([fes])
    form, keep = op.form(fes, 0, 0, grid.ctx)
    A.add_form(form, keep)
    A.build_pattern()
    A.assemble()
    ref = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_MASS, fe=O.FE_P1, quadorder=2, factor=2.0,
                                          coords=S.coords, cells=S.cells, celldofs=S.celldofs)])
    assert _check_csc(A, ref) == 1.0


def test_tet_parity_and_mesh(grmp, O):
    n = 6
    print(f"[case] tet n={n}", flush=True)
    grid = grmp.grid_unitcube(n, 0.1, 5)
    c, cn = grid.download()
    vol = grid.cellvolumes()
    assert np.all(vol > 0) and abs(vol.sum() - 1.0) < 1e-13
    # unjittered Kuhn mesh: every tet has volume h^3/6
    g0 = grmp.grid_unitcube(n, 0.0, 0)
    assert np.allclose(g0.cellvolumes(), 1.0 / (6 * n ** 3), rtol=1e-13)
    fes = grmp.FESpace(grmp.H1P1(1), grid)
    kap = grmp.DeviceVector(grid.ctx, grid.ncells)
    kap.lognormal(3, 0.0, 0.5)
    op = grmp.BilinearForm([grmp.Gradient, grmp.Gradient], grmp.CellCoefficient(kap))
    bd = fes.boundary_dofs()
    on_bnd = np.any((c == 0.0) | (c == 1.0), axis=1)
    assert np.array_equal(bd, np.nonzero(on_bnd)[0])
    A = grmp.FEMatrix([fes])
    form, keep = op.form(fes, 0, 0, grid.ctx)
    A.add_form(form, keep)
    A.set_fixed(bd)
    A.build_pattern()
    A.assemble()
    ref = O.assemble_matrix(fes.ndofs, [dict(kind=O.BLF_LAPLACE, geom=3, fe=O.FE_P1, quadorder=0, factor=1.0,
                                            coef=O.COEF_CELL, kappa=kap.download(), coords=c, cells=cn,
                                            celldofs=cn)], fixed=bd.astype(np.int32))
    assert _check_csc(A, ref) == 1.0
    # the unjittered Kuhn mesh: axis-aligned cells, whose maps have exact
    # zeros, and exactly zero local entries that the pattern leaves out;
    # values compared bit for bit (signs of zero included)
    c0, cn0 = g0.download()
    fes0 = grmp.FESpace(grmp.H1P1(1), g0)
    A0 = grmp.FEMatrix([fes0])
    form0, keep0 = op.form(fes0, 0, 0, g0.ctx)
    A0.add_form(form0, keep0)
    A0.build_pattern()
    A0.assemble()
    ref0 = O.assemble_matrix(fes0.ndofs, [dict(kind=O.BLF_LAPLACE, geom=3, fe=O.FE_P1, quadorder=0, factor=1.0,
                                              coef=O.COEF_CELL, kappa=kap.download(), coords=c0, cells=cn0,
                                              celldofs=cn0)])
    assert _check_csc(A0, ref0) == 1.0
    assert np.array_equal(A0.csc()[2].view(np.uint64), ref0.nzval.view(np.uint64))
    # LinearForm(Identity, f = 1) on the tetrahedra, bitwise
    b = grmp.FEVector([fes])
    grmp.assemble_linearform(b, 0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
    bref = np.zeros(fes.ndofs)
    O.assemble_lf_const(bref, O.FE_P1, 1, 1, [1.0], c, cn, cn, geom=3)
    assert np.array_equal(b.entries(), bref)
    assert abs(bref.sum() - 1.0) < 1e-13


@pytest.mark.parametrize("N,fe,ncomp,f", [(32, 1, 1, [1.0]), (32, 2, 1, [1.0]), (20, 1, 2, [0.0, -1.0])])
def test_linearform_parity(grmp, O, N, fe, ncomp, f):
    grid = grmp.grid_unitsquare(N)
    ft = grmp.H1P1(ncomp) if fe == 1 else grmp.H1P2(ncomp)
    fes = grmp.FESpace(ft, grid)
    b = grmp.FEVector([fes])
    grmp.assemble_linearform(b, 0, grmp.LinearForm(grmp.Identity, grmp.DataFunction(f)))
    S = O.Space(N, fe, ncomp=ncomp)
    bref = np.zeros(S.ndofs)
    O.assemble_lf_const(bref, fe, ncomp, fe, f, S.coords, S.cells, S.celldofs)
    got = b.entries()
    assert np.array_equal(got, bref)


def test_integrator_parity(grmp, O):
    N = 12
    grid = grmp.grid_unitsquare(N)
    fes = grmp.FESpace(grmp.H1P2(1), grid)
    S = O.Space(N, O.FE_P2)
    u = np.sin(np.arange(S.ndofs) * 0.37)
    v = grmp.DeviceVector(grid.ctx, S.ndofs, u)
    out = np.empty(grid.ncells)
    grid.ctx.check(grmp.lib().grmp_integrate_identity(grid.ctx.h, fes.h, 2, v.h, 0,
                                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    ref = O.integrate_identity(u, O.FE_P2, 1, 2, S.coords, S.cells, S.celldofs)[:, 0]
    assert np.array_equal(out, ref)


def test_spmv_matches_scipy(grmp, O):
    A, S, ref, _ = _laplace_case(grmp, O, 20, O.FE_P2, "cell")
    M = ref.to_scipy().tocsr()
    x = np.cos(np.arange(S.ndofs))
    xv = grmp.DeviceVector(A.ctx, S.ndofs, x)
    yv = grmp.DeviceVector(A.ctx, S.ndofs)
    A.spmv(xv, yv, grmp._abi.VALUES_FP64)
    y = yv.download()
    yr = M @ x
    assert np.max(np.abs(y - yr)) <= 1e-12 * np.max(np.abs(yr))
    A.spmv(xv, yv, grmp._abi.VALUES_FP32)
    y32 = yv.download()
    free = np.ones(S.ndofs, bool)
    free[S.fixed] = False
    M32 = M.copy()
    M32.data = M32.data.astype(np.float32).astype(np.float64)
    yr32 = M32 @ x
    assert np.max(np.abs(y32 - yr32)[free]) <= 1e-5 * np.max(np.abs(yr32[free]))


def test_nnz_limit_guard(grmp):
    """The pattern total is summed in 64 bits before the int32 row-pointer
    scan; a build above the context's limit fails cleanly and leaves the
    context usable (the limit stands in for the 2^31 - 1 bound)."""
    print("[case] nnz limit guard", flush=True)
    grid = grmp.grid_unitsquare(8)
    fes = grmp.FESpace(grmp.H1P1(1), grid)
    ctx = grid.ctx
    L = grmp.lib()
    A = grmp.FEMatrix([fes])
    form, keep = grmp.LaplaceOperator(1.0).form(fes, 0, 0, ctx)
    A.add_form(form, keep)
    try:
        ctx.check(L.grmp_ctx_set_max_nnz(ctx.h, 100))
        with pytest.raises(grmp.GrmpError) as ei:
            A.build_pattern()
        assert ei.value.code == grmp._abi.ERR_UNSUPPORTED and "exceed the limit 100" in str(ei.value)
    finally:
        ctx.check(L.grmp_ctx_set_max_nnz(ctx.h, 1 << 40))
    assert L.grmp_ctx_set_max_nnz(ctx.h, 0) == grmp._abi.ERR_ARG
    B = grmp.FEMatrix([fes])
    form, keep2 = grmp.LaplaceOperator(1.0).form(fes, 0, 0, ctx)
    B.add_form(form, keep2)
    B.build_pattern()
    assert B.info()[1] == 81 + 4 * 8 * 9  # the 5-point pattern: the diagonal couplings are exactly 0
