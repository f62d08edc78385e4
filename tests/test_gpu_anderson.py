"""Anderson acceleration on the device (grmp_anderson_*,
AndersonAccelerationManager / update_anderson!, src/solvers.jl:688-808)
against the oracle's restatement (oracle.Anderson).

Bar: the same sequence of fixpoint outputs fed to both gives coefficients
within 1e-9 and iterates within 1e-12 of the largest entry (the Gram entries
are summed in another order than the oracle's numpy dot products); given the
device's coefficients, the combination is the oracle formula's bit for bit
(both sum damping alpha_a tilde_a then (1 - damping) alpha_a last_a in the
reference's order, without contraction).  The l2 metric is tested on part of
the vector (the unknowns not accelerated keep the fixpoint output); the L2
metric with the device-assembled mass matrix of a P1 space.  In the Newton
loop of the nonlinear Poisson problem the accelerated device iterates stay
within 1e-9 of the oracle's accelerated loop."""
import ctypes

import numpy as np
import pytest

from meshes import shuffled_square

pytestmark = pytest.mark.gpu

C = (1.0, 0.0, 1.0)
F = [0.0] * 15
F[11], F[13] = -2.0, -2.0


def _create(grmp, ctx, n, depth, damping, metric, ranges, x0):
    lib = grmp._abi.lib()
    r = np.asarray(ranges, dtype=np.int64).ravel()
    h = ctypes.c_void_p()
    rc = lib.grmp_anderson_create(ctx.h, n, depth, damping, metric.h if metric is not None else None,
                                  len(r) // 2, r.ctypes.data, x0.h, ctypes.byref(h))
    return rc, h


@pytest.mark.parametrize("depth,damping,metric", [(2, 1.0, "l2"), (3, 0.7, "l2"), (2, 0.9, "L2"), (4, 1.0, "L2")])
def test_anderson_sequence_matches_oracle(grmp, O, depth, damping, metric):
    print(f"[case] anderson sequence depth={depth} damping={damping} metric={metric}", flush=True)
    lib = grmp._abi.lib()
    ctx = grmp.default_context()
    coords, cells = shuffled_square(O, 20, 3)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    fes = grmp.FESpace(grmp.H1P1(1), grid)
    n = fes.ndofs
    B, Bs = None, None
    if metric == "L2":
        B = grmp.FEMatrix([fes])
        form, keep = grmp.ReactionOperator(1.0).form(fes, 0, 0, B.ctx, fes)
        B.add_form(form, keep)
        B.build_pattern()
        B.assemble(grmp._abi.VALUES_FP64)
        Bs = B.to_scipy()
        ranges, mask = [[0, n]], np.ones(n, dtype=bool)
    else:
        ranges = [[0, n // 3], [n // 2, n - 7]]
        mask = np.zeros(n, dtype=bool)
        for lo, hi in ranges:
            mask[lo:hi] = True
    rng = np.random.default_rng(depth + 17)
    x0 = rng.standard_normal(n)
    X = grmp.DeviceVector(ctx, n)
    X.upload(x0)
    rc, h = _create(grmp, ctx, n, depth, damping, B, ranges, X)
    ctx.check(rc)
    ref = O.Anderson(x0, depth, damping, Bs, mask)
    G = rng.standard_normal((n, n)) * 0.3 / np.sqrt(n)
    c = rng.standard_normal(n)
    x = x0.copy()
    try:
        for k in range(1, 10):
            t = G @ x + c + 1e-3 * rng.standard_normal(n)
            X.upload(t)
            alpha = np.zeros(depth + 1)
            res = ctypes.c_double()
            ctx.check(lib.grmp_anderson_update(h, X.h, alpha.ctypes.data, ctypes.byref(res)))
            xd = X.download()
            # the stored iterates before this update (the oracle's equal the device's), for the bitwise check
            tilde_before = [v.copy() for v in ref.tilde]
            last_before = [v.copy() for v in ref.last]
            x = t.copy()
            aref, rref = ref.update(x)
            err = np.max(np.abs(xd - x)) / np.max(np.abs(x))
            aerr = np.max(np.abs(alpha - aref)) if k >= 3 else 0.0
            print(f"  update {k}: alpha {np.round(alpha, 6).tolist()}, max alpha diff {aerr:.1e}, max rel diff "
                  f"{err:.1e}, local residual {res.value:.1e}")
            assert err <= 1e-12 and aerr <= 1e-9
            assert lib.grmp_anderson_depth(h) == k - 1
            if k <= 2:
                assert np.array_equal(xd, t) and not alpha.any()
                continue
            # the combination with the device's coefficients, in the oracle's order, bit for bit
            m = depth
            tl = tilde_before[1:] + [t]
            cd = min(m, k - 2)
            out = np.zeros(n)
            for i in range(cd + 1):
                out += damping * alpha[i] * tl[m - i]
                out += (1.0 - damping) * alpha[i] * last_before[m - i]
            assert np.array_equal(xd[mask], out[mask]) and np.array_equal(xd[~mask], t[~mask])
            x = xd.copy()
            ref.last[m][:] = xd  # continue both from the device iterate
    finally:
        lib.grmp_anderson_destroy(h)


def test_anderson_argument_errors(grmp):
    print("[case] anderson argument errors", flush=True)
    ctx = grmp.default_context()
    X = grmp.DeviceVector(ctx, 100)
    X.fill(1.0)
    for depth, damping, ranges in [(0, 1.0, [[0, 100]]), (11, 1.0, [[0, 100]]), (2, 0.0, [[0, 100]]),
                                   (2, 1.5, [[0, 100]]), (2, 1.0, [[50, 101]]), (2, 1.0, [[60, 50]])]:
        rc, h = _create(grmp, ctx, 100, depth, damping, None, ranges, X)
        assert rc == grmp._abi.ERR_ARG and not h.value, (depth, damping, ranges)
    rc, h = _create(grmp, ctx, 100, 2, 1.0, None, [[0, 100]], X)
    assert rc == 0
    Y = grmp.DeviceVector(ctx, 99)
    lib = grmp._abi.lib()
    assert lib.grmp_anderson_update(h, Y.h, None, None) == grmp._abi.ERR_ARG
    assert lib.grmp_anderson_update(h, X.h, None, None) == 0
    lib.grmp_anderson_destroy(h)
    with pytest.raises(grmp._abi.GrmpError):
        grmp.SolverConfig(anderson_metric="H2")
    with pytest.raises(grmp._abi.GrmpError):
        grmp.SolverConfig(anderson_iterations=2, anderson_damping=0.0)


@pytest.mark.parametrize("depth,damping,metric", [(2, 1.0, "l2"), (1, 0.5, "L2"), (3, 1.0, "H1")])
def test_newton_with_anderson_matches_oracle(grmp, O, depth, damping, metric):
    print(f"[case] newton P1 with anderson depth={depth} damping={damping} metric={metric}", flush=True)
    coords, cells = shuffled_square(O, 12, 9)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    fes = grmp.FESpace(grmp.H1P1(1), grid)
    pde = grmp.PDEDescription("nonlinear Poisson problem")
    pde.add_unknown("u")
    nlf = grmp.NonlinearForm(grmp.Gradient, [grmp.Identity, grmp.Gradient], [0, 0], grmp.NonlinearDiffusion(*C),
                             [2, 3], bonus_quadorder=2)
    pde.add_operator((0, 0), nlf)
    pde.add_boundarydata(0, grmp.InterpolateDirichletBoundary, data=grmp.PolynomialData([0, 0, 0, 0, 1.0]))
    pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.PolynomialData(F)))
    sol = grmp.FEVector([fes])
    info, A, b = grmp.solve(sol, pde, target_residual=1e-11, nonlinear_maxiterations=30, rtol=1e-13,
                            anderson_iterations=depth, anderson_damping=damping, anderson_metric=metric,
                            anderson_unknowns=[0])
    Bs = None
    if metric != "l2":
        B = grmp.FEMatrix([fes])
        op = grmp.ReactionOperator(1.0) if metric == "L2" else grmp.LaplaceOperator(1.0)
        form, keep = op.form(fes, 0, 0, B.ctx, fes)
        B.add_form(form, keep)
        B.build_pattern()
        B.assemble(grmp._abi.VALUES_FP64)
        Bs = B.to_scipy()
    target = coords[:, 0] * coords[:, 1]
    xo, res = O.newton_solve(1, coords, cells, fes.celldofs(), fes.ndofs, fes.boundary_dofs(), target, C, F,
                             quadorder=3, lf_quadorder=5, maxit=30, target_residual=1e-11,
                             anderson=dict(depth=depth, damping=damping, B=Bs))
    err = np.max(np.abs(sol.entries() - xo)) / np.max(np.abs(xo))
    print(f"  device {info.iterations} steps {['%.1e' % r for r in info.residuals]}; oracle {len(res)} steps "
          f"{['%.1e' % r for r in res]}; max rel diff {err:.2e}")
    assert info.status == 0 and err <= 1e-9
    assert abs(info.iterations - len(res)) <= 1
    # the first steps agree closely: same iterates before the residual is at rounding level
    k = min(3, len(res), len(info.residuals))
    assert np.allclose(info.residuals[:k], res[:k], rtol=1e-5, atol=1e-12)
