"""Multigrid serving mode: FP64 CG preconditioned by a geometric V-cycle.

Stated tolerance: against the FP64 sparse-LU solution of the same penalized
system the results agree to 1e-8 relative (rtol 1e-12); a request's result is
bitwise the same whether it is solved alone or among others (each request
has its own stream and fixed-order reductions); the iteration count stays
bounded as N grows (at most 40 for rtol 1e-10)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reference(grmp, O, kind, fe, N, seed):
    grid = grmp.grid_unitsquare(N)
    kv = grmp.DeviceVector(grid.ctx, grid.ncells)
    kv.lognormal(seed, 0.0, 1.0)
    kap = kv.download()
    if kind == 1:
        Sp = O.Space(N, fe)
        form = dict(kind=O.BLF_LAPLACE, fe=fe, quadorder=2 * fe - 2, factor=1.0, coef=O.COEF_CELL, kappa=kap)
        f, q, nc = [1.0], fe, 1
    else:
        Sp = O.Space(N, O.FE_P1, ncomp=2)
        form = dict(kind=O.BLF_HOOKE2D, fe=O.FE_P1, ncomp=2, quadorder=0, factor=1.0, coef=O.COEF_CELL,
                    kappa=kap, params=np.array([0.3 / (1.3 * 0.4), 1.0 / 2.6]))
        f, q, nc = [0.0, -1.0], 1, 2
    A = O.assemble_matrix(Sp.ndofs, [dict(form, coords=Sp.coords, cells=Sp.cells, celldofs=Sp.celldofs)],
                          fixed=Sp.fixed)
    b = np.zeros(Sp.ndofs)
    O.assemble_lf_const(b, form["fe"], nc, q, f, Sp.coords, Sp.cells, Sp.celldofs)
    return O.direct_solve(A, O.penalized_rhs(b, Sp.fixed))


# single-level (N <= 4), two-level, non-nested (odd N) and P2 hierarchies
SPECS = [(1, 1, 1, 3), (1, 1, 4, 4), (2, 1, 3, 5), (1, 2, 2, 6), (1, 1, 5, 7), (1, 2, 9, 8), (2, 1, 13, 9),
         (1, 1, 37, 10), (1, 2, 21, 11), (2, 1, 30, 12)]


@pytest.mark.parametrize("block_rows,batch", [(0, 0), (64, 0), (0, 16), (64, 16)])
def test_mg_matches_lu(grmp, O, block_rows, batch):
    """block_rows 0: the default thresholds (these requests run whole inside
    one workgroup, or with their coarse levels there); 64: every level above
    the coarsest runs as grid kernels.  batch 16: the requests advance in
    lockstep, one launch per phase."""
    print(f"[case] multigrid vs LU, block_rows {block_rows}, batch {batch}", flush=True)
    with grmp.Server("mg", mg_block_rows=block_rows, mg_coarse_rows=block_rows, mg_batch=batch) as S:
        ts = [S.submit(*spec, rtol=1e-12) for spec in SPECS]
        for spec, t in zip(SPECS, ts):
            r = S.wait(t)
            assert r.status == 0, (spec, r.status)
            xref = _reference(grmp, O, *spec)
            # every dof is fixed on the 1 x 1 square: the solution is 0
            err = np.linalg.norm(S.solution(t) - xref) / max(np.linalg.norm(xref), 1e-300)
            print(f"  {spec}: {r.iterations} iterations, rel err {err:.1e}")
            assert err <= 1e-8, (spec, err)
            assert r.iterations <= 40, (spec, r.iterations)


def test_mg_iterations_bounded(grmp):
    """Mesh-independent convergence: the iteration count at N = 200 stays
    within a few of the one at N = 25 for every family."""
    print("[case] multigrid iteration counts", flush=True)
    with grmp.Server("mg") as S:
        out = {}
        for kind, fe in [(1, 1), (1, 2), (2, 1)]:
            for N in (25, 200):
                r = S.wait(S.submit(kind, fe, N, 21, rtol=1e-10, keep_solution=False))
                assert r.status == 0
                out[(kind, fe, N)] = r.iterations
        print("  ", out)
        for kind, fe in [(1, 1), (1, 2), (2, 1)]:
            assert out[(kind, fe, 200)] <= 40
            assert out[(kind, fe, 200)] <= out[(kind, fe, 25)] + 12


@pytest.mark.parametrize("block_rows", [0, 64])
def test_mg_concurrent_equals_alone_bitwise(grmp, block_rows):
    print(f"[case] multigrid concurrent vs alone, block_rows {block_rows}", flush=True)
    specs = SPECS[4:]
    alone = {}
    with grmp.Server("mg", mg_slots=1, mg_batch=0, mg_block_rows=block_rows, mg_coarse_rows=block_rows) as S:
        for spec in specs:
            t = S.submit(*spec)
            r = S.wait(t)
            alone[spec] = (r.iterations, S.solution(t).copy())
    with grmp.Server("mg", mg_slots=4, mg_batch=0, mg_block_rows=block_rows, mg_coarse_rows=block_rows) as S:
        ts = [S.submit(*spec) for spec in specs + specs]
        for spec, t in zip(specs + specs, ts):
            r = S.wait(t)
            assert r.status == 0
            assert r.iterations == alone[spec][0]
            assert np.array_equal(S.solution(t), alone[spec][1]), spec
    # lockstep batches run the same row bodies and fixed-order reductions
    with grmp.Server("mg", mg_batch=16, mg_block_rows=block_rows, mg_coarse_rows=block_rows) as S:
        ts = [S.submit(*spec) for spec in specs + specs]
        for spec, t in zip(specs + specs, ts):
            r = S.wait(t)
            assert r.status == 0
            assert r.iterations == alone[spec][0]
            assert np.array_equal(S.solution(t), alone[spec][1]), spec


def test_mg_failures_stay_local(grmp):
    print("[case] multigrid failure isolation", flush=True)
    good = SPECS[4:7]
    with grmp.Server("mg", mg_slots=1, mg_batch=0) as S:
        alone = {spec: S.solution(t).copy() for spec in good for t in [S.submit(*spec)] if S.wait(t).status == 0}
    for kw in ({"mg_slots": 3, "mg_batch": 0}, {"mg_batch": 8}, {"mg_batch": 4, "mg_slots": 1, "mg_threads": 1},
               {"mg_batch": 4, "mg_threads": -1}):
        _failures(grmp, good, alone, kw)


def _failures(grmp, good, alone, kw):
    with grmp.Server("mg", **kw) as S:
        tnan = S.submit(1, 1, 16, 9, sigma=float("nan"))
        ts = [S.submit(*spec) for spec in good]
        tbad = S.submit(1, 1, 12, 10, rtol=1e-30)
        assert S.wait(tnan).status == grmp._abi.ERR_NONFINITE
        rb = S.wait(tbad)
        assert rb.status == grmp._abi.ERR_NOCONV, rb.status
        for spec, t in zip(good, ts):
            assert S.wait(t).status == 0
            assert np.array_equal(S.solution(t), alone[spec])


def test_mg_profile_counts_residual_kernel(grmp):
    """Profiling times the residual kernel of levels with at least 2^21 rows
    (P2 on 724 x 724 has 2,099,601); results equal the unprofiled ones."""
    specs = [(1, 2, 724, 31), (1, 1, 40, 32)]
    with grmp.Server("mg") as S:
        ref = {spec: S.solution(t).copy() for spec in specs for t in [S.submit(*spec)] if S.wait(t).status == 0}
    with grmp.Server("mg", profile=True) as S:
        ts = [S.submit(*spec) for spec in specs]
        for spec, t in zip(specs, ts):
            assert S.wait(t).status == 0
            assert np.array_equal(S.solution(t), ref[spec])
        st = S.stats()
    assert st["spmv_launches"] > 0 and st["spmv_seconds"] > 0 and st["spmv_bytes"] > 0
