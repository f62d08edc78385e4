"""Serving path: continuous batching keeps each result identical to the
same request solved alone; results match the oracle; failures stay local.

Stated tolerance: served results equal their alone results bit for bit
(per-problem reductions run in a fixed order whatever else is in flight);
against the FP64 sparse-LU reference they agree to 1e-8 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIX = [(1, 1, 20, 101), (1, 2, 33, 102), (2, 1, 17, 103), (1, 1, 64, 104), (1, 2, 9, 105), (2, 1, 40, 106),
       (1, 2, 48, 107), (2, 1, 8, 108)]


def _alone(grmp, specs, **kw):
    out = {}
    with grmp.Server("served", **kw) as S:
        for spec in specs:
            t = S.submit(*spec)
            r = S.wait(t)
            out[spec] = (r.status, r.iterations, S.solution(t).copy())
    return out


def test_mix_equals_alone_bitwise(grmp):
    print("[case] server mix vs alone", flush=True)
    alone = _alone(grmp, MIX)
    with grmp.Server("served") as S:
        ts = [S.submit(*spec) for spec in MIX]
        for spec, t in zip(MIX, ts):
            r = S.wait(t)
            st, it, x = alone[spec]
            assert r.status == 0 and st == 0
            assert r.iterations == it
            assert np.array_equal(S.solution(t), x), spec


def test_served_matches_oracle(grmp, O):
    print("[case] server vs oracle", flush=True)
    specs = [(1, 1, 12, 5), (1, 2, 10, 6), (2, 1, 11, 7)]
    with grmp.Server("served") as S:
        ts = [S.submit(*spec, rtol=1e-12) for spec in specs]
        for (kind, fe, N, seed), t in zip(specs, ts):
            r = S.wait(t)
            assert r.status == 0
            x = S.solution(t)
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
            xref = O.direct_solve(A, O.penalized_rhs(b, Sp.fixed))
            err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
            assert err <= 1e-8, (kind, fe, N, err)


def test_plain_matches_served(grmp):
    print("[case] plain vs served", flush=True)
    specs = MIX[:4]
    served = _alone(grmp, specs)
    with grmp.Server("plain") as S:
        for spec in specs:
            t = S.submit(*spec)
            r = S.wait(t)
            assert r.status == 0
            x0 = served[spec][2]
            err = np.linalg.norm(S.solution(t) - x0) / np.linalg.norm(x0)
            assert err <= 1e-8, (spec, err)


def test_failures_stay_local(grmp):
    """A non-finite coefficient and an unreachable tolerance fail their own
    requests only; the neighbours in the batch are unchanged."""
    print("[case] server failure isolation", flush=True)
    good = MIX[:3]
    alone = _alone(grmp, good)
    with grmp.Server("served") as S:
        tnan = S.submit(1, 1, 16, 9, sigma=float("nan"))
        ts = [S.submit(*spec) for spec in good]
        tbad = S.submit(1, 1, 12, 10, rtol=1e-30)
        assert S.wait(tnan).status == grmp._abi.ERR_NONFINITE
        rb = S.wait(tbad)
        # the FP32 sweeps stall, FP64 CG takes over and stalls at its floor
        assert rb.status == grmp._abi.ERR_NOCONV and rb.fp64_fallback == 1
        for spec, t in zip(good, ts):
            assert S.wait(t).status == 0
            assert np.array_equal(S.solution(t), alone[spec][2])


def test_admission_rejects_oversize(grmp):
    with grmp.Server("served", memory_budget=1 << 20) as S:
        with pytest.raises(grmp.GrmpError) as ei:
            S.submit(1, 2, 200, 1)
        assert ei.value.code == grmp._abi.ERR_ADMISSION
        t = S.submit(1, 1, 8, 1)  # small enough
        assert S.wait(t).status == 0


def test_memory_budget_serializes(grmp):
    """With room for about one problem at a time every request still
    completes, with the same results."""
    specs = MIX[:4]
    alone = _alone(grmp, specs)
    budget = max(grmp.server.footprint(k, f, N) for k, f, N, _ in specs) + 1024
    with grmp.Server("served", memory_budget=budget) as S:
        ts = [S.submit(*spec) for spec in specs]
        for spec, t in zip(specs, ts):
            assert S.wait(t).status == 0
            assert np.array_equal(S.solution(t), alone[spec][2])


def test_profiled_engine_counts_spmv(grmp):
    """Profiled serving times every engine SpMV launch with HIP events and
    counts its algorithmic bytes; results equal the unprofiled ones."""
    specs = MIX[:3]
    alone = _alone(grmp, specs)
    with grmp.Server("served", profile=True) as S:
        ts = [S.submit(*spec) for spec in specs]
        for spec, t in zip(specs, ts):
            assert S.wait(t).status == 0
            assert np.array_equal(S.solution(t), alone[spec][2])
        st = S.stats()
    assert st["spmv_launches"] > 0 and st["spmv_seconds"] > 0 and st["spmv_bytes"] > 0
    assert st["steps"] >= st["spmv_launches"]


def test_fp64_engine_requests(grmp, O):
    """FP64 requests run plain FP64 CG inside the batched engine (the path
    stalled FP32 sweeps fall back to): they match sparse LU and leave their
    neighbours in the batch unchanged."""
    print("[case] server FP64 engine requests", flush=True)
    good = MIX[:2]
    alone = _alone(grmp, good)
    kind, fe, N, seed = 1, 2, 40, 21
    with grmp.Server("served") as S:
        ts = [S.submit(*spec) for spec in good]
        t = S.submit(kind, fe, N, seed, rtol=1e-11, precision="fp64")
        r = S.wait(t)
        assert r.status == 0 and r.fp64_fallback == 1
        x = S.solution(t)
        for spec, tt in zip(good, ts):
            assert S.wait(tt).status == 0
            assert np.array_equal(S.solution(tt), alone[spec][2])
    grid = grmp.grid_unitsquare(N)
    kv = grmp.DeviceVector(grid.ctx, grid.ncells)
    kv.lognormal(seed, 0.0, 1.0)
    Sp = O.Space(N, fe)
    A = O.assemble_matrix(Sp.ndofs, [dict(kind=O.BLF_LAPLACE, fe=fe, quadorder=2, factor=1.0, coef=O.COEF_CELL,
                                          kappa=kv.download(), coords=Sp.coords, cells=Sp.cells,
                                          celldofs=Sp.celldofs)], fixed=Sp.fixed)
    b = np.zeros(Sp.ndofs)
    O.assemble_lf_const(b, fe, 1, fe, [1.0], Sp.coords, Sp.cells, Sp.celldofs)
    xref = O.direct_solve(A, O.penalized_rhs(b, Sp.fixed))
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-9
