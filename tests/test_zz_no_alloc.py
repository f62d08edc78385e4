"""grmp_integrate_l2error_noalloc: the L2ErrorIntegrator entry point that
allocates nothing (its per-cell values and partials live in a file-scope
device array) against grmp_integrate_l2error, through the C ABI.

Bar: per-cell values and totals bitwise equal to the allocating entry
point's (same kernels, same reduction order), on two calls in a row, the
second on a smaller mesh; a mesh above the scratch limit is rejected with
GRMP_ERR_ARG."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CUB = [0.3, -0.2, 0.5, 0.1, 0.7, -0.4, 1.1, -0.6, 0.9, 0.25]


def _bind(grmp):
    return grmp.lib().grmp_integrate_l2error_noalloc


def _case(grmp, N, seed, fe):
    grid = grmp.grid_unitsquare(N)
    fes = grmp.FESpace(grmp.H1P2(1) if fe == 2 else grmp.H1P1(1), grid)
    X = grmp.FEVector([fes])
    X.set_entries(np.random.default_rng(seed).standard_normal(fes.ndofs))
    return grid, fes, X


def _both(grmp, f, grid, fes, X, q):
    c = np.ascontiguousarray(CUB)
    ref_cells = np.empty(grid.ncells)
    ref_tot = ctypes.c_double()
    grid.ctx.check(grmp.lib().grmp_integrate_l2error(grid.ctx.h, fes.h, X.vec.h, 0, grmp._abi.dptr(c), c.size, q, 0.5,
                                                     grmp._abi.dptr(ref_cells), ctypes.byref(ref_tot)))
    cells = np.full(grid.ncells, np.nan)
    tot = ctypes.c_double(np.nan)
    grid.ctx.check(f(grid.ctx.h, fes.h, X.vec.h, 0, grmp._abi.dptr(c), c.size, q, 0.5, grmp._abi.dptr(cells),
                     ctypes.byref(tot)))
    return ref_cells, ref_tot.value, cells, tot.value


def test_noalloc_equals_allocating_entry(grmp):
    f = _bind(grmp)
    for N, fe, q in [(300, 2, 4), (37, 1, 2)]:   # second call: a smaller input
        print(f"[case] no-alloc l2error N={N} fe={fe} q={q}", flush=True)
        grid, fes, X = _case(grmp, N, N, fe)
        ref_cells, ref_tot, cells, tot = _both(grmp, f, grid, fes, X, q)
        assert np.array_equal(cells, ref_cells)
        assert tot == ref_tot and tot > 0
    # total only (no per-cell output)
    tot = ctypes.c_double()
    c = np.ascontiguousarray(CUB)
    grid.ctx.check(f(grid.ctx.h, fes.h, X.vec.h, 0, grmp._abi.dptr(c), c.size, 2, 0.5, None, ctypes.byref(tot)))
    assert tot.value == ref_tot


def test_noalloc_rejects_oversize(grmp):
    f = _bind(grmp)
    print("[case] no-alloc l2error above the scratch limit", flush=True)
    grid = grmp.grid_unitsquare(1449)          # 4,199,202 cells > 2^22
    fes = grmp.FESpace(grmp.L2P0(1), grid)
    X = grmp.FEVector([fes])
    tot = ctypes.c_double()
    c = np.ascontiguousarray(CUB)
    rc = f(grid.ctx.h, fes.h, X.vec.h, 0, grmp._abi.dptr(c), c.size, 0, 1.0, None, ctypes.byref(tot))
    assert rc == grmp._abi.ERR_ARG
