"""Batched assembly of many meshes in one launch (grmp_matrix_assemble_batch:
a persistent kernel over a queue of (matrix, row block) tiles) against the
per-matrix assembly (src/assemblypatterns/bilinearform.jl:226-377 per item).

Bar: FP64 values and patterns bitwise equal for every matrix of the batch:
P1 / P2 Laplace with cell coefficients, P1{2} Hooke and P1 mass on generated
and shuffled meshes, and P1 on jittered tetrahedra, of uneven sizes."""
import ctypes

import numpy as np
import pytest

from meshes import shuffled_square

pytestmark = pytest.mark.gpu


def _mats(grmp, O):
    out, alive = [], []

    def add(fes, op):
        A = grmp.FEMatrix([fes])
        form, keep = op.form(fes, 0, 0, fes.ctx)
        A.add_form(form, keep)
        A.set_fixed(fes.boundary_dofs())
        A.build_pattern()
        out.append((A, keep, fes))

    for N, seed in [(37, 1), (5, 2), (64, 3)]:
        grid = grmp.grid_unitsquare(N)
        kap = grmp.DeviceVector(grid.ctx, grid.ncells)
        kap.lognormal(seed, 0.0, 1.0)
        alive += [grid, kap]
        add(grmp.FESpace(grmp.H1P1(1), grid), grmp.BilinearForm([grmp.Gradient, grmp.Gradient],
                                                                 grmp.CellCoefficient(kap)))
        add(grmp.FESpace(grmp.H1P2(1), grid), grmp.BilinearForm([grmp.Gradient, grmp.Gradient],
                                                                 grmp.CellCoefficient(kap)))
        add(grmp.FESpace(grmp.H1P1(2), grid), grmp.HookStiffnessOperator2D(1 / 2.6, 0.3 / 0.52,
                                                                           grmp.CellCoefficient(kap)))
    coords, cells = shuffled_square(O, 11, 5)
    up = grmp.ExtendableGrid.from_arrays(coords, cells)
    add(grmp.FESpace(grmp.H1P1(1), up), grmp.ReactionOperator(2.0))
    add(grmp.FESpace(grmp.H1P2(1), up), grmp.LaplaceOperator(0.7))
    tet = grmp.grid_unitcube(5, 0.1, 3)
    kt = grmp.DeviceVector(tet.ctx, tet.ncells)
    kt.lognormal(9, 0.0, 0.5)
    alive += [up, tet, kt]
    add(grmp.FESpace(grmp.H1P1(1), tet), grmp.BilinearForm([grmp.Gradient, grmp.Gradient],
                                                          grmp.CellCoefficient(kt)))
    return out, alive


def test_batch_equals_per_matrix(grmp, O):
    print("[case] batched assembly of 11 matrices", flush=True)
    one, alive1 = _mats(grmp, O)
    ref = []
    for A, keep, fes in one:
        A.assemble(grmp._abi.VALUES_FP64)
        ref.append(A.csc())
    batch, alive2 = _mats(grmp, O)
    arr = (grmp._abi.VP * len(batch))(*[A.h for A, _, _ in batch])
    batch[0][0].ctx.check(grmp.lib().grmp_matrix_assemble_batch(len(batch), arr))
    for (A, keep, fes), (cp, rv, nz) in zip(batch, ref):
        c2, r2, n2 = A.csc()
        assert np.array_equal(cp, c2) and np.array_equal(rv, r2)
        assert np.array_equal(nz, n2)
    # a second batch call on the same matrices gives the same values again
    batch[0][0].ctx.check(grmp.lib().grmp_matrix_assemble_batch(len(batch), arr))
    assert all(np.array_equal(A.csc()[2], r[2]) for (A, _, _), r in zip(batch, ref))


def test_batch_argument_errors(grmp):
    print("[case] batched assembly errors", flush=True)
    assert grmp.lib().grmp_matrix_assemble_batch(-1, None) == grmp._abi.ERR_ARG
    assert grmp.lib().grmp_matrix_assemble_batch(0, None) == 0
