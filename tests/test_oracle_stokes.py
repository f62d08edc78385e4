"""Stokes pieces of the oracle: Stroud rule exactness, block structure and
symmetry of the saddle point, and pressure robustness (the BR -> RT0
reconstructed right-hand side gives u_h = 0 for a gradient force, the
reference's own check, examples/Example222_PressureRobustness2D.jl:189-211)."""
from math import factorial

import numpy as np


def test_stroud_exactness(O):
    for order in [3, 4, 5, 6]:
        x, w = O.quad_stroud(order)
        assert abs(w.sum() - 1.0) < 1e-14
        for a in range(order + 1):
            for b in range(order + 1 - a):
                exact = factorial(a) * factorial(b) / factorial(a + b + 2) * 2.0
                got = float(np.sum(w * x[:, 0] ** a * x[:, 1] ** b))
                assert abs(got - exact) < 1e-14, (order, a, b)


def test_stokes_matrix_structure(O):
    S = O.StokesSpace(6)
    A = S.matrix(1e-2)
    M = A.to_scipy().toarray()
    nU = S.nU
    K, Bt, B, C = M[:nU, :nU], M[:nU, nU:], M[nU:, :nU], M[nU:, nU:]
    # [1,2] = -B^T, [2,1] = +B, bitwise negatives
    assert np.array_equal(Bt, -B.T)
    # velocity block symmetric (penalty on the diagonal)
    assert np.array_equal(K, K.T)
    # only the pinned pressure dof has a [2,2] entry
    assert np.count_nonzero(C) == 1 and C[0, 0] == O.PENALTY
    # discrete divergence of a constant velocity field is zero: B 1_x = 0
    ux = np.zeros(nU)
    ux[:S.nnodes] = 1.0
    assert np.max(np.abs(B @ ux)) < 1e-13


def test_pressure_robustness(O):
    """f = grad(phi) with polynomial phi integrated exactly: u_h = 0."""
    S = O.StokesSpace(8)
    A = S.matrix(1e-2)
    b = S.rhs(4, np.zeros((0, 4)), [1.0, 0.5, -2.0])
    b[S.fixed] = 0.0
    x = O.direct_solve(A, b)
    u, p = x[:S.nU], x[S.nU:]
    assert np.max(np.abs(u)) < 1e-12 * max(1.0, np.max(np.abs(p)))
    assert np.max(np.abs(p)) > 1e-3


def test_stokes_rhs_divergence_free_part(O):
    """A nonzero divergence-free force gives a nonzero velocity."""
    S = O.StokesSpace(8)
    A = S.matrix(1e-2)
    modes = np.array([[0.0, 1, 1, 0.0], [1.0, 1, 2, 0.3]])
    b = S.rhs(4, modes, [0.0, 0.0, 0.0])
    b[S.fixed] = 0.0
    x = O.direct_solve(A, b)
    assert np.max(np.abs(x[:S.nU])) > 1e-4


def test_stokes_exactness_br_p0(O):
    """The reference's Stokes test for H1BR{2} x L2P0 with orders (1, 0)
    (test/runtests.jl:521-540,581,606-640): u = (y + 1, x - 1), p = 0, f = 0,
    viscosity 1, best-approximated Dirichlet data, FixedIntegralMean; the
    L2 errors of u_h (quadorder 2 + 1) and p_h (0 + 0) are below the
    reference's tolerance 6e-12.  On the structured square and on a shuffled
    mesh (the reference uses ExtendableGrids' refined unit square, absent
    here)."""
    from meshes import shuffled_square
    lin = np.array([[1.0, 0.0, 1.0], [-1.0, 1.0, 0.0]])
    for seed in (None, 8):
        if seed is None:
            S = O.StokesSpace(8)
        else:
            coords, cells = shuffled_square(O, 8, seed)
            S = O.StokesSpace(8, coords, cells)
        A = S.matrix(1.0)
        t, bd = O.boundary_bestapprox_br(S, lin, 2 + 1)
        target = np.zeros(S.n)
        target[bd] = t[bd]
        x = S.pressure_mean_shift(O.direct_solve_fixed(A, np.zeros(S.n), S.fixed, target[S.fixed]))
        nrm = S.fnormal.reshape(-1, 2)[S.cellfaces]
        eu = O.l2error_cells(O.FE_BR, 2, x[:S.nU], lin, 3, 1.0, S.coords, S.cells, S.celldofs, normals=nrm).sum()
        ep = O.l2error_cells(O.FE_P0, 1, x[S.nU:], None, 0, 1.0, S.coords, S.cells,
                             np.arange(S.nc, dtype=np.int32)[:, None]).sum()
        assert np.sqrt(eu) < 6e-12 and np.sqrt(ep) < 6e-12, (seed, np.sqrt(eu), np.sqrt(ep))
