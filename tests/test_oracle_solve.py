"""End-to-end pins of the oracle: manufactured solutions that the discrete
spaces reproduce exactly, the Hooke kernel, and the CPU PCG vs sparse LU."""
import numpy as np


def _penalized(O, S, A_forms, f_consts, g_exact, fe, ncomp=1):
    A = O.assemble_matrix(S.ndofs, A_forms, fixed=S.fixed)
    b = np.zeros(S.ndofs)
    q = 1 if fe == O.FE_P1 else 2
    if f_consts is not None:
        O.assemble_lf_const(b, fe, ncomp, q, f_consts, S.coords, S.cells, S.celldofs)
    b[S.fixed] = O.PENALTY * g_exact[S.fixed]
    return A, b


def _dof_points(S):
    """coordinates of every scalar dof (nodes, then face midpoints)"""
    pts = S.coords
    if S.fe == 2:
        mids = 0.5 * (S.coords[S.facenodes[:, 0]] + S.coords[S.facenodes[:, 1]])
        pts = np.vstack([pts, mids])
    return pts


def test_p2_reproduces_quadratic(O):
    """-Delta u = 2 with u = x(1-x) + y: u is in P2, so u_h = u at every dof."""
    S = O.Space(8, O.FE_P2)
    P = _dof_points(S)
    u = P[:, 0] * (1 - P[:, 0]) + P[:, 1]
    A, b = _penalized(O, S, [dict(kind=O.BLF_LAPLACE, fe=O.FE_P2, quadorder=2, factor=1.0, coords=S.coords,
                                  cells=S.cells, celldofs=S.celldofs)], [2.0], u, O.FE_P2)
    x = O.direct_solve(A, b)
    assert np.max(np.abs(x - u)) < 1e-12


def test_p1_reproduces_linear(O):
    S = O.Space(8, O.FE_P1)
    u = 1 + 2 * S.coords[:, 0] - 3 * S.coords[:, 1]
    A, b = _penalized(O, S, [dict(kind=O.BLF_LAPLACE, fe=O.FE_P1, quadorder=0, factor=0.7, coords=S.coords,
                                  cells=S.cells, celldofs=S.celldofs)], [0.0], u, O.FE_P1)
    x = O.direct_solve(A, b)
    assert np.max(np.abs(x - u)) < 1e-12


def test_cell_action_equals_factor_for_constant_kappa(O):
    """Action path with kappa == c on every cell gives the NoAction matrix
    for factor c up to rounding."""
    S = O.Space(6, O.FE_P2)
    f1 = dict(kind=O.BLF_LAPLACE, fe=O.FE_P2, quadorder=2, factor=2.5, coords=S.coords, cells=S.cells,
              celldofs=S.celldofs)
    f2 = dict(f1, factor=1.0, coef=O.COEF_CELL, kappa=np.full(S.cells.shape[0], 2.5))
    A1 = O.assemble_matrix(S.ndofs, [f1]).to_scipy()
    A2 = O.assemble_matrix(S.ndofs, [f2]).to_scipy()
    assert abs(A1 - A2).max() <= 1e-15 * abs(A1).max()


def test_hooke_rigid_motions_and_linear_exactness(O):
    mu, lam = 1.0 / 2.6, 0.3 / (1.3 * 0.4)
    S = O.Space(6, O.FE_P1, ncomp=2)
    form = dict(kind=O.BLF_HOOKE2D, fe=O.FE_P1, ncomp=2, quadorder=0, factor=1.0, params=np.array([lam, mu]),
                coords=S.coords, cells=S.cells, celldofs=S.celldofs)
    K = O.assemble_matrix(S.ndofs, [form]).to_scipy()
    X, Y = S.coords[:, 0], S.coords[:, 1]
    for r in [np.concatenate([np.ones_like(X), 0 * X]), np.concatenate([0 * X, np.ones_like(X)]),
              np.concatenate([-Y, X])]:
        assert np.max(np.abs(K @ r)) < 1e-13
    # symmetric (full pattern type, computed twice)
    assert abs(K - K.T).max() < 1e-14
    # a linear displacement with zero body force is reproduced
    u = np.concatenate([0.1 + 0.2 * X - 0.3 * Y, -0.2 + 0.5 * X + 0.1 * Y])
    A, b = _penalized(O, S, [form], [0.0, 0.0], u, O.FE_P1, ncomp=2)
    x = O.direct_solve(A, b)
    assert np.max(np.abs(x - u)) < 1e-12


def test_pcg_matches_lu(O):
    S = O.Space(16, O.FE_P2)
    rng = np.random.default_rng(3)
    kappa = np.exp(rng.standard_normal(S.cells.shape[0]))
    A = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_LAPLACE, fe=O.FE_P2, quadorder=2, factor=1.0,
                                        coef=O.COEF_CELL, kappa=kappa, coords=S.coords, cells=S.cells,
                                        celldofs=S.celldofs)], fixed=S.fixed)
    b = np.zeros(S.ndofs)
    O.assemble_lf_const(b, O.FE_P2, 1, 2, [1.0], S.coords, S.cells, S.celldofs)
    b = O.penalized_rhs(b, S.fixed)
    x, it, rr = O.pcg(A, b, S.fixed, rtol=1e-13)
    xd = O.direct_solve(A, b)
    assert it > 0 and rr < 1e-13
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-11
