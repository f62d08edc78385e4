"""Pins of the oracle's BackwardEuler steps (advance!,
src/solvers_timedependent.jl:293-594) on the P1 heat equation
(u_t - div(kappa grad u) = f, homogeneous Dirichlet data).

* The P1 mass matrix integrates 1 x 1 to the domain's area.
* A generalized eigenvector v of (A, M) on the free dofs decays by exactly
  1/(1 + lambda dt) per step, also across a change of the step size: this
  fails for a wrong sign of A, a missing or misplaced 1/dt, the identity in
  place of M, or a system not refactored when dt changes.
* The stationary solution of A u = f is a fixed point of every step: this
  fails for a dropped or rescaled f.
* Fixed dofs keep their values (penalty rows)."""
import numpy as np
import scipy.linalg as sla


def _heat(O, N, kappa=1.0):
    S = O.Space(N, O.FE_P1)
    base = dict(fe=O.FE_P1, coords=S.coords, cells=S.cells, celldofs=S.celldofs)
    A = O.assemble_matrix(S.ndofs, [dict(base, kind=O.BLF_LAPLACE, quadorder=0, factor=kappa)])
    M = O.assemble_matrix(S.ndofs, [dict(base, kind=O.BLF_MASS, quadorder=2, factor=1.0)])
    free = np.setdiff1d(np.arange(S.ndofs), S.fixed)
    return S, A, M, free


def test_mass_matrix_integrates_area(O):
    S, A, M, free = _heat(O, 7)
    assert abs(M.to_scipy().sum() - 1.0) < 1e-13
    # Laplace rows of the unpenalized matrix sum to zero (constants in the kernel)
    assert np.max(np.abs(np.asarray(A.to_scipy().sum(axis=1)).ravel())) < 1e-12


def test_eigenvector_decay(O):
    S, A, M, free = _heat(O, 8, kappa=0.7)
    Af = A.to_scipy().toarray()[np.ix_(free, free)]
    Mf = M.to_scipy().toarray()[np.ix_(free, free)]
    lam, V = sla.eigh(Af, Mf)
    for k in (0, 5, 20):
        u0 = np.zeros(S.ndofs)
        u0[free] = V[:, k]
        dts = [0.05, 0.05, 0.01]
        out = O.backward_euler(A, M, np.zeros(S.ndofs), u0, dts, S.fixed)
        g = 1.0
        for dt, u in zip(dts, out):
            g /= 1.0 + lam[k] * dt
            assert np.max(np.abs(u[free] - g * V[:, k])) <= 1e-10 * np.max(np.abs(V[:, k])), (k, dt)
            assert np.max(np.abs(u[S.fixed])) <= 1e-40


def test_stationary_fixed_point(O):
    S, A, M, free = _heat(O, 9)
    f = np.zeros(S.ndofs)
    O.assemble_lf_const(f, O.FE_P1, 1, 1, [1.0], S.coords, S.cells, S.celldofs)
    us = np.zeros(S.ndofs)
    us[free] = np.linalg.solve(A.to_scipy().toarray()[np.ix_(free, free)], f[free])
    for dt in (1e-3, 0.1, 10.0):
        u1 = O.backward_euler(A, M, f, us, [dt], S.fixed)[0]
        assert np.max(np.abs(u1 - us)) <= 1e-12 * np.max(np.abs(us)), dt


def test_fixed_values_kept(O):
    S, A, M, free = _heat(O, 5)
    rng = np.random.default_rng(3)
    u0 = rng.standard_normal(S.ndofs)
    b = O.backward_euler_rhs(M, np.zeros(S.ndofs), u0, 0.1, S.fixed)
    assert np.array_equal(b[S.fixed], O.PENALTY * u0[S.fixed])
    u1 = O.backward_euler(A, M, np.zeros(S.ndofs), u0, [0.1], S.fixed)[0]
    assert np.allclose(u1[S.fixed], u0[S.fixed], rtol=1e-12, atol=0)


def test_crank_nicolson_eigenvector_decay(O):
    """CrankNicolson: a generalized eigenvector decays by exactly
    (1 - lambda dt / 2) / (1 + lambda dt / 2) per step, across a step size
    change (fails for a BackwardEuler factor, a missing A u^n or a wrong
    2/dt)."""
    S, A, M, free = _heat(O, 8, kappa=0.7)
    Af = A.to_scipy().toarray()[np.ix_(free, free)]
    Mf = M.to_scipy().toarray()[np.ix_(free, free)]
    lam, V = sla.eigh(Af, Mf)
    for k in (0, 5, 20):
        u0 = np.zeros(S.ndofs)
        u0[free] = V[:, k]
        dts = [0.05, 0.05, 0.01]
        out = O.crank_nicolson(A, M, np.zeros(S.ndofs), u0, dts, S.fixed)
        g = 1.0
        for dt, u in zip(dts, out):
            g *= (1.0 - 0.5 * lam[k] * dt) / (1.0 + 0.5 * lam[k] * dt)
            assert np.max(np.abs(u[free] - g * V[:, k])) <= 1e-10 * np.max(np.abs(V[:, k])), (k, dt)


def test_crank_nicolson_stationary_and_fixed(O):
    S, A, M, free = _heat(O, 9)
    f = np.zeros(S.ndofs)
    O.assemble_lf_const(f, O.FE_P1, 1, 1, [1.0], S.coords, S.cells, S.celldofs)
    us = np.zeros(S.ndofs)
    us[free] = np.linalg.solve(A.to_scipy().toarray()[np.ix_(free, free)], f[free])
    for dt in (1e-3, 0.1, 10.0):
        u1 = O.crank_nicolson(A, M, f, us, [dt], S.fixed)[0]
        assert np.max(np.abs(u1 - us)) <= 1e-12 * np.max(np.abs(us)), dt
    u0 = np.random.default_rng(4).standard_normal(S.ndofs)
    b = O.crank_nicolson_rhs(A, M, f, u0, 0.1, S.fixed)
    assert np.array_equal(b[S.fixed], O.PENALTY * u0[S.fixed])
