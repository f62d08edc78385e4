"""Pins of the oracle's Anderson acceleration (oracle.Anderson, a
restatement of AndersonAccelerationManager / update_anderson!,
src/solvers.jl:688-808) that do not come from the device:

* a history formulation: the k-th update pairs the fixpoint outputs t_k,
  t_(k-1), ... with the stored iterates x_(k-1), x_(k-2), ..., where the
  stored sequence is the initial iterate up to index 1 (the reference's
  stored ring does not shift during the first two updates, so x_1 is
  overwritten by x_2) and the coefficients minimize ||sum_j alpha_j d_j||_B
  subject to sum_j alpha_j = 1, solved here by least squares over the
  differences d_j - d_1 in the Cholesky factor of B (not the bordered
  system the reference solves);
* the closed form of depth 1 in the l2 metric,
  alpha_1 = (d_2.d_2 - d_1.d_2) / |d_1 - d_2|^2;
* the fixpoint loop with acceleration still reproduces u = x y with P2.
"""
import numpy as np
import pytest
import scipy.sparse as sp


def _history(O, x0, targets, depth, damping, B, mask):
    """the k-th output of the accelerated sequence, from the history lists"""
    t_hist, x_hist = [], []
    init = np.array(x0, dtype=np.float64)
    L = None if B is None else np.linalg.cholesky(B.toarray()).T  # B = L' L
    out = []
    for k, t in enumerate(targets, start=1):
        t = np.array(t, dtype=np.float64)
        t_hist.append(t)

        def tseq(i):
            return init if i <= 0 else t_hist[i - 1]

        def sseq(i):
            return init if i <= 1 else x_hist[i - 1]

        cd = min(depth, k - 2)
        x = t.copy()
        if cd > 0:
            d = [np.where(mask, tseq(k + 1 - j) - sseq(k - j), 0.0) for j in range(1, cd + 2)]
            W = [di if L is None else L @ di for di in d]
            E = np.stack([W[j] - W[0] for j in range(1, cd + 1)], axis=1)
            beta = np.linalg.lstsq(E, -W[0], rcond=None)[0]
            alpha = np.concatenate([[1.0 - beta.sum()], beta])
            comb = sum(alpha[j] * (damping * tseq(k - j) + (1 - damping) * sseq(k - 1 - j)) for j in range(cd + 1))
            x[mask] = comb[mask]
        x_hist.append(x)
        out.append(x)
    return out


@pytest.mark.parametrize("depth,damping,metric", [(1, 1.0, "l2"), (2, 1.0, "l2"), (3, 0.6, "B"), (2, 0.8, "B")])
def test_anderson_matches_history_formulation(O, depth, damping, metric):
    rng = np.random.default_rng(depth * 10 + int(damping * 10))
    n = 40
    mask = np.ones(n, dtype=bool)
    mask[25:] = metric == "l2"  # a partly accelerated vector with the metric
    B = None
    if metric == "B":
        R = sp.random(n, n, density=0.15, random_state=5)
        B = (R @ R.T + sp.identity(n) * 2.0).tocsr()
    x0 = rng.standard_normal(n)
    # a contraction's iterates plus noise, so that the differences are independent
    G = rng.standard_normal((n, n)) * 0.3 / np.sqrt(n)
    c = rng.standard_normal(n)
    aam = O.Anderson(x0, depth, damping, B, mask)
    x = x0.copy()
    targets, outs = [], []
    for _ in range(8):
        t = G @ x + c + 1e-3 * rng.standard_normal(n)
        targets.append(t.copy())
        x = t.copy()
        aam.update(x)
        outs.append(x.copy())
    ref = _history(O, x0, targets, depth, damping, B, mask)
    for k, (a, b) in enumerate(zip(outs, ref), start=1):
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b)), k
    # updates one and two leave the fixpoint output as it is
    assert np.array_equal(outs[0], targets[0]) and np.array_equal(outs[1], targets[1])


def test_anderson_depth1_closed_form(O):
    rng = np.random.default_rng(7)
    n = 30
    x0 = rng.standard_normal(n)
    aam = O.Anderson(x0, 1)
    ts = [rng.standard_normal(n) for _ in range(3)]
    xs = []
    for t in ts:
        x = t.copy()
        alpha, res = aam.update(x)
        xs.append(x)
    # third update: d_1 = t_3 - x_2, d_2 = t_2 - x0 (x_2 = t_2)
    d1, d2 = ts[2] - ts[1], ts[1] - x0
    a1 = (d2 @ d2 - d1 @ d2) / ((d1 - d2) @ (d1 - d2))
    assert abs(alpha[0] - a1) <= 1e-12 and abs(alpha[1] - (1 - a1)) <= 1e-12 and res < 1e-12
    assert np.allclose(xs[2], a1 * ts[2] + (1 - a1) * ts[1], rtol=0, atol=1e-12)


def test_newton_with_anderson_reproduces_xy(O):
    S = O.Space(6, O.FE_P2)
    P = np.vstack([S.coords, 0.5 * (S.coords[S.facenodes[:, 0]] + S.coords[S.facenodes[:, 1]])])
    F = [0.0] * 15
    F[11], F[13] = -2.0, -2.0
    u_ex = P[:, 0] * P[:, 1]
    for aa in (None, dict(depth=2, damping=1.0), dict(depth=1, damping=0.5)):
        x, res = O.newton_solve(O.FE_P2, S.coords, S.cells, S.celldofs, S.ndofs, S.fixed, u_ex, (1.0, 0.0, 1.0), F,
                                quadorder=6, lf_quadorder=6, maxit=30, target_residual=1e-11, anderson=aa)
        assert res[-1] < 1e-11 and np.max(np.abs(x - u_ex)) < 1e-10, (aa, res)
