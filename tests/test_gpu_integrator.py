"""L2ErrorIntegrator / L2NormIntegrator on the device
(src/assemblypatterns/itemintegrator.jl:40-128, evaluate! :169-307) against
the oracle.

Bar: per-cell values bitwise equal to the oracle's for the closed-form rules
(quadrature order <= 2); within 1e-13 relative for Stroud rules (the device
and the oracle compute the Gauss-Jacobi points each by their own eigen
solver); the total (a fixed-order device reduction; the reference sums
sequentially) within 1e-13 relative of the oracle's cell sum; data in the
discrete space gives a squared error below 1e-26."""
import numpy as np
import pytest

from meshes import shuffled_square

pytestmark = pytest.mark.gpu

CUB = [[0.3, -0.2, 0.5, 0.1, 0.7, -0.4, 1.1, -0.6, 0.9, 0.25],
       [0.1, 0.4, -0.3, 0.2, -0.5, 0.6, -0.7, 0.3, 0.2, -0.1]]


def _fe(grmp, name):
    return {"p1": (grmp.H1P1(1), 1, 1), "p1v": (grmp.H1P1(2), 1, 2), "p2": (grmp.H1P2(1), 2, 1),
            "p2v": (grmp.H1P2(2), 2, 2), "br": (grmp.H1BR(2), 3, 2), "p0": (grmp.L2P0(1), 4, 1)}[name]


@pytest.mark.parametrize("name", ["p1", "p1v", "p2", "p2v", "br", "p0"])
@pytest.mark.parametrize("bonus", [0, 1, 3])
def test_l2error_matches_oracle(grmp, O, name, bonus):
    print(f"[case] l2error {name} bonus={bonus}", flush=True)
    coords, cells = shuffled_square(O, 9, 4)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    ft, fe, ncomp = _fe(grmp, name)
    fes = grmp.FESpace(ft, grid)
    rng = np.random.default_rng(1)
    u = rng.standard_normal(fes.ndofs)
    X = grmp.FEVector([fes])
    X.set_entries(u)
    data = grmp.PolynomialData(np.array(CUB[:ncomp]))
    I = grmp.L2ErrorIntegrator(data, quadorder=bonus, factor=0.5)
    got = grmp.evaluate_cells(I, X[0])
    tot = grmp.evaluate(I, X[0])
    cd = fes.celldofs()
    nrm = None
    if fe == 3:
        S = O.StokesSpace(9, coords, cells)
        assert np.array_equal(S.celldofs, cd)
        nrm = S.fnormal.reshape(-1, 2)[S.cellfaces]
    q = ft.order + bonus
    ref = O.l2error_cells(fe, ncomp, u, CUB[:ncomp], q, 0.5, coords, cells, cd, normals=nrm)
    if q <= 2:
        assert np.array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref) / ref) <= 1e-13
    assert abs(tot - ref.sum()) <= 1e-13 * ref.sum()


def test_l2norm_and_zero_error(grmp, O):
    print("[case] l2norm / zero error", flush=True)
    grid = grmp.grid_unitsquare(16)
    fes = grmp.FESpace(grmp.H1P2(1), grid)
    X = grmp.FEVector([fes])
    coords, cells = grid.download()
    cf, fn, fc = O.tri_faces(cells, coords.shape[0])
    P = np.vstack([coords, 0.5 * (coords[fn[:, 0]] + coords[fn[:, 1]])])
    q2 = [1.0, 1.0, -2.0, 0.5, 1.0, 1.5]
    X.set_entries(O.poly_eval(q2, P[:, 0], P[:, 1]))
    assert grmp.evaluate(grmp.L2ErrorIntegrator(grmp.PolynomialData(q2)), X[0]) < 1e-26
    # ||1||^2 = area
    X.set_entries(np.ones(fes.ndofs))
    assert abs(grmp.evaluate(grmp.L2NormIntegrator(1), X[0]) - 1.0) < 1e-14
