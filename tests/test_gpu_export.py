"""Node values and CSV export on the device (nodevalues!, writeCSV!,
src/interpolations.jl:704-812, src/fedefs/l2_p0.jl:52-68,
src/dataexport.jl:22-77) against the oracle.

Bar: node values bitwise equal to the oracle's (same cell-order sums) for
P1, P1{2}, P2, BR{2} and P0 on a shuffled mesh; the CSV file equals the
reference layout filled with the oracle's values."""
import numpy as np
import pytest

from meshes import shuffled_square

pytestmark = pytest.mark.gpu


def test_nodevalues_match_oracle(grmp, O):
    N, seed = 7, 2
    print(f"[case] nodevalues N={N} seed={seed}", flush=True)
    coords, cells = shuffled_square(O, N, seed)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    rng = np.random.default_rng(5)
    for name, ft, fe in (("P1", grmp.H1P1(1), O.FE_P1), ("P1v", grmp.H1P1(2), O.FE_P1),
                         ("P2", grmp.H1P2(1), O.FE_P2), ("BR", grmp.H1BR(2), O.FE_BR),
                         ("P0", grmp.L2P0(1), O.FE_P0)):
        fes = grmp.FESpace(ft, grid)
        other = grmp.FESpace(grmp.H1P1(1), grid)
        X = grmp.FEVector([other, fes])  # second block: a nonzero offset
        u = rng.standard_normal(X.n)
        X.set_entries(u)
        nv = grmp.nodevalues(X[1])
        ref = O.nodevalues(fe, ft.ncomp, u[X.offsets[1]:], cells, fes.celldofs(), coords.shape[0])
        assert nv.shape == ref.shape, name
        assert np.array_equal(nv, ref), name


def test_write_csv(grmp, O, tmp_path):
    N, seed = 5, 3
    print(f"[case] writeCSV N={N} seed={seed}", flush=True)
    coords, cells = shuffled_square(O, N, seed)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    V = grmp.FESpace(grmp.H1BR(2), grid)
    Q = grmp.FESpace(grmp.L2P0(1), grid)
    X = grmp.FEVector([V, Q])
    u = np.random.default_rng(7).standard_normal(X.n)
    X.set_entries(u)
    path = tmp_path / "sol.csv"
    grmp.writeCSV(str(path), [X[0], X[1]], names=["velocity", "pressure"])
    nn = coords.shape[0]
    vv = O.nodevalues(O.FE_BR, 2, u, cells, V.celldofs(), nn)
    pv = O.nodevalues(O.FE_P0, 1, u[X.offsets[1]:], cells, Q.celldofs(), nn)
    lines = ["X1\tX2\tvelocity \tpressure \t"]
    for n in range(nn):
        row = [coords[n, 0], coords[n, 1], vv[0, n], vv[1, n], pv[0, n]]
        lines.append("".join("%.6f\t" % v for v in row))
    assert path.read_text() == "\n".join(lines) + "\n"
