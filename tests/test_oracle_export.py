"""Pins of the oracle's node values (nodevalues!, src/interpolations.jl:704-812,
src/fedefs/l2_p0.jl:52-68), the data of writeCSV! (src/dataexport.jl:22-77).

* Continuous P1 / P2 data: the node value is the vertex dof's value, up to
  the rounding of summing it once per cell and dividing by the count.
* P0 data: the mean over the cells around the node equals a sparse
  node-cell incidence product divided by the node degree.
* Node degrees on the unit square: 1 or 2 at corners, 3 on edges, 6 inside."""
import numpy as np
import scipy.sparse as sp


def test_continuous_values(O):
    rng = np.random.default_rng(1)
    for fe in (O.FE_P1, O.FE_P2):
        S = O.Space(6, fe)
        u = rng.standard_normal(S.ndofs)
        nv = O.nodevalues(fe, 1, u, S.cells, S.celldofs, S.nnodes)
        assert np.allclose(nv[0], u[:S.nnodes], rtol=4e-16, atol=0)


def test_p0_mean_and_degrees(O):
    coords, cells = O.unitsquare(5)
    nn, nc = coords.shape[0], cells.shape[0]
    rng = np.random.default_rng(2)
    u = rng.standard_normal(nc)
    cd = np.arange(nc, dtype=np.int32).reshape(nc, 1)
    nv = O.nodevalues(O.FE_P0, 1, u, cells, cd, nn)
    inc = sp.csr_matrix((np.ones(3 * nc), (cells.ravel(), np.repeat(np.arange(nc), 3))), shape=(nn, nc))
    deg = np.asarray(inc.sum(axis=1)).ravel()
    assert np.allclose(nv[0], (inc @ u) / deg, rtol=1e-14, atol=1e-15)
    corners = [0, 5, 30, 35]
    assert sorted(deg[corners].tolist()) == [1, 1, 2, 2]
    inner = [j * 6 + i for j in range(1, 5) for i in range(1, 5)]
    assert np.all(deg[inner] == 6)
