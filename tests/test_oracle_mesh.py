"""The oracle's face enumeration on arbitrary triangle meshes, pinned by a
brute-force restatement and by Euler's formula.

First appearance: walking cells in order and local faces (0,1), (1,2), (2,0)
in order, a face gets the next id the first time its node pair is met and
keeps that appearance's orientation."""
import numpy as np

from meshes import shuffled_square


def _brute_faces(cells):
    ids, nodes, owners = {}, [], []
    cf = np.zeros(cells.shape, dtype=np.int32)
    for c, tri in enumerate(cells.tolist()):
        for f in range(3):
            a, b = tri[f], tri[(f + 1) % 3]
            key = (min(a, b), max(a, b))
            if key not in ids:
                ids[key] = len(nodes)
                nodes.append((a, b))
                owners.append([c, -1])
            else:
                owners[ids[key]][1] = c
            cf[c, f] = ids[key]
    return cf, np.array(nodes, dtype=np.int32), np.array(owners, dtype=np.int32)


def test_faces_match_brute_force(O):
    for N, seed in [(1, 0), (4, 1), (9, 2)]:
        coords, cells = shuffled_square(O, N, seed)
        cf, fn, fc = O.tri_faces(cells, coords.shape[0])
        bcf, bfn, bfc = _brute_faces(cells)
        assert np.array_equal(cf, bcf)
        assert np.array_equal(fn, bfn)
        assert np.array_equal(fc, bfc)
        # Euler's formula for a triangulated disk: V - E + F = 1
        assert coords.shape[0] - fn.shape[0] + cells.shape[0] == 1
        # 4N boundary faces, each interior face in exactly two cells
        assert int(np.sum(fc[:, 1] < 0)) == 4 * N
        assert set(np.bincount(cf.ravel()).tolist()) <= {1, 2}


def test_face_signs_and_geometry(O):
    """CellFaceSigns is +1 exactly where the cell's local face runs the way
    the face's first appearance does; the face normal is the unit tangent
    turned clockwise, so it points out of the first cell when that cell is
    counter-clockwise."""
    coords, cells = shuffled_square(O, 6, 3)
    S = O.StokesSpace(6, coords, cells)
    sign = S.csign.reshape(-1, 3)
    for c in range(S.nc):
        for f in range(3):
            a, b = cells[c, f], cells[c, (f + 1) % 3]
            fa, fb = S.facenodes[S.cellfaces[c, f]]
            assert sign[c, f] == (1 if (a, b) == (fa, fb) else -1)
    fnorm = S.fnormal.reshape(-1, 2)
    assert np.allclose(np.linalg.norm(fnorm, axis=1), 1.0, atol=1e-14)
    t = coords[S.facenodes[:, 1]] - coords[S.facenodes[:, 0]]
    assert np.allclose(S.fvol, np.linalg.norm(t, axis=1), rtol=1e-14)
    assert np.allclose(np.sum(fnorm * t, axis=1), 0.0, atol=1e-14)
    # a counter-clockwise first cell has the normal pointing away from its third node
    for fidx in range(0, S.nf, 7):
        c = S.facecells[fidx, 0]
        P = coords[cells[c]]
        det = (P[1, 0] - P[0, 0]) * (P[2, 1] - P[0, 1]) - (P[2, 0] - P[0, 0]) * (P[1, 1] - P[0, 1])
        third = [v for v in cells[c] if v not in S.facenodes[fidx]][0]
        out = np.dot(fnorm[fidx], coords[S.facenodes[fidx, 0]] - coords[third])
        assert (out > 0) == (det > 0)
