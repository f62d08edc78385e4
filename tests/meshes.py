"""Triangle meshes for the mesh and face-adjacency tests."""
import numpy as np


def shuffled_square(O, N, seed, flip=True):
    """Unit-square triangulation with jittered interior nodes, permuted nodes
    and cells, cyclically shifted local node order and (optionally) some
    cells with reversed orientation."""
    rng = np.random.default_rng(seed)
    coords, cells = O.unitsquare(N)
    nn, nc = coords.shape[0], cells.shape[0]
    inner = np.all((coords > 0) & (coords < 1), axis=1)
    coords = coords.copy()
    coords[inner] += rng.uniform(-0.2, 0.2, size=(int(inner.sum()), 2)) / N
    perm = rng.permutation(nn)       # old node -> new node
    newc = np.empty_like(coords)
    newc[perm] = coords
    cells = perm[cells][rng.permutation(nc)]
    shift = rng.integers(0, 3, size=nc)
    cells = np.stack([np.roll(cells[c], shift[c]) for c in range(nc)])
    if flip:
        rev = rng.random(nc) < 0.3
        cells[rev] = cells[rev][:, [0, 2, 1]]
    return newc, np.ascontiguousarray(cells, dtype=np.int32)
