"""Face adjacency built on the device for uploaded triangle meshes
(ExtendableGrids' CellFaces, CellFaceSigns, boundary faces) and the P2 / BR
spaces, boundary dofs and forms that use it, against the oracle.

Bar: face numbering, signs, boundary flags, dofmaps, boundary dofs and
sparsity patterns identical; matrix values and load vectors as in the
assembly tests (P2 Laplace bitwise; Stokes blocks within 1e-12 of the row's
largest entry; reconstructed right-hand side within 1e-12 of max|b|);
MINRES within 1e-8 of the FP64 sparse-LU solution."""
import numpy as np
import pytest

from meshes import shuffled_square

pytestmark = pytest.mark.gpu

NU = 1e-2
MODES = np.array([[0.3, 1, 0, 0.1], [-0.2, 1, 2, 1.3], [0.5, 1, 1, 0.4], [0.1, 2, -1, 2.0]])
POLY = [0.4, -0.3, 0.2]


def _spaces(grmp, grid):
    return {"p1": grmp.FESpace(grmp.H1P1(1), grid), "p1v": grmp.FESpace(grmp.H1P1(2), grid),
            "p2": grmp.FESpace(grmp.H1P2(1), grid), "p2v": grmp.FESpace(grmp.H1P2(2), grid),
            "br": grmp.FESpace(grmp.H1BR(2), grid)}


@pytest.mark.parametrize("N", [1, 3, 16])
def test_uploaded_square_equals_generated(grmp, O, N):
    """The sort-based enumeration on the uploaded unit square reproduces the
    generated mesh's closed-form face numbering, signs and boundary."""
    print(f"[case] uploaded unit square N={N}", flush=True)
    gen = grmp.grid_unitsquare(N)
    c, cn = gen.download()
    up = grmp.ExtendableGrid.from_arrays(c, cn)
    fg, fu = gen.faces(), up.faces()
    for a, b in zip(fg, fu):
        assert np.array_equal(a, b)
    S = O.StokesSpace(N)
    assert np.array_equal(fu[0], S.cellfaces)
    assert np.array_equal(fu[1], S.csign.reshape(-1, 3))
    assert np.array_equal(fu[2], S.facecells[:, 1] < 0)
    sg, su = _spaces(grmp, gen), _spaces(grmp, up)
    for k in sg:
        assert sg[k].ndofs == su[k].ndofs, k
        assert np.array_equal(sg[k].celldofs(), su[k].celldofs()), k
        assert np.array_equal(sg[k].boundary_dofs(), su[k].boundary_dofs()), k


def _stokes_on(grmp, grid):
    V = grmp.FESpace(grmp.H1BR(2), grid)
    Q = grmp.FESpace(grmp.L2P0(1), grid)
    pde = grmp.IncompressibleNavierStokesProblem(2, viscosity=NU)
    K = len(MODES) // 2
    f = grmp.GradientCurlData(MODES[:K], MODES[K:], POLY, bonus_quadorder=2)
    pde.add_rhsdata(0, grmp.LinearForm(grmp.ReconstructionIdentity(grmp.HDIVRT0), f))
    return V, Q, pde, grmp.FEVector([V, Q])


@pytest.mark.parametrize("N,seed", [(5, 1), (12, 2)])
def test_shuffled_mesh_faces_and_p2(grmp, O, N, seed):
    print(f"[case] shuffled mesh faces + P2 N={N} seed={seed}", flush=True)
    coords, cells = shuffled_square(O, N, seed)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    cf, cs, bf = grid.faces()
    ocf, ofn, ofc = O.tri_faces(cells, coords.shape[0])
    S = O.StokesSpace(N, coords, cells)
    assert np.array_equal(cf, ocf)
    assert np.array_equal(cs, S.csign.reshape(-1, 3))
    assert np.array_equal(bf, ofc[:, 1] < 0)
    nf = ofn.shape[0]
    # P2 with a cell coefficient: dofmap, boundary dofs, matrix bitwise
    fes = grmp.FESpace(grmp.H1P2(1), grid)
    cd, nd = O.celldofs(O.FE_P2, 1, cells, coords.shape[0], ocf, nf)
    fixed = O.boundary_dofs(O.FE_P2, 1, coords.shape[0], ofn, ofc, nf)
    assert fes.ndofs == nd
    assert np.array_equal(fes.celldofs(), cd)
    assert np.array_equal(fes.boundary_dofs(), fixed)
    kap = grmp.DeviceVector(grid.ctx, grid.ncells)
    kap.lognormal(seed, 0.0, 1.0)
    op = grmp.BilinearForm([grmp.Gradient, grmp.Gradient], grmp.CellCoefficient(kap))
    A = grmp.FEMatrix([fes])
    form, keep = op.form(fes, 0, 0, grid.ctx)
    A.add_form(form, keep)
    A.set_fixed(fixed)
    A.build_pattern()
    A.assemble()
    ref = O.assemble_matrix(nd, [dict(kind=O.BLF_LAPLACE, fe=O.FE_P2, quadorder=op.quadorder(fes), factor=1.0,
                                      coef=O.COEF_CELL, kappa=kap.download(), coords=coords, cells=cells,
                                      celldofs=cd)], fixed=fixed)
    colptr, rowval, nzval = A.csc()
    assert np.array_equal(colptr, ref.colptr) and np.array_equal(rowval, ref.rowval)
    assert np.array_equal(nzval, ref.nzval)
    # P2 load vector, bitwise
    b = grmp.FEVector([fes])
    grmp.assemble_linearform(b, 0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
    bref = np.zeros(nd)
    O.assemble_lf_const(bref, O.FE_P2, 1, 2, [1.0], coords, cells, cd)
    assert np.array_equal(b.entries(), bref)


def test_shuffled_mesh_stokes(grmp, O):
    """Bernardi-Raugel Stokes on an uploaded, shuffled, partly clockwise mesh:
    matrix, reconstructed right-hand side and MINRES solution."""
    N, seed = 10, 4
    print(f"[case] shuffled mesh stokes N={N} seed={seed}", flush=True)
    coords, cells = shuffled_square(O, N, seed)
    grid = grmp.ExtendableGrid.from_arrays(coords, cells)
    V, Q, pde, sol = _stokes_on(grmp, grid)
    S = O.StokesSpace(N, coords, cells)
    assert V.ndofs == S.nU and Q.ndofs == S.nP
    assert np.array_equal(V.celldofs(), S.celldofs)
    A, b, fixed = grmp.assemble(pde, sol)
    assert np.array_equal(np.unique(fixed), S.fixed)
    R = S.matrix(NU)
    colptr, rowval, nzval = A.csc()
    assert np.array_equal(colptr, R.colptr) and np.array_equal(rowval, R.rowval)
    Ar = R.to_scipy().tocsr()
    scale = np.maximum(np.abs(Ar).max(axis=1).toarray().ravel(), 1e-300)
    diff = np.abs(A.to_scipy().tocsr() - Ar).tocoo()
    rel = diff.data / scale[diff.row] if diff.nnz else np.zeros(1)
    print(f"  matrix max rel {rel.max():.2e}")
    assert rel.max() <= 1e-12
    bl = grmp.FEVector([V, Q])
    grmp.assemble_linearform(bl, 0, pde.RHS[0][0])
    rhs = S.rhs(4, MODES, POLY)
    err = np.max(np.abs(bl.entries() - rhs)) / np.max(np.abs(rhs))
    print(f"  rhs max rel {err:.2e}")
    assert err <= 1e-12
    info, A, b = grmp.solve(sol, pde, rtol=1e-11)
    assert info.status == 0
    xref = S.pressure_mean_shift(O.direct_solve(R, O.penalized_rhs(rhs, S.fixed)))
    x = sol.entries()
    err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    print(f"  MINRES iterations {info.iterations} rel err vs LU {err:.2e}")
    assert err <= 1e-8


def test_nonmanifold_mesh_rejected(grmp):
    """Three triangles on one edge: the upload reports the mesh as invalid."""
    coords = np.array([[0.0, 0.0], [1.0, 0.0], [0.5, 1.0], [0.5, -1.0], [0.5, 0.5]])
    cells = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], dtype=np.int32)
    print("[case] non-manifold upload", flush=True)
    with pytest.raises(grmp.GrmpError):
        grmp.ExtendableGrid.from_arrays(coords, cells)
