"""Python front end of the CPU oracle (test infrastructure only).

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this
module; it is the checker, never the product.  It wraps libgrmp_oracle.so, the
C restatement of GradientRobustMultiPhysics.jl's element loops (see the header
of grmp_oracle.c for the reference file:line each function follows), and adds
the host-side bookkeeping the reference does in Julia:

* ``boundary_dofs`` -- BFaceDofs of all boundary faces, the fixed-dof list of
  HomogeneousDirichletBoundary (src/boundarydata.jl:198-257,
  src/dofmaps.jl:201-363).
* ``assemble_system`` -- the system assemble + penalty of solve_direct!
  (src/solvers.jl:622-660; penalty src/solvers.jl:647-652 and
  src/fematrix.jl:349-355).
* ``backward_euler`` -- the BackwardEuler steps of advance! for a linear
  problem (src/solvers_timedependent.jl:337-359, :452-475), with sparse LU.
* ``crank_nicolson`` -- the CrankNicolson steps of the same loop
  (src/solvers_timedependent.jl:255, :376, :521).
* ``boundary_interpolate`` / ``boundary_bestapprox`` (and the ``_br``
  variants for H1BR{2}) -- boundarydata! (src/boundarydata.jl:198-352).
* ``l2error_cells`` -- evaluate! of L2ErrorIntegrator / L2NormIntegrator
  (src/assemblypatterns/itemintegrator.jl:40-95, :169-307).
* ``nodevalues`` -- nodevalues! with continuous = false
  (src/interpolations.jl:704-812).
* ``newton_assemble`` / ``newton_solve`` / ``assemble_lf_poly`` -- the
  Newton system of NonlinearForm (src/assemblypatterns/nonlinearform.jl:44-246)
  and the fixpoint loop of solve_fixpoint_full! (src/solvers.jl:812-1017),
  and polynomial LinearForms (src/assemblypatterns/linearform.jl:181-220).
* ``Anderson`` -- AndersonAccelerationManager / update_anderson!
  (src/solvers.jl:688-808).

What pins each function independently of the device code is in the tests
that exercise it (tests/test_oracle_*.py): exact-rational fixtures, closed
forms, invariants and brute force.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FE_P1, FE_P2, FE_BR, FE_P0 = 1, 2, 3, 4
BLF_LAPLACE, BLF_HOOKE2D, BLF_MASS = 1, 2, 3
COEF_FACTOR, COEF_CELL = 0, 1
PENALTY = 1e60

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_I64 = ctypes.c_int64
_VP = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libgrmp_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.or_unitsquare.argtypes = [ctypes.c_int, _d, _i32]
        L.or_tri_faces.argtypes = [_I64, _I64, _i32, _i32, _i32, _i32]
        L.or_tri_faces.restype = _I64
        L.or_celldofs.argtypes = [ctypes.c_int, ctypes.c_int, _I64, _I64, _I64, _i32, _i32, _i32]
        L.or_matrix_new.argtypes = [_I64]
        L.or_matrix_new.restype = _VP
        L.or_matrix_free.argtypes = [_VP]
        L.or_matrix_flush.argtypes = [_VP]
        L.or_matrix_flush.restype = _I64
        L.or_matrix_export.argtypes = [_VP, _i64, _i32, _d]
        L.or_matrix_penalize.argtypes = [_VP, _I64, _i32, ctypes.c_double]
        L.or_assemble_blf.argtypes = [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double, _d, _d, _I64,
                                      _d, _i32, _i32, ctypes.c_int32, ctypes.c_int32]
        L.or_assemble_lf_const.argtypes = [_d, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           _d, ctypes.c_double, _I64, _d, _i32, _i32, ctypes.c_int32]
        L.or_integrate_identity.argtypes = [_d, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            _d, _I64, _d, _i32, _i32]
        L.or_pcg.argtypes = [_I64, _i64, _i32, _d, _d, _d, _u8, ctypes.c_double, _I64,
                             ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.or_pcg.restype = _I64
        L.or_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


# ------------------------------------------------------------------ meshes

def unitsquare(N):
    """Structured N x N unit square, two triangles per square (0-based)."""
    nn, nc = (N + 1) ** 2, 2 * N * N
    coords = np.zeros((nn, 2))
    cells = np.zeros((nc, 3), dtype=np.int32)
    lib().or_unitsquare(N, coords, cells)
    return coords, cells


def tri_faces(cells, nnodes):
    nc = cells.shape[0]
    cellfaces = np.zeros((nc, 3), dtype=np.int32)
    facenodes = np.zeros((3 * nc, 2), dtype=np.int32)
    facecells = np.zeros((3 * nc, 2), dtype=np.int32)
    nf = lib().or_tri_faces(nc, nnodes, np.ascontiguousarray(cells, dtype=np.int32),
                            cellfaces, facenodes, facecells)
    return cellfaces, facenodes[:nf].copy(), facecells[:nf].copy()


def celldofs(fe, ncomp, cells, nnodes, cellfaces=None, nfaces=0):
    nloc = {FE_P1: 3, FE_P2: 6}[fe]
    nc = cells.shape[0]
    out = np.zeros((nc, nloc * ncomp), dtype=np.int32)
    cf = cellfaces if cellfaces is not None else np.zeros((nc, 3), dtype=np.int32)
    lib().or_celldofs(fe, ncomp, nc, nnodes, nfaces, np.ascontiguousarray(cells, dtype=np.int32),
                      np.ascontiguousarray(cf, dtype=np.int32), out)
    ndofs = (nnodes + (nfaces if fe == FE_P2 else 0)) * ncomp
    return out, ndofs


def boundary_dofs(fe, ncomp, nnodes, facenodes, facecells, nfaces):
    """Fixed dofs of a homogeneous Dirichlet condition on every boundary face."""
    bf = np.nonzero(facecells[:, 1] < 0)[0]
    nodes = np.unique(facenodes[bf].ravel())
    per = nnodes + (nfaces if fe == FE_P2 else 0)
    out = []
    for k in range(ncomp):
        out.append(nodes + k * per)
        if fe == FE_P2:
            out.append(nnodes + bf + k * per)
    return np.unique(np.concatenate(out)).astype(np.int32)


class Space:
    """Host FE space for the oracle: structured unit square + P1/P2 dofmaps."""

    def __init__(self, N, fe, ncomp=1):
        self.N, self.fe, self.ncomp = N, fe, ncomp
        self.coords, self.cells = unitsquare(N)
        self.nnodes = self.coords.shape[0]
        self.cellfaces, self.facenodes, self.facecells = tri_faces(self.cells, self.nnodes)
        self.nfaces = self.facenodes.shape[0]
        self.celldofs, self.ndofs = celldofs(fe, ncomp, self.cells, self.nnodes, self.cellfaces,
                                             self.nfaces)
        self.fixed = boundary_dofs(fe, ncomp, self.nnodes, self.facenodes, self.facecells,
                                   self.nfaces)


# ---------------------------------------------------------------- assembly

class CSC:
    def __init__(self, n, colptr, rowval, nzval):
        self.n, self.colptr, self.rowval, self.nzval = n, colptr, rowval, nzval

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.nzval, self.rowval, self.colptr), shape=(self.n, self.n))


def assemble_matrix(n, forms, fixed=None, penalty=PENALTY):
    """Assembles the BilinearForms in ``forms`` (sequentially, as the reference
    assembles the operators of one block) into a fresh matrix, applies the
    penalty on ``fixed`` and returns the flushed CSC (0-based)."""
    L = lib()
    h = L.or_matrix_new(n)
    try:
        for f in forms:
            geom = f.get("geom", 2)
            kappa = np.ascontiguousarray(f.get("kappa", np.zeros(1)), dtype=np.float64)
            params = np.ascontiguousarray(f.get("params", np.zeros(2)), dtype=np.float64)
            L.or_assemble_blf(h, f["kind"], geom, f["fe"], f.get("ncomp", 1), f["quadorder"],
                              f.get("coef", COEF_FACTOR), float(f.get("factor", 1.0)), kappa,
                              params, f["cells"].shape[0],
                              np.ascontiguousarray(f["coords"], dtype=np.float64),
                              np.ascontiguousarray(f["cells"], dtype=np.int32),
                              np.ascontiguousarray(f["celldofs"], dtype=np.int32),
                              f.get("ro", 0), f.get("co", 0))
        if fixed is not None and len(fixed):
            fx = np.ascontiguousarray(fixed, dtype=np.int32)
            L.or_matrix_penalize(h, fx.size, fx, penalty)
        nnz = L.or_matrix_flush(h)
        colptr = np.zeros(n + 1, dtype=np.int64)
        rowval = np.zeros(max(nnz, 1), dtype=np.int32)
        nzval = np.zeros(max(nnz, 1))
        L.or_matrix_export(h, colptr, rowval, nzval)
        return CSC(n, colptr, rowval[:nnz], nzval[:nnz])
    finally:
        L.or_matrix_free(h)


def assemble_lf_const(b, fe, ncomp, quadorder, f, coords, cells, celldofs, factor=1.0, geom=2,
                      offset=0):
    fv = np.ascontiguousarray(np.atleast_1d(np.asarray(f, dtype=np.float64)))
    lib().or_assemble_lf_const(b, geom, fe, ncomp, quadorder, fv, factor, cells.shape[0],
                               np.ascontiguousarray(coords, dtype=np.float64),
                               np.ascontiguousarray(cells, dtype=np.int32),
                               np.ascontiguousarray(celldofs, dtype=np.int32), offset)
    return b


def integrate_identity(u, fe, ncomp, quadorder, coords, cells, celldofs, geom=2):
    nc = cells.shape[0]
    out = np.zeros(nc * ncomp)
    lib().or_integrate_identity(out, geom, fe, ncomp, quadorder,
                                np.ascontiguousarray(u, dtype=np.float64), nc,
                                np.ascontiguousarray(coords, dtype=np.float64),
                                np.ascontiguousarray(cells, dtype=np.int32),
                                np.ascontiguousarray(celldofs, dtype=np.int32))
    return out.reshape(nc, ncomp)


def penalized_rhs(b, fixed, penalty=PENALTY, values=None):
    b = b.copy()
    b[fixed] = penalty * (0.0 if values is None else values)
    return b


def pcg(A, b, fixed, rtol=1e-10, maxit=100000, threads=0):
    """FP64 Jacobi-PCG with fixed dofs eliminated; returns (x, iters, relres)."""
    n = A.n
    mask = np.zeros(n, dtype=np.uint8)
    mask[fixed] = 1
    x = np.zeros(n)
    rr = ctypes.c_double(0.0)
    it = lib().or_pcg(n, A.colptr, A.rowval, A.nzval, np.ascontiguousarray(b, dtype=np.float64), x,
                      mask, rtol, maxit, threads, ctypes.byref(rr))
    return x, int(it), rr.value


def direct_solve(A, b):
    """FP64 sparse LU (SuperLU) of the penalized CSC: the direct-solve stand-in
    for UMFPACK (src/solvers.jl:55,656-660)."""
    import scipy.sparse.linalg as spla
    return spla.spsolve(A.to_scipy().tocsc(), b)



def direct_solve_fixed(A, b, fixed, values):
    """The penalized solve with the fixed dofs eliminated: x_d = values_d,
    A_ff x_f = b_f - A_fd values (sparse LU).  Equal to the penalized
    system's solution up to O(1/penalty) in exact arithmetic; used where the
    1e60 rows with inhomogeneous values leave SuperLU's pivoting without
    accuracy in the saddle point's pressure."""
    import scipy.sparse.linalg as spla
    M = A.to_scipy().tocsr()
    n = M.shape[0]
    fixed = np.asarray(fixed, dtype=np.int64)
    free = np.setdiff1d(np.arange(n), fixed)
    x = np.zeros(n)
    x[fixed] = values
    r = b[free] - M[free][:, fixed] @ x[fixed]
    x[free] = spla.spsolve(M[free][:, free].tocsc(), r)
    return x

# ------------------------------------------------------- Dirichlet data

def poly_eval(c, x, y):
    """sum_k c_k m_k(x, y) over 1, x, y, x^2, xy, y^2, x^3, x^2 y, x y^2, y^3."""
    mons = (1.0, x, y, x * x, x * y, y * y, x * x * x, x * x * y, x * y * y, y * y * y)
    return sum(ck * m for ck, m in zip(c, mons))


def boundary_faces(cells, cellfaces, facecells):
    """(cell, local face) of every boundary face, walked in cell order and
    local face order; a face is on the boundary when one cell holds it."""
    out = []
    for c in range(cells.shape[0]):
        for k in range(3):
            if facecells[cellfaces[c, k], 1] < 0:
                out.append((c, k))
    return out


def _face_dofs(fe, celldofs, c, k):
    d = [celldofs[c, k], celldofs[c, (k + 1) % 3]]
    if fe == FE_P2:
        d.append(celldofs[c, 3 + k])
    return d


def _trace_basis(fe, s):
    if fe == FE_P1:
        return np.array([1.0 - s, s])
    return np.array([(1.0 - s) * (1.0 - 2.0 * s), s * (2.0 * s - 1.0), 4.0 * s * (1.0 - s)])


def boundary_interpolate(fe, coef, coords, cells, celldofs, cellfaces, facecells, ndofs):
    """InterpolateDirichletBoundary (src/boundarydata.jl:198-257): the data at
    the points of the boundary faces' dofs; NaN off the boundary."""
    t = np.full(ndofs, np.nan)
    for c, k in boundary_faces(cells, cellfaces, facecells):
        a, b = cells[c, k], cells[c, (k + 1) % 3]
        pts = [coords[a], coords[b]] + ([0.5 * (coords[a] + coords[b])] if fe == FE_P2 else [])
        for d, p in zip(_face_dofs(fe, celldofs, c, k), pts):
            t[d] = poly_eval(coef, p[0], p[1])
    return t


def boundary_bestapprox(fe, coef, quadorder, coords, cells, celldofs, cellfaces, facecells, ndofs):
    """BestapproxDirichletBoundary (src/boundarydata.jl:259-352): mass matrix
    and load of the data on the boundary faces (Gauss-Legendre with
    ceil((quadorder+1)/2) points, at least 2 (P1) / 3 (P2)), compressed to
    the boundary dofs and solved by sparse LU (:322-344).  Returns
    (target with NaN off the boundary, compressed matrix, boundary dofs)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    nq = max((quadorder + 2) // 2, 3 if fe == FE_P2 else 2)
    xg, wg = np.polynomial.legendre.leggauss(nq)
    sq, wq = 0.5 * (xg + 1.0), 0.5 * wg
    faces = boundary_faces(cells, cellfaces, facecells)
    bd = sorted({d for c, k in faces for d in _face_dofs(fe, celldofs, c, k)})
    comp = {d: r for r, d in enumerate(bd)}
    rows, cols, vals = [], [], []
    load = np.zeros(len(bd))
    for c, k in faces:
        a, b = cells[c, k], cells[c, (k + 1) % 3]
        p0, p1 = coords[a], coords[b]
        L = float(np.hypot(*(p1 - p0)))
        dofs = _face_dofs(fe, celldofs, c, k)
        phi = np.array([_trace_basis(fe, s) for s in sq])            # [nq, nloc]
        Mloc = L * (phi.T * wq) @ phi
        gq = np.array([poly_eval(coef, *(p0 + s * (p1 - p0))) for s in sq])
        floc = L * (phi.T @ (wq * gq))
        for i, di in enumerate(dofs):
            load[comp[di]] += floc[i]
            for j, dj in enumerate(dofs):
                rows.append(comp[di])
                cols.append(comp[dj])
                vals.append(Mloc[i, j])
    Mb = sp.csc_matrix((vals, (rows, cols)), shape=(len(bd), len(bd)))
    t = np.full(ndofs, np.nan)
    t[bd] = spla.spsolve(Mb, load)
    return t, Mb, np.array(bd, dtype=np.int64)


def boundary_values_vec(kind, fe, ncomp, coef, quadorder, coords, cells, celldofs, cellfaces, facecells, ndofs):
    """Vector-valued data: each component's dofs (the component's block of
    the local dofs) get the scalar routine with that component's polynomial;
    the boundary systems of the components are independent."""
    per = 6 if fe == FE_P2 else 3
    coef = np.atleast_2d(coef)
    t = np.full(ndofs, np.nan)
    for m in range(ncomp):
        cd = np.ascontiguousarray(celldofs[:, m * per:(m + 1) * per])
        if kind == "bestapprox":
            tm = boundary_bestapprox(fe, coef[m], quadorder, coords, cells, cd, cellfaces, facecells, ndofs)[0]
        else:
            tm = boundary_interpolate(fe, coef[m], coords, cells, cd, cellfaces, facecells, ndofs)
        sel = ~np.isnan(tm)
        t[sel] = tm[sel]
    return t


# ------------------------------------------------------------- export

def nodevalues(fe, ncomp, u, cells, celldofs, nnodes):
    """nodevalues!(target, source, FE, Identity) with continuous = false
    (src/interpolations.jl:704-812; L2P0 src/fedefs/l2_p0.jl:52-68): per node
    the cells' values at that vertex, summed in cell order, over their number.
    Vertex values of P1/P2/BR are the vertex dof's coefficient, of P0 the
    cell's.  Returns [ncomp, nnodes]."""
    per = 6 if fe == FE_P2 else 3
    out = np.zeros((ncomp, nnodes))
    cnt = np.zeros(nnodes, dtype=np.int64)
    for c in range(cells.shape[0]):
        for k in range(3):
            n = cells[c, k]
            cnt[n] += 1
            for m in range(ncomp):
                loc = m if fe == FE_P0 else m * per + k
                out[m, n] += u[celldofs[c, loc]]
    return out / np.maximum(cnt, 1)


# ------------------------------------------------------- time stepping

def backward_euler_rhs(M, f, u, dt, fixed, penalty=PENALTY):
    """Right-hand side of one BackwardEuler step of advance!
    (src/solvers_timedependent.jl:349-359: b = f + (1/dt) AM u; penalty rows
    b_d = fixed_penalty * x_d at :452-461)."""
    b = f + (M.to_scipy().tocsr() @ u) * (1.0 / dt)
    b[fixed] = penalty * u[fixed]
    return b


def backward_euler_system(A, M, dt, fixed, penalty=PENALTY):
    """S = A + (1/dt) AM (add!, src/solvers_timedependent.jl:345-346) with the
    fixed diagonals overwritten by the penalty (:457-458); A unpenalized."""
    S = (A.to_scipy() + M.to_scipy() * (1.0 / dt)).tolil()
    for d in np.asarray(fixed).tolist():
        S[d, d] = penalty
    return S.tocsc()


def backward_euler(A, M, f, u0, dts, fixed, penalty=PENALTY):
    """advance! with BackwardEuler for a linear problem, one step per entry of
    ``dts``; S is refactored only when the step size changes (:337).
    Returns the list of states after each step (sparse LU solves)."""
    import scipy.sparse.linalg as spla
    u, out, lu, last = np.array(u0, dtype=np.float64), [], None, None
    for dt in dts:
        if lu is None or dt != last:
            lu = spla.splu(backward_euler_system(A, M, dt, fixed, penalty))
            last = dt
        u = lu.solve(backward_euler_rhs(M, f, u, dt, fixed, penalty))
        out.append(u.copy())
    return out


# ------------------------------------------------------------------ Stokes

_i8 = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")


def _stokes_lib():
    L = lib()
    if not hasattr(L, "_stokes_ready"):
        L.or_face_geometry.argtypes = [_I64, _I64, _d, _i32, _i32, _i32, _d, _d, _i8]
        L.or_br_celldofs.argtypes = [_I64, _I64, _i32, _i32, _i32]
        L.or_assemble_stokes.argtypes = [_VP, ctypes.c_double, _I64, _I64, _I64, _d, _i32, _i32, _d]
        L.or_quad_stroud.argtypes = [ctypes.c_int, _d, _d]
        L.or_quad_stroud.restype = ctypes.c_int
        L.or_assemble_stokes_rhs.argtypes = [_d, ctypes.c_int, ctypes.c_int, _d, _d, _I64, _I64, _d, _i32, _i32, _d,
                                             _d, _i8]
        L._stokes_ready = True
    return L


def quad_stroud(order):
    xr = np.zeros(2 * 64)
    w = np.zeros(64)
    nq = _stokes_lib().or_quad_stroud(order, xr, w)
    return xr[:2 * nq].reshape(nq, 2), w[:nq]


class StokesSpace:
    """H1BR{2} x L2P0 on the structured unit square (or on the triangle mesh
    ``coords``, ``cells`` when given): dofs, face geometry, no-slip fixed
    dofs and the pinned first pressure dof (FixedIntegralMean)."""

    def __init__(self, N, coords=None, cells=None):
        L = _stokes_lib()
        self.N = N
        if coords is None:
            self.coords, self.cells = unitsquare(N)
        else:
            self.coords = np.ascontiguousarray(coords, dtype=np.float64)
            self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.nnodes = self.coords.shape[0]
        self.nc = self.cells.shape[0]
        self.cellfaces, self.facenodes, self.facecells = tri_faces(self.cells, self.nnodes)
        self.nf = self.facenodes.shape[0]
        self.fnormal = np.zeros(2 * self.nf)
        self.fvol = np.zeros(self.nf)
        self.csign = np.zeros(3 * self.nc, dtype=np.int8)
        L.or_face_geometry(self.nc, self.nf, self.coords, self.cells, np.ascontiguousarray(self.cellfaces),
                           np.ascontiguousarray(self.facenodes), self.fnormal, self.fvol, self.csign)
        self.celldofs = np.zeros((self.nc, 9), dtype=np.int32)
        L.or_br_celldofs(self.nc, self.nnodes, self.cells, np.ascontiguousarray(self.cellfaces), self.celldofs)
        self.nU = 2 * self.nnodes + self.nf
        self.nP = self.nc
        self.n = self.nU + self.nP
        bf = np.nonzero(self.facecells[:, 1] < 0)[0]
        bn = np.unique(self.facenodes[bf].ravel())
        self.fixed = np.unique(np.concatenate([bn, bn + self.nnodes, 2 * self.nnodes + bf, [self.nU]])).astype(np.int32)

    def matrix(self, nu):
        L = _stokes_lib()
        h = L.or_matrix_new(self.n)
        try:
            L.or_assemble_stokes(h, nu, self.nc, self.nnodes, self.nf, self.coords, self.cells,
                                 np.ascontiguousarray(self.cellfaces), self.fnormal)
            L.or_matrix_penalize(h, self.fixed.size, self.fixed, PENALTY)
            nnz = L.or_matrix_flush(h)
            colptr = np.zeros(self.n + 1, dtype=np.int64)
            rowval = np.zeros(max(nnz, 1), dtype=np.int32)
            nzval = np.zeros(max(nnz, 1))
            L.or_matrix_export(h, colptr, rowval, nzval)
            return CSC(self.n, colptr, rowval[:nnz], nzval[:nnz])
        finally:
            L.or_matrix_free(h)

    def rhs(self, order, modes, poly):
        K = len(modes) // 2
        b = np.zeros(self.n)
        _stokes_lib().or_assemble_stokes_rhs(b, order, K, np.ascontiguousarray(np.asarray(modes, dtype=np.float64)
                                                                                  .reshape(-1)),
                                             np.ascontiguousarray(np.asarray(poly, dtype=np.float64)), self.nc,
                                             self.nnodes, self.coords, self.cells,
                                             np.ascontiguousarray(self.cellfaces), self.fnormal, self.fvol,
                                             self.csign)
        return b

    def pressure_mean_shift(self, x):
        """realize_constraint!(FixedIntegralMean(2, 0)) on P0: p -= mean(p)."""
        P = self.coords[self.cells]
        vol = 0.5 * np.abs((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1])
                           - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1]))
        p = x[self.nU:]
        mean = float(np.sum(p * vol) / np.sum(vol))
        y = x.copy()
        y[self.nU:] -= mean
        return y


# ------------------------------------------------------ L2ErrorIntegrator

def tri_rule(order):
    """The reference's triangle rule (src/quadrature.jl:173-195): midpoint,
    face midpoints, else the Stroud rule; points (nq, 2), weights (sum 1)."""
    if order <= 1:
        return np.array([[1.0 / 3.0, 1.0 / 3.0]]), np.array([1.0])
    if order == 2:
        return np.array([[0.5, 0.5], [0.0, 0.5], [0.5, 0.0]]), np.full(3, 1.0 / 3.0)
    return quad_stroud(order)


def _shape_vals(fe, x0, x1):
    """scalar shape functions at one reference point, op order of the
    reference's get_basis (P1 h1_p1.jl:64-75, P2 h1_p2.jl:134-145, BR
    h1v_br.jl:117-130: P1 then 6 x l, 6 y x, 6 l y)"""
    l = 1.0
    l -= x0
    l -= x1
    if fe == FE_P1:
        return [l, x0, x1]
    if fe == FE_P2:
        l2 = 1.0 - x0 - x1
        return [2.0 * l2 * (l2 - 0.5), 2.0 * x0 * (x0 - 0.5), 2.0 * x1 * (x1 - 0.5), 4.0 * l2 * x0, 4.0 * x0 * x1,
                4.0 * x1 * l2]
    return [l, x0, x1, 6.0 * x0 * l, 6.0 * x1 * x0, 6.0 * l * x1]


def l2error_cells(fe, ncomp, u, coef, quadorder, factor, coords, cells, celldofs, normals=None):
    """evaluate!(b, L2ErrorIntegrator(g, Identity; quadorder, factor), u)
    (src/assemblypatterns/itemintegrator.jl:40-89, :243-307) per cell:
    sum_q (sum_k (g_k(x_q) - factor u_k(x_q))^2) w_q vol, with
    u_k(x_q) = sum_dofs u_dof cvals[k, dof, q] in dof order
    (src/feevaluator.jl:445-452) and x_q = x_1 + A xref_q.  coef: (ncomp,
    ncoef) polynomial coefficients (None: zero data).  normals: (ncells, 3,
    2) face normals of H1BR{2}."""
    nc = cells.shape[0]
    X = coords[cells]
    X0, Y0 = X[:, 0, 0], X[:, 0, 1]
    A11, A12 = X[:, 1, 0] - X0, X[:, 2, 0] - X0
    A21, A22 = X[:, 1, 1] - Y0, X[:, 2, 1] - Y0
    vol = np.abs(A11 * A22 - A12 * A21) * 0.5
    xq, wq = tri_rule(quadorder)
    c = np.zeros((ncomp, 10)) if coef is None else np.atleast_2d(np.asarray(coef, dtype=np.float64))
    acc = np.zeros(nc)
    for (x0, x1), w in zip(xq, wq):
        sv = _shape_vals(fe, x0, x1)
        inp = [np.zeros(nc) for _ in range(ncomp)]
        if fe == FE_P0:
            inp[0] += u[celldofs[:, 0]] * 1.0
        elif fe == FE_BR:
            for dof in range(9):
                ud = u[celldofs[:, dof]]
                for k in range(2):
                    if dof < 6:
                        cv = sv[dof % 3] * 1.0 if dof // 3 == k else 0.0
                    else:
                        cv = sv[dof - 3] * normals[:, dof - 6, k]
                    inp[k] += ud * cv * 1.0
        else:
            per = 6 if fe == FE_P2 else 3
            for cc in range(ncomp):
                for s in range(per):
                    ud = u[celldofs[:, cc * per + s]]
                    for k in range(ncomp):
                        inp[k] += ud * (sv[s] if cc == k else 0.0) * 1.0
        xp = np.zeros(nc)
        xp += A11 * x0
        xp += A12 * x1
        xp += X0
        yp = np.zeros(nc)
        yp += A21 * x0
        yp += A22 * x1
        yp += Y0
        r = np.zeros(nc)
        for k in range(ncomp):
            val = _poly_ordered(c[k], xp, yp)
            val -= inp[k] * factor
            r += val * val
        acc += r * w * vol
    return acc


def _poly_ordered(c, x, y):
    """g(x, y) summed term by term in monomial order (as the device does)"""
    c = list(c) + [0.0] * (10 - len(c))
    s = np.zeros_like(x)
    s += c[0]
    s += c[1] * x
    s += c[2] * y
    s += c[3] * (x * x)
    s += c[4] * (x * y)
    s += c[5] * (y * y)
    s += c[6] * (x * x * x)
    s += c[7] * (x * x * y)
    s += c[8] * (x * y * y)
    s += c[9] * (y * y * y)
    return s


# ------------------------------------------ H1BR{2} Dirichlet data

def _br_face_units(S):
    """boundary faces of a StokesSpace in cell order: (cell, local face,
    5 dofs [a_x, b_x, a_y, b_y, bubble], face normal)"""
    out = []
    nrm = S.fnormal.reshape(-1, 2)
    for c, k in boundary_faces(S.cells, S.cellfaces, S.facecells):
        d = S.celldofs[c]
        dofs = [d[k], d[(k + 1) % 3], d[3 + k], d[3 + (k + 1) % 3], d[6 + k]]
        out.append((c, k, dofs, nrm[S.cellfaces[c, k]]))
    return out


def _br_trace(s, n):
    """vector values of the five face dofs of H1BR{2} at s
    (src/fedefs/h1v_br.jl:104-113: P1{2} on the edge, then 6 s (1-s) n)"""
    b = 6.0 * s * (1.0 - s)
    return np.array([[1.0 - s, 0.0], [s, 0.0], [0.0, 1.0 - s], [0.0, s], [b * n[0], b * n[1]]])


def boundary_bestapprox_br(S, coef, quadorder):
    """BestapproxDirichletBoundary on H1BR{2} (src/boundarydata.jl:259-352):
    the boundary mass matrix of the vector-valued trace basis and the load
    of g = (g_x, g_y) on the boundary faces, solved by sparse LU.  Returns
    (target with NaN off the boundary, boundary dofs)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    coef = np.atleast_2d(coef)
    nq = max((quadorder + 2) // 2, 3)
    xg, wg = np.polynomial.legendre.leggauss(nq)
    sq, wq = 0.5 * (xg + 1.0), 0.5 * wg
    units = _br_face_units(S)
    bd = sorted({d for _, _, dofs, _ in units for d in dofs})
    comp = {d: r for r, d in enumerate(bd)}
    rows, cols, vals = [], [], []
    load = np.zeros(len(bd))
    for c, k, dofs, n in units:
        a, b = S.cells[c, k], S.cells[c, (k + 1) % 3]
        p0, p1 = S.coords[a], S.coords[b]
        L = float(np.hypot(*(p1 - p0)))
        for s, w in zip(sq, wq):
            V = _br_trace(s, n)
            x, y = p0 + s * (p1 - p0)
            g = np.array([poly_eval(coef[0], x, y), poly_eval(coef[1], x, y)])
            for i, di in enumerate(dofs):
                load[comp[di]] += L * w * float(V[i] @ g)
                for j, dj in enumerate(dofs):
                    rows.append(comp[di])
                    cols.append(comp[dj])
                    vals.append(L * w * float(V[i] @ V[j]))
    Mb = sp.csc_matrix((vals, (rows, cols)), shape=(len(bd), len(bd)))
    t = np.full(S.n, np.nan)
    t[bd] = spla.spsolve(Mb, load)
    return t, np.array(bd, dtype=np.int64)


def boundary_interpolate_br(S, coef):
    """InterpolateDirichletBoundary on H1BR{2} (src/fedefs/h1v_br.jl:37-81):
    node values of g, and the bubble that keeps the normal flux of g through
    each boundary face: sum_c (int_F g_c - int_F I g_c) n_c / |F|."""
    coef = np.atleast_2d(coef)
    xg, wg = np.polynomial.legendre.leggauss(6)
    sq, wq = 0.5 * (xg + 1.0), 0.5 * wg
    t = np.full(S.n, np.nan)
    for c, k, dofs, n in _br_face_units(S):
        a, b = S.cells[c, k], S.cells[c, (k + 1) % 3]
        p0, p1 = S.coords[a], S.coords[b]
        L = float(np.hypot(*(p1 - p0)))
        g0 = [poly_eval(coef[m], *p0) for m in range(2)]
        g1 = [poly_eval(coef[m], *p1) for m in range(2)]
        t[dofs[0]], t[dofs[1]], t[dofs[2]], t[dofs[3]] = g0[0], g1[0], g0[1], g1[1]
        val = 0.0
        for m in range(2):
            mean = L * sum(w * poly_eval(coef[m], *(p0 + s * (p1 - p0))) for s, w in zip(sq, wq))
            val += (mean - 0.5 * L * (g0[m] + g1[m])) * n[m]
        t[dofs[4]] = val / L
    return t


def crank_nicolson_rhs(A, M, f, u, dt, fixed, penalty=PENALTY):
    """Right-hand side of one CrankNicolson step of advance! for
    time-independent data (src/solvers_timedependent.jl:376-433):
    rhs = b^n - A u^n + b^{n+1} + (2/dt) AM u^n with b^n = b^{n+1} = f;
    penalty rows rhs_d = fixed_penalty * x_d (:452-461)."""
    Ms, As = M.to_scipy().tocsr(), A.to_scipy().tocsr()
    b = f - As @ u + f + (Ms @ u) * (2.0 / dt)
    b[fixed] = penalty * u[fixed]
    return b


def crank_nicolson(A, M, f, u0, dts, fixed, penalty=PENALTY):
    """advance! with CrankNicolson for a linear problem without algebraic
    constraints: S = A + (2/dt) AM (:408-413), refactored when the step
    size changes; returns the states after each step (sparse LU)."""
    import scipy.sparse.linalg as spla
    u, out, lu, last = np.array(u0, dtype=np.float64), [], None, None
    for dt in dts:
        if lu is None or dt != last:
            S = (A.to_scipy() + M.to_scipy() * (2.0 / dt)).tolil()
            for d in np.asarray(fixed).tolist():
                S[d, d] = penalty
            lu = spla.splu(S.tocsc())
            last = dt
        u = lu.solve(crank_nicolson_rhs(A, M, f, u, dt, fixed, penalty))
        out.append(u.copy())
    return out


# ------------------------------------------------------- NonlinearForm (Newton)

def _tri_shape(fe, x0, x1):
    """values and reference derivatives of the P1/P2 basis at one point"""
    l = 1.0 - x0 - x1
    if fe == FE_P1:
        return np.array([l, x0, x1]), np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])
    v = np.array([2 * l * (l - 0.5), 2 * x0 * (x0 - 0.5), 2 * x1 * (x1 - 0.5), 4 * l * x0, 4 * x0 * x1, 4 * x1 * l])
    d = np.array([[-(4 * l - 1), -(4 * l - 1)], [4 * x0 - 1, 0.0], [0.0, 4 * x1 - 1],
                  [4 * (l - x0), -4 * x0], [4 * x1, 4 * x0], [-4 * x1, 4 * (l - x1)]])
    return v, d


def _tri_geometry(coords, cells):
    X = coords[cells]
    A11, A12 = X[:, 1, 0] - X[:, 0, 0], X[:, 2, 0] - X[:, 0, 0]
    A21, A22 = X[:, 1, 1] - X[:, 0, 1], X[:, 2, 1] - X[:, 0, 1]
    det = A11 * A22 - A12 * A21
    M = np.empty((cells.shape[0], 2, 2))
    M[:, 0, 0], M[:, 0, 1], M[:, 1, 0], M[:, 1, 1] = A22 / det, -A21 / det, -A12 / det, A11 / det
    return M, np.abs(det) * 0.5, X[:, 0], np.stack([[A11, A12], [A21, A22]]).transpose(2, 0, 1)


def newton_assemble(fe, coords, cells, celldofs, ndofs, u, c, quadorder, factor=1.0):
    """full_assemble! of NonlinearForm(Gradient, [Identity, Gradient], [1,1],
    kernel, [2,3]) with kernel = alpha(u) grad u, alpha = c0 + c1 u + c2 u^2
    (src/assemblypatterns/nonlinearform.jl:44-246, transposed_assembly as
    src/pdeoperators.jl:630 sets it): returns (A, b) with
    A[test a, ansatz b] = sum_T vol sum_q w (J [phi_b, grad phi_b]) . grad phi_a,
    b[a] = sum_T vol sum_q w ((J [u, grad u]) - kernel) . grad phi_a; A keeps
    every local pair (also exact zeros)."""
    import scipy.sparse as sp
    nd = 3 if fe == FE_P1 else 6
    xq, wq = tri_rule(quadorder)
    M, vol, _, _ = _tri_geometry(coords, cells)
    uc = u[celldofs[:, :nd]]
    nc = cells.shape[0]
    L = np.zeros((nc, nd, nd))      # [cell, test a, ansatz b]
    lb = np.zeros((nc, nd))
    for (x0, x1), w in zip(xq, wq):
        phi, dref = _tri_shape(fe, x0, x1)
        g = np.einsum("ckj,dj->cdk", M, dref)          # [cell, dof, k] gradients
        uu = uc @ phi
        gu = np.einsum("cd,cdk->ck", uc, g)
        alpha = c[0] + c[1] * uu + c[2] * uu * uu
        dalpha = c[1] + 2.0 * c[2] * uu
        # J [phi_b, grad phi_b] = dalpha grad u phi_b + alpha grad phi_b
        Jb = dalpha[:, None, None] * gu[:, None, :] * phi[None, :, None] + alpha[:, None, None] * g
        L += w * np.einsum("cbk,cak->cab", Jb, g)
        Jin = dalpha[:, None] * gu * uu[:, None] + alpha[:, None] * gu
        val = alpha[:, None] * gu
        lb += w * np.einsum("ck,cak->ca", Jin - val, g)
    itf = vol * factor
    rows = np.repeat(celldofs[:, :nd], nd, axis=1).ravel()
    cols = np.tile(celldofs[:, :nd], (1, nd)).ravel()
    A = sp.csr_matrix(((L * itf[:, None, None]).ravel(), (rows, cols)), shape=(ndofs, ndofs))
    b = np.zeros(ndofs)
    np.add.at(b, celldofs[:, :nd].ravel(), (lb * itf[:, None]).ravel())
    return A, b


def assemble_lf_poly(fe, coef, quadorder, coords, cells, celldofs, ndofs, factor=1.0):
    """LinearForm(Identity, f) with polynomial f (degree <= 4, monomials
    1, x, y, x^2, xy, y^2, x^3, x^2 y, x y^2, y^3, x^4, x^3 y, x^2 y^2, x y^3,
    y^4) on scalar P1/P2 (src/assemblypatterns/linearform.jl:181-220)."""
    mons = ((0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3),
            (4, 0), (3, 1), (2, 2), (1, 3), (0, 4))
    nd = 3 if fe == FE_P1 else 6
    xq, wq = tri_rule(quadorder)
    M, vol, X0, Amat = _tri_geometry(coords, cells)
    lb = np.zeros((cells.shape[0], nd))
    for (x0, x1), w in zip(xq, wq):
        phi, _ = _tri_shape(fe, x0, x1)
        xp = X0[:, 0] + Amat[:, 0, 0] * x0 + Amat[:, 0, 1] * x1
        yp = X0[:, 1] + Amat[:, 1, 0] * x0 + Amat[:, 1, 1] * x1
        f = sum(ck * xp ** i * yp ** j for ck, (i, j) in zip(coef, mons))
        lb += w * f[:, None] * phi[None, :]
    b = np.zeros(ndofs)
    np.add.at(b, celldofs[:, :nd].ravel(), (lb * (factor * vol)[:, None]).ravel())
    return b


class Anderson:
    """AndersonAccelerationManager + update_anderson! (src/solvers.jl:688-808),
    restated with the reference's array shifts: LastIteratesTilde and
    LastIterates of depth + 1 vectors each (all the initial iterate at
    construction), the (depth + 2) system Matrix (identity on the first
    depth + 1 diagonal entries) and Rhs = e_{depth+2}, the depth counter
    starting at -1.  B is the norm matrix (None: the identity of "l2"), mask
    the accelerated entries.  update(x) changes x in place and returns
    (alpha[:depth + 1], local residual)."""

    def __init__(self, x0, depth, damping=1.0, B=None, mask=None):
        self.m, self.damping, self.B = int(depth), float(damping), B
        self.mask = np.ones(len(x0), dtype=bool) if mask is None else np.asarray(mask, dtype=bool)
        self.last = [np.array(x0, dtype=np.float64) for _ in range(self.m + 1)]
        self.tilde = [np.array(x0, dtype=np.float64) for _ in range(self.m + 1)]
        K = self.m + 2
        self.M = np.zeros((K, K))
        self.rhs = np.zeros(K)
        self.rhs[K - 1] = 1.0
        for j in range(self.m + 1):
            self.M[j, j] = 1.0
        self.cdepth = -1

    def _inner(self, a1, a2, b1, b2):
        """ldrdmatmul (src/fematrix.jl:524-535): (a1 - a2)' B (b1 - b2)"""
        da = np.where(self.mask, a1 - a2, 0.0)
        db = np.where(self.mask, b1 - b2, 0.0)
        return float(da @ (db if self.B is None else self.B @ db))

    def update(self, x):
        m = self.m
        cd = min(m, self.cdepth)
        for j in range(m):
            self.tilde[j][:] = self.tilde[j + 1]
        self.tilde[m][:] = x
        alpha, res = np.zeros(m + 1), 0.0
        if cd > 0:
            K = m + 2
            for j in range(cd + 1):
                for k in range(j, cd + 1):
                    g = self._inner(self.tilde[m - j], self.last[m - j], self.tilde[m - k], self.last[m - k])
                    self.M[j, k] = self.M[k, j] = g
                    self.M[j, K - 1] = self.M[K - 1, j] = 1.0
            a = np.linalg.solve(self.M, self.rhs)
            res = float(np.sqrt(np.sum((self.M @ a - self.rhs) ** 2)))
            out = np.zeros(len(x))
            for i in range(cd + 1):
                out += self.damping * a[i] * self.tilde[m - i]
                out += (1.0 - self.damping) * a[i] * self.last[m - i]
            x[self.mask] = out[self.mask]
            for j in range(m):
                self.last[j][:] = self.last[j + 1]
            alpha = a[:m + 1].copy()
        self.last[m][:] = x
        self.cdepth += 1
        return alpha, res


def newton_solve(fe, coords, cells, celldofs, ndofs, fixed, target, c, f_coef, quadorder, lf_quadorder,
                 maxit=10, target_residual=1e-12, u0=None, anderson=None):
    """solve_fixpoint_full! (src/solvers.jl:812-1017) for the nonlinear
    Poisson problem: assemble at the initial iterate, set the boundary
    values, then per iteration: penalized direct solve (fixed dofs
    eliminated), Anderson update when anderson = dict(depth, damping, B)
    (src/solvers.jl:849-859, :923-929), reassemble, residual ||A x - b|| on
    the free rows.  Returns (x, residuals)."""
    import scipy.sparse.linalg as spla
    u = np.zeros(ndofs) if u0 is None else np.array(u0, dtype=np.float64)
    f = assemble_lf_poly(fe, f_coef, lf_quadorder, coords, cells, celldofs, ndofs)
    A, bn = newton_assemble(fe, coords, cells, celldofs, ndofs, u, c, quadorder)
    b = f + bn
    u[fixed] = target[fixed]
    free = np.setdiff1d(np.arange(ndofs), fixed)
    aam = None
    if anderson is not None:
        aam = Anderson(u, anderson["depth"], anderson.get("damping", 1.0), anderson.get("B"))
    res = []
    for _ in range(maxit):
        rhs = b[free] - A[free][:, fixed] @ u[fixed]
        u[free] = spla.spsolve(A[free][:, free].tocsc(), rhs)
        if aam is not None:
            aam.update(u)
        A, bn = newton_assemble(fe, coords, cells, celldofs, ndofs, u, c, quadorder)
        b = f + bn
        r = A @ u - b
        r[fixed] = 0.0
        res.append(float(np.linalg.norm(r)))
        if res[-1] < target_residual:
            break
    return u, res
