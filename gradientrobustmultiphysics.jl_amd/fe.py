"""Grids, finite element spaces, FEVector and FEMatrix on the device.

Python mirror of the reference's L0/L1 containers for the assembly path:
ExtendableGrid (coordinates + CellNodes), FESpace (src/finiteelements.jl:42-112)
with its CellDofs (src/dofmaps.jl:201-363), FEVector (src/fevector.jl:13-150)
and FEMatrix (src/fematrix.jl:14-213).  Every object owns a handle of the
native library; no arithmetic happens in Python.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import VP, default_context, dptr, i32ptr, i64ptr, lib


# ----------------------------------------------------------------- FE types

class FEType:
    """H1P1{ncomp}, H1P2{ncomp,2}: the element families on the hot path."""

    def __init__(self, code, ncomp, order, name):
        self.code, self.ncomp, self.order, self.name = code, ncomp, order, name

    def __repr__(self):
        return self.name


def H1P1(ncomp=1):
    return FEType(_abi.FE_H1P1, ncomp, 1, f"H1P1{{{ncomp}}}")


def H1P2(ncomp=1, edim=2):
    return FEType(_abi.FE_H1P2, ncomp, 2, f"H1P2{{{ncomp},{edim}}}")


def H1BR(ncomp=2):
    """Bernardi-Raugel H1BR{2}: P1 vector field plus one normal face bubble per
    face (src/fedefs/h1v_br.jl); polynomial order 2 on triangles."""
    if ncomp != 2:
        raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "H1BR is built for ncomp = 2")
    return FEType(_abi.FE_H1BR, 2, 2, "H1BR{2}")


def L2P0(ncomp=1):
    """Piecewise constants L2P0{1} (src/fedefs/l2v_p0.jl)."""
    if ncomp != 1:
        raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "L2P0 is built for ncomp = 1")
    return FEType(_abi.FE_L2P0, 1, 0, "L2P0{1}")


# -------------------------------------------------------------------- grids

class ExtendableGrid:
    """Device mesh.  Host arrays follow Julia's layout transposed to numpy:
    coordinates (nnodes, dim), cellnodes (ncells, nloc), both 0-based here;
    the C ABI receives them 1-based in Julia's column-major order."""

    def __init__(self, handle, ctx, kind="generic", N=0):
        self.h, self.ctx, self.kind, self.N = handle, ctx, kind, N
        dim, nn, nc = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        ctx.check(lib().grmp_mesh_info(self.h, ctypes.byref(dim), ctypes.byref(nn), ctypes.byref(nc)))
        self.dim, self.nnodes, self.ncells = dim.value, nn.value, nc.value

    @classmethod
    def from_arrays(cls, coords, cellnodes, ctx=None):
        ctx = ctx or default_context()
        c = np.ascontiguousarray(coords, dtype=np.float64)
        cn = np.ascontiguousarray(np.asarray(cellnodes) + 1, dtype=np.int32)
        h = VP()
        ctx.check(lib().grmp_mesh_upload(ctx.h, c.shape[1], c.shape[0], dptr(c), cn.shape[0], cn.shape[1],
                                         i32ptr(cn), ctypes.byref(h)))
        return cls(h, ctx)

    def download(self):
        """(coordinates (nnodes, dim), cellnodes (ncells, nloc) 0-based)"""
        nloc = self.dim + 1
        c = np.empty((self.nnodes, self.dim))
        cn = np.empty((self.ncells, nloc), dtype=np.int32)
        self.ctx.check(lib().grmp_mesh_download(self.h, dptr(c), i32ptr(cn)))
        return c, cn - 1

    def cellvolumes(self):
        out = np.empty(self.ncells)
        self.ctx.check(lib().grmp_mesh_cellvolumes(self.h, dptr(out)))
        return out

    def faces(self):
        """Triangle meshes: (CellFaces (ncells, 3) 0-based, CellFaceSigns
        (ncells, 3), boundary flag per face), numbered by first appearance."""
        nf = ctypes.c_int64()
        self.ctx.check(lib().grmp_mesh_faces(self.h, ctypes.byref(nf), None, None, None))
        cf = np.empty((self.ncells, 3), dtype=np.int32)
        cs = np.empty((self.ncells, 3), dtype=np.int8)
        bf = np.empty(nf.value, dtype=np.uint8)
        self.ctx.check(lib().grmp_mesh_faces(self.h, ctypes.byref(nf), i32ptr(cf), cs.ctypes.data, bf.ctypes.data))
        return cf - 1, cs, bf.astype(bool)

    def __del__(self):
        try:
            if self.h and _abi.alive():
                lib().grmp_mesh_destroy(self.h)
            self.h = None
        except Exception:
            pass


def grid_unitsquare(N, ctx=None):
    """N x N squares of the unit square, each split along its (0,0)-(1,1)
    diagonal (generated on the device)."""
    ctx = ctx or default_context()
    h = VP()
    ctx.check(lib().grmp_mesh_unitsquare(ctx.h, int(N), ctypes.byref(h)))
    return ExtendableGrid(h, ctx, "unitsquare", N)


def grid_unitcube(n, jitter=0.0, seed=0, ctx=None):
    """n^3 cubes of the unit cube, 6 Kuhn tetrahedra each, interior nodes
    jittered by +-jitter*h (generated on the device)."""
    ctx = ctx or default_context()
    h = VP()
    ctx.check(lib().grmp_mesh_unitcube(ctx.h, int(n), float(jitter), int(seed), ctypes.byref(h)))
    return ExtendableGrid(h, ctx, "unitcube", n)


# ------------------------------------------------------------------ FESpace

class FESpace:
    """FESpace{FEType}(xgrid).  The dofmap is generated on the device, or
    taken from ``celldofs`` ((ncells, ndofs_loc), 0-based) when given --
    the path a Julia caller uses with FES[CellDofs]."""

    def __init__(self, fetype, grid, celldofs=None, ndofs=None, name=""):
        self.FEType, self.xgrid, self.ctx, self.name = fetype, grid, grid.ctx, name
        h = VP()
        if celldofs is None:
            self.ctx.check(lib().grmp_space_create(self.ctx.h, grid.h, fetype.code, fetype.ncomp, ctypes.byref(h)))
        else:
            cd = np.ascontiguousarray(np.asarray(celldofs) + 1, dtype=np.int32)
            nd = int(ndofs if ndofs is not None else cd.max())
            self.ctx.check(lib().grmp_space_upload(self.ctx.h, grid.h, fetype.code, fetype.ncomp, nd, cd.shape[1],
                                                   i32ptr(cd), ctypes.byref(h)))
        self.h = h
        nd, nl = ctypes.c_int64(), ctypes.c_int()
        self.ctx.check(lib().grmp_space_info(self.h, ctypes.byref(nd), ctypes.byref(nl)))
        self.ndofs, self.ndofs_loc = nd.value, nl.value

    def celldofs(self):
        out = np.empty((self.xgrid.ncells, self.ndofs_loc), dtype=np.int32)
        self.ctx.check(lib().grmp_space_download_celldofs(self.h, i32ptr(out)))
        return out - 1

    def boundary_dofs(self):
        """0-based dofs on all boundary faces (BFaceDofs of regions 1:4 / 1:6)."""
        cnt = ctypes.c_int64()
        self.ctx.check(lib().grmp_space_boundary_dofs(self.h, ctypes.byref(cnt), None))
        out = np.empty(cnt.value, dtype=np.int64)
        self.ctx.check(lib().grmp_space_boundary_dofs(self.h, ctypes.byref(cnt), i64ptr(out)))
        return out - 1

    def __del__(self):
        try:
            if self.h and _abi.alive():
                lib().grmp_space_destroy(self.h)
            self.h = None
        except Exception:
            pass


# ------------------------------------------------------------------ FEVector

class FEVectorBlock:
    def __init__(self, parent, FES, offset, name=""):
        self.parent, self.FES, self.offset, self.name = parent, FES, offset, name

    def __len__(self):
        return self.FES.ndofs

    def values(self):
        return self.parent.entries()[self.offset:self.offset + self.FES.ndofs]


class FEVector:
    """One device FP64 array with one block per FESpace (src/fevector.jl:125-150)."""

    def __init__(self, spaces, name="", ctx=None):
        spaces = spaces if isinstance(spaces, (list, tuple)) else [spaces]
        self.ctx = ctx or spaces[0].ctx
        self.FES = list(spaces)
        self.offsets = np.concatenate([[0], np.cumsum([s.ndofs for s in spaces])]).astype(np.int64)
        self.n = int(self.offsets[-1])
        self.vec = _abi.DeviceVector(self.ctx, self.n)
        self.name = name
        self.blocks = [FEVectorBlock(self, s, int(o), getattr(s, "name", "") or f"{name or 'u'}[{k + 1}]")
                       for k, (s, o) in enumerate(zip(spaces, self.offsets[:-1]))]

    def __getitem__(self, i):
        return self.blocks[i]

    def entries(self):
        return self.vec.download()

    def set_entries(self, values):
        self.vec.upload(values)


# ------------------------------------------------------------------ FEMatrix

class FEMatrix:
    """Square block matrix over FESpaces backed by one device CSR
    (src/fematrix.jl:186-213)."""

    def __init__(self, spaces, name=""):
        spaces = spaces if isinstance(spaces, (list, tuple)) else [spaces]
        self.FES, self.ctx = list(spaces), spaces[0].ctx
        arr = (VP * len(spaces))(*[s.h for s in spaces])
        h = VP()
        self.ctx.check(lib().grmp_matrix_create(self.ctx.h, len(spaces), arr, ctypes.byref(h)))
        self.h = h
        self.offsets = np.concatenate([[0], np.cumsum([s.ndofs for s in spaces])]).astype(np.int64)
        self.n = int(self.offsets[-1])
        self._keep = []  # coefficient vectors referenced by registered forms

    def add_form(self, form, keep=None):
        if keep is not None:
            self._keep.append(keep)
        self.ctx.check(lib().grmp_matrix_add_form(self.h, ctypes.byref(form)))

    def set_fixed(self, fixed0, penalty=1e60):
        f = np.ascontiguousarray(np.asarray(fixed0, dtype=np.int64) + 1)
        self.ctx.check(lib().grmp_matrix_set_fixed(self.h, f.size, i64ptr(f), float(penalty)))

    def build_pattern(self):
        self.ctx.check(lib().grmp_matrix_build_pattern(self.h))

    def assemble(self, values=_abi.VALUES_FP64):
        self.ctx.check(lib().grmp_matrix_assemble(self.h, int(values)))

    def info(self):
        n, nnz, mr = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        self.ctx.check(lib().grmp_matrix_info(self.h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(mr)))
        return n.value, nnz.value, mr.value

    def csc(self, values=True):
        """A.entries.cscmatrix as 0-based (colptr, rowval, nzval)."""
        n, nnz, _ = self.info()
        colptr = np.empty(n + 1, dtype=np.int64)
        rowval = np.empty(nnz, dtype=np.int64)
        nzval = np.empty(nnz) if values else None
        self.ctx.check(lib().grmp_matrix_export_csc(self.h, i64ptr(colptr), i64ptr(rowval),
                                                    dptr(nzval) if values else None))
        return colptr - 1, rowval - 1, nzval

    def to_scipy(self):
        import scipy.sparse as sp
        colptr, rowval, nzval = self.csc()
        return sp.csc_matrix((nzval, rowval, colptr), shape=(self.n, self.n))

    def spmv(self, x, y, values=_abi.VALUES_FP64):
        self.ctx.check(lib().grmp_matrix_spmv(self.h, int(values), x.h, y.h))

    def last_assembly_seconds(self):
        return lib().grmp_matrix_last_assembly_seconds(self.h)

    def __getitem__(self, ij):
        return FEMatrixBlock(self, ij[0], ij[1])

    def __del__(self):
        try:
            if self.h and _abi.alive():
                lib().grmp_matrix_destroy(self.h)
            self.h = None
        except Exception:
            pass


class FEMatrixBlock:
    def __init__(self, parent, i, j):
        self.parent, self.i, self.j = parent, i, j
        self.FESX, self.FESY = parent.FES[i], parent.FES[j]
        self.offsetX, self.offsetY = int(parent.offsets[i]), int(parent.offsets[j])
