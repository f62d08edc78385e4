"""ctypes binding of libgrmp_hip.so (the C ABI declared in include/grmp.h).

This module is the only place the host code touches the native library.
There is no fallback: if the library is missing or a call fails, a
``GrmpError`` is raised.
"""
import atexit
import ctypes
import os

import numpy as np

# Handles are not released once the interpreter shuts down: the HIP runtime
# may already be tearing down, and the process releases the device anyway.
_finalizing = [False]
atexit.register(lambda: _finalizing.__setitem__(0, True))


def alive():
    return not _finalizing[0]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libgrmp_hip.so")

OK, ERR_ARG, ERR_OOM, ERR_HIP, ERR_NOCONV, ERR_NONFINITE, ERR_ADMISSION, ERR_UNSUPPORTED, ERR_BUSY = range(9)
ERR_NAMES = {0: "ok", 1: "bad argument", 2: "out of device memory", 3: "HIP error", 4: "no convergence",
             5: "non-finite values", 6: "admission refused", 7: "unsupported", 8: "busy"}

FE_H1P1, FE_H1P2, FE_H1BR, FE_L2P0 = 1, 2, 3, 4
BLF_LAPLACE, BLF_HOOKE2D, BLF_MASS, BLF_DIVP0, BLF_NLDIFF = 1, 2, 3, 4, 5
COEF_FACTOR, COEF_CELL = 0, 1
LF_CONST, LF_RECO_RT0, LF_POLY, LF_NEWTON = 1, 2, 3, 4
VALUES_FP64, VALUES_FP32 = 1, 2
SOLVE_CG, SOLVE_MINRES, SOLVE_BICGSTAB = 1, 2, 3
PREC_JACOBI = 1
PRECISION_FP64, PRECISION_FP32_IR = 1, 2
REQ_POISSON, REQ_ELASTICITY = 1, 2

I64 = ctypes.c_int64
VP = ctypes.c_void_p
DP = ctypes.POINTER(ctypes.c_double)
I32P = ctypes.POINTER(ctypes.c_int32)
I64P = ctypes.POINTER(ctypes.c_int64)


class GrmpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"grmp error {code} ({ERR_NAMES.get(code, '?')}): {msg}")
        self.code = code


class Form(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("coef_mode", ctypes.c_int), ("quadorder", ctypes.c_int),
                ("row_block", ctypes.c_int), ("col_block", ctypes.c_int), ("factor", ctypes.c_double),
                ("params", ctypes.c_double * 4), ("cell_coef", VP), ("current", VP), ("current_offset", I64)]


class LForm(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("quadorder", ctypes.c_int), ("factor", ctypes.c_double),
                ("f", ctypes.c_double * 4), ("nmodes", ctypes.c_int), ("pad_", ctypes.c_int), ("modes", DP),
                ("poly", ctypes.c_double * 4), ("current", VP), ("current_offset", I64), ("coef", DP),
                ("ncoef", ctypes.c_int), ("pad2_", ctypes.c_int)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("precond", ctypes.c_int), ("precision", ctypes.c_int),
                ("rtol", ctypes.c_double), ("maxit", I64), ("inner_rtol", ctypes.c_double),
                ("inner_maxit", I64), ("check_every", ctypes.c_int),
                ("initial_guess", ctypes.c_int)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("outer", ctypes.c_int), ("iterations", I64),
                ("relres", ctypes.c_double), ("seconds", ctypes.c_double)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("fe", ctypes.c_int), ("N", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("sigma", ctypes.c_double), ("E", ctypes.c_double), ("nu", ctypes.c_double),
                ("rtol", ctypes.c_double), ("x_out", DP), ("precision", ctypes.c_int), ("pad_", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("ndofs", I64), ("iterations", I64), ("outer", ctypes.c_int),
                ("relres", ctypes.c_double), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("fp64_fallback", ctypes.c_int), ("pad_", ctypes.c_int)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("memory_budget", I64), ("max_active", ctypes.c_int),
                ("graph", ctypes.c_int), ("profile", ctypes.c_int), ("tile_rows", ctypes.c_int),
                ("fp64_first_sweep", ctypes.c_int), ("mg_slots", ctypes.c_int), ("mg_threads", ctypes.c_int),
                ("mg_block_rows", ctypes.c_int), ("mg_coarse_rows", ctypes.c_int),
                ("mg_batch", ctypes.c_int), ("pad_", ctypes.c_int)]


class KernelStats(ctypes.Structure):
    _fields_ = [("spmv_seconds", ctypes.c_double), ("spmv_launches", I64), ("spmv_bytes", ctypes.c_double),
                ("steps", I64), ("templates", I64), ("template_seconds", ctypes.c_double)]


# name -> (restype, argtypes); every symbol include/grmp.h declares
SIGNATURES = {
    "grmp_version": (ctypes.c_int, []),
    "grmp_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(VP)]),
    "grmp_ctx_destroy": (ctypes.c_int, [VP]),
    "grmp_last_error": (ctypes.c_char_p, [VP]),
    "grmp_ctx_synchronize": (ctypes.c_int, [VP]),
    "grmp_vector_create": (ctypes.c_int, [VP, I64, ctypes.POINTER(VP)]),
    "grmp_vector_destroy": (ctypes.c_int, [VP]),
    "grmp_vector_length": (I64, [VP]),
    "grmp_vector_upload": (ctypes.c_int, [VP, DP]),
    "grmp_vector_download": (ctypes.c_int, [VP, DP]),
    "grmp_vector_fill": (ctypes.c_int, [VP, ctypes.c_double]),
    "grmp_vector_scatter": (ctypes.c_int, [VP, I64, I64P, DP]),
    "grmp_vector_lognormal": (ctypes.c_int, [VP, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]),
    "grmp_mesh_upload": (ctypes.c_int, [VP, ctypes.c_int, I64, DP, I64, ctypes.c_int, I32P, ctypes.POINTER(VP)]),
    "grmp_mesh_unitsquare": (ctypes.c_int, [VP, ctypes.c_int, ctypes.POINTER(VP)]),
    "grmp_mesh_unitcube": (ctypes.c_int, [VP, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(VP)]),
    "grmp_mesh_destroy": (ctypes.c_int, [VP]),
    "grmp_mesh_info": (ctypes.c_int, [VP, ctypes.POINTER(ctypes.c_int), I64P, I64P]),
    "grmp_mesh_download": (ctypes.c_int, [VP, DP, I32P]),
    "grmp_mesh_cellvolumes": (ctypes.c_int, [VP, DP]),
    "grmp_mesh_faces": (ctypes.c_int, [VP, ctypes.POINTER(I64), I32P, ctypes.c_void_p, ctypes.c_void_p]),
    "grmp_space_upload": (ctypes.c_int, [VP, VP, ctypes.c_int, ctypes.c_int, I64, ctypes.c_int, I32P,
                                         ctypes.POINTER(VP)]),
    "grmp_space_create": (ctypes.c_int, [VP, VP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(VP)]),
    "grmp_space_destroy": (ctypes.c_int, [VP]),
    "grmp_space_info": (ctypes.c_int, [VP, I64P, ctypes.POINTER(ctypes.c_int)]),
    "grmp_space_download_celldofs": (ctypes.c_int, [VP, I32P]),
    "grmp_space_boundary_dofs": (ctypes.c_int, [VP, I64P, I64P]),
    "grmp_matrix_create": (ctypes.c_int, [VP, ctypes.c_int, ctypes.POINTER(VP), ctypes.POINTER(VP)]),
    "grmp_matrix_destroy": (ctypes.c_int, [VP]),
    "grmp_matrix_add_form": (ctypes.c_int, [VP, ctypes.POINTER(Form)]),
    "grmp_matrix_set_fixed": (ctypes.c_int, [VP, I64, I64P, ctypes.c_double]),
    "grmp_matrix_build_pattern": (ctypes.c_int, [VP]),
    "grmp_matrix_assemble": (ctypes.c_int, [VP, ctypes.c_int]),
    "grmp_matrix_info": (ctypes.c_int, [VP, I64P, I64P, ctypes.POINTER(ctypes.c_int)]),
    "grmp_matrix_export_csc": (ctypes.c_int, [VP, I64P, I64P, DP]),
    "grmp_matrix_spmv": (ctypes.c_int, [VP, ctypes.c_int, VP, VP]),
    "grmp_matrix_last_assembly_seconds": (ctypes.c_double, [VP]),
    "grmp_timestep_rhs": (ctypes.c_int, [VP, VP, VP, VP, ctypes.c_double, VP]),
    "grmp_boundary_interpolate": (ctypes.c_int, [VP, DP, ctypes.c_int, VP, I64]),
    "grmp_boundary_bestapprox": (ctypes.c_int, [VP, DP, ctypes.c_int, ctypes.c_int, ctypes.c_double, VP, I64,
                                                I64P, DP]),
    "grmp_matrix_penalize_rhs": (ctypes.c_int, [VP, VP, VP]),
    "grmp_nodevalues": (ctypes.c_int, [VP, VP, I64, DP]),
    "grmp_mg_solve": (ctypes.c_int, [VP, ctypes.c_int, ctypes.POINTER(VP), ctypes.POINTER(VP), VP, VP, VP,
                                     ctypes.c_double, I64, ctypes.POINTER(SolveInfo)]),
    "grmp_mg_create": (ctypes.c_int, [VP, ctypes.c_int, ctypes.POINTER(VP), ctypes.POINTER(VP), VP,
                                      ctypes.POINTER(VP)]),
    "grmp_mg_apply": (ctypes.c_int, [VP, VP, VP, ctypes.c_double, I64, ctypes.POINTER(SolveInfo)]),
    "grmp_mg_destroy": (ctypes.c_int, [VP]),
    "grmp_vector_axpy": (ctypes.c_int, [ctypes.c_double, VP, VP]),
    "grmp_assemble_linearform": (ctypes.c_int, [VP, VP, ctypes.POINTER(LForm), VP, I64]),
    "grmp_integrate_identity": (ctypes.c_int, [VP, VP, ctypes.c_int, VP, I64, DP]),
    "grmp_host_alloc": (ctypes.c_int, [VP, ctypes.c_size_t, ctypes.POINTER(VP)]),
    "grmp_matrix_set_tune": (ctypes.c_int, [VP, VP]),
    "grmp_matrix_assemble_batch": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(VP)]),
    "grmp_solve_minres_mg": (ctypes.c_int, [VP, VP, VP, VP, VP, ctypes.c_int, VP]),
    "grmp_ctx_set_max_nnz": (ctypes.c_int, [VP, I64]),
    "grmp_amg_create": (ctypes.c_int, [VP, ctypes.POINTER(VP)]),
    "grmp_amg_apply": (ctypes.c_int, [VP, VP, VP, ctypes.c_double, I64, VP]),
    "grmp_amg_info": (ctypes.c_int, [VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(I64),
                                     ctypes.POINTER(ctypes.c_double)]),
    "grmp_amg_destroy": (ctypes.c_int, [VP]),
    "grmp_solve_bicgstab_amg": (ctypes.c_int, [VP, VP, VP, VP, VP, VP]),
    "grmp_anderson_create": (ctypes.c_int, [VP, ctypes.c_int64, ctypes.c_int, ctypes.c_double, VP, ctypes.c_int64,
                                            VP, VP, VP]),
    "grmp_anderson_update": (ctypes.c_int, [VP, VP, VP, VP]),
    "grmp_anderson_depth": (ctypes.c_int, [VP]),
    "grmp_anderson_destroy": (ctypes.c_int, [VP]),
    "grmp_matrix_tune_key": (ctypes.c_int, [VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_int)]),
    "grmp_tune_load": (ctypes.c_int, [VP, ctypes.c_char_p]),
    "grmp_tune_count": (ctypes.c_int, [VP]),
    "grmp_host_free": (ctypes.c_int, [VP]),
    "grmp_integrate_l2error": (ctypes.c_int, [VP, VP, VP, I64, DP, ctypes.c_int, ctypes.c_int, ctypes.c_double, DP,
                                              ctypes.POINTER(ctypes.c_double)]),
    "grmp_integrate_l2error_noalloc": (ctypes.c_int, [VP, VP, VP, I64, DP, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                      DP, ctypes.POINTER(ctypes.c_double)]),
    "grmp_solve_opts_default": (None, [ctypes.POINTER(SolveOpts)]),
    "grmp_solve": (ctypes.c_int, [VP, VP, VP, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    "grmp_solve_csc": (ctypes.c_int, [VP, I64, I64P, I64P, DP, DP, DP, I64, I64P, ctypes.c_int, ctypes.c_double]),
    "grmp_server_create": (ctypes.c_int, [VP, ctypes.POINTER(ServerConfig), ctypes.POINTER(VP)]),
    "grmp_server_submit": (ctypes.c_int, [VP, ctypes.POINTER(Request), I64P]),
    "grmp_server_poll": (ctypes.c_int, [VP, I64, ctypes.POINTER(Result)]),
    "grmp_server_wait": (ctypes.c_int, [VP, I64, ctypes.POINTER(Result)]),
    "grmp_server_clock": (ctypes.c_double, [VP]),
    "grmp_server_destroy": (ctypes.c_int, [VP]),
    "grmp_server_stats": (ctypes.c_int, [VP, ctypes.POINTER(KernelStats)]),
    "grmp_request_footprint": (I64, [ctypes.POINTER(Request)]),
    "grmp_request_ndofs": (I64, [ctypes.POINTER(Request)]),
}

_lib = None


def lib():
    """Loads libgrmp_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise GrmpError(ERR_ARG, f"{LIBPATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(DP)


def i32ptr(a):
    return a.ctypes.data_as(I32P)


def i64ptr(a):
    return a.ctypes.data_as(I64P)


class Context:
    """A device context (one HIP stream on one GPU)."""

    def __init__(self, device=0):
        h = VP()
        rc = lib().grmp_ctx_create(device, ctypes.byref(h))
        if rc:
            raise GrmpError(rc, f"grmp_ctx_create(device={device}) failed")
        self.h = h

    def check(self, rc):
        if rc:
            msg = lib().grmp_last_error(self.h)
            raise GrmpError(rc, msg.decode() if msg else "")
        return rc

    def sync(self):
        self.check(lib().grmp_ctx_synchronize(self.h))

    def close(self):
        if self.h and alive():
            lib().grmp_ctx_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class AsmTune(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int32), ("rows_per_block", ctypes.c_int32), ("xcd_remap", ctypes.c_int32)]


class PinnedBuffer:
    """Page-locked host memory (grmp_host_alloc) holding ``n`` doubles;
    ``view(off, m)`` is a numpy array over part of it."""

    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, int(n)
        p = VP()
        ctx.check(lib().grmp_host_alloc(ctx.h, 8 * max(self.n, 1), ctypes.byref(p)))
        self.p = p
        self.array = np.ctypeslib.as_array((ctypes.c_double * max(self.n, 1)).from_address(p.value))[:self.n]

    def view(self, off, m):
        return self.array[off:off + m]

    def __del__(self):
        try:
            if self.p and alive():
                lib().grmp_host_free(self.p)
            self.p = None
        except Exception:
            pass


class DeviceVector:
    def __init__(self, ctx, n, values=None):
        self.ctx, self.n = ctx, int(n)
        h = VP()
        ctx.check(lib().grmp_vector_create(ctx.h, self.n, ctypes.byref(h)))
        self.h = h
        if values is not None:
            self.upload(values)

    def upload(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        assert a.size == self.n
        self.ctx.check(lib().grmp_vector_upload(self.h, dptr(a)))

    def download(self):
        out = np.empty(self.n)
        self.ctx.check(lib().grmp_vector_download(self.h, dptr(out)))
        return out

    def fill(self, v):
        self.ctx.check(lib().grmp_vector_fill(self.h, float(v)))

    def scatter(self, idx1, values):
        i = np.ascontiguousarray(idx1, dtype=np.int64)
        v = np.ascontiguousarray(np.broadcast_to(values, i.shape), dtype=np.float64)
        self.ctx.check(lib().grmp_vector_scatter(self.h, i.size, i64ptr(i), dptr(v)))

    def lognormal(self, seed, mu=0.0, sigma=1.0):
        self.ctx.check(lib().grmp_vector_lognormal(self.h, int(seed), float(mu), float(sigma)))

    def __del__(self):
        try:
            if self.h and alive():
                lib().grmp_vector_destroy(self.h)
            self.h = None
        except Exception:
            pass
