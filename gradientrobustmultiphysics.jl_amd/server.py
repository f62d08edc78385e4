"""Request server: continuous batching of PDE solves on one GPU.

Thin wrapper of grmp_server_* (include/grmp.h).  ``Server(mode="served")``
admits requests by HBM footprint, keeps meshes and patterns resident,
assembles new requests while others iterate, and solves all problems in
flight with one batched FP32-in-FP64 Krylov engine.  ``mode="plain"`` serves
one request at a time with the host-driven FP64 PCG, built from scratch.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import VP, default_context, lib

POISSON, ELASTICITY = _abi.REQ_POISSON, _abi.REQ_ELASTICITY


def request(kind, fe, N, seed, sigma=1.0, rtol=1e-10, E=1.0, nu=0.3, x_out=None, precision="fp32_ir"):
    r = _abi.Request()
    r.kind, r.fe, r.N, r.seed = int(kind), int(fe), int(N), int(seed)
    r.sigma, r.E, r.nu, r.rtol = float(sigma), float(E), float(nu), float(rtol)
    r.precision = _abi.PRECISION_FP64 if precision == "fp64" else _abi.PRECISION_FP32_IR
    r.x_out = x_out.ctypes.data_as(_abi.DP) if x_out is not None else None
    return r


def ndofs(kind, fe, N):
    return int(lib().grmp_request_ndofs(ctypes.byref(request(kind, fe, N, 0))))


def footprint(kind, fe, N):
    return int(lib().grmp_request_footprint(ctypes.byref(request(kind, fe, N, 0))))


class Server:
    MODES = {"served": 1, "plain": 2, "mg": 3}

    def __init__(self, mode="served", memory_budget=64 << 30, max_active=256, graph=True, profile=False, ctx=None,
                 tile_rows=0, fp64_first_sweep=0, mg_slots=0, mg_threads=0, mg_block_rows=0,
                 mg_coarse_rows=0, mg_batch=None):
        """mode "served": batched Jacobi-PCG engine with FP32 sweeps; "plain":
        one request at a time, FP64 Jacobi-PCG; "mg": FP64 CG preconditioned
        by a geometric multigrid V-cycle, ``mg_slots`` requests at once."""
        self.ctx = ctx or default_context()
        cfg = _abi.ServerConfig()
        if mode not in self.MODES:
            raise ValueError(f"unknown server mode {mode!r}")
        cfg.mode = self.MODES[mode]
        cfg.mg_slots = int(mg_slots)
        cfg.mg_threads = int(mg_threads)
        cfg.mg_block_rows = int(mg_block_rows)
        cfg.mg_coarse_rows = int(mg_coarse_rows)
        cfg.mg_batch = -1 if mg_batch is None else int(mg_batch)  # -1: library default
        cfg.memory_budget = int(memory_budget)
        cfg.max_active = int(max_active)
        cfg.graph = 1 if graph else 0
        cfg.profile = 1 if profile else 0
        cfg.tile_rows = int(tile_rows)
        cfg.fp64_first_sweep = int(fp64_first_sweep)
        h = VP()
        self.ctx.check(lib().grmp_server_create(self.ctx.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._outs = {}

    def submit(self, kind, fe, N, seed, sigma=1.0, rtol=1e-10, keep_solution=True, out=None, **kw):
        """Returns a ticket; the solution lands in a host array on completion:
        ``out`` (e.g. a PinnedBuffer view of ndofs doubles) or a new array."""
        if out is not None:
            if out.dtype != np.float64 or out.size != ndofs(kind, fe, N) or not out.flags.c_contiguous:
                raise _abi.GrmpError(_abi.ERR_ARG, "submit: out must be a contiguous float64 array of ndofs entries")
            x = out
        else:
            x = np.zeros(ndofs(kind, fe, N)) if keep_solution else None
        r = request(kind, fe, N, seed, sigma, rtol, x_out=x, **kw)
        t = ctypes.c_int64()
        self.ctx.check(lib().grmp_server_submit(self.h, ctypes.byref(r), ctypes.byref(t)))
        self._outs[t.value] = x
        return t.value

    def poll(self, ticket):
        res = _abi.Result()
        rc = lib().grmp_server_poll(self.h, int(ticket), ctypes.byref(res))
        if rc == _abi.ERR_BUSY:
            return None
        self.ctx.check(rc)
        return res

    def wait(self, ticket):
        res = _abi.Result()
        self.ctx.check(lib().grmp_server_wait(self.h, int(ticket), ctypes.byref(res)))
        return res

    def solution(self, ticket):
        return self._outs.get(ticket)

    def stats(self):
        st = _abi.KernelStats()
        self.ctx.check(lib().grmp_server_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def clock(self):
        return lib().grmp_server_clock(self.h)

    def close(self):
        if self.h and _abi.alive():
            lib().grmp_server_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
