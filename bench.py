#!/usr/bin/env python3
"""Served-stream benchmark on one MI355X.

Workload (one step): a stream of 4096 requests arriving as a Poisson
process (default 5000 req/s, seeded).  Each request is a 2D Poisson or P1
linear-elasticity problem (50:50; E=1, nu=0.3) on an N x N triangulated unit
square with N = round(exp(U(ln 16, ln 768))) and a per-request seeded
log-normal cell coefficient (sigma = 1).  Poisson requests use H1P1 or H1P2
(50:50), elasticity uses H1P1{2}.  Every request is solved to relative
FP64 residual 1e-9 and its solution copied into page-locked host memory
before the request counts as done (the copy is inside the timed region and
inside each latency).  The error target (relative error <= 1e-8) is checked
against the oracle's FP64 solves of the same requests (same seeded
coefficients) and against device FP64 solves to 1e-13 on requests spanning
the size range.

Solver (--solver): mg (default) serves every request with FP64 CG
preconditioned by a geometric multigrid V-cycle, requests in lockstep
batches; engine serves them with the batched Jacobi-PCG engine (FP32 sweeps
in FP64 refinement).

Reported:
  value           DoF-solves/s of the served stream (sum of ndofs / wall)
  p50/p99         request latency (arrival -> solution on the host)
  plain           the same metric for plain one-at-a-time FP64 GPU solves
                  (host-driven Jacobi-PCG, every request built from scratch),
                  measured on a systematic sample of the stream
  plain_mg        the same sample one at a time with the served solver
                  (multigrid V-cycle CG, one request in flight, no batching)
  rate_sweep      the first 512 requests at 100 ... 2000 req/s
  roofline        mg: the batched residual over all requests' rows of a level
                  (k_mgb_rows<OP_RESID>, FP32 values and FP64 vectors; with
                  --mg-batch 0 the residual on levels of >= 2^21 rows,
                  k_mg_resid_fine); engine: the engine SpMV.  Algorithmic
                  bytes / summed durations from HIP events around every
                  launch of that kernel in the timed run; beside the
                  published peak, the streaming ceiling measured after the
                  timed region (k_axpy over two 2 GiB vectors)
  cpu_baseline    the oracle (single-thread assembly + OpenMP PCG) on a
                  bounded sample of the same stream
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__ as GE  # noqa: E402

METRIC = "DoF-solves/s + p50/p99 latency at 1e-8 rel err, served vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0


def make_stream(nreq, rate, seed=2024):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, nreq))
    N = np.rint(np.exp(rng.uniform(np.log(16), np.log(768), nreq))).astype(int)
    kind = np.where(rng.random(nreq) < 0.5, 1, 2)
    fe = np.where(rng.random(nreq) < 0.5, 1, 2)
    fe[kind == 2] = 1
    seeds = rng.integers(1, 2 ** 62, nreq)
    return [dict(t=float(t[i]), kind=int(kind[i]), fe=int(fe[i]), N=int(N[i]), seed=int(seeds[i]))
            for i in range(nreq)]


def stream_ndofs(grmp, stream):
    return np.array([grmp.server.ndofs(r["kind"], r["fe"], r["N"]) for r in stream], dtype=np.int64)


def serve(grmp, mode, stream, rtol, keep=(), profile=False, respect_arrivals=True, tile_rows=0, precision="fp32_ir",
          fp64_first_sweep=0, max_active=512, mg_slots=0, mg_threads=0, mg_block_rows=0, mg_graph=False,
          mg_coarse_rows=0, mg_batch=None, pinned=None):
    """Submits the stream at its arrival times and waits for every result.
    With ``pinned`` (a PinnedBuffer of at least the stream's total dofs),
    every request's solution is copied into its slice of it before the
    request completes; otherwise only the ``keep`` requests are copied."""
    graph = mg_graph if mode == "mg" else True
    S = grmp.Server(mode, profile=profile, memory_budget=96 << 30, max_active=max_active, tile_rows=tile_rows,
                    graph=graph,
                    fp64_first_sweep=fp64_first_sweep, mg_slots=mg_slots, mg_threads=mg_threads,
                    mg_block_rows=mg_block_rows, mg_coarse_rows=mg_coarse_rows, mg_batch=mg_batch)
    keep = set(keep)
    tickets = []
    offs = None
    if pinned is not None:
        nd = stream_ndofs(grmp, stream)
        offs = np.concatenate([[0], np.cumsum(nd)])
        assert offs[-1] <= pinned.n, "pinned buffer too small for the stream"
    t0 = time.perf_counter()
    for k, r in enumerate(stream):
        if respect_arrivals:
            dt = r["t"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
        out = pinned.view(int(offs[k]), int(offs[k + 1] - offs[k])) if pinned is not None else None
        tickets.append(S.submit(r["kind"], r["fe"], r["N"], r["seed"], rtol=rtol, keep_solution=k in keep,
                                precision=precision, out=out))
    res = []
    for k, t in enumerate(tickets):
        res.append(S.wait(t))
        if (k + 1) % 256 == 0 or k + 1 == len(tickets):
            print(f"[bench] {mode}: {k + 1}/{len(tickets)} results, {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)
    wall = time.perf_counter() - t0
    sols = {k: S.solution(tickets[k]).copy() for k in keep}
    stats = S.stats()
    S.close()
    lat = np.array([r.t_done - r.t_submit for r in res])
    ok = np.array([r.status == 0 for r in res])
    dofs = np.array([r.ndofs for r in res])
    return dict(wall=wall, latency=lat, ok=ok, dofs=dofs, iters=np.array([r.iterations for r in res]),
                status=np.array([r.status for r in res]), fallback=np.array([r.fp64_fallback for r in res]),
                outer=np.array([r.outer for r in res]), sols=sols, stats=stats,
                service=np.array([r.t_done - r.t_admit for r in res]))


def request_kappa(grmp, r, sigma=1.0):
    """the request's per-cell coefficient, generated as the server does
    (Philox log-normal keyed by the request's seed)"""
    kv = grmp.DeviceVector(grmp.default_context(), 2 * r["N"] * r["N"])
    kv.lognormal(r["seed"], 0.0, sigma)
    return kv.download()


def _oracle_request(O, r, kap, rtol):
    """assembles and solves one request with the oracle (single-thread
    assembly, OpenMP FP64 Jacobi-PCG); returns (ndofs, solution)"""
    if r["kind"] == 1:
        S = O.Space(r["N"], r["fe"])
        A = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_LAPLACE, fe=r["fe"], quadorder=2 * r["fe"] - 2,
                                            factor=1.0, coef=O.COEF_CELL, kappa=kap, coords=S.coords,
                                            cells=S.cells, celldofs=S.celldofs)], fixed=S.fixed)
        b = np.zeros(S.ndofs)
        O.assemble_lf_const(b, r["fe"], 1, r["fe"], [1.0], S.coords, S.cells, S.celldofs)
    else:
        S = O.Space(r["N"], O.FE_P1, ncomp=2)
        A = O.assemble_matrix(S.ndofs, [dict(kind=O.BLF_HOOKE2D, fe=O.FE_P1, ncomp=2, quadorder=0, factor=1.0,
                                            coef=O.COEF_CELL, kappa=kap,
                                            params=np.array([0.3 / (1.3 * 0.4), 1.0 / 2.6]),
                                            coords=S.coords, cells=S.cells, celldofs=S.celldofs)],
                              fixed=S.fixed)
        b = np.zeros(S.ndofs)
        O.assemble_lf_const(b, O.FE_P1, 2, 1, [0.0, -1.0], S.coords, S.cells, S.celldofs)
    b = O.penalized_rhs(b, S.fixed)
    x, it, rr = O.pcg(A, b, S.fixed, rtol=rtol)
    return S.ndofs, x


def oracle_errors(grmp, stream, sols, idx, rtol=1e-12):
    """relative errors of served solutions against the oracle's FP64 solves
    of the same requests (same coefficient fields) to relative residual rtol"""
    O = GE.load_oracle()
    errs = []
    for k in idx:
        r = stream[k]
        _, x = _oracle_request(O, r, request_kappa(grmp, r), rtol)
        errs.append(float(np.linalg.norm(sols[k] - x) / np.linalg.norm(x)))
    return errs


def cpu_baseline(grmp, stream, budget_s=15.0, rtol=1e-9):
    """The oracle (single-thread assembly + OpenMP FP64 Jacobi-PCG) on a
    bounded sample of the stream: requests in a fixed random order, each
    taken only if its predicted time (cost ~ ndofs * N, calibrated on the
    requests already run) fits the remaining budget.  The coefficient field
    is generated outside the timing."""
    O = GE.load_oracle()
    threads = O.lib().or_num_threads()
    order = np.random.default_rng(11).permutation(len(stream))
    done_dofs, t_all, n, unit = 0, 0.0, 0, None
    nmax = 0
    for k in order:
        r = stream[k]
        nd = (r["N"] + 1) ** 2 * (2 if r["kind"] == 2 else 1)
        if r["kind"] == 1 and r["fe"] == 2:
            nd = (r["N"] + 1) ** 2 + 3 * r["N"] ** 2 + 2 * r["N"]
        cost = nd * r["N"]
        if unit is not None and t_all + unit * cost > budget_s:
            continue
        if unit is None and r["N"] > 64:
            continue  # calibrate on a small request first
        kap = request_kappa(grmp, r)
        t0 = time.perf_counter()
        _oracle_request(O, r, kap, rtol)
        dt = time.perf_counter() - t0
        t_all += dt
        done_dofs += nd
        n += 1
        nmax = max(nmax, r["N"])
        unit = dt / cost if unit is None else 0.5 * unit + 0.5 * dt / cost
        if t_all >= budget_s:
            break
    return dict(value=done_dofs / t_all, unit="DoF-solves/s", cores=int(threads), kind="port",
                sample=f"{n} requests of the stream (random order, N <= {nmax}, {done_dofs} DoFs, {t_all:.1f} s "
                       f"CPU): single-thread assembly + {threads}-thread FP64 Jacobi-PCG to rel. residual {rtol:g}; "
                       "requests too large for the time budget are skipped, which favours the CPU")


# ------------------------------------------------- single-instance workloads
def _spmv_rate(grmp, A, values, reps=20):
    """Wall time per SpMV over ``reps`` back-to-back launches on the context's
    stream (launch gaps are hidden behind kernels of this size) and the
    algorithmic bytes nnz (s_v + 4) + 4 (n + 1) + 16 n (FP64 x and y)."""
    n, nnz, _ = A.info()
    ctx = A.ctx
    x = grmp.DeviceVector(ctx, n)
    y = grmp.DeviceVector(ctx, n)
    x.fill(1.0)
    A.spmv(x, y, values)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        A.spmv(x, y, values)
    ctx.sync()
    dt = (time.perf_counter() - t0) / reps
    sv = 8 if values == grmp._abi.VALUES_FP64 else 4
    nbytes = nnz * (sv + 4) + 4 * (n + 1) + 16 * n
    gbs = nbytes / dt / 1e9
    return dict(us=round(dt * 1e6, 1), bytes=int(nbytes), gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))


def _assembly_rate(grmp, A, values, ncells, nloc, nnodes, dim, reps=3):
    """Element loop time (HIP events inside the library, pattern reused; best
    of ``reps``) and the design-independent minimum traffic
    C (4 n_loc + 8) + 8 d V + nnz s_v + 8 n (connectivity, coefficient,
    coordinates once, values and inverse diagonal written once)."""
    ts = []
    for _ in range(reps):
        A.assemble(values)
        ts.append(A.last_assembly_seconds())
    t = min(ts)
    n, nnz, _ = A.info()
    sv = (8 if values & grmp._abi.VALUES_FP64 else 0) + (4 if values & grmp._abi.VALUES_FP32 else 0)
    nbytes = ncells * (4 * nloc + 8) + 8 * dim * nnodes + nnz * sv + 8 * n
    gbs = nbytes / t / 1e9
    return dict(ms=round(t * 1e3, 3), elements_per_s=round(ncells / t, 1), min_bytes=int(nbytes), gbs=round(gbs, 1),
                frac=round(gbs / HBM_PEAK_GBS, 4))


def _timestep_rhs_rate(grmp, tcs, dt, reps=20):
    """Wall time per step right-hand side b = f + M u / dt (one kernel,
    back-to-back launches) and its algorithmic bytes
    nnz_M (8 + 4) + 4 (n + 1) + 8 n (u) + 8 n (f) + 8 n (b) + n (fixed mask)."""
    n, nnz, _ = tcs.AM.info()
    tcs.rhs(dt)
    tcs.X.ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        tcs.rhs(dt)
    tcs.X.ctx.sync()
    t = (time.perf_counter() - t0) / reps
    nbytes = nnz * 12 + 4 * (n + 1) + 25 * n
    gbs = nbytes / t / 1e9
    return dict(us=round(t * 1e6, 1), bytes=int(nbytes), gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))


def _heat_cpu_baseline(dt, N=512, steps=5):
    """The oracle's BackwardEuler steps (the reference's scheme: S = A + M/dt
    factorized once while dt is fixed, then one triangular solve per step;
    SuperLU in place of UMFPACK) on the heat problem at N x N, timed on one
    host thread after the assembly."""
    O = GE.load_oracle()
    S = O.Space(N, O.FE_P1)
    kap = np.exp(np.random.default_rng(2024).standard_normal(S.cells.shape[0]))
    base = dict(fe=O.FE_P1, coords=S.coords, cells=S.cells, celldofs=S.celldofs)
    A = O.assemble_matrix(S.ndofs, [dict(base, kind=O.BLF_LAPLACE, quadorder=0, factor=1.0, coef=O.COEF_CELL,
                                         kappa=kap)])
    M = O.assemble_matrix(S.ndofs, [dict(base, kind=O.BLF_MASS, quadorder=2, factor=1.0)])
    f = np.zeros(S.ndofs)
    O.assemble_lf_const(f, O.FE_P1, 1, 1, [1.0], S.coords, S.cells, S.celldofs)
    t0 = time.perf_counter()
    O.backward_euler(A, M, f, np.zeros(S.ndofs), [dt] * steps, S.fixed)
    t = time.perf_counter() - t0
    return dict(value=round(S.ndofs * steps / t, 1), unit="DoF-steps/s", cores=1, kind="port",
                sample=f"{N}x{N} P1 heat ({S.ndofs} dofs), {steps} steps including one sparse-LU factorization, "
                       f"{t:.1f} s")


def copy_ceiling(grmp, ctx, nbytes=2048 << 20, reps=20):
    """The streaming ceiling the kernels' rates are read against, and the
    calibration kernel of the FETCH_SIZE / WRITE_SIZE counters: the library's
    y += a x over two buffers far larger than the 256 MB MALL (reads 2 n
    doubles, writes n doubles per launch)."""
    n = nbytes // 8
    x = grmp.DeviceVector(ctx, n)
    y = grmp.DeviceVector(ctx, n)
    x.fill(1.0)
    y.fill(0.0)
    L = grmp.lib()
    for _ in range(3):
        ctx.check(L.grmp_vector_axpy(1.0, x.h, y.h))
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.check(L.grmp_vector_axpy(1.0, x.h, y.h))
    ctx.sync()
    t = (time.perf_counter() - t0) / reps
    moved = 3 * nbytes
    return {"us": round(t * 1e6, 1), "gbs": round(moved / t / 1e9, 1), "frac": round(moved / t / 1e9 / HBM_PEAK_GBS, 4),
            "reps": reps}


def run_workload(grmp, which, args, log):
    ctx = grmp.default_context()
    if args.rtol == 1e-9 and "--rtol" not in sys.argv:
        args.rtol = 1e-10  # the single-instance configs are quoted at rel. residual 1e-10
    V64, V32 = grmp._abi.VALUES_FP64, grmp._abi.VALUES_FP32
    out = {"workload": which, "n_gpus": 1, "data": "synthetic"}
    if which in ("c2-p1", "c2-p2", "c3"):
        if which == "c3":
            n3 = args.size or 160
            grid = grmp.grid_unitcube(n3, jitter=0.1, seed=3)
            fes = grmp.FESpace(grmp.H1P1(1), grid)
            sigma, nloc, dim = 0.5, 4, 3
            out["config"] = {"workload": f"3D P1 Poisson, {n3}^3 x 6 Kuhn tets (interior nodes jittered +-0.1h), "
                                         "log-normal per-cell kappa sigma=0.5", "cells": grid.ncells}
        else:
            fe = 1 if which == "c2-p1" else 2
            N = args.size or 2048
            grid = grmp.grid_unitsquare(N)
            fes = grmp.FESpace(grmp.H1P1(1) if fe == 1 else grmp.H1P2(1), grid)
            sigma, nloc, dim = 1.0, 3 * fe, 2
            out["config"] = {"workload": f"2D P{fe} Poisson, {N}x{N} triangulated square, log-normal per-cell "
                                         "kappa sigma=1, f=1", "cells": grid.ncells}
        kap = grmp.DeviceVector(ctx, grid.ncells)
        kap.lognormal(2024, 0.0, sigma)
        pde = grmp.PoissonProblem(cell_coefficient=grmp.CellCoefficient(kap))
        pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
        pde.add_boundarydata(0)
        sol = grmp.FEVector(fes)
        log(f"[bench] {which}: pattern + assembly, {fes.ndofs} dofs")
        t0 = time.perf_counter()
        A, b, fixed = grmp.assemble(pde, sol, values=V64 | V32)
        ctx.sync()
        t_first = time.perf_counter() - t0
        n, nnz, maxrow = A.info()
        out["config"].update(ndofs=int(n), nnz=int(nnz))
        out["first_assemble_s"] = round(t_first, 3)
        out["assembly_fp64"] = _assembly_rate(grmp, A, V64, grid.ncells, nloc, grid.nnodes, dim)
        out["assembly_fp32"] = _assembly_rate(grmp, A, V32, grid.ncells, nloc, grid.nnodes, dim)
        out["assembly_fp64_fp32"] = _assembly_rate(grmp, A, V64 | V32, grid.ncells, nloc, grid.nnodes, dim)
        log(f"[bench] {which}: spmv")
        out["spmv_fp64"] = _spmv_rate(grmp, A, V64)
        out["spmv_fp32"] = _spmv_rate(grmp, A, V32)
        if (which != "c3" or args.solve) and not args.assembly_only:
            log(f"[bench] {which}: FP64 CG")
            x64 = grmp.FEVector(fes)
            i64 = grmp.solve_system(A, b, x64, grmp.SolverConfig(precision="fp64", rtol=args.rtol))
            log(f"[bench] {which}: FP32-inner / FP64-refined CG")
            x32 = grmp.FEVector(fes)
            i32 = grmp.solve_system(A, b, x32, grmp.SolverConfig(precision="fp32_ir", rtol=args.rtol))
            a, c = x64.entries(), x32.entries()
            out["solve"] = {"rtol": args.rtol,
                            "fp64": {"status": i64.status, "iterations": int(i64.iterations), "s": round(i64.seconds, 3),
                                     "dof_solves_per_s": round(n / i64.seconds, 1)},
                            "fp32_ir": {"status": i32.status, "iterations": int(i32.iterations), "outer": int(i32.outer),
                                        "s": round(i32.seconds, 3), "dof_solves_per_s": round(n / i32.seconds, 1)},
                            "rel_diff_fp32_ir_vs_fp64": float(np.linalg.norm(c - a) / np.linalg.norm(a))}
            # the same assembled system with the algebraic multigrid CG (any mesh): hierarchy, then solve
            log(f"[bench] {which}: algebraic multigrid CG")
            t0 = time.perf_counter()
            amg = grmp.AlgebraicMultigridSolver(A)
            ctx.sync()
            t_setup = time.perf_counter() - t0
            xa = grmp.FEVector(fes)
            ia = amg.solve(b, xa, args.rtol)
            nl, coarse, cx = amg.info()
            out["solve"]["amg"] = {"status": int(ia.status), "iterations": int(ia.iterations),
                                   "setup_s": round(t_setup, 3), "s": round(ia.seconds, 3),
                                   "dof_solves_per_s": round(n / ia.seconds, 1), "levels": nl,
                                   "operator_complexity": round(cx, 3),
                                   "rel_diff_vs_fp64": float(np.linalg.norm(xa.entries() - a) / np.linalg.norm(a))}
            del amg
            if which != "c3":
                # the same problem (same seeded coefficient field) as one multigrid request:
                # shapes of every level, level assembly, V-cycle CG and the copy back, end to end
                log(f"[bench] {which}: multigrid-preconditioned CG (one served request)")
                with grmp.Server("mg", memory_budget=96 << 30) as S:
                    t0 = time.perf_counter()
                    t = S.submit(1, fe, N, 2024, sigma=sigma, rtol=args.rtol)
                    r = S.wait(t)
                    dt = time.perf_counter() - t0
                    xm = S.solution(t)
                out["solve"]["multigrid"] = {"status": int(r.status), "iterations": int(r.iterations), "s": round(dt, 3),
                                             "dof_solves_per_s": round(n / dt, 1),
                                             "includes": "shapes, assembly and setup of all levels, solve, copy back",
                                             "rel_diff_vs_fp64": float(np.linalg.norm(xm - a) / np.linalg.norm(a))}
                # the same system through solve(..., solver="mg"): the assembled level 0 plus coarse levels
                # built for it, one V-cycle CG
                log(f"[bench] {which}: solve(solver='mg') on the assembled system")
                xs = grmp.FEVector(fes)
                t0 = time.perf_counter()
                levels, kappa0 = grmp.mg_hierarchy(pde, fes, A)
                ctx.sync()
                t_h = time.perf_counter() - t0
                t0 = time.perf_counter()
                im = grmp.mg_solve(levels, kappa0, b, xs, args.rtol)
                t_s = time.perf_counter() - t0
                out["solve"]["multigrid_solve"] = {
                    "status": int(im.status), "iterations": int(im.iterations), "hierarchy_s": round(t_h, 3),
                    "solve_s": round(t_s, 3), "dof_solves_per_s": round(n / t_s, 1),
                    "includes": "level assembly, Gershgorin bounds, V-cycle CG (coarse patterns in hierarchy_s)",
                    "rel_diff_vs_fp64": float(np.linalg.norm(xs.entries() - a) / np.linalg.norm(a))}
    elif which == "c4":
        N = args.size or 512
        grid = grmp.grid_unitsquare(N)
        Vs = grmp.FESpace(grmp.H1BR(2), grid)
        Qs = grmp.FESpace(grmp.L2P0(1), grid)
        pde = grmp.IncompressibleNavierStokesProblem(2, viscosity=1e-2)
        rng = np.random.default_rng(4)
        K = 8
        modes = np.column_stack([rng.normal(0, 0.3, 2 * K), rng.integers(-3, 4, 2 * K), rng.integers(-3, 4, 2 * K),
                                 rng.uniform(0, 2 * np.pi, 2 * K)])
        f = grmp.GradientCurlData(modes[:K], modes[K:], (0.5, -0.25, 0.3), bonus_quadorder=2)
        pde.add_rhsdata(0, grmp.LinearForm(grmp.ReconstructionIdentity(grmp.HDIVRT0), f))
        out["config"] = {"workload": f"Stokes H1BR{{2}} x L2P0, {N}x{N} triangulated square, nu=1e-2, "
                                     "reconstructed BR->RT0 RHS of grad phi + curl psi (K=8 seeded modes)",
                         "cells": grid.ncells}
        res = {}
        xs = {}
        for prec in ("fp64", "fp32_ir"):
            sol = grmp.FEVector([Vs, Qs])
            log(f"[bench] c4: {prec}")
            t0 = time.perf_counter()
            info, A, b = grmp.solve(sol, pde, precision=prec, rtol=args.rtol, minres_inner_rtol=args.minres_inner)
            ctx.sync()
            wall = time.perf_counter() - t0
            n, nnz, _ = A.info()
            res[prec] = {"status": info.status, "iterations": int(info.iterations), "outer": int(info.outer),
                         "solve_s": round(info.seconds, 3), "wall_s": round(wall, 3),
                         "dof_solves_per_s": round(n / info.seconds, 1)}
            xs[prec] = sol.entries()
            if prec == "fp64":
                out["config"].update(ndofs=int(n), nnz=int(nnz))
                out["assembly_fp64"] = _assembly_rate(grmp, A, V64, grid.ncells, 9, grid.nnodes, 2)
                out["spmv_fp64"] = _spmv_rate(grmp, A, V64)
        res["rel_diff_fp32_ir_vs_fp64"] = float(np.linalg.norm(xs["fp32_ir"] - xs["fp64"]) / np.linalg.norm(xs["fp64"]))
        # MINRES with the velocity's node block preconditioned by a multigrid V-cycle per component
        sol = grmp.FEVector([Vs, Qs])
        log("[bench] c4: fp64, multigrid velocity block")
        t0 = time.perf_counter()
        info, A, b = grmp.solve(sol, pde, solver="mg", precision="fp64", rtol=args.rtol)
        ctx.sync()
        wall = time.perf_counter() - t0
        n, nnz, _ = A.info()
        res["fp64_mg_block"] = {"status": info.status, "iterations": int(info.iterations),
                                "solve_s": round(info.seconds, 3), "wall_s": round(wall, 3),
                                "dof_solves_per_s": round(n / info.seconds, 1),
                                "includes": "solve_s: MINRES with two V-cycles per iteration; wall_s adds system "
                                            "assembly and the velocity hierarchy",
                                "rel_diff_vs_fp64_diagonal": float(np.linalg.norm(sol.entries() - xs["fp64"]) /
                                                                   np.linalg.norm(xs["fp64"]))}
        out["solve"] = res
    elif which == "heat":
        N = args.size or 2048
        dt = args.dt
        grid = grmp.grid_unitsquare(N)
        fes = grmp.FESpace(grmp.H1P1(1), grid)
        kap = grmp.DeviceVector(ctx, grid.ncells)
        kap.lognormal(2024, 0.0, 1.0)
        pde = grmp.PoissonProblem(cell_coefficient=grmp.CellCoefficient(kap))
        pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
        pde.add_boundarydata(0)
        X = grmp.FEVector(fes)
        out["config"] = {"workload": f"2D P1 heat equation u_t - div(kappa grad u) = 1, {N}x{N} triangulated square, "
                                     f"log-normal per-cell kappa sigma=1, BackwardEuler dt={dt:g}, u0 = 0",
                         "cells": grid.ncells, "ndofs": fes.ndofs}
        log(f"[bench] heat: solver setup (mass matrix, rhs data), {fes.ndofs} dofs")
        t0 = time.perf_counter()
        tcs = grmp.TimeControlSolver(pde, X, grmp.BackwardEuler, timedependent_equations=[0],
                                     precision=args.precision, rtol=args.rtol, solver=args.ts_solver)
        ctx.sync()
        t_setup = time.perf_counter() - t0
        log("[bench] heat: first step (system assembly + solve)")
        t0 = time.perf_counter()
        i1 = tcs.advance(dt)
        ctx.sync()
        t_first = time.perf_counter() - t0
        log(f"[bench] heat: {args.timesteps} steps at fixed dt (right-hand side + solve per step)")
        its, t0 = 0, time.perf_counter()
        for _ in range(args.timesteps):
            info = tcs.advance(dt)
            its += int(info.iterations)
        ctx.sync()
        t_steps = time.perf_counter() - t0
        per = t_steps / args.timesteps
        n, nnz, _ = tcs.S.info()
        out["config"].update(nnz=int(nnz), precision=args.precision if tcs.solver == "krylov" else "fp64",
                             rtol=args.rtol, step_solver=tcs.solver)
        out["setup_s"] = round(t_setup, 3)
        out["first_step_s"] = round(t_first, 3)
        out["first_step_iterations"] = int(i1.iterations)
        out["steps"] = {"count": args.timesteps, "s_per_step": round(per, 5), "steps_per_s": round(1.0 / per, 2),
                        "dof_steps_per_s": round(n / per, 1), "iterations_per_step": round(its / args.timesteps, 1),
                        "system_assemblies": tcs.assemblies}
        out["timestep_rhs"] = _timestep_rhs_rate(grmp, tcs, dt)
        out["spmv_fp64_system"] = _spmv_rate(grmp, tcs.S, V64)
        if not args.no_cpu:
            log("[bench] heat: cpu baseline (oracle: sparse LU of S once, then one solve per step)")
            out["cpu_baseline"] = _heat_cpu_baseline(dt)
    elif which == "dirichlet":
        N = args.size or 2048
        grid = grmp.grid_unitsquare(N)
        fes = grmp.FESpace(grmp.H1P2(1), grid)
        kap = grmp.DeviceVector(ctx, grid.ncells)
        kap.lognormal(2024, 0.0, 1.0)
        coef = [0.3, -0.2, 0.5, 0.1, 0.7, -0.4, 1.1, -0.6, 0.9, 0.25]
        res = {}
        for name, kind in (("interpolate", grmp.InterpolateDirichletBoundary),
                           ("bestapprox", grmp.BestapproxDirichletBoundary)):
            pde = grmp.PoissonProblem(cell_coefficient=grmp.CellCoefficient(kap))
            pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0])))
            pde.add_boundarydata(0, kind, data=grmp.PolynomialData(coef))
            sol = grmp.FEVector(fes)
            log(f"[bench] dirichlet: {name} boundary values")
            grmp.boundary_values(pde, sol)
            ctx.sync()
            reps = 5
            t0 = time.perf_counter()
            for _ in range(reps):
                grmp.boundary_values(pde, sol)
            ctx.sync()
            tb = (time.perf_counter() - t0) / reps
            log(f"[bench] dirichlet: {name} solve")
            t0 = time.perf_counter()
            info, A, b = grmp.solve(sol, pde, precision=args.precision, rtol=args.rtol, solver="auto")
            ctx.sync()
            res[name] = {"boundary_values_ms": round(tb * 1e3, 3), "solve_status": info.status,
                         "solve_iterations": int(info.iterations), "solve_s": round(time.perf_counter() - t0, 3),
                         "solve": "multigrid V-cycle CG, data lifted, assembly and hierarchy included"}
        nb = int(fes.boundary_dofs().size)
        out["config"] = {"workload": f"2D P2 Poisson, {N}x{N} triangulated square, log-normal per-cell kappa sigma=1, "
                                     "f=1, cubic polynomial Dirichlet data on all boundary faces",
                         "cells": grid.ncells, "ndofs": fes.ndofs, "boundary_dofs": nb, "precision": args.precision,
                         "rtol": args.rtol}
        res["boundary_dofs_per_s_bestapprox"] = round(nb / (res["bestapprox"]["boundary_values_ms"] * 1e-3), 1)
        out["dirichlet"] = res
        log("[bench] dirichlet: node values of the solution (export)")
        grmp.nodevalues(sol[0])
        t0 = time.perf_counter()
        nv = grmp.nodevalues(sol[0])
        te = time.perf_counter() - t0
        out["nodevalues"] = {"ms": round(te * 1e3, 3), "nodes": int(nv.shape[1]),
                             "includes": "node-cell adjacency, cell-order averages, copy to the host"}
        if not args.no_cpu:
            log("[bench] dirichlet: cpu baseline (oracle best approximation at 256x256)")
            O = GE.load_oracle()
            S2 = O.Space(256, O.FE_P2)
            t0 = time.perf_counter()
            _, _, bdo = O.boundary_bestapprox(O.FE_P2, coef, 2 + 3, S2.coords, S2.cells, S2.celldofs, S2.cellfaces,
                                              S2.facecells, S2.ndofs)
            tc = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(bdo.size / tc, 1), "unit": "boundary dofs/s (best approximation)",
                                   "cores": 1, "kind": "port",
                                   "sample": f"256x256 P2 ({bdo.size} boundary dofs): boundary walk, face matrices, "
                                             f"sparse LU; {tc:.2f} s"}
            S1 = O.Space(256, O.FE_P1)
            u1 = np.random.default_rng(3).standard_normal(S1.ndofs)
            t0 = time.perf_counter()
            O.nodevalues(O.FE_P1, 1, u1, S1.cells, S1.celldofs, S1.nnodes)
            tn = time.perf_counter() - t0
            out["cpu_baseline_nodevalues"] = {"value": round(S1.nnodes / tn, 1), "unit": "nodes/s", "cores": 1,
                                              "kind": "port",
                                              "sample": f"256x256 P1 ({S1.nnodes} nodes), cell-order averages; "
                                                        f"{tn:.2f} s"}
    elif which == "nonlinear":
        # -div((1 + u^2) grad u) = f with u = x y (examples/Example205_NonlinearPoisson2D.jl, q = 1) on the
        # N x N square: Newton steps of the fixpoint loop, each a Newton-matrix and right-hand-side
        # assembly at the iterate plus a BiCGStab solve
        N = args.size or 256
        fe = 2
        grid = grmp.grid_unitsquare(N)
        fes = grmp.FESpace(grmp.H1P2(1), grid)
        pde = grmp.PDEDescription("nonlinear Poisson problem")
        pde.add_unknown("u")
        nlf = grmp.NonlinearForm(grmp.Gradient, [grmp.Identity, grmp.Gradient], [0, 0],
                                 grmp.NonlinearDiffusion(1.0, 0.0, 1.0), [2, 3], bonus_quadorder=2)
        pde.add_operator((0, 0), nlf)
        pde.add_boundarydata(0, grmp.InterpolateDirichletBoundary, data=grmp.PolynomialData([0, 0, 0, 0, 1.0]))
        F = [0.0] * 15
        F[11], F[13] = -2.0, -2.0
        pde.add_rhsdata(0, grmp.LinearForm(grmp.Identity, grmp.PolynomialData(F)))
        log(f"[bench] nonlinear: P2 {N}x{N}, {fes.ndofs} dofs")
        # one untimed loop first: the first launches of each kernel and the first allocations of the
        # stream-ordered pool are one-time costs of the process (0.104 s cold against 0.077 s warm)
        grmp.solve(grmp.FEVector([fes]), pde, target_residual=1e-10, nonlinear_maxiterations=12, rtol=1e-12)
        ctx.sync()
        sol = grmp.FEVector([fes])
        t0 = time.perf_counter()
        info, A, b = grmp.solve(sol, pde, target_residual=1e-10, nonlinear_maxiterations=12, rtol=1e-12)
        ctx.sync()
        t = time.perf_counter() - t0
        A.assemble()
        ts = []
        for _ in range(3):
            A.assemble()
            ts.append(A.last_assembly_seconds())
        n, nnz, _ = A.info()
        coords, _ = grid.download()
        err = float(grmp.evaluate(grmp.L2ErrorIntegrator(grmp.PolynomialData([0, 0, 0, 0, 1.0]), quadorder=2),
                                  sol[0])) ** 0.5
        out["config"] = {"workload": f"nonlinear Poisson -div((1+u^2) grad u) = f, u = xy, P2 on {N}x{N}",
                         "ndofs": int(n), "nnz": int(nnz), "cells": grid.ncells}
        out["nonlinear"] = {"status": info.status, "newton_steps": info.iterations,
                            "residuals": [float("%.3e" % r) for r in info.residuals],
                            "bicgstab_iterations": info.linear_iterations,
                            "bicgstab_restarts": info.linear_restarts,
                            "bicgstab_preconditioner": info.preconditioner, "s": round(t, 3),
                            "newton_matrix_ms": round(min(ts) * 1e3, 3),
                            "newton_matrix_elements_per_s": round(grid.ncells / min(ts), 1),
                            "l2_error_vs_xy": err}
        if not args.no_cpu:
            log("[bench] nonlinear: cpu baseline (oracle Newton loop, P2 on 48x48)")
            O = GE.load_oracle()
            S = O.Space(48, O.FE_P2)
            P = np.vstack([S.coords, 0.5 * (S.coords[S.facenodes[:, 0]] + S.coords[S.facenodes[:, 1]])])
            t0 = time.perf_counter()
            xo, res = O.newton_solve(O.FE_P2, S.coords, S.cells, S.celldofs, S.ndofs, S.fixed, P[:, 0] * P[:, 1],
                                     (1.0, 0.0, 1.0), F, quadorder=6, lf_quadorder=6, maxit=12,
                                     target_residual=1e-10)
            tc = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(S.ndofs / tc, 1), "unit": "DoF-solves/s (whole Newton loop)",
                                   "cores": 1, "kind": "port",
                                   "sample": f"P2 48x48 ({S.ndofs} dofs), {len(res)} Newton steps with sparse LU, "
                                             f"{tc:.2f} s"}
            out["dof_solves_per_s"] = round(n / t, 1)
    elif which == "copy":
        nbytes = (args.size or 2048) << 20
        out["config"] = {"workload": f"y += a x over two {nbytes >> 20} MiB FP64 vectors (k_axpy)",
                         "bytes_per_launch": 3 * nbytes, "read_bytes": 2 * nbytes, "write_bytes": nbytes}
        out["copy"] = copy_ceiling(grmp, ctx, nbytes)
    else:
        raise SystemExit(f"unknown workload {which}")
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0, help="untimed runs of the first 256 requests")
    ap.add_argument("--requests", type=int, default=4096)
    ap.add_argument("--rate", type=float, default=5000.0)
    ap.add_argument("--rtol", type=float, default=1e-9,
                    help="relative FP64 residual; the served error target is 1e-8 and is verified")
    ap.add_argument("--plain-every", type=int, default=32, help="plain leg: every k-th request of the stream")
    ap.add_argument("--verify", type=int, default=24, help="requests re-solved in FP64 for the error check")
    ap.add_argument("--oracle-verify", type=int, default=12,
                    help="requests (N <= 256) re-solved by the oracle on the CPU for the error check")
    ap.add_argument("--sweep-requests", type=int, default=512, help="rate sweep: requests per rate (0: off)")
    ap.add_argument("--sweep-rates", default="100,200,500,1000,2000", help="rate sweep: arrival rates (req/s)")
    ap.add_argument("--max-active", type=int, default=512, help="served problems in flight")
    ap.add_argument("--minres-inner", type=float, default=1e-4, help="c4: FP32 inner MINRES relative tolerance")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="served",
                    help="served (the headline stream) | c2-p1 | c2-p2 | c3 | c4 (single large instances) | "
                         "heat (BackwardEuler time stepping) | dirichlet (inhomogeneous Dirichlet data) | "
                         "nonlinear (Newton fixpoint loop) | copy (the streaming ceiling)")
    ap.add_argument("--size", type=int, default=0, help="single-instance workloads: N (default 2048 / 160 / 512)")
    ap.add_argument("--solve", action="store_true", help="c3: also solve")
    ap.add_argument("--assembly-only", action="store_true", help="c2-p1/c2-p2/c3: assembly and SpMV rates only")
    ap.add_argument("--dt", type=float, default=1e-4, help="heat: time step size")
    ap.add_argument("--timesteps", type=int, default=20, help="heat: timed steps after the first")
    ap.add_argument("--ts-solver", default="auto", choices=["auto", "krylov", "mg"],
                    help="heat: step solver (auto = multigrid V-cycle CG where it applies)")
    ap.add_argument("--tile-rows", type=int, default=0, help="engine rows per work tile (0 = library default)")
    ap.add_argument("--fp64-first-sweep", type=int, default=0,
                    help="served requests whose first FP32 sweep takes more iterations continue in FP64 (0: off)")
    ap.add_argument("--precision", default="fp32_ir", choices=["fp32_ir", "fp64"],
                    help="engine requests: FP32 sweeps with FP64 refinement and fallback, or FP64 CG in the engine")
    ap.add_argument("--solver", default="mg", choices=["mg", "engine"],
                    help="served stream: FP64 CG with a multigrid V-cycle per request (mg), or the batched "
                         "Jacobi-PCG engine (engine)")
    ap.add_argument("--mg-slots", type=int, default=0, help="mg: requests solved concurrently (0 = library default)")
    ap.add_argument("--mg-threads", type=int, default=0, help="mg: host threads launching (0 = library default)")
    ap.add_argument("--mg-block-rows", type=int, default=0,
                    help="mg: levels up to this many rows run in one workgroup (0 = library default)")
    ap.add_argument("--mg-coarse-rows", type=int, default=0,
                    help="mg: levels of larger requests up to this many rows run in one workgroup (0 = default)")
    ap.add_argument("--mg-batch", type=int, default=None,
                    help="mg: requests advanced in lockstep per launch (0: one stream per request; default: library)")
    ap.add_argument("--mg-graph", action="store_true",
                    help="mg: replay captured chunks of CG iterations (profiles every 4th request only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the served stream shards by request "
              "and is not measured across GPUs here", file=sys.stderr)
        sys.exit(2)
    grmp = GE.load()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    if args.workload != "served":
        run_workload(grmp, args.workload, args, log)
        return
    stream = make_stream(args.requests, args.rate)
    ctx = grmp.default_context()
    # one page-locked arena for every request's solution, allocated before the timed runs
    nd_stream = stream_ndofs(grmp, stream)
    pinned = grmp._abi.PinnedBuffer(ctx, int(nd_stream.sum()))

    mode = "mg" if args.solver == "mg" else "served"
    # warmup: the first 256 requests of the stream, untimed
    for w in range(args.warmup):
        log(f"[bench] warmup {w}")
        serve(grmp, mode, stream[:256], args.rtol, profile=True, mg_slots=args.mg_slots, pinned=pinned)

    # verification sample: spread over the size range, largest included
    order = np.argsort([r["N"] * (2 if r["fe"] == 2 or r["kind"] == 2 else 1) for r in stream])
    vidx = sorted(set(order[np.linspace(0, len(order) - 1, args.verify).astype(int)].tolist()))
    # oracle sample: the same spread over the requests the CPU solves in seconds (N <= 256)
    small = [k for k in order.tolist() if stream[k]["N"] <= 256]
    oidx = sorted(set(small[j] for j in np.linspace(0, len(small) - 1, args.oracle_verify).astype(int)))
    offs = np.concatenate([[0], np.cumsum(nd_stream)])

    runs = []
    for s in range(args.steps):
        log(f"[bench] step {s}: {len(stream)} requests at {args.rate:g}/s")
        ctx.sync()
        runs.append(serve(grmp, mode, stream, args.rtol, profile=True, tile_rows=args.tile_rows,
                          precision=args.precision, fp64_first_sweep=args.fp64_first_sweep,
                          max_active=args.max_active, mg_slots=args.mg_slots, mg_threads=args.mg_threads,
                          mg_block_rows=args.mg_block_rows, mg_graph=args.mg_graph,
                          mg_coarse_rows=args.mg_coarse_rows, mg_batch=args.mg_batch, pinned=pinned))
        ctx.sync()
    run = runs[-1]
    walls = [r["wall"] for r in runs]
    if not run["ok"].all():
        for k in np.nonzero(~run["ok"])[0]:
            r = stream[k]
            log(f"[bench] request {k} failed: kind {r['kind']} fe {r['fe']} N {r['N']} seed {r['seed']} "
                f"status {run['status'][k]} iterations {run['iters'][k]} outer {run['outer'][k]}")
        raise SystemExit(f"{int((~run['ok']).sum())} requests failed")
    value = float(np.mean([r["dofs"].sum() / r["wall"] for r in runs]))
    # the last run's solutions, as delivered to the host
    sols = {k: pinned.view(int(offs[k]), int(nd_stream[k])).copy() for k in set(vidx) | set(oidx)}

    log("[bench] verification against device FP64 solves (rel. residual 1e-13)")
    vstream = [dict(stream[k], t=0.0) for k in vidx]
    ref = serve(grmp, "plain", vstream, 1e-13, keep=range(len(vidx)), respect_arrivals=False)
    errs = []
    for j, k in enumerate(vidx):
        xr, xs = ref["sols"][j], sols[k]
        errs.append(float(np.linalg.norm(xs - xr) / np.linalg.norm(xr)))
    oerrs = []
    if not args.no_cpu:
        log(f"[bench] verification against the oracle's FP64 solves ({len(oidx)} requests, N <= 256)")
        oerrs = oracle_errors(grmp, stream, sols, oidx)

    log("[bench] plain one-at-a-time FP64 leg (Jacobi-PCG, every request built from scratch)")
    pstream = [dict(r, t=0.0) for r in stream[::args.plain_every]]
    plain = serve(grmp, "plain", pstream, args.rtol, respect_arrivals=False, pinned=pinned)
    plain_value = float(plain["dofs"].sum() / plain["wall"])
    plain_mg = None
    if mode == "mg":
        log("[bench] plain one-at-a-time leg with the served solver (multigrid CG, one request in flight)")
        pm = serve(grmp, "mg", pstream, args.rtol, respect_arrivals=False, mg_slots=1, mg_threads=1, mg_batch=0,
                   pinned=pinned)
        plain_mg = float(pm["dofs"].sum() / pm["wall"])

    sweep = []
    if args.sweep_requests > 0:
        for rate in [float(x) for x in args.sweep_rates.split(",") if x]:
            sstream = make_stream(args.sweep_requests, rate)
            log(f"[bench] rate sweep: {len(sstream)} requests at {rate:g}/s")
            sr = serve(grmp, mode, sstream, args.rtol, max_active=args.max_active, mg_batch=args.mg_batch,
                       pinned=pinned)
            if not sr["ok"].all():
                raise SystemExit(f"rate sweep at {rate:g}/s: {int((~sr['ok']).sum())} requests failed")
            sweep.append({"rate_per_s": rate, "requests": len(sstream), "wall_s": round(sr["wall"], 3),
                          "arrival_span_s": round(sstream[-1]["t"], 3),
                          "solves_per_s": round(len(sstream) / sr["wall"], 2),
                          "dof_solves_per_s": round(float(sr["dofs"].sum() / sr["wall"]), 1),
                          "p50_latency_ms": round(1e3 * float(np.percentile(sr["latency"], 50)), 2),
                          "p99_latency_ms": round(1e3 * float(np.percentile(sr["latency"], 99)), 2)})

    st = run["stats"]
    achieved = st["spmv_bytes"] / st["spmv_seconds"] / 1e9 if st["spmv_seconds"] > 0 else 0.0
    avg_launch = st["spmv_seconds"] / max(st["spmv_launches"], 1)
    batched = mode == "mg" and (args.mg_batch is None or args.mg_batch > 0)
    kernel = ("k_mgb_rows<OP_RESID>" if batched else "k_mg_resid_fine") if mode == "mg" else "k_eng_spmv"
    roof = {"bound": "hbm", "kernel": kernel,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_us": round(avg_launch * 1e6, 2), "launches": int(st["spmv_launches"]),
            "bytes_per_launch": round(st["spmv_bytes"] / max(st["spmv_launches"], 1), 1)}
    if mode == "served":
        roof["traffic_over_algorithmic"] = 1.19  # FETCH_SIZE + WRITE_SIZE passes, profiles/engine_spmv_hbm_traffic.md
    elif batched:
        # FETCH_SIZE + WRITE_SIZE passes on the same kernel (profiles/mgb_resid_hbm_traffic.md): HBM bytes
        # are 1.08 x the algorithmic bytes; the measured ratio scales this run's bytes per launch
        roof["traffic_over_algorithmic"] = 1.08
        roof["traffic"] = round(1.08 * roof["bytes_per_launch"], 1)
        roof["traffic_basis"] = "PMC ratio 1.08 (256-request stream) x this run's algorithmic bytes per launch"
    # the measured streaming ceiling beside the published peak (after the timed region)
    cc = copy_ceiling(grmp, grmp.default_context())
    roof["measured_copy_gbs"] = cc["gbs"]
    roof["frac_of_measured_copy"] = round(achieved / cc["gbs"], 4) if cc["gbs"] > 0 else None

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle)")
        cpu = cpu_baseline(grmp, stream, args.cpu_seconds, args.rtol)

    lat = run["latency"]
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "DoF-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * float(np.mean(walls)), 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if mode == "mg" else "f32 inner / f64 refined",
        "data": "synthetic",
        "config": {"workload": f"served stream: {args.requests} requests, Poisson arrivals {args.rate:g}/s, "
                               "Poisson(P1|P2)/elasticity(P1) 50:50, N log-uniform [16,768], log-normal sigma=1",
                   "requests": args.requests, "rate_per_s": args.rate, "rtol": args.rtol,
                   "total_dofs": int(run["dofs"].sum()),
                   "delivery": "every solution copied into page-locked host memory before its request completes "
                               "(inside the timed region and the latency)"},
        "solves_per_s": round(len(stream) / float(np.mean(walls)), 2),
        "p50_latency_ms": round(1e3 * float(np.percentile(lat, 50)), 2),
        "p99_latency_ms": round(1e3 * float(np.percentile(lat, 99)), 2),
        "max_rel_err_vs_fp64": max(errs),
        "max_rel_err_vs_oracle": max(oerrs) if oerrs else None,
        "oracle_verified": {"requests": len(oerrs), "rtol": 1e-12,
                            "sample": "requests with N <= 256 spread over sizes and kinds; the oracle assembles and "
                                      "solves each with the server's coefficient field"} if oerrs else None,
        "plain": {"value": round(plain_value, 1), "unit": "DoF-solves/s", "requests": len(pstream),
                  "sample": f"every {args.plain_every}th request of the stream, served one at a time "
                            "(host-driven FP64 Jacobi-PCG, mesh, dofmap and pattern built per request)",
                  "p50_service_ms": round(1e3 * float(np.percentile(plain["service"], 50)), 2),
                  "p99_service_ms": round(1e3 * float(np.percentile(plain["service"], 99)), 2)},
        "served_vs_plain": round(value / plain_value, 3),
        "plain_mg": {"value": round(plain_mg, 1), "unit": "DoF-solves/s", "requests": len(pstream),
                     "sample": "the same requests one at a time with the served solver (multigrid V-cycle CG, one "
                               "request in flight, no lockstep batching)"} if plain_mg else None,
        "served_vs_plain_same_solver": round(value / plain_mg, 3) if plain_mg else None,
        "rate_sweep": sweep,
        "solver": ("FP64 CG + geometric multigrid V-cycle (FP32-stored level values, FP64 arithmetic), " +
                   ("requests in lockstep batches" if batched else
                   "one stream per request in flight")) if mode == "mg"
                  else "batched Jacobi-PCG engine, FP32 sweeps in FP64 refinement",
        "engine": {"steps": int(st["steps"]), "templates": int(st["templates"]),
                   "template_seconds": round(st["template_seconds"], 3),
                   "mean_outer_sweeps": round(float(run["outer"].mean()), 2),
                   "mean_iterations": round(float(run["iters"].mean()), 1),
                   "p99_iterations": int(np.percentile(run["iters"], 99)),
                   "max_iterations": int(run["iters"].max()),
                   "fp64_fallbacks": int(run["fallback"].sum()),
                   "slowest": {k: v for k, v in stream[int(np.argmax(run["iters"]))].items() if k != "t"} | {
                       "iterations": int(run["iters"].max()), "outer": int(run["outer"][int(np.argmax(run["iters"]))]),
                       "fp64_fallback": int(run["fallback"][int(np.argmax(run["iters"]))])},
                   "tile_rows": args.tile_rows or 2048},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
