"""Autotuner of the element-assembly launch shape on the GPU it runs on.

For each single-form kind of the element loop (P1 / P2 Laplace on
triangles, P1 Laplace on tetrahedra, P1{2} Hooke, P1 mass) it measures every
launch shape -- kernel variant (1: one thread walks a row's cells; 2: the
row block's cell work staged over the workgroup; 3: as 2 with several
cells' gathers in flight per thread), rows per workgroup, and
XCD-contiguous row blocks on or off -- on a mesh of the benchmark's scale,
checks on a small mesh that each shape gives bitwise the values of the
built-in shape, and writes the fastest per (form kind, quadrature points,
nnz/row bucket) to a JSON table.  The library reads that table from
tuning_gfx950.json beside libgrmp_hip.so when a context is created
(grmp_tune_load), so a tuned shape applies to every later assembly with the
same key.  Element loop of the reference: src/assemblypatterns/bilinearform.jl:226-377.

    python gradientrobustmultiphysics.jl_amd/autotune.py [--out PATH] [--quick]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuning_gfx950.json")


def _load():
    sys.path.insert(0, os.path.dirname(HERE))
    import __graft_entry__ as GE  # noqa: E402
    return GE.load()


def _cases(g, quick):
    """(name, builder(scale) -> (matrix, keep-alive objects)) per tuned form kind"""
    N2 = 256 if quick else 2048
    N3 = 24 if quick else 128

    def tri(fe, op_of, N):
        grid = g.grid_unitsquare(N)
        ft = {1: g.H1P1(1), 2: g.H1P2(1), 3: g.H1P1(2)}[fe]
        fes = g.FESpace(ft, grid)
        kap = g.DeviceVector(grid.ctx, grid.ncells)
        kap.lognormal(2024, 0.0, 1.0)
        A = g.FEMatrix([fes])
        form, keep = op_of(kap).form(fes, 0, 0, grid.ctx)
        A.add_form(form, keep)
        A.set_fixed(fes.boundary_dofs())
        A.build_pattern()
        return A, (grid, fes, kap, keep)

    def tet(n):
        grid = g.grid_unitcube(n, 0.1, 3)
        fes = g.FESpace(g.H1P1(1), grid)
        kap = g.DeviceVector(grid.ctx, grid.ncells)
        kap.lognormal(2024, 0.0, 0.5)
        A = g.FEMatrix([fes])
        form, keep = g.BilinearForm([g.Gradient, g.Gradient], g.CellCoefficient(kap)).form(fes, 0, 0, grid.ctx)
        A.add_form(form, keep)
        A.set_fixed(fes.boundary_dofs())
        A.build_pattern()
        return A, (grid, fes, kap, keep)

    lap = lambda kap: g.BilinearForm([g.Gradient, g.Gradient], g.CellCoefficient(kap))  # noqa: E731
    hooke = lambda kap: g.HookStiffnessOperator2D(1.0 / 2.6, 0.3 / (1.3 * 0.4), g.CellCoefficient(kap))  # noqa: E731
    mass = lambda kap: g.ReactionOperator(1.0)  # noqa: E731
    return [
        ("laplace_p1_tri", lambda small: tri(1, lap, 64 if small else N2)),
        ("laplace_p2_tri", lambda small: tri(2, lap, 64 if small else N2)),
        ("laplace_p1_tet", lambda small: tet(6 if small else N3)),
        ("hooke_p1v_tri", lambda small: tri(3, hooke, 64 if small else N2 // 2)),
        ("mass_p1_tri", lambda small: tri(1, mass, 64 if small else N2)),
    ]


def _shapes():
    out = []
    for remap in (0, 1):
        for rows in (64, 128, 256):
            out.append((1, rows, remap))
        for rows in (32, 64, 128, 256):
            out.append((2, rows, remap))
        for rows in (32, 64, 128, 256):
            out.append((3, rows, remap))
    return out


def _set(g, A, shape):
    if shape is None:
        A.ctx.check(g.lib().grmp_matrix_set_tune(A.h, None))
        return
    t = g._abi.AsmTune(*shape)
    A.ctx.check(g.lib().grmp_matrix_set_tune(A.h, ctypes.byref(t)))


def _time(g, A, reps):
    ts = []
    for _ in range(reps):
        A.assemble(g._abi.VALUES_FP64)
        ts.append(g.lib().grmp_matrix_last_assembly_seconds(A.h))
    return min(ts)


def tune(g, quick=False, reps=3, log=print):
    entries = []
    for name, build in _cases(g, quick):
        # bitwise check of every shape on a small mesh
        As, keep_s = build(True)
        _set(g, As, None)
        As.assemble(g._abi.VALUES_FP64)
        ref = As.csc()[2]
        for shape in _shapes():
            _set(g, As, shape)
            As.assemble(g._abi.VALUES_FP64)
            if not np.array_equal(As.csc()[2], ref):
                raise RuntimeError(f"{name}: shape {shape} changes the values")
        # timing at scale
        A, keep = build(False)
        kind, nq, bucket = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        A.ctx.check(g.lib().grmp_matrix_tune_key(A.h, ctypes.byref(kind), ctypes.byref(nq), ctypes.byref(bucket)))
        # the library's built-in shape, whatever table the context loaded
        _set(g, A, (0, 0, -1))
        A.assemble(g._abi.VALUES_FP64)
        t_builtin = _time(g, A, reps)
        best, t_best = None, t_builtin
        timings = {}
        for shape in _shapes():
            _set(g, A, shape)
            t = _time(g, A, reps)
            timings["v%d_r%d_x%d" % shape] = round(t * 1e3, 4)
            if t < t_best:
                best, t_best = shape, t
        n, nnz, mr = A.info()
        log(f"[autotune] {name}: built-in {t_builtin * 1e3:.3f} ms, best {best} {t_best * 1e3:.3f} ms "
            f"({n} rows, {nnz} nnz)")
        e = {"name": name, "kind": kind.value, "nq": nq.value, "bucket": bucket.value, "rows_total": n, "nnz": nnz,
             "builtin_ms": round(t_builtin * 1e3, 4), "best_ms": round(t_best * 1e3, 4), "timings_ms": timings}
        if best is not None:
            e.update(variant=best[0], rows=best[1], remap=best[2])
        else:
            e.update(variant=0, rows=0, remap=-1)
        entries.append(e)
        del A, keep, As, keep_s
    return entries


def write(entries, path):
    """one entry per line: the library's reader takes each {...} object"""
    with open(path, "w") as fh:
        fh.write('{"device": "gfx950", "written": "%s", "entries": [\n' % time.strftime("%Y-%m-%d %H:%M:%S"))
        for k, e in enumerate(entries):
            flat = {key: v for key, v in e.items() if key != "timings_ms"}
            fh.write("  " + json.dumps(flat) + ("," if k + 1 < len(entries) else "") + "\n")
        fh.write("]}\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--quick", action="store_true", help="small meshes (a functional check, not a tuning)")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    g = _load()
    entries = tune(g, quick=args.quick, reps=args.reps, log=lambda *a: print(*a, file=sys.stderr, flush=True))
    write(entries, args.out)
    print(json.dumps({"out": args.out, "entries": entries}))


if __name__ == "__main__":
    main()
