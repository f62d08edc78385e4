"""Autotuned launch shapes of the element assembly: every shape gives the
built-in shape's values bit for bit; the tuning table is read by the
library (grmp_tune_load) and rejected when out of range; the autotuner runs
end to end on small meshes (--quick) and writes a table the library reads.

Bar: bitwise equality of the FP64 values for every (variant, rows per block,
XCD remap) shape and every tuned form kind."""
import ctypes
import importlib.util
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _autotune():
    spec = importlib.util.spec_from_file_location("grmp_autotune", os.path.join(ROOT, "gradientrobustmultiphysics.jl_amd",
                                                                                "autotune.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize("case", range(5))
def test_every_shape_bitwise(grmp, case):
    at = _autotune()
    name, build = at._cases(grmp, True)[case]
    print(f"[case] tune shapes {name}", flush=True)
    A, keep = build(True)
    at._set(grmp, A, None)
    A.assemble(grmp._abi.VALUES_FP64)
    ref = A.csc()[2]
    for shape in at._shapes():
        at._set(grmp, A, shape)
        A.assemble(grmp._abi.VALUES_FP64)
        assert np.array_equal(A.csc()[2], ref), shape


def test_table_load_and_errors(grmp, tmp_path):
    print("[case] tuning table load", flush=True)
    at = _autotune()
    name, build = at._cases(grmp, True)[1]
    A, keep = build(True)
    kind, nq, bucket = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    A.ctx.check(grmp.lib().grmp_matrix_tune_key(A.h, ctypes.byref(kind), ctypes.byref(nq), ctypes.byref(bucket)))
    A.assemble(grmp._abi.VALUES_FP64)
    ref = A.csc()[2]
    good = tmp_path / "t.json"
    at.write([{"name": name, "kind": kind.value, "nq": nq.value, "bucket": bucket.value, "variant": 2, "rows": 64,
               "remap": 1}], str(good))
    A.ctx.check(grmp.lib().grmp_tune_load(A.ctx.h, str(good).encode()))
    assert grmp.lib().grmp_tune_count(A.ctx.h) == 1
    A.assemble(grmp._abi.VALUES_FP64)
    assert np.array_equal(A.csc()[2], ref)
    bad = tmp_path / "b.json"
    bad.write_text('{"entries": [{"kind": 1, "nq": 1, "bucket": 0, "variant": 7}]}')
    assert grmp.lib().grmp_tune_load(A.ctx.h, str(bad).encode()) == grmp._abi.ERR_ARG
    assert grmp.lib().grmp_tune_load(A.ctx.h, str(tmp_path / "missing.json").encode()) == grmp._abi.ERR_ARG
    # the failed loads kept the table loaded before
    assert grmp.lib().grmp_tune_count(A.ctx.h) == 1
    _empty(grmp, A.ctx, tmp_path)


def _empty(grmp, ctx, tmp_path):
    e = tmp_path / "empty.json"
    e.write_text('{"entries": []}')
    ctx.check(grmp.lib().grmp_tune_load(ctx.h, str(e).encode()))
    assert grmp.lib().grmp_tune_count(ctx.h) == 0


def test_autotuner_quick_run(grmp, tmp_path):
    print("[case] autotuner --quick", flush=True)
    at = _autotune()
    entries = at.tune(grmp, quick=True, reps=1, log=lambda *a: None)
    assert len(entries) == 5 and all(e["best_ms"] <= e["builtin_ms"] for e in entries)
    out = tmp_path / "tuning.json"
    at.write(entries, str(out))
    json.loads(out.read_text())
    ctx = grmp.default_context()
    ctx.check(grmp.lib().grmp_tune_load(ctx.h, str(out).encode()))
    assert grmp.lib().grmp_tune_count(ctx.h) == 5
    _empty(grmp, ctx, tmp_path)
