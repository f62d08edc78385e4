"""The C-ABI library loads and exports every function include/grmp.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "grmp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(grmp_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(grmp):
    L = ctypes.CDLL(grmp._abi.LIBPATH)
    names = _declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header(grmp):
    names = set(_declared())
    bound = set(grmp._abi.SIGNATURES)
    assert names == bound, (names - bound, bound - names)


def test_version(grmp):
    assert grmp.lib().grmp_version() >= 100


def test_quadrature_orders(grmp):
    """prepare_assembly! quadrature orders (src/assemblypatterns.jl:559-565)."""
    class _G:
        ncells = 1

    class _S:
        def __init__(self, ft):
            self.FEType, self.xgrid = ft, _G()

    p1, p2, p1v = _S(grmp.H1P1(1)), _S(grmp.H1P2(1)), _S(grmp.H1P1(2))
    assert grmp.LaplaceOperator(1.0).quadorder(p1) == 0
    assert grmp.LaplaceOperator(1.0).quadorder(p2) == 2
    assert grmp.ReactionOperator(1.0).quadorder(p1) == 2
    assert grmp.HookStiffnessOperator2D(1.0, 1.0).quadorder(p1v) == 0
    lf = grmp.LinearForm(grmp.Identity, grmp.DataFunction([1.0]))
    assert lf.lform(p1).quadorder == 1 and lf.lform(p2).quadorder == 2


def test_solver_config_validation(grmp):
    """The fixpoint keywords are checked on the host before any device work:
    the Newton preconditioner and the Anderson keywords of
    default_solver_kwargs (src/solvers.jl:37-59)."""
    import pytest
    cfg = grmp.SolverConfig(anderson_iterations=3, anderson_metric="H1", anderson_damping=0.5,
                            anderson_unknowns=[0, 1], newton_preconditioner="jacobi")
    assert (cfg.anderson_iterations, cfg.anderson_metric, cfg.anderson_damping, cfg.anderson_unknowns) == \
        (3, "H1", 0.5, (0, 1))
    assert grmp.SolverConfig().anderson_iterations == 0 and grmp.SolverConfig().anderson_unknowns == (0,)
    for kw in (dict(anderson_metric="H2"), dict(anderson_iterations=11), dict(anderson_iterations=-1),
               dict(anderson_damping=0.0), dict(anderson_damping=1.5), dict(newton_preconditioner="ilu")):
        with pytest.raises(grmp.GrmpError) as ei:
            grmp.SolverConfig(**kw)
        assert ei.value.code == grmp._abi.ERR_ARG, kw
