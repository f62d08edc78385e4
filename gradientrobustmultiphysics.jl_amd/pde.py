"""PDE operators, PDEDescription, assemble and solve.

Python mirror of the reference's operator layer for the hot path
(src/pdeoperators.jl, src/pdedescription.jl, src/pdeprototypes.jl) and of the
system assemble + solve_direct! of src/solvers.jl:397-686.  Operators only
describe a form; assembly and solve run in libgrmp_hip.so.

Quadrature orders follow prepare_assembly! (src/assemblypatterns.jl:559-565):
sum of the polynomial orders of the FE types, minus one per derivative
(QuadratureOrderShift4Operator, src/functionoperators.jl:294), plus the
action's bonus order.
"""
import ctypes

import numpy as np

from . import _abi
from .fe import FEMatrix, FEVector
from ._abi import lib

Identity, Gradient, SymmetricGradient = "Identity", "Gradient", "SymmetricGradient"
APT_SymmetricBilinearForm, APT_BilinearForm = "symmetric", "full"
_DERIV = {Identity: 0, Gradient: 1, SymmetricGradient: 1}


class DataFunction:
    """Constant data DataFunction(c) (src/userdata.jl:51-63)."""

    def __init__(self, c, bonus_quadorder=0):
        self.c = np.atleast_1d(np.asarray(c, dtype=np.float64))
        self.bonus_quadorder = bonus_quadorder


class CellCoefficient:
    """The item-dependent action result = kappa[item[1]] .* input
    (src/actions.jl:53-92 with dependencies = "I"); kappa is a device
    vector of length ncells or a host array."""

    def __init__(self, kappa, ctx=None, bonus_quadorder=0):
        self.kappa, self.ctx, self.bonus_quadorder = kappa, ctx, bonus_quadorder

    def device(self, ctx, ncells):
        if isinstance(self.kappa, _abi.DeviceVector):
            return self.kappa
        v = _abi.DeviceVector(ctx, ncells, np.asarray(self.kappa, dtype=np.float64))
        self.kappa = v
        return v


class PDEOperator:
    pass


class _BLF(PDEOperator):
    def __init__(self, kind, operators, factor=1.0, action=None, apt=APT_SymmetricBilinearForm, params=None,
                 name=""):
        self.kind, self.operators, self.factor, self.action = kind, operators, factor, action
        self.APT, self.params, self.name = apt, params or [0.0, 0.0], name

    def quadorder(self, fes):
        q = sum(fes.FEType.order - _DERIV[o] for o in self.operators)
        if self.action is not None:
            q += self.action.bonus_quadorder
        return max(q, 0)

    def form(self, fes, rb, cb, ctx):
        f = _abi.Form()
        f.kind = self.kind
        f.quadorder = self.quadorder(fes)
        f.row_block, f.col_block = rb, cb
        f.factor = float(self.factor)
        for k, p in enumerate(self.params[:4]):
            f.params[k] = float(p)
        keep = None
        if self.action is not None:
            keep = self.action.device(ctx, fes.xgrid.ncells)
            f.coef_mode = _abi.COEF_CELL
            f.cell_coef = keep.h
        else:
            f.coef_mode = _abi.COEF_FACTOR
            f.cell_coef = None
        return f, keep


def LaplaceOperator(kappa=1.0, name="auto"):
    """(kappa grad u, grad v), APT_SymmetricBilinearForm, NoAction
    (src/pdeoperators.jl:154-169)."""
    return _BLF(_abi.BLF_LAPLACE, [Gradient, Gradient], factor=kappa, name=name)


def ReactionOperator(alpha=1.0, name="auto"):
    """(alpha u, v) (src/pdeoperators.jl:176-203)."""
    return _BLF(_abi.BLF_MASS, [Identity, Identity], factor=alpha, name=name)


def HookStiffnessOperator2D(mu, lam, cell_coefficient=None, name="auto"):
    """(C(mu,lambda) eps(u), eps(v)), APT_BilinearForm
    (src/pdeoperators.jl:256-276); cell_coefficient scales mu and lambda per
    cell."""
    return _BLF(_abi.BLF_HOOKE2D, [SymmetricGradient, SymmetricGradient], factor=1.0, action=cell_coefficient,
                apt=APT_BilinearForm, params=[lam, mu], name=name)


def BilinearForm(operators, action=None, APT=APT_SymmetricBilinearForm, factor=1.0, name="auto"):
    """BilinearForm([Gradient, Gradient], CellCoefficient(kappa); APT)
    (src/pdeoperators.jl:355-409) for the operator pairs built on the device."""
    if list(operators) == [Gradient, Gradient]:
        return _BLF(_abi.BLF_LAPLACE, operators, factor=factor, action=action, apt=APT, name=name)
    if list(operators) == [Identity, Identity]:
        return _BLF(_abi.BLF_MASS, operators, factor=factor, action=action, apt=APT, name=name)
    raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, f"BilinearForm{tuple(operators)} is not built")


class LinearForm(PDEOperator):
    """LinearForm(Identity, DataFunction(f)) (src/pdeoperators.jl:694-705)."""

    def __init__(self, operator, data, factor=1.0, name="auto"):
        if operator != Identity or not isinstance(data, DataFunction):
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "only LinearForm(Identity, constant DataFunction) is built")
        self.operator, self.data, self.factor, self.name = operator, data, factor, name

    def lform(self, fes):
        lf = _abi.LForm()
        lf.kind = _abi.LF_CONST
        lf.quadorder = fes.FEType.order + self.data.bonus_quadorder
        lf.factor = float(self.factor)
        for k, v in enumerate(self.data.c[:4]):
            lf.f[k] = float(v)
        return lf


def assemble_linearform(b, j, lf):
    """assemble!(b[j], LinearForm) (src/assemblypatterns/linearform.jl:239-251)."""
    fes = b.FES[j]
    form = lf.lform(fes)
    b.ctx.check(lib().grmp_assemble_linearform(b.ctx.h, fes.h, ctypes.byref(form), b.vec.h, int(b.offsets[j])))


class HomogeneousDirichletBoundary:
    """BoundaryData with zero values on all boundary faces
    (src/boundarydata.jl:198-257)."""


class PDEDescription:
    """Unknowns, LHS operators per block, RHS operators per block and
    boundary data (src/pdedescription.jl:40-284)."""

    def __init__(self, name="PDE"):
        self.name = name
        self.unknowns = []
        self.LHS = {}
        self.RHS = {}
        self.BoundaryOperators = {}

    def add_unknown(self, name):
        self.unknowns.append(name)
        return len(self.unknowns) - 1

    def add_operator(self, pos, op):
        if isinstance(op, LinearForm):
            self.RHS.setdefault(pos, []).append(op)
        else:
            self.LHS.setdefault(tuple(pos), []).append(op)

    def add_rhsdata(self, j, op):
        self.RHS.setdefault(j, []).append(op)

    def add_boundarydata(self, j, data=HomogeneousDirichletBoundary):
        self.BoundaryOperators[j] = data


def PoissonProblem(diffusion=1.0, cell_coefficient=None):
    """PoissonProblem (src/pdeprototypes.jl:105-115): LaplaceOperator(kappa),
    or the per-cell action form when cell_coefficient is given."""
    pde = PDEDescription("Poisson problem")
    pde.add_unknown("u")
    if cell_coefficient is None:
        pde.add_operator((0, 0), LaplaceOperator(diffusion))
    else:
        pde.add_operator((0, 0), BilinearForm([Gradient, Gradient], cell_coefficient, APT=APT_SymmetricBilinearForm))
    return pde


def LinearElasticityProblem(E=1.0, nu=0.3, cell_coefficient=None):
    """LinearElasticityProblem in 2D (src/pdeprototypes.jl:74-93) with
    mu = E/(2(1+nu)), lambda = E nu/((1+nu)(1-2nu))."""
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    pde = PDEDescription("linear elasticity problem")
    pde.add_unknown("u")
    pde.add_operator((0, 0), HookStiffnessOperator2D(mu, lam, cell_coefficient))
    return pde


class SolverConfig:
    """Keyword configuration of solve (src/solvers.jl:38-59)."""

    def __init__(self, precision="fp64", rtol=1e-10, maxiterations=1000000, inner_rtol=0.1,
                 fixed_penalty=1e60):
        self.precision, self.rtol, self.maxiterations = precision, rtol, maxiterations
        self.inner_rtol, self.fixed_penalty = inner_rtol, fixed_penalty


def fixed_dofs(pde, Solution):
    out = []
    for j in sorted(pde.BoundaryOperators):
        fes = Solution.FES[j]
        out.append(fes.boundary_dofs() + int(Solution.offsets[j]))
    return np.unique(np.concatenate(out)) if out else np.zeros(0, dtype=np.int64)


def assemble(pde, Solution, values=_abi.VALUES_FP64, penalty=1e60):
    """System assemble! (src/solvers.jl:397-610) plus the penalty of
    solve_direct! (:647-652): returns (A::FEMatrix, b::FEVector, fixed)."""
    A = FEMatrix(Solution.FES)
    for (i, j), ops in sorted(pde.LHS.items()):
        for op in ops:
            form, keep = op.form(Solution.FES[i], i, j, A.ctx)
            A.add_form(form, keep)
    fixed = fixed_dofs(pde, Solution)
    A.set_fixed(fixed, penalty)
    A.build_pattern()
    A.assemble(values)
    b = FEVector(Solution.FES)
    for j, ops in sorted(pde.RHS.items()):
        for op in ops:
            assemble_linearform(b, j, op)
    if fixed.size:
        b.vec.scatter(fixed + 1, penalty * 0.0)
    return A, b, fixed


def solve(Solution, pde, config=None, **kwargs):
    """solve!(Solution, PDE) for linear stationary problems
    (src/solvers.jl:1354-1402 -> solve_direct! :622-686).  Returns the
    SolveInfo of the device Krylov solve."""
    cfg = config or SolverConfig(**kwargs)
    fp32 = cfg.precision == "fp32_ir"
    values = _abi.VALUES_FP64 | (_abi.VALUES_FP32 if fp32 else 0)
    A, b, fixed = assemble(pde, Solution, values, cfg.fixed_penalty)
    return solve_system(A, b, Solution, cfg), A, b


def solve_system(A, b, Solution, cfg):
    o = _abi.SolveOpts()
    lib().grmp_solve_opts_default(ctypes.byref(o))
    o.precision = _abi.PRECISION_FP32_IR if cfg.precision == "fp32_ir" else _abi.PRECISION_FP64
    o.rtol = cfg.rtol
    o.maxit = int(cfg.maxiterations)
    o.inner_rtol = cfg.inner_rtol
    info = _abi.SolveInfo()
    rc = lib().grmp_solve(A.h, b.vec.h, Solution.vec.h, ctypes.byref(o), ctypes.byref(info))
    if rc not in (_abi.OK,):
        A.ctx.check(rc)
    return info
