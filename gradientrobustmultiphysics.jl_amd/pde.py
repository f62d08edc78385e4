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
import copy
import ctypes
import warnings

import numpy as np

from . import _abi
from .fe import FEMatrix, FESpace, FEVector, H1P1, grid_unitsquare
from ._abi import lib

Identity, Gradient, SymmetricGradient, Divergence = "Identity", "Gradient", "SymmetricGradient", "Divergence"
APT_SymmetricBilinearForm, APT_BilinearForm = "symmetric", "full"
_DERIV = {Identity: 0, Gradient: 1, SymmetricGradient: 1, Divergence: 1}
HDIVRT0 = "HDIVRT0{2}"


class ReconstructionIdentity:
    """ReconstructionIdentity{FER}: the test function's Identity replaced by
    its reconstruction in FER (src/functionoperators.jl, evaluated by
    src/feevaluator_h1.jl:342-381); HDIVRT0 on H1BR{2} is built."""

    def __init__(self, target=HDIVRT0):
        if target != HDIVRT0:
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, f"ReconstructionIdentity{{{target}}} is not built")
        self.target = target


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

    def quadorder(self, fes, fes_col=None):
        """Sum of the polynomial orders minus one per derivative plus the
        action's bonus order (src/assemblypatterns.jl:559-565)."""
        fes_col = fes if fes_col is None else fes_col
        q = (fes.FEType.order - _DERIV[self.operators[0]]) + (fes_col.FEType.order - _DERIV[self.operators[1]])
        if self.action is not None:
            q += self.action.bonus_quadorder
        return max(q, 0)

    def form(self, fes, rb, cb, ctx, fes_col=None):
        f = _abi.Form()
        f.kind = self.kind
        f.quadorder = self.quadorder(fes, fes_col)
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


def LagrangeMultiplier(operator=Divergence, factor=-1.0, name="auto"):
    """LagrangeMultiplier(Divergence) (src/pdeoperators.jl:214-226): the
    block (velocity rows, pressure columns) gets factor * (div u, q) and the
    transposed block gets the copy with the opposite sign
    (src/assemblypatterns/bilinearform.jl:358-363)."""
    if operator != Divergence:
        raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, f"LagrangeMultiplier({operator}) is not built")
    return _BLF(_abi.BLF_DIVP0, [Divergence, Identity], factor=factor, apt=APT_BilinearForm, name=name)


def BilinearForm(operators, action=None, APT=APT_SymmetricBilinearForm, factor=1.0, name="auto"):
    """BilinearForm([Gradient, Gradient], CellCoefficient(kappa); APT)
    (src/pdeoperators.jl:355-409) for the operator pairs built on the device."""
    if list(operators) == [Gradient, Gradient]:
        return _BLF(_abi.BLF_LAPLACE, operators, factor=factor, action=action, apt=APT, name=name)
    if list(operators) == [Identity, Identity]:
        return _BLF(_abi.BLF_MASS, operators, factor=factor, action=action, apt=APT, name=name)
    raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, f"BilinearForm{tuple(operators)} is not built")


class GradientCurlData:
    """Right-hand side f = grad(phi) + curl(psi) in closed form, the device's
    stand-in for a DataFunction closure (src/userdata.jl:3-63):
    phi = sum_m A_m sin(2 pi (p_m x + q_m y) + t_m) + c0 x^3 + c1 x y^2 + c2 y^2,
    psi = the same sum over the second set of modes.  ``modes_phi`` and
    ``modes_psi`` are (K, 4) arrays (A, p, q, t) with the same K."""

    def __init__(self, modes_phi, modes_psi, poly=(0.0, 0.0, 0.0), bonus_quadorder=2):
        a = np.asarray(modes_phi, dtype=np.float64).reshape(-1, 4)
        b = np.asarray(modes_psi, dtype=np.float64).reshape(-1, 4)
        if a.shape != b.shape:
            raise _abi.GrmpError(_abi.ERR_ARG, "GradientCurlData: phi and psi need the same number of modes")
        self.modes = np.ascontiguousarray(np.concatenate([a, b]))
        self.K = a.shape[0]
        self.poly = [float(v) for v in poly] + [0.0] * (4 - len(poly))
        self.bonus_quadorder = bonus_quadorder


class LinearForm(PDEOperator):
    """LinearForm(Identity, DataFunction(f)) (src/pdeoperators.jl:694-705)
    with constant or polynomial (PolynomialData, degree <= 4) data, and
    LinearForm(ReconstructionIdentity(HDIVRT0), GradientCurlData) on
    H1BR{2}, the pressure-robust Stokes right-hand side."""

    def __init__(self, operator, data, factor=1.0, name="auto"):
        reco = isinstance(operator, ReconstructionIdentity)
        if not ((operator == Identity and isinstance(data, (DataFunction, PolynomialData))) or
                (reco and isinstance(data, GradientCurlData))):
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED,
                                 "LinearForm(Identity, DataFunction or PolynomialData) and "
                                 "LinearForm(ReconstructionIdentity(HDIVRT0), GradientCurlData) are built")
        self.operator, self.data, self.factor, self.name = operator, data, factor, name

    def lform(self, fes):
        lf = _abi.LForm()
        lf.quadorder = fes.FEType.order + self.data.bonus_quadorder
        lf.factor = float(self.factor)
        if isinstance(self.operator, ReconstructionIdentity):
            lf.kind = _abi.LF_RECO_RT0
            lf.nmodes = self.data.K
            lf.modes = self.data.modes.ctypes.data_as(_abi.DP)
            for k, v in enumerate(self.data.poly[:4]):
                lf.poly[k] = v
        elif isinstance(self.data, PolynomialData):
            if self.data.ncomp != fes.FEType.ncomp:
                raise _abi.GrmpError(_abi.ERR_ARG, "LinearForm: data components differ from the space's")
            lf.kind = _abi.LF_POLY
            lf.coef = self.data.c.ctypes.data_as(_abi.DP)
            lf.ncoef = self.data.ncoef
        else:
            lf.kind = _abi.LF_CONST
            for k, v in enumerate(self.data.c[:4]):
                lf.f[k] = float(v)
        return lf


class NonlinearDiffusion:
    """The kernel (result, input) -> result = alpha(u) grad u, input =
    [u, grad u], with alpha(u) = c0 + c1 u + c2 u^2: a closed-form stand-in
    for the Julia closure of examples/Example205_NonlinearPoisson2D.jl:57-61,
    72-76 (q = 1: alpha(u) = 1 + u^2), whose Jacobian ForwardDiff forms there."""

    def __init__(self, c0=1.0, c1=0.0, c2=1.0):
        self.c = (float(c0), float(c1), float(c2))

    def alpha(self, u):
        return self.c[0] + self.c[1] * u + self.c[2] * u * u


class NonlinearForm(PDEOperator):
    """NonlinearForm(Gradient, [Identity, Gradient], [0, 0], kernel, [2, 3];
    bonus_quadorder, factor) (src/pdeoperators.jl:590-635) with a
    NonlinearDiffusion kernel: assembled by full_assemble! with Newton's
    linearization at the current iterate (src/assemblypatterns/nonlinearform.jl:
    44-246): the matrix (J [phi, grad phi]) . grad v and the right-hand side
    (J [u, grad u] - kernel) . grad v.  Unknown indices are 0-based."""

    nonlinear = True

    def __init__(self, operator_test, operators_current, coeff_from, action_kernel, argsizes, bonus_quadorder=0,
                 factor=1.0, name="auto"):
        if operator_test != Gradient or list(operators_current) != [Identity, Gradient] or \
                not isinstance(action_kernel, NonlinearDiffusion) or list(argsizes)[:2] != [2, 3]:
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "NonlinearForm(Gradient, [Identity, Gradient], [j, j], "
                                 "NonlinearDiffusion, [2, 3]) is built")
        if len(set(int(c) for c in coeff_from)) != 1:
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "NonlinearForm: both arguments from the same unknown")
        self.coeff_from = [int(c) for c in coeff_from]
        self.kernel, self.bonus_quadorder, self.factor, self.name = action_kernel, int(bonus_quadorder), factor, name

    def quadorder(self, fes, fes_col=None):
        """polynomial orders of Identity, Gradient and the test Gradient plus
        the bonus (src/assemblypatterns.jl:559-563)"""
        p = fes.FEType.order
        return p + (p - 1) + (p - 1) + self.bonus_quadorder

    def form(self, fes, rb, cb, ctx, fes_col=None, current=None):
        if current is None:
            raise _abi.GrmpError(_abi.ERR_ARG, "NonlinearForm: the current iterate is needed")
        f = _abi.Form()
        f.kind = _abi.BLF_NLDIFF
        f.quadorder = self.quadorder(fes)
        f.row_block, f.col_block = rb, cb
        f.factor = float(self.factor)
        for k in range(3):
            f.params[k] = self.kernel.c[k]
        f.coef_mode = _abi.COEF_FACTOR
        f.current = current.vec.h
        f.current_offset = int(current.offsets[self.coeff_from[0]])
        return f, None

    def newton_lform(self, fes, current):
        lf = _abi.LForm()
        lf.kind = _abi.LF_NEWTON
        lf.quadorder = self.quadorder(fes)
        lf.factor = float(self.factor)
        for k in range(3):
            lf.f[k] = self.kernel.c[k]
        lf.current = current.vec.h
        lf.current_offset = int(current.offsets[self.coeff_from[0]])
        return lf


def assemble_linearform(b, j, lf):
    """assemble!(b[j], LinearForm) (src/assemblypatterns/linearform.jl:239-251)."""
    fes = b.FES[j]
    form = lf.lform(fes)
    b.ctx.check(lib().grmp_assemble_linearform(b.ctx.h, fes.h, ctypes.byref(form), b.vec.h, int(b.offsets[j])))


class HomogeneousDirichletBoundary:
    """BoundaryData with zero values on all boundary faces
    (src/boundarydata.jl:198-257)."""


class InterpolateDirichletBoundary:
    """Boundary dofs take the data's values at their points, vertices and P2
    face midpoints (src/boundarydata.jl:198-257)."""


class BestapproxDirichletBoundary:
    """Boundary dofs from the L2 best approximation of the data on the
    boundary faces (src/boundarydata.jl:259-352)."""


class PolynomialData:
    """Closed-form stand-in for a DataFunction of (x, y) on the device
    (src/userdata.jl:3-63): g = sum_k c_k m_k over the monomials
    1, x, y, x^2, xy, y^2, x^3, x^2 y, x y^2, y^3, x^4, x^3 y, x^2 y^2, x y^3,
    y^4 (at most 15 coefficients; boundary data takes up to degree 3, 10
    coefficients); a (ncomp, ncoef) array gives vector-valued data with 2
    components at most.  bonus_quadorder defaults to the degree."""

    MONOMIALS = ((0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3),
                 (4, 0), (3, 1), (2, 2), (1, 3), (0, 4))

    def __init__(self, coeffs, bonus_quadorder=None):
        c = np.asarray(coeffs, dtype=np.float64)
        c = c.reshape(1, -1) if c.ndim <= 1 else c
        if c.ndim != 2 or not 1 <= c.shape[0] <= 2 or not 1 <= c.shape[1] <= 15:
            raise _abi.GrmpError(_abi.ERR_ARG, "PolynomialData: 1..15 coefficients for each of 1 or 2 components")
        self.ncomp, self.ncoef = c.shape
        self.c = np.ascontiguousarray(c.ravel())
        nz = [k for k in range(self.ncoef) if np.any(c[:, k] != 0.0)]
        self.degree = max((sum(self.MONOMIALS[k]) for k in nz), default=0)
        self.bonus_quadorder = self.degree if bonus_quadorder is None else int(bonus_quadorder)

    def __call__(self, x, y, comp=0):
        row = self.c[comp * self.ncoef:(comp + 1) * self.ncoef]
        return sum(ck * x ** i * y ** j for ck, (i, j) in zip(row, self.MONOMIALS))


class L2ErrorIntegrator:
    """L2ErrorIntegrator(compare_data, Identity; quadorder = "auto", factor = 1)
    (src/assemblypatterns/itemintegrator.jl:40-89): integrates
    |compare_data(x) - factor u_h(x)|^2 over the cells; ``evaluate`` returns
    the sum (the squared L2 error), ``evaluate_cells`` the per-cell values of
    evaluate!.  compare_data is a PolynomialData (a Julia closure cannot run
    on the device).  The quadrature order is the element's polynomial order
    plus quadorder, "auto" meaning twice the data's bonus_quadorder
    (src/assemblypatterns.jl:559-563)."""

    def __init__(self, compare_data, operator=Identity, quadorder="auto", factor=1.0, name="auto"):
        if operator != Identity:
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, f"L2ErrorIntegrator: operator {operator} is not built")
        self.data = compare_data
        self.ncomp = compare_data.ncomp if compare_data is not None else None
        self.quadorder = 2 * compare_data.bonus_quadorder if quadorder == "auto" else int(quadorder)
        self.factor = float(factor)
        self.name = f"L2 error" if name == "auto" else name

    def _run(self, block, cells):
        fes = block.FES
        if self.ncomp is not None and self.ncomp != fes.FEType.ncomp:
            raise _abi.GrmpError(_abi.ERR_ARG, f"{self.name}: data with {self.ncomp} components for a space with "
                                 f"{fes.FEType.ncomp}")
        coef = self.data.c if self.data is not None else np.zeros(1)
        ncoef = self.data.ncoef if self.data is not None else 0
        out = np.empty(fes.xgrid.ncells) if cells else None
        tot = ctypes.c_double()
        block.parent.ctx.check(lib().grmp_integrate_l2error(
            block.parent.ctx.h, fes.h, block.parent.vec.h, block.offset, _abi.dptr(coef), ncoef,
            fes.FEType.order + self.quadorder, self.factor, _abi.dptr(out) if cells else None, ctypes.byref(tot)))
        return out if cells else tot.value


def L2NormIntegrator(ncomponents, operator=Identity, quadorder=2, name="L2 norm"):
    """L2NormIntegrator(ncomponents, Identity; quadorder = 2)
    (src/assemblypatterns/itemintegrator.jl:96-128): the squared L2 norm,
    the L2 error against zero data."""
    I = L2ErrorIntegrator(None, operator, quadorder=quadorder, name=name)
    I.ncomp = int(ncomponents)
    return I


def evaluate(integrator, block):
    """evaluate(ItemIntegrator, FEVectorBlock) (itemintegrator.jl:323-360):
    the sum over all cells."""
    return integrator._run(block, cells=False)


def evaluate_cells(integrator, block):
    """evaluate!(b, ItemIntegrator, FEVectorBlock) (itemintegrator.jl:169-307):
    one value per cell."""
    return integrator._run(block, cells=True)


class FixedIntegralMean:
    """FixedIntegralMean(unknown, value) (src/globalconstraints.jl:23-44):
    the unknown's first dof is fixed during the solve (apply, :259-283) and
    the integral mean is shifted to ``value`` afterwards (realize, :337-351)."""

    def __init__(self, unknown, value=0.0):
        self.unknown, self.value = unknown, float(value)


class PDEDescription:
    """Unknowns, LHS operators per block, RHS operators per block, boundary
    data and global constraints (src/pdedescription.jl:40-284)."""

    def __init__(self, name="PDE"):
        self.name = name
        self.unknowns = []
        self.LHS = {}
        self.RHS = {}
        self.BoundaryOperators = {}
        self.GlobalConstraints = []

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

    def add_boundarydata(self, j, kind=HomogeneousDirichletBoundary, data=None):
        """add_boundarydata!(PDE, j, regions, kind; data) on all boundary
        regions (src/pdedescription.jl); ``data`` is a PolynomialData for the
        interpolated and best-approximated kinds."""
        if kind is not HomogeneousDirichletBoundary and not isinstance(data, PolynomialData):
            raise _abi.GrmpError(_abi.ERR_ARG, f"{kind.__name__} needs PolynomialData")
        self.BoundaryOperators[j] = (kind, data)

    def add_constraint(self, c):
        self.GlobalConstraints.append(c)


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


def IncompressibleNavierStokesProblem(dim=2, viscosity=1.0, nonlinear=False):
    """The linear Stokes prototype (src/pdeprototypes.jl:17-55): unknowns
    velocity and pressure, LaplaceOperator(viscosity) on the velocity,
    LagrangeMultiplier(Divergence) coupling, no-slip boundary data and
    FixedIntegralMean(pressure, 0)."""
    if dim != 2 or nonlinear:
        raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "only the linear 2D Stokes problem is built")
    pde = PDEDescription("incompressible Stokes problem")
    pde.add_unknown("velocity")
    pde.add_unknown("pressure")
    pde.add_operator((0, 0), LaplaceOperator(viscosity))
    pde.add_operator((0, 1), LagrangeMultiplier(Divergence))
    pde.add_boundarydata(0, HomogeneousDirichletBoundary)
    pde.add_constraint(FixedIntegralMean(1, 0.0))
    return pde


class SolverConfig:
    """Keyword configuration of solve (src/solvers.jl:38-59)."""

    def __init__(self, precision="fp64", rtol=1e-10, maxiterations=1000000, inner_rtol=0.1,
                 fixed_penalty=1e60, minres_inner_rtol=1e-4, target_residual=1e-12, nonlinear_maxiterations=10,
                 damping=0.0, newton_preconditioner="auto", anderson_iterations=0, anderson_metric="l2",
                 anderson_damping=1.0, anderson_unknowns=(0,)):
        self.precision, self.rtol, self.maxiterations = precision, rtol, maxiterations
        self.inner_rtol, self.fixed_penalty = inner_rtol, fixed_penalty
        self.minres_inner_rtol = minres_inner_rtol
        # fixpoint iterations of nonlinear problems (default_solver_kwargs, src/solvers.jl:37-59)
        self.target_residual, self.nonlinear_maxiterations = target_residual, nonlinear_maxiterations
        self.damping = damping
        # the Newton steps' BiCGStab preconditioner: "amg" (one algebraic multigrid V-cycle, hierarchy built
        # from the first Newton matrix), "jacobi", or "auto" (amg, jacobi where the hierarchy cannot be built)
        if newton_preconditioner not in ("auto", "amg", "jacobi"):
            raise _abi.GrmpError(_abi.ERR_ARG, f"newton_preconditioner {newton_preconditioner!r}: "
                                 "'auto', 'amg' or 'jacobi'")
        self.newton_preconditioner = newton_preconditioner
        # Anderson acceleration of the fixpoint loop (src/solvers.jl:49-52); unknowns are 0-based here
        if anderson_metric not in ("l2", "L2", "H1"):
            raise _abi.GrmpError(_abi.ERR_ARG, f"anderson_metric {anderson_metric!r}: 'l2', 'L2' or 'H1'")
        if not (0 <= int(anderson_iterations) <= 10) or not (0.0 < anderson_damping <= 1.0):
            raise _abi.GrmpError(_abi.ERR_ARG, f"anderson_iterations {anderson_iterations} not in 0..10 or "
                                 f"anderson_damping {anderson_damping} not in (0, 1]")
        self.anderson_iterations, self.anderson_metric = int(anderson_iterations), anderson_metric
        self.anderson_damping = float(anderson_damping)
        self.anderson_unknowns = tuple(int(u) for u in anderson_unknowns)


def fixed_dofs(pde, Solution):
    """Boundary dofs of every unknown with boundary data, plus the first dof
    of every unknown under a FixedIntegralMean."""
    out = []
    for j in sorted(pde.BoundaryOperators):
        fes = Solution.FES[j]
        out.append(fes.boundary_dofs() + int(Solution.offsets[j]))
    for c in pde.GlobalConstraints:
        out.append(np.array([int(Solution.offsets[c.unknown])], dtype=np.int64))
    return np.unique(np.concatenate(out)) if out else np.zeros(0, dtype=np.int64)


def realize_constraints(pde, Solution):
    """realize_constraint!(FixedIntegralMean) (src/globalconstraints.jl:337-351):
    shift the unknown by (value - integral mean); the cell integrals come from
    the device ItemIntegrator."""
    for c in pde.GlobalConstraints:
        fes = Solution.FES[c.unknown]
        off = int(Solution.offsets[c.unknown])
        ints = np.empty(fes.xgrid.ncells * fes.FEType.ncomp)
        Solution.ctx.check(lib().grmp_integrate_identity(Solution.ctx.h, fes.h, 2 * fes.FEType.order, Solution.vec.h,
                                                         off, _abi.dptr(ints)))
        area = float(np.sum(fes.xgrid.cellvolumes()))
        shift = c.value - float(np.sum(ints)) / area
        x = Solution.entries()
        x[off:off + fes.ndofs] += shift
        Solution.set_entries(x)


def boundary_values(pde, Solution, rtol=1e-14):
    """Target of boundarydata! (src/boundarydata.jl:198-352): a device vector
    with the Dirichlet values on the boundary dofs of every unknown with
    non-homogeneous data, or None when all data is homogeneous."""
    target = None
    for j, (kind, data) in sorted(pde.BoundaryOperators.items()):
        if kind is HomogeneousDirichletBoundary:
            continue
        if target is None:
            target = _abi.DeviceVector(Solution.ctx, Solution.n)
            target.fill(0.0)
        fes, off = Solution.FES[j], int(Solution.offsets[j])
        if data.ncomp != fes.FEType.ncomp:
            raise _abi.GrmpError(_abi.ERR_ARG, f"boundary data with {data.ncomp} components for a space with "
                                 f"{fes.FEType.ncomp}")
        if kind is InterpolateDirichletBoundary:
            rc = lib().grmp_boundary_interpolate(fes.h, _abi.dptr(data.c), data.ncoef, target.h, off)
        elif kind is BestapproxDirichletBoundary:
            it, rr = ctypes.c_int64(), ctypes.c_double()
            q = fes.FEType.order + data.bonus_quadorder
            rc = lib().grmp_boundary_bestapprox(fes.h, _abi.dptr(data.c), data.ncoef, q, rtol, target.h, off,
                                                ctypes.byref(it), ctypes.byref(rr))
        else:
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, f"boundary data {kind} is not built")
        Solution.ctx.check(rc)
    return target


def _saddle_point(pde, Solution):
    """True when an unknown has no operator on its diagonal block (a Lagrange
    multiplier), which makes the system indefinite."""
    return any(not pde.LHS.get((j, j)) for j in range(len(Solution.FES)))


def assemble(pde, Solution, values=_abi.VALUES_FP64, penalty=1e60):
    """System assemble! (src/solvers.jl:397-610) plus the penalty of
    solve_direct! (:647-652): returns (A::FEMatrix, b::FEVector, fixed)."""
    A = FEMatrix(Solution.FES)
    for (i, j), ops in sorted(pde.LHS.items()):
        for op in ops:
            form, keep = op.form(Solution.FES[i], i, j, A.ctx, Solution.FES[j])
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
        target = boundary_values(pde, Solution)
        if target is not None:
            A.ctx.check(lib().grmp_matrix_penalize_rhs(A.h, target.h, b.vec.h))
    return A, b, fixed


def mg_eligible(pde, Solution):
    """True when grmp_mg_solve applies: one unknown, H1P1 (1 or 2
    components) or scalar H1P2 on a generated unit square, Dirichlet data on
    the whole boundary (inhomogeneous data is lifted: the solve runs with
    homogeneous data), no global constraints, at most one cell-coefficient
    operator."""
    if len(Solution.FES) != 1 or pde.GlobalConstraints or list(pde.LHS) != [(0, 0)]:
        return False
    fes = Solution.FES[0]
    ft = fes.FEType
    if fes.xgrid.kind != "unitsquare" or not ((ft.code == _abi.FE_H1P1 and ft.ncomp in (1, 2)) or
                                              (ft.code == _abi.FE_H1P2 and ft.ncomp == 1)):
        return False
    if not pde.BoundaryOperators:
        return False
    return sum(getattr(op, "action", None) is not None for op in pde.LHS[(0, 0)]) <= 1


def stokes_mg_eligible(pde, Solution):
    """True when grmp_solve_minres_mg applies: velocity H1BR{2} on a
    generated unit square and a pressure, the velocity block one
    LaplaceOperator with a constant factor (the viscosity), no-slip data."""
    if len(Solution.FES) != 2 or not _saddle_point(pde, Solution):
        return False
    V = Solution.FES[0]
    if V.FEType.code != _abi.FE_H1BR or V.xgrid.kind != "unitsquare":
        return False
    ops = pde.LHS.get((0, 0), [])
    if len(ops) != 1 or ops[0].kind != _abi.BLF_LAPLACE or ops[0].action is not None:
        return False
    bd = pde.BoundaryOperators.get(0)
    return bd is not None and bd[0] is HomogeneousDirichletBoundary


def stokes_velocity_mg(pde, V, penalty=1e60):
    """The velocity block's preconditioner hierarchy: the viscosity times the
    scalar P1 Laplacian on the velocity's square (the node part of H1BR{2}),
    boundary dofs fixed, and its coarser levels; assembled once."""
    nu = float(pde.LHS[(0, 0)][0].factor)
    fes = FESpace(H1P1(1), V.xgrid)
    A0 = FEMatrix([fes])
    op = LaplaceOperator(nu)
    form, keep = op.form(fes, 0, 0, fes.ctx)
    A0.add_form(form, keep)
    A0.set_fixed(fes.boundary_dofs(), penalty)
    A0.build_pattern()

    class _Velocity:
        LHS = {(0, 0): [op]}
    levels, kappa0 = mg_hierarchy(_Velocity, fes, A0, penalty=penalty)
    return MultigridSolver(levels, kappa0)


def mg_hierarchy(pde, fes0, A0, extra_ops=(), penalty=1e60):
    """Levels for grmp_mg_solve: level 0 is the assembled system A0 over
    fes0; then H1P1 on the same square when fes0 is H1P2, and H1P1 on the
    squares of ceil(N/2) down to N <= 4, each with pde's operators (a cell
    coefficient as a placeholder: the solve restricts the fine one) and
    ``extra_ops``, its boundary dofs fixed.  Returns (levels, kappa0)."""
    ctx = fes0.ctx
    ncomp = fes0.FEType.ncomp
    N = fes0.xgrid.N
    Ns = [N] if fes0.FEType.code == _abi.FE_H1P2 else []
    while N > 4:
        N = (N + 1) // 2
        Ns.append(N)
    ops = list(pde.LHS[(0, 0)]) + list(extra_ops)
    levels = [(A0, fes0)]
    for Nl in Ns:
        grid = fes0.xgrid if Nl == fes0.xgrid.N else grid_unitsquare(Nl, ctx)
        fes = FESpace(H1P1(ncomp), grid)
        A = FEMatrix([fes])
        for op in ops:
            if getattr(op, "action", None) is not None:
                ph = _abi.DeviceVector(ctx, grid.ncells)
                ph.fill(1.0)
                op = copy.copy(op)
                op.action = CellCoefficient(ph, bonus_quadorder=op.action.bonus_quadorder)
            form, keep = op.form(fes, 0, 0, ctx)
            A.add_form(form, keep)
        A.set_fixed(fes.boundary_dofs(), penalty)
        A.build_pattern()
        levels.append((A, fes, grid))
    act = [op.action for op in ops if getattr(op, "action", None) is not None]
    if act:
        kappa0 = act[0].device(ctx, fes0.xgrid.ncells)
    else:
        kappa0 = _abi.DeviceVector(ctx, fes0.xgrid.ncells)
        kappa0.fill(1.0)
    return levels, kappa0


class MultigridSolver:
    """A hierarchy from mg_hierarchy assembled once (grmp_mg_create), then
    solved for any number of right-hand sides (grmp_mg_apply), as the
    reference reuses its factorization while the matrix is unchanged."""

    def __init__(self, levels, kappa0):
        self.levels, self.kappa0 = levels, kappa0  # kept alive: the handle reads them
        self.ctx = levels[0][1].ctx
        n = len(levels)
        mats = (_abi.VP * n)(*[lv[0].h for lv in levels])
        sps = (_abi.VP * n)(*[lv[1].h for lv in levels])
        h = _abi.VP()
        self.ctx.check(lib().grmp_mg_create(self.ctx.h, n, mats, sps, kappa0.h, ctypes.byref(h)))
        self.h = h

    def solve(self, b, x, rtol, maxit=500):
        info = _abi.SolveInfo()
        self.ctx.check(lib().grmp_mg_apply(self.h, b.vec.h, x.vec.h, float(rtol), int(maxit), ctypes.byref(info)))
        return info

    def __del__(self):
        try:
            if self.h and _abi.alive():
                lib().grmp_mg_destroy(self.h)
            self.h = None
        except Exception:
            pass


class AlgebraicMultigridSolver:
    """Algebraic multigrid (grmp_amg_create) for an assembled FP64 system on
    any mesh, then FP64 CG with its V-cycle per right-hand side
    (grmp_amg_apply); b is the penalized right-hand side, fixed dofs are
    eliminated.  A must stay alive and unchanged while the handle lives."""

    def __init__(self, A):
        self.A, self.ctx = A, A.ctx
        h = _abi.VP()
        self.ctx.check(lib().grmp_amg_create(A.h, ctypes.byref(h)))
        self.h = h

    def solve(self, b, x, rtol, maxit=2000):
        info = _abi.SolveInfo()
        self.ctx.check(lib().grmp_amg_apply(self.h, b.vec.h, x.vec.h, float(rtol), int(maxit), ctypes.byref(info)))
        return info

    def info(self):
        """(levels, rows of the coarsest level, operator complexity)"""
        nl, cr, cx = ctypes.c_int(), ctypes.c_int64(), ctypes.c_double()
        self.ctx.check(lib().grmp_amg_info(self.h, ctypes.byref(nl), ctypes.byref(cr), ctypes.byref(cx)))
        return nl.value, cr.value, cx.value

    def __del__(self):
        try:
            if self.h and _abi.alive():
                lib().grmp_amg_destroy(self.h)
            self.h = None
        except Exception:
            pass


def amg_eligible(pde, Solution):
    """True when the algebraic multigrid CG applies: a linear stationary
    problem without Lagrange multipliers (symmetric positive definite on the
    free dofs)."""
    return not _saddle_point(pde, Solution) and not is_nonlinear(pde)


def lift_dirichlet(A, b, target, out):
    """out = b - A target on the free rows of A (penalty rows as target
    times the penalty; the multigrid solve zeroes them): the right-hand side
    for the correction x - target, which has homogeneous data."""
    A.ctx.check(lib().grmp_timestep_rhs(A.h, A.h, b.vec.h, target.h, -1.0, out.vec.h))
    return out


def mg_solve(levels, kappa0, b, x, rtol, maxit=500):
    """grmp_mg_solve on a hierarchy from mg_hierarchy: x = S^{-1} b."""
    n = len(levels)
    mats = (_abi.VP * n)(*[lv[0].h for lv in levels])
    sps = (_abi.VP * n)(*[lv[1].h for lv in levels])
    info = _abi.SolveInfo()
    rc = lib().grmp_mg_solve(x.ctx.h, n, mats, sps, kappa0.h, b.vec.h, x.vec.h, float(rtol), int(maxit),
                             ctypes.byref(info))
    x.ctx.check(rc)
    return info


def solve(Solution, pde, config=None, solver="krylov", **kwargs):
    """solve!(Solution, PDE) for linear stationary problems
    (src/solvers.jl:1354-1402 -> solve_direct! :622-686).  Returns the
    SolveInfo of the device solve.  solver="krylov": the device Krylov
    solvers (CG / MINRES, FP64 or FP32 inner); "mg": the multigrid V-cycle CG
    (FP64) where mg_eligible holds, and for the Stokes system MINRES with the
    multigrid velocity block (stokes_mg_eligible); "amg": CG with algebraic
    multigrid on any mesh (amg_eligible); "auto": the geometric multigrid where
    it applies, else (FP64) the algebraic one, else Krylov."""
    cfg = config or SolverConfig(**kwargs)
    if solver not in ("auto", "krylov", "mg", "amg"):
        raise _abi.GrmpError(_abi.ERR_ARG, f"solver {solver!r}: 'auto', 'krylov', 'mg' or 'amg'")
    if is_nonlinear(pde):
        if solver == "mg":
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "nonlinear problems are solved by the Newton fixpoint loop")
        return solve_fixpoint(Solution, pde, cfg)
    use_mg = solver in ("auto", "mg") and mg_eligible(pde, Solution)
    use_amg = (solver == "amg" or (solver == "auto" and not use_mg and cfg.precision == "fp64")) and \
        amg_eligible(pde, Solution)
    if solver == "amg" and not use_amg:
        raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "algebraic multigrid needs a linear problem without Lagrange "
                             "multipliers")
    if use_amg:
        A, b, fixed = assemble(pde, Solution, _abi.VALUES_FP64, cfg.fixed_penalty)
        amg = AlgebraicMultigridSolver(A)
        info = amg.solve(b, Solution, cfg.rtol, min(cfg.maxiterations, 2000))
        realize_constraints(pde, Solution)
        return info, A, b
    if solver == "mg" and not use_mg and not stokes_mg_eligible(pde, Solution):
        raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "multigrid needs one H1P1 / H1P2 unknown on a generated unit "
                             "square, Dirichlet data and no global constraints")
    fp32 = cfg.precision == "fp32_ir" and not use_mg
    values = _abi.VALUES_FP64 | (_abi.VALUES_FP32 if fp32 else 0)
    A, b, fixed = assemble(pde, Solution, values, cfg.fixed_penalty)
    if use_mg:
        levels, kappa0 = mg_hierarchy(pde, Solution.FES[0], A, penalty=cfg.fixed_penalty)
        target = boundary_values(pde, Solution)
        rhs = b if target is None else lift_dirichlet(A, b, target, FEVector(Solution.FES))
        info = mg_solve(levels, kappa0, rhs, Solution, cfg.rtol, min(cfg.maxiterations, 500))
        if target is not None:
            A.ctx.check(lib().grmp_vector_axpy(1.0, target.h, Solution.vec.h))
        return info, A, b
    method = _abi.SOLVE_MINRES if _saddle_point(pde, Solution) else _abi.SOLVE_CG
    if solver != "krylov" and method == _abi.SOLVE_MINRES and cfg.precision == "fp64" and \
            stokes_mg_eligible(pde, Solution):
        mg = stokes_velocity_mg(pde, Solution.FES[0], cfg.fixed_penalty)
        info = solve_system(A, b, Solution, cfg, method, velocity_mg=mg)
    else:
        info = solve_system(A, b, Solution, cfg, method)
    realize_constraints(pde, Solution)
    return info, A, b


def _newton_solve_amg(A, b, Solution, cfg, pc):
    """grmp_solve_bicgstab_amg from the current iterate: BiCGStab with one
    algebraic multigrid V-cycle of pc as the right preconditioner."""
    o = _abi.SolveOpts()
    lib().grmp_solve_opts_default(ctypes.byref(o))
    o.method = _abi.SOLVE_BICGSTAB
    o.precision = _abi.PRECISION_FP64
    o.rtol = cfg.rtol
    o.maxit = int(cfg.maxiterations)
    o.initial_guess = 1
    info = _abi.SolveInfo()
    rc = lib().grmp_solve_bicgstab_amg(A.h, b.vec.h, Solution.vec.h, ctypes.byref(o), pc.h, ctypes.byref(info))
    if rc not in (_abi.OK,):
        A.ctx.check(rc)
    return info


def solve_system(A, b, Solution, cfg, method=_abi.SOLVE_CG, initial_guess=False, velocity_mg=None):
    o = _abi.SolveOpts()
    lib().grmp_solve_opts_default(ctypes.byref(o))
    o.method = method
    o.precision = _abi.PRECISION_FP32_IR if cfg.precision == "fp32_ir" else _abi.PRECISION_FP64
    o.rtol = cfg.rtol
    o.maxit = int(cfg.maxiterations)
    o.inner_rtol = cfg.inner_rtol if method == _abi.SOLVE_CG else cfg.minres_inner_rtol
    o.check_every = 16 if method == _abi.SOLVE_MINRES else 1
    o.initial_guess = 1 if initial_guess else 0
    info = _abi.SolveInfo()
    if velocity_mg is not None:  # MINRES with the multigrid velocity block (Stokes)
        ncomp = Solution.FES[0].FEType.ncomp
        rc = lib().grmp_solve_minres_mg(A.h, b.vec.h, Solution.vec.h, ctypes.byref(o), velocity_mg.h, ncomp,
                                        ctypes.byref(info))
    else:
        rc = lib().grmp_solve(A.h, b.vec.h, Solution.vec.h, ctypes.byref(o), ctypes.byref(info))
    if rc not in (_abi.OK,):
        A.ctx.check(rc)
    return info


# ------------------------------------------------------------ nonlinear
def is_nonlinear(pde):
    return any(getattr(op, "nonlinear", False) for ops in pde.LHS.values() for op in ops)


class FixpointInfo:
    """Statistics of solve_fixpoint: iterations, the nonlinear residual of
    each, and the linear solves' iterations."""

    def __init__(self):
        self.status, self.iterations, self.residuals, self.linear_iterations = 0, 0, [], []
        self.linear_restarts = []  # BiCGStab restarts after near-breakdowns, per Newton step
        self.preconditioner = None  # "amg" or "jacobi": the Newton steps' BiCGStab preconditioner

    @property
    def relres(self):
        return self.residuals[-1] if self.residuals else float("nan")


def _nl_system(pde, Solution, A):
    """values of A and the right-hand side at the current iterate: linear
    forms, then the Newton right-hand sides of the nonlinear forms"""
    A.assemble(_abi.VALUES_FP64)
    b = FEVector(Solution.FES)
    b.vec.fill(0.0)
    for j, ops in sorted(pde.RHS.items()):
        for op in ops:
            assemble_linearform(b, j, op)
    for (i, j), ops in sorted(pde.LHS.items()):
        for op in ops:
            if getattr(op, "nonlinear", False):
                lf = op.newton_lform(Solution.FES[i], Solution)
                b.ctx.check(lib().grmp_assemble_linearform(b.ctx.h, Solution.FES[i].h, ctypes.byref(lf), b.vec.h,
                                                           int(b.offsets[i])))
    return b


class AndersonAccelerator:
    """AndersonAccelerationManager (src/solvers.jl:688-746) on the device:
    grmp_anderson_create with the norm matrix of the metric ("L2": mass,
    "H1": gradient stiffness, assembled on the device for the accelerated
    unknowns; "l2": none) and the entry ranges of those unknowns.  update()
    is update_anderson! (:748-808) on the FEVector given at construction."""

    def __init__(self, Solution, depth, metric="l2", damping=1.0, unknowns=(0,)):
        self.ctx = Solution.vec.ctx
        unknowns = sorted(set(int(u) for u in unknowns))
        if not unknowns or unknowns[0] < 0 or unknowns[-1] >= len(Solution.FES):
            raise _abi.GrmpError(_abi.ERR_ARG, f"anderson_unknowns {unknowns} outside 0..{len(Solution.FES) - 1}")
        self.metric = None
        if metric != "l2":
            B = FEMatrix(Solution.FES)
            op = ReactionOperator(1.0) if metric == "L2" else LaplaceOperator(1.0)
            for u in unknowns:
                form, keep = op.form(Solution.FES[u], u, u, B.ctx, Solution.FES[u])
                B.add_form(form, keep)
            B.build_pattern()
            B.assemble(_abi.VALUES_FP64)
            self.metric = B
        r = []
        for u in unknowns:
            r += [int(Solution.offsets[u]), int(Solution.offsets[u]) + int(Solution.FES[u].ndofs)]
        self.ranges = np.array(r, dtype=np.int64)
        self.depth = int(depth)
        self.last_alpha = np.zeros(self.depth + 1)
        self.last_local_residual = 0.0
        h = _abi.VP()
        self.ctx.check(lib().grmp_anderson_create(
            self.ctx.h, int(Solution.n), self.depth, float(damping), self.metric.h if self.metric else None,
            len(unknowns), self.ranges.ctypes.data, Solution.vec.h, ctypes.byref(h)))
        self.h = h

    def update(self, Solution):
        alpha = np.zeros(self.depth + 1)
        res = ctypes.c_double()
        self.ctx.check(lib().grmp_anderson_update(self.h, Solution.vec.h, alpha.ctypes.data, ctypes.byref(res)))
        self.last_alpha, self.last_local_residual = alpha, res.value
        if res.value > 1e-4:
            warnings.warn(f"Anderson acceleration local system residual unsatisfactory ({res.value} > 1e-4)")
        return alpha

    def __del__(self):
        try:
            if self.h and _abi.alive():
                lib().grmp_anderson_destroy(self.h)
            self.h = None
        except Exception:
            pass


def solve_fixpoint(Solution, pde, cfg):
    """solve_fixpoint_full! (src/solvers.jl:812-1017) for PDEs with
    NonlinearForms: assemble at the initial iterate, set the boundary data,
    then per iteration penalize the fixed dofs, solve the Newton system on
    the device (BiCGStab: the Newton matrix is nonsymmetric), optionally
    accelerate (Anderson, grmp_anderson_update) and damp, realize
    constraints, reassemble at the new iterate and stop when the nonlinear
    residual ||A x - b|| (fixed rows zeroed) drops below target_residual.
    Returns (FixpointInfo, A, b)."""
    if cfg.precision != "fp64":
        raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "nonlinear problems are solved in FP64")
    A = FEMatrix(Solution.FES)
    for (i, j), ops in sorted(pde.LHS.items()):
        for op in ops:
            if getattr(op, "nonlinear", False):
                form, keep = op.form(Solution.FES[i], i, j, A.ctx, Solution.FES[j], current=Solution)
            else:
                form, keep = op.form(Solution.FES[i], i, j, A.ctx, Solution.FES[j])
            A.add_form(form, keep)
    fixed = fixed_dofs(pde, Solution)
    A.set_fixed(fixed, cfg.fixed_penalty)
    A.build_pattern()
    b = _nl_system(pde, Solution, A)
    if fixed.size:
        target = boundary_values(pde, Solution)
        x = Solution.entries()
        x[fixed] = 0.0 if target is None else target.download()[fixed]
        Solution.set_entries(x)
    info = FixpointInfo()
    aa = None
    if cfg.anderson_iterations > 0:  # after the boundary data, as the reference builds it (src/solvers.jl:849-859)
        aa = AndersonAccelerator(Solution, cfg.anderson_iterations, cfg.anderson_metric, cfg.anderson_damping,
                                 cfg.anderson_unknowns)
    use_amg = cfg.newton_preconditioner in ("auto", "amg")
    pc = None
    last = Solution.entries() if cfg.damping > 0 else None
    res = FEVector(Solution.FES)
    mask = np.zeros(Solution.n, dtype=bool)
    mask[fixed] = True
    for it in range(1, cfg.nonlinear_maxiterations + 1):
        if fixed.size:  # b[fixed] = penalty * Target[fixed]
            A.ctx.check(lib().grmp_matrix_penalize_rhs(A.h, Solution.vec.h, b.vec.h))
        # BiCGStab from the current iterate (the Newton step is the full system for the new iterate),
        # preconditioned by the multigrid hierarchy of the first Newton matrix.  Its V-cycle uses the
        # values copied when it was built; later Newton matrices differ by a bounded coefficient
        # (1 + u^2) and a lower-order nonsymmetric term, so it stays a good preconditioner, while
        # hierarchies rebuilt from them degrade: their coarse Galerkin levels turn convection-dominated.
        li = None
        if use_amg and pc is None:
            try:
                pc = AlgebraicMultigridSolver(A)
            except _abi.GrmpError as e:
                if cfg.newton_preconditioner == "amg" or e.code != _abi.ERR_UNSUPPORTED:
                    raise
                use_amg = False
        if pc is not None:
            li = _newton_solve_amg(A, b, Solution, cfg, pc)
        if li is None:
            li = solve_system(A, b, Solution, cfg, _abi.SOLVE_BICGSTAB, initial_guess=True)
        info.preconditioner = "amg" if use_amg else "jacobi"
        info.linear_iterations.append(int(li.iterations))
        info.linear_restarts.append(int(li.outer) - 1)
        if aa is not None:  # POSTPROCESS: ANDERSON ITERATE, before the damping (src/solvers.jl:923-938)
            aa.update(Solution)
        if cfg.damping > 0:
            x = cfg.damping * last + (1.0 - cfg.damping) * Solution.entries()
            Solution.set_entries(x)
            last = x
        realize_constraints(pde, Solution)
        b = _nl_system(pde, Solution, A)
        # residual A x - b on the free rows (the step kernel forms b - A x)
        A.ctx.check(lib().grmp_timestep_rhs(A.h, A.h, b.vec.h, Solution.vec.h, -1.0, res.vec.h))
        r = res.entries()
        r[mask] = 0.0
        resnorm = float(np.linalg.norm(r))
        info.residuals.append(resnorm)
        info.iterations = it
        if resnorm < cfg.target_residual:
            break
    else:
        info.status = _abi.ERR_NOCONV
    del pc, aa
    return info, A, b
