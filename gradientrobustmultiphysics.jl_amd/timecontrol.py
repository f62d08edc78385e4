"""TimeControlSolver with the BackwardEuler and CrankNicolson rules on the device.

Python mirror of src/solvers_timedependent.jl for linear problems (one
subiteration holding every unknown, time-independent data):

* TimeControlSolver (:116-281): fixed dofs and the Dirichlet values of the
  initial state, the right-hand side f assembled once, the mass matrix M
  of the time-dependent equations (Identity x Identity, dt_operator's default)
  assembled once.
* advance! (:293-594): the system S = A + M/dt is assembled on the first step
  and whenever the step size changes (update_matrix, :337); every step builds
  b = f + M u^n / dt with the penalty rows b_d = penalty * u^n_d
  (:349-359, :452-461) in one device kernel (grmp_timestep_rhs), solves
  S u^{n+1} = b with the device Krylov solver and realizes global
  constraints (:578-581).
* CrankNicolson (:255-258, :376-433): S = A + (2/dt) M is kept beside the
  spatial matrix A; every step builds rhs = 2 f - A u^n + (2/dt) M u^n
  (time-independent f: b^n = b^{n+1} = f) with the penalty rows, as two
  passes of the step kernel.  Every unknown must be time-dependent (the
  reference's weighting of algebraic constraints, :384-394, is not built).
* advance_until_stationarity! (:605-665): steps until the summed per-unknown
  change norms drop below the threshold.
* advance_until_time! (:683-760): steps of the given size while the time is
  below finaltime - finaltime_tolerance; warns when the end time is missed.

Each step's system is solved by the served path's multigrid V-cycle CG
(FP64, grmp_mg_create / grmp_mg_apply) when the problem is one scalar H1P1
unknown on a generated unit square with Dirichlet data (inhomogeneous values
lifted): the coarse levels
rediscretize A + M/dt on squares of ceil(N/2) down to N <= 4.  Other
symmetric positive definite systems (uploaded meshes; no Lagrange
multipliers) are solved by the algebraic multigrid CG (grmp_amg_create once
per step size, grmp_amg_apply per step).  Otherwise, or with
solver="krylov", the device Krylov solver of solve_system runs.

Dirichlet data (homogeneous, interpolated or best-approximated) is fixed in
time.  Nonlinear operators, dt_action / dt_lump and time-dependent data are
not built and raise GrmpError(ERR_UNSUPPORTED).
"""
import warnings

import numpy as np

from . import _abi
from ._abi import lib
from .fe import FEMatrix, FEVector
from .pde import (AlgebraicMultigridSolver, ReactionOperator, SolverConfig, _saddle_point, amg_eligible,
                  assemble_linearform, boundary_values, fixed_dofs, MultigridSolver, lift_dirichlet, mg_eligible,
                  mg_hierarchy, realize_constraints, solve_system)


class BackwardEuler:
    """(u^{n+1} - u^n)/dt + A u^{n+1} = f (src/solvers_timedependent.jl:8)."""


class CrankNicolson:
    """(u^{n+1} - u^n)/dt + (A u^{n+1} + A u^n)/2 = (f^{n+1} + f^n)/2
    (src/solvers_timedependent.jl:9), scaled by 2 as the reference does."""


class TimeControlSolver:
    """TimeControlSolver(PDE, InitialValues, BackwardEuler;
    timedependent_equations) (src/solvers_timedependent.jl:116-281).
    ``InitialValues`` holds the initial state and, after each advance, the
    current solution."""

    def __init__(self, pde, InitialValues, rule=BackwardEuler, timedependent_equations=(), config=None,
                 solver="auto", **kwargs):
        if rule is not BackwardEuler and rule is not CrankNicolson:
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, f"time integration rule {getattr(rule, '__name__', rule)} "
                                 "is not built (BackwardEuler and CrankNicolson are)")
        which = sorted(set(int(d) for d in timedependent_equations))
        nu = len(InitialValues.FES)
        if not which or which[0] < 0 or which[-1] >= nu:
            raise _abi.GrmpError(_abi.ERR_ARG, "timedependent_equations must name at least one unknown "
                                 f"in 0..{nu - 1}")
        if rule is CrankNicolson and len(which) != nu:
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "CrankNicolson needs every unknown time-dependent "
                                 "(algebraic constraints are not built)")
        self.rule = rule
        self.PDE, self.X, self.which = pde, InitialValues, which
        self.SC = config or SolverConfig(**kwargs)
        self.values = _abi.VALUES_FP64 | (_abi.VALUES_FP32 if self.SC.precision == "fp32_ir" else 0)
        self.method = _abi.SOLVE_MINRES if _saddle_point(pde, InitialValues) else _abi.SOLVE_CG
        self.fixed = fixed_dofs(pde, InitialValues)
        # boundarydata!: the Dirichlet values on the boundary dofs of the initial state (:194-200);
        # the penalty rows of every step keep them
        bnd = [InitialValues.FES[j].boundary_dofs() + int(InitialValues.offsets[j])
               for j in sorted(pde.BoundaryOperators)]
        self.target = None
        if bnd:
            bnd = np.concatenate(bnd)
            self.target = boundary_values(pde, InitialValues)
            x = InitialValues.entries()
            x[bnd] = 0.0 if self.target is None else self.target.download()[bnd]
            InitialValues.set_entries(x)
        FES = InitialValues.FES
        # mass matrix of the time derivative, assembled once (dt_is_nonlinear = false)
        self.AM = FEMatrix(FES, name="mass matrix")
        for d in which:
            form, keep = ReactionOperator(1.0).form(FES[d], d, d, self.AM.ctx)
            self.AM.add_form(form, keep)
        self.AM.build_pattern()
        self.AM.assemble(_abi.VALUES_FP64)
        # right-hand side data, constant in time
        self.f = FEVector(FES, name="rhs data")
        for j, ops in sorted(pde.RHS.items()):
            for op in ops:
                assemble_linearform(self.f, j, op)
        self.b = FEVector(FES, name="system rhs")
        if rule is CrankNicolson:
            # the spatial matrix A (no penalty: its fixed rows are overwritten) and 2 f, kept for every step
            self.A = FEMatrix(FES, name="spatial matrix")
            for (i, j), ops in sorted(pde.LHS.items()):
                for op in ops:
                    form, keep = op.form(FES[i], i, j, self.A.ctx, FES[j])
                    self.A.add_form(form, keep)
            self.A.build_pattern()
            self.A.assemble(_abi.VALUES_FP64)
            self.f2 = FEVector(FES, name="2 f")
            self.f2.vec.fill(0.0)
            for _ in range(2):
                self.f2.ctx.check(lib().grmp_vector_axpy(1.0, self.f.vec.h, self.f2.vec.h))
            self._tmp = FEVector(FES, name="2 f - A u")
        # step solver: the multigrid V-cycle CG where the served hierarchy applies, else the algebraic
        # multigrid CG for symmetric positive definite systems, else the Krylov solver
        if solver not in ("auto", "krylov", "mg", "amg"):
            raise _abi.GrmpError(_abi.ERR_ARG, f"solver {solver!r}: 'auto', 'krylov', 'mg' or 'amg'")
        mg_ok = self._mg_eligible()
        amg_ok = amg_eligible(pde, InitialValues) and not pde.GlobalConstraints
        if solver == "mg" and not mg_ok:
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "multigrid steps need one scalar H1P1 unknown on a generated "
                                 "unit square with Dirichlet data on its whole boundary (inhomogeneous values are "
                                 "lifted) and no global constraints")
        if solver == "amg" and not amg_ok:
            raise _abi.GrmpError(_abi.ERR_UNSUPPORTED, "algebraic multigrid steps need a symmetric positive definite "
                                 "system: no Lagrange multipliers and no global constraints")
        if solver == "mg" or (solver == "auto" and mg_ok):
            self.solver = "mg"
        elif solver == "amg" or (solver == "auto" and amg_ok):
            self.solver = "amg"
        else:
            self.solver = "krylov"
        self.levels = None
        self.amg = None
        self.S = None
        self.cstep, self.ctime, self.last_timestep = 0, 0.0, None
        self.assemblies = 0
        self.last_info = None

    def _mg_eligible(self):
        """mg_eligible, with the mass matrix's restriction to scalar H1P1."""
        fes = self.X.FES[0]
        return mg_eligible(self.PDE, self.X) and fes.FEType.code == _abi.FE_H1P1 and fes.FEType.ncomp == 1

    def _mg_solve(self):
        """The step's solve; inhomogeneous Dirichlet values are lifted (the
        multigrid solve has homogeneous data) and added back."""
        maxit = min(self.SC.maxiterations, 500)
        if self.target is None:
            return self.levels.solve(self.b, self.X, self.SC.rtol, maxit)
        if getattr(self, "_lifted", None) is None:
            self._lifted = FEVector(self.X.FES, name="lifted rhs")
        lift_dirichlet(self.S, self.b, self.target, self._lifted)
        info = self.levels.solve(self._lifted, self.X, self.SC.rtol, maxit)
        self.X.ctx.check(lib().grmp_vector_axpy(1.0, self.target.h, self.X.vec.h))
        return info

    def _mfactor(self, timestep):
        """the mass matrix's factor in S: 1/dt (BackwardEuler), 2/dt (CrankNicolson)"""
        return (2.0 if self.rule is CrankNicolson else 1.0) / timestep

    def _system(self, timestep):
        """S = A + (1/dt) M (BackwardEuler, :337-351) or A + (2/dt) M
        (CrankNicolson, :408-413) with the fixed dofs penalized."""
        FES = self.X.FES
        S = FEMatrix(FES, name="system matrix")
        for (i, j), ops in sorted(self.PDE.LHS.items()):
            for op in ops:
                form, keep = op.form(FES[i], i, j, S.ctx, FES[j])
                S.add_form(form, keep)
        for d in self.which:
            form, keep = ReactionOperator(self._mfactor(timestep)).form(FES[d], d, d, S.ctx)
            S.add_form(form, keep)
        S.set_fixed(self.fixed, self.SC.fixed_penalty)
        S.build_pattern()
        S.assemble(self.values)
        self.assemblies += 1
        return S

    def rhs(self, timestep):
        """BackwardEuler: b = f + M u^n / dt; CrankNicolson: b = 2 f - A u^n +
        (2/dt) M u^n; b_d = penalty * u^n_d on fixed dofs (on the device)."""
        X = self.X
        if self.rule is CrankNicolson:
            X.ctx.check(lib().grmp_timestep_rhs(self.S.h, self.A.h, self.f2.vec.h, X.vec.h, -1.0, self._tmp.vec.h))
            X.ctx.check(lib().grmp_timestep_rhs(self.S.h, self.AM.h, self._tmp.vec.h, X.vec.h, 2.0 / timestep,
                                                self.b.vec.h))
            return self.b
        X.ctx.check(lib().grmp_timestep_rhs(self.S.h, self.AM.h, self.f.vec.h, X.vec.h, 1.0 / timestep,
                                            self.b.vec.h))
        return self.b

    def advance(self, timestep=1e-1):
        """advance!(TCS, timestep) (src/solvers_timedependent.jl:293-594).
        Returns the SolveInfo of the step's Krylov solve."""
        timestep = float(timestep)
        if not np.isfinite(timestep) or timestep <= 0.0:
            raise _abi.GrmpError(_abi.ERR_ARG, f"advance: timestep {timestep} must be positive and finite")
        self.cstep += 1
        self.ctime += timestep
        if self.S is None or timestep != self.last_timestep:
            self.amg = None
            self.S = self.levels = None
            self.S = self._system(timestep)
            if self.solver == "mg":  # assembled once per step size, reused by every step
                self.levels = MultigridSolver(*mg_hierarchy(self.PDE, self.X.FES[0], self.S,
                                                            [ReactionOperator(self._mfactor(timestep))],
                                                            self.SC.fixed_penalty))
            elif self.solver == "amg":  # built once per step size from S's FP64 values
                self.amg = AlgebraicMultigridSolver(self.S)
        self.rhs(timestep)
        if self.solver == "mg":
            info = self._mg_solve()
        elif self.solver == "amg":
            info = self.amg.solve(self.b, self.X, self.SC.rtol, min(self.SC.maxiterations, 2000))
        else:
            info = solve_system(self.S, self.b, self.X, self.SC, self.method)
        realize_constraints(self.PDE, self.X)
        self.last_timestep = timestep
        self.last_info = info
        return info

    def advance_until_stationarity(self, timestep, stationarity_threshold=1e-11, maxtimesteps=100,
                                   do_after_each_timestep=None):
        """advance_until_stationarity! (src/solvers_timedependent.jl:605-665):
        the change of a step is the sum over unknowns of ||u^{n+1} - u^n||
        on that unknown's dofs.  Returns the number of steps taken."""
        offs = self.X.offsets
        for it in range(1, maxtimesteps + 1):
            last = self.X.entries()
            self.advance(timestep)
            d = self.X.entries() - last
            change = sum(float(np.linalg.norm(d[offs[j]:offs[j + 1]])) for j in range(len(offs) - 1))
            self.last_change = change
            if do_after_each_timestep is not None:
                do_after_each_timestep(self)
            if change < stationarity_threshold:
                return it
        return maxtimesteps

    def advance_until_time(self, timestep, finaltime, finaltime_tolerance=1e-15, do_after_each_timestep=None):
        """advance_until_time!(TCS, timestep, finaltime; finaltime_tolerance,
        do_after_each_timestep) (src/solvers_timedependent.jl:683-760): steps
        of size ``timestep`` while the time is below finaltime minus the
        tolerance; warns when the final time is not met within it.  Returns
        the number of steps taken."""
        steps = 0
        while self.ctime < finaltime - finaltime_tolerance:
            self.advance(timestep)
            steps += 1
            if do_after_each_timestep is not None:
                do_after_each_timestep(self)
        if abs(self.ctime - finaltime) > finaltime_tolerance:
            warnings.warn("final time not reached within tolerance! (consider another timestep)")
        return steps
