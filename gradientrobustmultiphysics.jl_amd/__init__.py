"""MI355X (gfx950) assembly-and-solve engine for GradientRobustMultiPhysics.jl's
hot path, driven from Python.

The package keeps the reference's vocabulary (ExtendableGrid, FESpace,
FEVector, FEMatrix, PDEDescription, LaplaceOperator, LinearForm, assemble,
solve) and routes all work through libgrmp_hip.so (include/grmp.h).  The
directory name contains a dot, so it is loaded by path, see ``load()`` in
__graft_entry__.py; inside the package, imports are relative.
"""
from ._abi import (Context, DeviceVector, GrmpError, default_context, lib)  # noqa: F401
from . import _abi  # noqa: F401
from .fe import (ExtendableGrid, FEMatrix, FEType, FESpace, FEVector, H1BR, H1P1, H1P2, L2P0,  # noqa: F401
                 grid_unitcube, grid_unitsquare)
from .pde import (APT_BilinearForm, APT_SymmetricBilinearForm, BestapproxDirichletBoundary,  # noqa: F401
                  BilinearForm, CellCoefficient, DataFunction, Divergence, FixedIntegralMean, Gradient,
                  GradientCurlData, HDIVRT0, HomogeneousDirichletBoundary, HookStiffnessOperator2D, Identity,
                  IncompressibleNavierStokesProblem, InterpolateDirichletBoundary, PolynomialData, boundary_values,
                  L2ErrorIntegrator, L2NormIntegrator, evaluate, evaluate_cells, NonlinearDiffusion,
                  NonlinearForm, solve_fixpoint, is_nonlinear,
                  LagrangeMultiplier, LaplaceOperator, LinearElasticityProblem, LinearForm, PDEDescription,
                  PoissonProblem, ReactionOperator, ReconstructionIdentity, SolverConfig, SymmetricGradient, assemble,
                  assemble_linearform, MultigridSolver, mg_eligible, mg_hierarchy, mg_solve, realize_constraints,
                  solve, solve_system, stokes_mg_eligible, stokes_velocity_mg, AlgebraicMultigridSolver,
                  AndersonAccelerator, amg_eligible)

from .dataexport import nodevalues, writeCSV  # noqa: F401
from .timecontrol import BackwardEuler, CrankNicolson, TimeControlSolver  # noqa: F401
from . import server  # noqa: F401
from .server import Server  # noqa: F401

__version__ = "0.1.0"
