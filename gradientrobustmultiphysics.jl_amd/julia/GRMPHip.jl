#=
GRMPHip.jl -- Julia binding of libgrmp_hip.so for GradientRobustMultiPhysics.jl.

Loaded next to GradientRobustMultiPhysics, it
  * adds `GRMPKrylov`, a LinearSolve algorithm that solve! accepts through its
    `linsolver` keyword (src/solvers.jl:55, 89-101, 656-660): `init` builds a
    LinearCache with it (`init_cacheval`), and `LinearSolve.solve(cache)`
    reaches `SciMLBase.solve!(cache, alg::GRMPKrylov)` below;
  * adds `HIPOperator <: AbstractPDEOperator` (src/pdeoperators.jl:23-56), an
    operator a PDEDescription holds like LaplaceOperator: `assemble!`,
    `check_PDEoperator` and `check_dependency` are implemented, so
    solve!/assemble! of the PDE (src/solvers.jl:533,568-573,595) assemble its
    block on the GPU;
  * keeps `hip_assemble!` as the block-level call both use.

Every call passes plain pointers: xgrid[Coordinates] ([dim, nnodes]),
xgrid[CellNodes] ([nloc, ncells]) and FES[CellDofs] ([ndofs_loc, ncells]) in
Julia's own 1-based column-major layout, which is what include/grmp.h expects.
This file has not been run: the build container has no Julia toolchain.
=#
module GRMPHip

using GradientRobustMultiPhysics
using ExtendableGrids
using ExtendableSparse
using SparseArrays
using LinearSolve

const libgrmp = joinpath(@__DIR__, "..", "libgrmp_hip.so")

# ---------------------------------------------------------------- ABI
const GRMP_OK = Cint(0)
const GRMP_FE = Dict(H1P1 => Cint(1), H1P2 => Cint(2), H1BR => Cint(3), L2P0 => Cint(4))
const GRMP_BLF_LAPLACE, GRMP_BLF_HOOKE2D, GRMP_BLF_MASS = Cint(1), Cint(2), Cint(3)
const GRMP_COEF_FACTOR, GRMP_COEF_CELL = Cint(0), Cint(1)
const GRMP_VALUES_FP64, GRMP_VALUES_FP32 = Cint(1), Cint(2)

struct GrmpForm
    kind::Cint
    coef_mode::Cint
    quadorder::Cint
    row_block::Cint
    col_block::Cint
    factor::Cdouble
    params::NTuple{4,Cdouble}
    cell_coef::Ptr{Cvoid}
end

struct GrmpLForm
    kind::Cint                # GRMP_LF_CONST = 1, GRMP_LF_RECO_RT0 = 2
    quadorder::Cint
    factor::Cdouble
    f::NTuple{4,Cdouble}
    nmodes::Cint              # RECO_RT0: gradient + curl force modes
    pad_::Cint
    modes::Ptr{Cdouble}       # [2*nmodes][4] = (amplitude, p, q, phase)
    poly::NTuple{4,Cdouble}
end
GrmpLForm(kind, quadorder, factor, f) = GrmpLForm(kind, quadorder, factor, f, 0, 0, C_NULL, (0.0, 0.0, 0.0, 0.0))

struct GrmpSolveOpts
    method::Cint
    precond::Cint
    precision::Cint
    rtol::Cdouble
    maxit::Int64
    inner_rtol::Cdouble
    inner_maxit::Int64
    check_every::Cint
    initial_guess::Cint
end

mutable struct GrmpSolveInfo
    status::Cint
    outer::Cint
    iterations::Int64
    relres::Cdouble
    seconds::Cdouble
end
GrmpSolveInfo() = GrmpSolveInfo(0, 0, 0, 0.0, 0.0)

function check(ctx::Ptr{Cvoid}, rc::Cint)
    rc == GRMP_OK && return nothing
    msg = unsafe_string(ccall((:grmp_last_error, libgrmp), Cstring, (Ptr{Cvoid},), ctx))
    error("libgrmp_hip error $rc: $msg")
end

# one context per process (device 0)
const CTX = Ref{Ptr{Cvoid}}(C_NULL)
function context()
    if CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:grmp_ctx_create, libgrmp), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h)
        rc == GRMP_OK || error("grmp_ctx_create failed ($rc)")
        CTX[] = h[]
    end
    return CTX[]
end

# ------------------------------------------------- grids and spaces
# device images are cached per xgrid / FESpace object
const MESHES = IdDict{Any,Ptr{Cvoid}}()
const SPACES = IdDict{Any,Ptr{Cvoid}}()

function device_mesh(xgrid::ExtendableGrid)
    get!(MESHES, xgrid) do
        coords = xgrid[Coordinates]                 # [dim, nnodes]
        cells = Int32.(xgrid[CellNodes])            # [nloc, ncells], 1-based
        h = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve coords cells begin
            check(context(), ccall((:grmp_mesh_upload, libgrmp), Cint,
                (Ptr{Cvoid}, Cint, Int64, Ptr{Cdouble}, Int64, Cint, Ptr{Int32}, Ref{Ptr{Cvoid}}),
                context(), size(coords, 1), size(coords, 2), coords, size(cells, 2), size(cells, 1), cells, h))
        end
        h[]
    end
end

"""
    device_faces(xgrid) -> (CellFaces, CellFaceSigns, bface)

Face adjacency of a triangle grid as the device builds it on upload:
CellFaces [3, ncells] (1-based) and CellFaceSigns [3, ncells] in
ExtendableGrids' layout, and a 0/1 flag per face for the boundary.  Faces are
numbered by first appearance; a grid whose faces ExtendableGrids numbered
differently keeps passing FES[CellDofs] through device_space.
"""
function device_faces(xgrid::ExtendableGrid)
    m = device_mesh(xgrid)
    nf = Ref{Int64}(0)
    check(context(), ccall((:grmp_mesh_faces, libgrmp), Cint,
        (Ptr{Cvoid}, Ref{Int64}, Ptr{Int32}, Ptr{Int8}, Ptr{UInt8}), m, nf, C_NULL, C_NULL, C_NULL))
    nc = size(xgrid[CellNodes], 2)
    cf = zeros(Int32, 3, nc)
    cs = zeros(Int8, 3, nc)
    bf = zeros(UInt8, nf[])
    check(context(), ccall((:grmp_mesh_faces, libgrmp), Cint,
        (Ptr{Cvoid}, Ref{Int64}, Ptr{Int32}, Ptr{Int8}, Ptr{UInt8}), m, nf, cf, cs, bf))
    return cf, cs, bf
end

function device_space(FES::FESpace{Tv,Ti,FEType}) where {Tv,Ti,FEType}
    get!(SPACES, FES) do
        dofs = Int32.(Matrix(FES[CellDofs]))        # [ndofs_loc, ncells], 1-based
        fe = GRMP_FE[FEType.name.wrapper]
        h = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve dofs begin
            check(context(), ccall((:grmp_space_upload, libgrmp), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Int64, Cint, Ptr{Int32}, Ref{Ptr{Cvoid}}),
                context(), device_mesh(FES.xgrid), fe, get_ncomponents(FEType), FES.ndofs, size(dofs, 1), dofs, h))
        end
        h[]
    end
end

# ---------------------------------------------- pattern-level assemble!
"""
    hip_assemble!(A::FEMatrixBlock, O::PDEOperator; kappa = nothing)

Device replacement of assemble!(A, AssemblyPattern) for LaplaceOperator,
ReactionOperator, HookStiffnessOperator2D and the per-cell-coefficient
BilinearForm([Gradient,Gradient], action; APT_SymmetricBilinearForm): the
block is assembled on the GPU into the reference's CSC pattern (_addnz zero
skip) and added into A.entries.
"""
function hip_assemble!(A::FEMatrixBlock, kind::Cint, quadorder::Int; factor = 1.0, kappa = nothing,
                       params = (0.0, 0.0, 0.0, 0.0))
    ctx = context()
    sx = device_space(A.FESX)
    m = Ref{Ptr{Cvoid}}(C_NULL)
    check(ctx, ccall((:grmp_matrix_create, libgrmp), Cint, (Ptr{Cvoid}, Cint, Ref{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}),
                     ctx, 1, Ref(sx), m))
    kv = Ref{Ptr{Cvoid}}(C_NULL)
    if kappa !== nothing
        check(ctx, ccall((:grmp_vector_create, libgrmp), Cint, (Ptr{Cvoid}, Int64, Ref{Ptr{Cvoid}}), ctx, length(kappa), kv))
        check(ctx, ccall((:grmp_vector_upload, libgrmp), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), kv[], Vector{Float64}(kappa)))
    end
    form = GrmpForm(kind, kappa === nothing ? GRMP_COEF_FACTOR : GRMP_COEF_CELL, quadorder, 0, 0, factor, params, kv[])
    check(ctx, ccall((:grmp_matrix_add_form, libgrmp), Cint, (Ptr{Cvoid}, Ref{GrmpForm}), m[], form))
    check(ctx, ccall((:grmp_matrix_assemble, libgrmp), Cint, (Ptr{Cvoid}, Cint), m[], GRMP_VALUES_FP64))
    n, nnz, mr = Ref{Int64}(0), Ref{Int64}(0), Ref{Cint}(0)
    check(ctx, ccall((:grmp_matrix_info, libgrmp), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Cint}), m[], n, nnz, mr))
    colptr, rowval, nzval = Vector{Int64}(undef, n[] + 1), Vector{Int64}(undef, nnz[]), Vector{Float64}(undef, nnz[])
    check(ctx, ccall((:grmp_matrix_export_csc, libgrmp), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Cdouble}),
                     m[], colptr, rowval, nzval))
    ccall((:grmp_matrix_destroy, libgrmp), Cint, (Ptr{Cvoid},), m[])
    kv[] != C_NULL && ccall((:grmp_vector_destroy, libgrmp), Cint, (Ptr{Cvoid},), kv[])
    # add into the block: the device's pattern is the one _addnz would create;
    # one sparse addition at the CSC level (shifted into the block's offsets)
    # instead of an update call per entry
    flush!(A.entries)
    N = size(A.entries.cscmatrix, 1)
    rows, cols, vals = findnz(SparseMatrixCSC(n[], n[], colptr, rowval, nzval))
    B = sparse(rows .+ A.offsetX, cols .+ A.offsetY, vals, N, size(A.entries.cscmatrix, 2))
    A.entries.cscmatrix = A.entries.cscmatrix + B
    return nothing
end

# ------------------------------------------ operator plug-in (PDEDescription)
"""
    HIPOperator(kind; factor = 1.0, kappa = nothing, params = (0,0,0,0), name)

An AbstractPDEOperator (src/pdeoperators.jl:23-56) whose matrix block is
assembled by libgrmp_hip.so.  `kind`: `:laplace` (LaplaceOperator(factor),
or the per-cell action BilinearForm([Gradient, Gradient], kappa) when
`kappa` is a vector over the cells), `:reaction` (ReactionOperator(factor)),
`:hooke2d` (HookStiffnessOperator2D with params = (lambda, mu, 0, 0),
optionally scaled per cell by `kappa`).  Linear and time-independent:
assembled once (AssemblyInitial).
"""
struct HIPOperator <: GradientRobustMultiPhysics.AbstractPDEOperator
    name::String
    kind::Symbol
    factor::Float64
    kappa::Union{Nothing,Vector{Float64}}
    params::NTuple{4,Float64}
end
HIPOperator(kind::Symbol; factor = 1.0, kappa = nothing, params = (0.0, 0.0, 0.0, 0.0), name = "HIP $(kind)") =
    HIPOperator(name, kind, factor, kappa === nothing ? nothing : Vector{Float64}(kappa), params)

GradientRobustMultiPhysics.check_PDEoperator(O::HIPOperator, involved_equations) =
    (false, false, GradientRobustMultiPhysics.AssemblyInitial)
GradientRobustMultiPhysics.check_dependency(O::HIPOperator, arg::Int) = false

# quadrature order of prepare_assembly! (src/assemblypatterns.jl:559-563):
# polynomial orders of both arguments plus the operators' shifts (-1 per gradient)
function _quadorder(O::HIPOperator, FES)
    EG = FES.xgrid[UniqueCellGeometries][1]
    p = GradientRobustMultiPhysics.get_polynomialorder(eltype(FES), EG)
    return O.kind == :reaction ? 2 * p : 2 * (p - 1)
end

function GradientRobustMultiPhysics.assemble!(A::FEMatrixBlock, SC, j::Int, k::Int, o::Int, O::HIPOperator,
                                              CurrentSolution::FEVector; time = 0, At = nothing)
    kind = O.kind == :laplace ? GRMP_BLF_LAPLACE : O.kind == :reaction ? GRMP_BLF_MASS :
           O.kind == :hooke2d ? GRMP_BLF_HOOKE2D : error("HIPOperator: unknown kind $(O.kind)")
    hip_assemble!(A, kind, _quadorder(O, A.FESX); factor = O.factor, kappa = O.kappa, params = O.params)
    At === nothing || error("HIPOperator: transposed copies are not built")
    return nothing
end

# ---------------------------------------------------- Dirichlet data
"""
    hip_boundary_bestapprox!(space, coef, quadorder, target, offset; rtol = 1e-14)

BestapproxDirichletBoundary (src/boundarydata.jl:259-352) on the device for
the polynomial data sum(coef[k] m_k), m = (1, x, y, x^2, xy, y^2, x^3, x^2 y,
x y^2, y^3): the boundary values go to the `grmp_vector` handle `target` at
`offset` (0-based dof numbering of the space).
"""
function hip_boundary_bestapprox!(space::Ptr{Cvoid}, coef::Vector{Float64}, quadorder::Int, target::Ptr{Cvoid},
                                  offset::Int; rtol = 1e-14)
    check(context(), ccall((:grmp_boundary_bestapprox, libgrmp), Cint,
                           (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Cint, Cdouble, Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Cdouble}),
                           space, coef, length(coef), quadorder, rtol, target, offset, C_NULL, C_NULL))
    return target
end

# ---------------------------------------------------- time stepping
"""
    hip_timestep_rhs!(S, M, f, u, dt, b)

Right-hand side update of `advance!` with `BackwardEuler`
(src/solvers_timedependent.jl:339-359, :452-461) on device handles:
`b = f + M u / dt`, and `b[d] = fixed_penalty * u[d]` on the fixed rows of the
system matrix `S`.  `S`, `M` are `grmp_matrix` handles (M from the mass form
of the time-dependent block), `f`, `u`, `b` are `grmp_vector` handles.
"""
function hip_timestep_rhs!(S::Ptr{Cvoid}, M::Ptr{Cvoid}, f::Ptr{Cvoid}, u::Ptr{Cvoid}, dt::Real, b::Ptr{Cvoid})
    check(context(), ccall((:grmp_timestep_rhs, libgrmp), Cint,
                           (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}),
                           S, M, f, u, 1.0 / dt, b))
    return b
end

# ---------------------------------------------------- ItemIntegrators
"""
    hip_l2error(space, u, offset, coef, ncoef, quadorder; factor = 1.0, cellvals = nothing)

evaluate(L2ErrorIntegrator(g, Identity; quadorder, factor), u) for polynomial
data g (per component `ncoef` coefficients over 1, x, y, x^2, xy, y^2, x^3,
x^2 y, x y^2, y^3; `ncoef = 0` gives L2NormIntegrator) on the device
(src/assemblypatterns/itemintegrator.jl:40-95, 169-360).  `u` is a
`grmp_vector` handle holding the FEVector's entries, `offset` the block's.
Returns the squared L2 error; `cellvals` (length ncells) gets evaluate!'s
per-item values.
"""
function hip_l2error(space::Ptr{Cvoid}, u::Ptr{Cvoid}, offset::Int, coef::Vector{Float64}, ncoef::Int,
                     quadorder::Int; factor = 1.0, cellvals = nothing)
    total = Ref{Cdouble}(0.0)
    check(context(), ccall((:grmp_integrate_l2error, libgrmp), Cint,
                           (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cdouble}, Cint, Cint, Cdouble, Ptr{Cdouble},
                            Ref{Cdouble}),
                           context(), space, u, offset, coef, ncoef, quadorder, factor,
                           cellvals === nothing ? C_NULL : cellvals, total))
    return total[]
end

# ---------------------------------------------------- batched assembly
"""
    hip_assemble_batch!(mats::Vector{Ptr{Cvoid}})

FP64 values of many single-form `grmp_matrix` handles (any meshes, spaces
and form kinds) in one launch of the persistent assembly kernel; each equals
assembling it alone bit for bit.  For callers that assemble many small
problems, e.g. one operator per mesh of a parameter sweep.
"""
hip_assemble_batch!(mats::Vector{Ptr{Cvoid}}) =
    check(context(), ccall((:grmp_matrix_assemble_batch, libgrmp), Cint, (Cint, Ptr{Ptr{Cvoid}}),
                           length(mats), mats))

# ---------------------------------------------------- LinearSolve plug-in
"""
    GRMPKrylov(; precision = :fp32_ir, rtol = 1e-10)

LinearSolve algorithm for solve!(Solution, PDE; linsolver = GRMPKrylov()):
Jacobi-PCG on the GPU, FP32 inner sweeps inside FP64 refinement
(`precision = :fp32_ir`) or plain FP64 (`:fp64`).  Penalized rows
(diagonal >= 1e59, src/solvers.jl:649-652) are eliminated.
"""
struct GRMPKrylov <: LinearSolve.SciMLLinearSolveAlgorithm
    precision::Symbol
    rtol::Float64
end
GRMPKrylov(; precision = :fp32_ir, rtol = 1e-10) = GRMPKrylov(precision, rtol)

# no factorization or Krylov state lives on the host: the cache value is empty
LinearSolve.init_cacheval(alg::GRMPKrylov, A, b, u, Pl, Pr, maxiters::Int, abstol, reltol, verbose::Bool,
                          assumptions::LinearSolve.OperatorAssumptions) = nothing

# LinearSolve 2 dispatches solve(cache) -> SciMLBase.solve!(cache, cache.alg)
function LinearSolve.SciMLBase.solve!(cache::LinearSolve.LinearCache, alg::GRMPKrylov; kwargs...)
    A::SparseMatrixCSC{Float64,Int64} = cache.A
    b::Vector{Float64} = cache.b
    x = cache.u
    fill!(x, 0.0)
    n = size(A, 1)
    fixed = [j for j in 1:n if A[j, j] >= 1e59]
    # CSC of A is CSR of A^T; the systems solved here are symmetric
    rc = ccall((:grmp_solve_csc, libgrmp), Cint,
               (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Int64, Ptr{Int64},
                Cint, Cdouble), context(), n, A.colptr, A.rowval, A.nzval, b, x, length(fixed), fixed,
               alg.precision == :fp64 ? Cint(1) : Cint(2), alg.rtol)
    check(context(), rc)
    return LinearSolve.SciMLBase.build_linear_solution(alg, x, nothing, cache)
end

# ---------------------------------------------------- algebraic multigrid
"""
    hip_amg_solve!(A, b, x; rtol = 1e-10, maxit = 2000) -> GrmpSolveInfo

CG preconditioned by the device's algebraic multigrid for an SPD device
matrix `A` (a `grmp_matrix` handle with FP64 values, penalized rows set by
`grmp_matrix_set_fixed`) on any mesh; `b` and `x` are `grmp_vector` handles.
The hierarchy is built once per call here; a caller that solves many
right-hand sides with one matrix keeps the `grmp_amg` handle of
`grmp_amg_create` and calls `grmp_amg_apply` for each.
"""
function hip_amg_solve!(A::Ptr{Cvoid}, b::Ptr{Cvoid}, x::Ptr{Cvoid}; rtol = 1e-10, maxit = 2000)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(context(), ccall((:grmp_amg_create, libgrmp), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), A, h))
    info = GrmpSolveInfo()
    try
        check(context(), ccall((:grmp_amg_apply, libgrmp), Cint,
                               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Int64, Ref{GrmpSolveInfo}),
                               h[], b, x, rtol, maxit, info))
    finally
        ccall((:grmp_amg_destroy, libgrmp), Cint, (Ptr{Cvoid},), h[])
    end
    return info
end

"""
    hip_anderson(x, n, depth; damping = 1.0, metric = C_NULL, ranges = [(0, n)]) -> Ptr{Cvoid}
    hip_update_anderson!(aa, x) -> Vector{Float64}

The device counterpart of `AndersonAccelerationManager` / `update_anderson!`
(src/solvers.jl:688-808) for a fixpoint loop whose iterate lives in the
`grmp_vector` `x` of length `n`: create it after the boundary data are set
and call `hip_update_anderson!` after each linear solve, before the damping,
as solve_fixpoint_full! does (src/solvers.jl:923-929).  `metric` is the
`grmp_matrix` of the "L2" / "H1" norm (C_NULL for "l2"); `ranges` are the
0-based `[offset, offset + ndofs)` of the blocks in `anderson_unknowns`.
Returns the coefficients (zeros for the first two calls).  Free the handle
with `ccall((:grmp_anderson_destroy, libgrmp), Cint, (Ptr{Cvoid},), aa)`.
"""
function hip_anderson(x::Ptr{Cvoid}, n::Integer, depth::Integer; damping = 1.0, metric::Ptr{Cvoid} = C_NULL,
                      ranges = [(0, n)])
    r = Int64[v for rg in ranges for v in rg]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(context(), ccall((:grmp_anderson_create, libgrmp), Cint,
                           (Ptr{Cvoid}, Int64, Cint, Cdouble, Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Cvoid},
                            Ref{Ptr{Cvoid}}),
                           context(), n, depth, damping, metric, length(ranges), r, x, h))
    return h[]
end

function hip_update_anderson!(aa::Ptr{Cvoid}, x::Ptr{Cvoid}; depth::Integer)
    alpha = zeros(Float64, depth + 1)
    res = Ref{Cdouble}(0.0)
    check(context(), ccall((:grmp_anderson_update, libgrmp), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble},
                                                                      Ref{Cdouble}), aa, x, alpha, res))
    res[] > 1e-4 && @warn "Anderson acceleration local system residual unsatisfactory ($(res[]) > 1e-4)"
    return alpha
end

export GRMPKrylov, HIPOperator, hip_l2error, hip_assemble_batch!, hip_amg_solve!, hip_anderson, hip_update_anderson!

end # module
